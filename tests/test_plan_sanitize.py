"""The plan builder (validation, bins, pixel lists, slab layout and every work list of
dxg2_plan_create) under AddressSanitizer and UndefinedBehaviorSanitizer, off the GPU.

The library sources are compiled again with the sanitizers on the host code only (the
device code is built as usual and never runs), linked with a C++ driver,
tests/sanitize/plan_fuzz.cpp, that sends randomised descriptions through
dxg2_plan_dry_run and checks its summary against counts made independently.  Any
out-of-bounds access, use after free, signed overflow or bad shift in the host code aborts
the driver."""
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "dynamix_amd", "csrc")
INCLUDE = os.path.join(ROOT, "include")
DRIVER = os.path.join(ROOT, "tests", "sanitize", "plan_fuzz.cpp")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CLANGXX = "/opt/rocm/llvm/bin/clang++"

# Sanitizers on the host half of each translation unit only.
HOST_SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-sanitize-recover=all", "-Xarch_host", "-g"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("%s failed:\n%s" % (os.path.basename(cmd[0]), r.stdout[-4000:]))
    return r.stdout


@pytest.mark.skipif(not (os.path.exists(HIPCC) and os.path.exists(CLANGXX)),
                    reason="hipcc / clang++ not found")
def test_plan_builder_under_sanitizers(tmp_path):
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))

    def compile_one(f):
        obj = str(tmp_path / (f[:-4] + ".o"))
        _run([HIPCC, "--offload-arch=gfx950", "-O1", "-std=c++17", "-fPIC",
              "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
              "-I" + INCLUDE] + HOST_SAN + ["-c", os.path.join(CSRC, f), "-o", obj])
        return obj

    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(compile_one, srcs))
    drv = str(tmp_path / "plan_fuzz.o")
    _run([CLANGXX, "-O1", "-g", "-std=c++17", "-I" + INCLUDE,
          "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-c", DRIVER,
          "-o", drv])
    exe = str(tmp_path / "plan_fuzz")
    _run([HIPCC, "--offload-arch=gfx950", "-Xarch_host", "-fsanitize=address",
          "-Xarch_host", "-fsanitize=undefined", drv] + objs + ["-o", exe])

    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe, "200"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stdout[-6000:]
    assert "200 cases, 0 failed" in r.stdout
    shutil.rmtree(str(tmp_path), ignore_errors=True)
