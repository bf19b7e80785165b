"""Write golden input/output vectors by running the reference's pure-numpy functions.

Run in the development container only (it reads /root/reference, which does not
exist on GPU machines).  The reference modules import pyopencl/silx at module
level, which are not installed, so the script compiles only the needed top-level
definitions out of the reference source files (`ast`) and runs them with numpy:

  dynamix/correlator/dense.py   py_dense_correlator          (lines 32-70)
  dynamix/correlator/common.py  BaseCorrelator               (lines 10-80)
  dynamix/correlator/event.py   FramesCompressor             (lines 207-321)
  dynamix/utils.py              generate_powers, get_next_power (lines 29-60)
  dynamix/tools/tools.py        radi, cftomt, format_result  (lines 14-50, 146-172, 313-346)

Output: tests/golden/reference_vectors.npz (data only: seeded inputs and the
reference's outputs).  Nothing of the reference's source is stored.
"""
import ast
import os
import sys
from collections import namedtuple
from math import sqrt

import numpy as np

REF = "/root/reference/dynamix"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_vectors.npz")


def load_defs(path, names, extra_globals=None):
    with open(path) as f:
        tree = ast.parse(f.read(), filename=path)
    keep = [node for node in tree.body
            if isinstance(node, (ast.FunctionDef, ast.ClassDef)) and node.name in names]
    missing = set(names) - {n.name for n in keep}
    if missing:
        raise RuntimeError("not found in %s: %s" % (path, missing))
    mod = ast.Module(body=keep, type_ignores=[])
    g = {"np": np, "__name__": "golden_" + os.path.basename(path)}
    if extra_globals:
        g.update(extra_globals)
    exec(compile(mod, path, "exec"), g)
    return g


def speckle(rng, n, shape, mean, rho, dtype):
    """Small AR(1) speckle Poisson stack (inputs only)."""
    h, w = shape
    e = (rng.standard_normal((h, w)) + 1j * rng.standard_normal((h, w))) / np.sqrt(2)
    out = np.zeros((n, h, w), dtype=np.int64)
    c = np.sqrt(1 - rho * rho)
    for t in range(n):
        if t:
            e = rho * e + c * (rng.standard_normal((h, w)) + 1j * rng.standard_normal((h, w))) / np.sqrt(2)
        out[t] = rng.poisson(mean * np.abs(e) ** 2)
    info = np.iinfo(dtype)
    return np.clip(out, 0, info.max).astype(dtype)


def main():
    if not os.path.isdir(REF):
        sys.exit("reference tree not present; golden vectors are generated in the dev container")
    CorrelationResult = namedtuple("CorrelationResult", "res dev")
    dense = load_defs(os.path.join(REF, "correlator", "dense.py"), ["py_dense_correlator"],
                      {"sqrt": sqrt, "CorrelationResult": CorrelationResult})
    common = load_defs(os.path.join(REF, "correlator", "common.py"), ["BaseCorrelator"])
    event = load_defs(os.path.join(REF, "correlator", "event.py"), ["FramesCompressor"])
    utils = load_defs(os.path.join(REF, "utils.py"), ["generate_powers", "get_next_power"],
                      {"product": __import__("itertools").product,
                       "bisect": __import__("bisect").bisect})
    py_dense_correlator = dense["py_dense_correlator"]
    BaseCorrelator = common["BaseCorrelator"]
    FramesCompressor = event["FramesCompressor"]
    get_next_power = utils["get_next_power"]

    rng = np.random.default_rng(20261202)
    out = {}

    # ---- dense MatMul correlator: several small cases, uint8 / int8 / uint16
    cases = [
        ("u8_small", (6, 7), 16, 3.0, 0.8, np.uint8, 3),
        ("u8_big_values", (5, 6), 12, 150.0, 0.9, np.uint8, 2),
        ("i8_sparse", (9, 8), 40, 0.3, 0.95, np.int8, 4),
        ("u16_mixed", (7, 5), 20, 400.0, 0.7, np.uint16, 3),
        ("u8_onepix", (4, 4), 10, 2.0, 0.5, np.uint8, 2),
    ]
    for name, shape, n, mean, rho, dt, nbins in cases:
        frames = speckle(rng, n, shape, mean, rho, dt)
        q = rng.integers(0, nbins + 1, size=shape).astype(np.int32)
        if name == "u8_onepix":
            q[:] = 2
            q[1, 2] = 1          # bin 1 holds a single pixel
        for b in range(1, nbins + 1):
            if not (q == b).any():
                q.flat[b] = b
        res = np.zeros((nbins, n), np.float32)
        dev = np.zeros((nbins, n), np.float32)
        for b in range(1, nbins + 1):
            r = py_dense_correlator(frames, q == b, calc_std=True)
            res[b - 1], dev[b - 1] = r.res, r.dev
            res_nostd = py_dense_correlator(frames, q == b, calc_std=False)
            assert np.array_equal(np.float32(res_nostd), res[b - 1], equal_nan=True)
        out["dense_%s_frames" % name] = frames
        out["dense_%s_qmask" % name] = q
        out["dense_%s_res" % name] = res
        out["dense_%s_dev" % name] = dev

    # ---- BaseCorrelator conventions: bins, n_bins, scale factors
    class Probe(BaseCorrelator):
        pass

    q = np.array([[0, 1, 1, 3], [2, 2, 0, 3], [3, 3, 3, 0]], dtype=np.int32)
    p = Probe()
    p._set_parameters(q.shape, 5, q, None, {})
    out["base_qmask"] = q
    out["base_bins"] = np.asarray(p.bins)
    out["base_n_bins"] = np.int64(p.n_bins)
    out["base_scale"] = np.array([p.scale_factors[b] for b in p.bins], dtype=np.float64)
    p = Probe()
    p._set_parameters(q.shape, 5, q, [10.0, 20.0, 30.0], {})
    out["base_scale_given"] = np.array([p.scale_factors[b] for b in p.bins], dtype=np.float64)
    p = Probe()
    p._set_parameters((3, 4), 5, None, None, {})
    out["base_none_bins"] = np.asarray(p.bins)
    out["base_none_scale"] = np.array([p.scale_factors[b] for b in p.bins], dtype=np.float64)
    out["base_none_qmask"] = np.asarray(p.qmask)
    p = Probe()
    p._set_parameters(4, 5, None, 7.5, {})
    out["base_scalar_shape"] = np.asarray(p.shape)
    out["base_scalar_scale"] = np.array([p.scale_factors[b] for b in p.bins], dtype=np.float64)

    # ---- FramesCompressor: one-shot and progressive compaction
    for name, shape, n, mean, dt in [("evt_u8", (6, 5), 30, 0.4, np.uint8),
                                     ("evt_i8", (4, 7), 25, 0.8, np.int8)]:
        frames = speckle(rng, n, shape, mean, 0.9, dt)
        d, t, o = FramesCompressor.compress_all_stack(frames)
        fc = FramesCompressor(shape, n, int((frames > 0).sum(axis=0).max()), dtype=dt)
        for f in frames:
            fc.process_frame(f)
        d2, t2, o2 = fc.get_compacted_events()
        out["%s_frames" % name] = frames
        out["%s_data" % name] = d
        out["%s_times" % name] = t
        out["%s_offsets" % name] = o
        out["%s_prog_data" % name] = d2
        out["%s_prog_times" % name] = t2
        out["%s_prog_offsets" % name] = o2

    # ---- FFT length rule
    ns = np.array([1, 2, 3, 4, 5, 7, 9, 17, 100, 333, 1000, 1998, 2000, 2001, 4097,
                   19999, 20000, 20001, 65536, 99999, 200000, 123457], dtype=np.int64)
    out["nextpow_n"] = ns
    out["nextpow_val"] = np.array([int(get_next_power(int(v))) for v in ns], dtype=np.int64)

    # ---- host tools: radial average, multitau resampling, result formatting
    tl = load_defs(os.path.join(REF, "tools", "tools.py"), ["radi", "cftomt", "format_result"])
    for name, shape, cx, cy in [("a", (23, 31), 12.3, 9.7), ("b", (40, 37), 5.6, 30.2)]:
        saxs = rng.random(shape) * 100
        mask = (rng.random(shape) < 0.05).astype(np.uint8)
        mask[shape[0] // 2 - 1: shape[0] // 2 + 2, :] = 1
        rad, q, new_saxs = tl["radi"](saxs, mask, cx, cy)
        out["radi_%s_saxs" % name] = saxs
        out["radi_%s_mask" % name] = mask
        out["radi_%s_center" % name] = np.array([cx, cy])
        out["radi_%s_rad" % name] = rad
        out["radi_%s_q" % name] = q
        out["radi_%s_model" % name] = new_saxs
    for name, n, par, dt in [("f32_16", 1000, 16, np.float32), ("f64_8", 333, 8, np.float64),
                             ("f32_32", 97, 32, np.float32)]:
        d = np.zeros((n, 3), dt)
        d[:, 0] = np.arange(1, n + 1)
        d[:, 1] = 1 + 0.3 * np.exp(-np.arange(n) / 50.0) + rng.normal(0, 1e-3, n)
        d[:, 2] = rng.random(n) * 1e-2
        out["cftomt_%s_in" % name] = d
        out["cftomt_%s_par" % name] = np.int64(par)
        out["cftomt_%s_out" % name] = tl["cftomt"](d, par)
    Result = namedtuple("Result", "res dev trc")
    n, nq, shape = 300, 3, (12, 14)
    qq = rng.integers(0, nq + 1, size=shape).astype(np.int32)
    for b in range(1, nq + 1):
        qq.flat[b] = b
    flat = (1 + 0.1 * rng.random(shape)).astype(np.float32)
    cdata = (5 + rng.random(shape)).astype(np.float32)
    res = (1 + 0.2 * rng.random((nq, n))).astype(np.float32)
    dev = (0.01 * rng.random((nq, n))).astype(np.float32)
    trc = rng.random((n, n)).astype(np.float32)
    per_q, table, trc_out = tl["format_result"](Result(res, dev, trc), qq, flat, cdata, 0.01, 2)
    out["fmt_qmask"] = qq
    out["fmt_flat"] = flat
    out["fmt_cdata"] = cdata
    out["fmt_res"] = res
    out["fmt_dev"] = dev
    out["fmt_trc"] = trc
    out["fmt_table"] = table
    for i, m in enumerate(per_q):
        out["fmt_q%d" % (i + 1)] = m
    out["fmt_trc_out"] = trc_out

    np.savez_compressed(OUT, **out)
    print("wrote", OUT, "with", len(out), "arrays")


if __name__ == "__main__":
    main()
