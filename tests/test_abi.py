"""CPU checks of the C ABI surface and host-side logic (no GPU compute)."""
import numpy as np
import pytest

import dynamix_amd._lib as L
from dynamix_amd import synth


def test_library_loads_and_exports_header_symbols():
    lib = L.load()
    names = L.header_symbols()
    assert len(names) >= 10
    for n in names:
        assert hasattr(lib, n), n
    assert lib.dxg2_version().startswith(b"dxg2")


def test_signatures_cover_header():
    assert set(L.header_symbols()) <= set(L._SIGNATURES), \
        set(L.header_symbols()) - set(L._SIGNATURES)


def test_error_mapping_on_bad_desc():
    """Invalid plan descriptions fail on the host before any device call."""
    import ctypes
    lib = L.load()
    d = L.PlanDesc()
    d.height, d.width, d.nframes, d.dtype = 0, 4, 10, L.DXG2_U8
    h = ctypes.c_void_p()
    rc = lib.dxg2_plan_create(ctypes.byref(d), ctypes.byref(h))
    assert rc == L.DXG2_EINVAL
    with pytest.raises(ValueError, match="shape"):
        L.check(rc)
    d.height, d.dtype = 4, 99
    rc = lib.dxg2_plan_create(ctypes.byref(d), ctypes.byref(h))
    with pytest.raises(ValueError, match="dtype"):
        L.check(rc)


def test_plan_dry_run_headline_layout():
    """The host half of plan creation on the headline geometry, with no device call:
    20 bins, each padded to whole 64-pixel k-blocks, slab rows rounded up to 256."""
    import ctypes
    lib = L.load()
    q = np.ascontiguousarray(synth.workload_qmask("headline"), dtype=np.int32)
    d = L.PlanDesc()
    d.height, d.width, d.nframes, d.dtype = q.shape[0], q.shape[1], 1000, L.DXG2_U8
    d.qmask = q.ctypes.data
    s = (ctypes.c_int64 * 12)()
    L.check(lib.dxg2_plan_dry_run(ctypes.byref(d), s, 12))
    counts = np.bincount(q.ravel()[q.ravel() > 0], minlength=21)[1:]
    padded = int(((counts + 63) // 64 * 64).sum())
    assert s[0] == 20 and s[1] == padded and s[2] == padded * 1024
    assert min(s[3:12]) >= 0 and s[7] > 0
    d.dtype = L.DXG2_U16
    L.check(lib.dxg2_plan_dry_run(ctypes.byref(d), s, 12))
    assert s[2] == 2 * padded * 1024


def test_headline_qmask_geometry():
    q = synth.workload_qmask("headline")
    assert q.shape == (514, 1030)
    assert q.max() == 20
    assert (q > 0).sum() == 479980          # 466 unmasked rows x 1030
    assert (q == 0).all(axis=1).sum() == 48  # six 8-row gap bands


def test_annulus_widths():
    q = synth.annular_qmask((101, 101), 5, center=(50, 50), r_max=50.0)
    assert q[50, 50] == 1 and q[50, 99] == 5 and q[50, 100] == 0 and q[0, 0] == 0
    ql = synth.annular_qmask((201, 201), 4, center=(100, 100), r_max=100.0, widths="linear")
    # linear widths: edges at 10, 30, 60, 100
    assert ql[100, 109] == 1 and ql[100, 111] == 2 and ql[100, 131] == 3 and ql[100, 161] == 4


def test_host_speckle_statistics():
    f = synth.speckle_frames((32, 32), 400, 2.0, 10.0, seed=3)
    assert f.dtype == np.uint8
    assert abs(f.mean() - 2.0) < 0.1


def test_serving_job_stream_is_seeded_and_in_range():
    from dynamix_amd.serving import make_jobs, SERVE_SHAPES
    a = make_jobs(300, rate=20.0, seed=9)
    b = make_jobs(300, rate=20.0, seed=9)
    assert [(x.shape, x.nframes, x.n_bins, x.layout, x.tau_c, x.seed, x.arrival) for x in a] == \
           [(x.shape, x.nframes, x.n_bins, x.layout, x.tau_c, x.seed, x.arrival) for x in b]
    assert all(x.shape in SERVE_SHAPES for x in a)
    assert all(1000 <= x.nframes <= 20000 and 10 <= x.n_bins <= 100 for x in a)
    assert len({x.seed for x in a}) == len(a)
    assert all(10.0 <= x.tau_c <= 500.0 for x in a)
    dense = sum(x.layout == "dense" for x in a) / len(a)
    assert 0.6 < dense < 0.8
    arr = np.array([x.arrival for x in a])
    assert arr[0] == 0.0 and np.all(np.diff(arr) >= 0)
    rate = (len(a) - 1) / arr[-1]
    assert 15.0 < rate < 25.0


def test_autotune_key_and_cache_roundtrip(tmp_path):
    from dynamix_amd import autotune
    k1 = autotune.tuning_key("uint8", 10000, [26000, 24000, 100, 0, 64, 65])
    assert k1 == "uint8|N10000|kb0:1,1:2,9:2"
    # bin order and empty bins do not change the key; sizes in another bucket do
    assert autotune.tuning_key("uint8", 10000, [65, 64, 0, 24000, 100, 26000]) == k1
    assert autotune.tuning_key("uint8", 10000, [26000, 24000, 100, 0, 64, 129]) != k1
    assert autotune.tuning_key("uint16", 10000, [26000, 24000, 100, 64, 65]) != k1
    p = str(tmp_path / "sub" / "t.json")
    assert autotune.load_cache(p) == {}
    autotune.save_entry(k1, {"tile": 256, "band": 4, "group": 32, "queue": 1}, p)
    autotune.save_entry("other", {"tile": 128, "band": 2, "group": 16, "queue": 0}, p)
    c = autotune.load_cache(p)
    assert c[k1]["band"] == 4 and c["other"]["tile"] == 128


def test_noalloc_window_rejects_oversize_on_host():
    """dxg2_events_window_noalloc refuses pixel counts beyond its static workspace before
    any device call (the pointers are never touched)."""
    import ctypes
    lib = L.load()
    fake = ctypes.c_void_p(256)
    too_many = L.DXG2_STATIC_SCRATCH_WORDS * 1024 + 1
    rc = lib.dxg2_events_window_noalloc(fake, fake, fake, 10, too_many, L.DXG2_U8, 0, 5, fake,
                                        fake, fake, 10, None)
    assert rc == L.DXG2_EINVAL
    assert b"static scratch limit" in lib.dxg2_last_error()
    rc = lib.dxg2_events_window_noalloc(None, fake, fake, 10, 100, L.DXG2_U8, 0, 5, fake, fake,
                                        fake, 10, None)
    assert rc == L.DXG2_EINVAL
    assert L.DXG2_STATIC_SCRATCH_WORDS == int(
        [ln for ln in open(L.HEADER) if "#define DXG2_STATIC_SCRATCH_WORDS" in ln][0].split()[2])
