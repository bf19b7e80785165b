"""HIP graph replay of the dense chain and Gram launch tuning on the GPU: every replay
and every tuning gives the exact numerators of the eager default path.  Marked gpu."""
import ctypes
import os
import sys

import numpy as np
import pytest

from oracle import g2 as og

pytestmark = pytest.mark.gpu


def _announce(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def _stack(seed, n=300, shape=(20, 24), mean=0.6):
    from dynamix_amd.synth import speckle_frames
    return speckle_frames(shape, n, mean, 25.0, seed, np.uint8)


def _qmask(seed, shape=(20, 24), nbins=4):
    rng = np.random.default_rng(seed)
    q = rng.integers(0, nbins + 1, size=shape).astype(np.int32)
    for b in range(1, nbins + 1):
        q.flat[b] = b
    return q


def test_graph_replays_equal_eager():
    """Several replays of one captured graph, with eager work and synchronisations in
    between, each against the eager result of the same frames."""
    import torch
    from dynamix_amd.correlator.dense import MatMulCorrelator
    _announce("graph replays")
    q = _qmask(1)
    eager = MatMulCorrelator((20, 24), 300, q)
    graph = MatMulCorrelator((20, 24), 300, q, extra_options={"graph": True})
    for rep, seed in enumerate([3, 4, 3, 5, 6]):
        f = _stack(seed)
        g_e = eager.correlate(f)
        n_e = eager.last_numerators.clone()
        g_g = graph.correlate(torch.from_numpy(f).cuda())
        torch.cuda.synchronize()
        assert torch.equal(n_e, graph.last_numerators), "replay %d" % rep
        np.testing.assert_array_equal(g_e.view(np.uint32), g_g.view(np.uint32))
        # unrelated eager work between replays
        x = torch.randn(1 << 16, device="cuda")
        (x * 2).sum().item()
    ref = og.dense_g2(_stack(6), q)
    np.testing.assert_array_equal(graph.last_numerators.cpu().numpy(), ref["num"])


def test_graph_recaptures_for_new_buffers_and_uint16():
    import torch
    from dynamix_amd import _lib
    from dynamix_amd.correlator.common import Plan
    _announce("graph recapture / uint16")
    rng = np.random.default_rng(9)
    f = rng.integers(0, 3000, size=(260, 16, 18)).astype(np.uint16)
    q = _qmask(9, (16, 18), 3)
    plan = Plan((16, 18), 260, q, np.uint16)
    ref = og.dense_g2(f, q)
    d = torch.from_numpy(f.astype(np.int32)).to(torch.int16).cuda()
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for k in range(3):  # new output buffers each time: a capture per call
        num = torch.empty((plan.n_bins, 260), dtype=torch.int64, device="cuda")
        sums = torch.empty_like(num)
        den = torch.empty((plan.n_bins, 260), dtype=torch.float64, device="cuda")
        g2 = torch.empty((plan.n_bins, 260), dtype=torch.float32, device="cuda")
        for _ in range(2):  # and a replay of it
            _lib.check(plan.lib.dxg2_correlate_dense_graph(
                plan.handle, _lib.ptr(d), _lib.ptr(num), _lib.ptr(sums), _lib.ptr(den),
                _lib.ptr(g2), st))
            torch.cuda.synchronize()
            np.testing.assert_array_equal(num.cpu().numpy(), ref["num"])
            np.testing.assert_array_equal(sums.cpu().numpy(), ref["sums"])
    plan.close()


@pytest.mark.parametrize("tile,band,group,queue,kchunk", [
    (256, 4, 32, 1, 2047), (256, 4, 32, 0, 2047), (128, 4, 32, 1, 2047), (128, 1, 0, 0, 2047),
    (256, 3, 8, 1, 2047), (128, 16, 64, 1, 2047), (256, 4, 32, 1, 1), (128, 2, 16, 0, 2)])
def test_every_tuning_gives_exact_numerators(tile, band, group, queue, kchunk):
    import torch
    from dynamix_amd import _lib
    from dynamix_amd.correlator.common import Plan
    _announce("tuning %s" % ((tile, band, group, queue, kchunk),))
    f = _stack(12, n=700, shape=(30, 32), mean=0.4)
    q = _qmask(12, (30, 32), 5)
    ref = og.dense_g2(f, q)
    plan = Plan((30, 32), 700, q, np.uint8)
    plan.set_tuning(tile, band, group, queue, kchunk)
    assert plan.get_tuning() == {"tile": tile, "band": band, "group": group, "queue": queue,
                                 "kchunk": kchunk}
    d = torch.from_numpy(f).cuda()
    nb = plan.n_bins
    num = torch.empty((nb, 700), dtype=torch.int64, device="cuda")
    sums = torch.empty_like(num)
    den = torch.empty((nb, 700), dtype=torch.float64, device="cuda")
    g2 = torch.empty((nb, 700), dtype=torch.float32, device="cuda")
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for _ in range(2):
        _lib.check(plan.lib.dxg2_correlate_dense(plan.handle, _lib.ptr(d), _lib.ptr(num),
                                                 _lib.ptr(sums), _lib.ptr(den), _lib.ptr(g2), st))
        torch.cuda.synchronize()
        np.testing.assert_array_equal(num.cpu().numpy(), ref["num"])
        np.testing.assert_array_equal(sums.cpu().numpy(), ref["sums"])
    plan.close()


@pytest.mark.parametrize("kchunk", [1, 2, 0])
def test_uint16_kchunks_give_exact_numerators(kchunk):
    """uint16 with wide pixels handled apart, its HH and cross items split into k-chunks
    (0: the uint16 default) and clipped to the wide k-blocks: exact against the oracle."""
    import torch
    from dynamix_amd import _lib
    from dynamix_amd.correlator.common import Plan
    _announce("uint16 kchunk %d" % kchunk)
    rng = np.random.default_rng(31)
    shape, n = (40, 48), 330
    f = rng.poisson(0.7, size=(n,) + shape).astype(np.uint16)
    hot = rng.choice(shape[0] * shape[1], size=300, replace=False)  # several wide k-blocks
    f.reshape(n, -1)[:, hot] += rng.integers(0, 900, size=(n, hot.size)).astype(np.uint16)
    q = _qmask(31, shape, 3)
    ref = og.dense_g2(f, q)
    plan = Plan(shape, n, q, np.uint16)
    t = plan.get_tuning()
    plan.set_tuning(t["tile"], t["band"], t["group"], t["queue"], kchunk)
    assert plan.get_tuning()["kchunk"] == (kchunk or 512)
    d = torch.from_numpy(f.astype(np.int32)).to(torch.int16).cuda()
    nb = plan.n_bins
    num = torch.empty((nb, n), dtype=torch.int64, device="cuda")
    sums = torch.empty_like(num)
    den = torch.empty((nb, n), dtype=torch.float64, device="cuda")
    g2 = torch.empty((nb, n), dtype=torch.float32, device="cuda")
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(plan.lib.dxg2_correlate_dense(plan.handle, _lib.ptr(d), _lib.ptr(num),
                                             _lib.ptr(sums), _lib.ptr(den), _lib.ptr(g2), st))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(num.cpu().numpy(), ref["num"])
    np.testing.assert_array_equal(sums.cpu().numpy(), ref["sums"])
    plan.close()


def test_bad_tuning_is_rejected():
    from dynamix_amd.correlator.common import Plan
    plan = Plan((8, 8), 40, None, np.uint8)
    for bad in [(64, 4, 32, 1), (256, 0, 32, 1), (256, 4, -1, 1), (256, 4, 32, 2),
                (256, 4, 32, 1, -1), (256, 4, 32, 1, 2048)]:
        with pytest.raises(ValueError):
            plan.set_tuning(*bad)
    plan.close()


def test_autotune_records_and_reapplies(tmp_path, monkeypatch):
    from dynamix_amd import autotune
    from dynamix_amd.correlator.dense import MatMulCorrelator
    _announce("autotune")
    monkeypatch.setenv("DXG2_TUNING_CACHE", str(tmp_path / "tune.json"))
    f = _stack(21, n=520, shape=(24, 30), mean=0.5)
    q = _qmask(21, (24, 30), 3)
    c = MatMulCorrelator((24, 30), 520, q, extra_options={"autotune": True})
    g = c.correlate(f)
    ref = og.dense_g2(f, q)
    np.testing.assert_array_equal(c.last_numerators.cpu().numpy(), ref["num"])
    cache = autotune.load_cache()
    assert len(cache) == 1
    (key, entry), = cache.items()
    assert entry["gram_ms"] <= entry["default_ms"]
    assert len(entry["trials"]) >= 10
    plan = c._plans[np.dtype(np.uint8)]
    assert key == autotune.plan_key(plan, 520)
    assert plan.get_tuning() == {k: entry[k] for k in ("tile", "band", "group", "queue",
                                                        "kchunk")}
    c2 = MatMulCorrelator((24, 30), 520, q, extra_options={"autotune": True})
    g2 = c2.correlate(f)
    p2 = c2._plans[np.dtype(np.uint8)]
    assert p2.get_tuning() == plan.get_tuning()
    np.testing.assert_array_equal(g.view(np.uint32), g2.view(np.uint32))


def test_set_tuning_allocation_failure_keeps_plan_whole():
    """A set_tuning whose work-list allocation fails returns ENOMEM (MemoryError) and
    leaves the plan on its previous tuning, still exact; the same tuning succeeds once
    allocations work again."""
    import torch
    from dynamix_amd import _lib
    from dynamix_amd.correlator.common import Plan
    _announce("set_tuning allocation failure")
    lib = _lib.load()
    state = {"fail": False, "live": {}}

    def alloc(nbytes, stream, ctx):
        if state["fail"]:
            return None
        t = torch.empty(int(nbytes) + 256, dtype=torch.uint8, device="cuda")
        addr = (t.data_ptr() + 255) // 256 * 256
        state["live"][addr] = t
        return addr

    def free(ptr, stream, ctx):
        state["live"].pop(int(ptr), None)

    cb_a, cb_f = _lib.ALLOC_FN(alloc), _lib.FREE_FN(free)
    _lib.check(lib.dxg2_set_allocator(cb_a, cb_f, None))
    try:
        q = _qmask(7)
        f = _stack(8)
        d = torch.from_numpy(f).cuda()
        plan = Plan((20, 24), 300, q, np.uint8)
        ref = og.dense_g2(f, q)

        def run():
            num = torch.empty((plan.n_bins, 300), dtype=torch.int64, device="cuda")
            g2 = torch.empty((plan.n_bins, 300), dtype=torch.float32, device="cuda")
            st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
            _lib.check(lib.dxg2_correlate_dense(plan.handle, _lib.ptr(d), _lib.ptr(num), None,
                                                None, _lib.ptr(g2), st))
            torch.cuda.synchronize()
            return num.cpu().numpy(), g2.cpu().numpy()

        num0, g20 = run()
        np.testing.assert_array_equal(num0, ref["num"])
        before = plan.get_tuning()
        new = {"tile": 128, "band": 2, "group": 16, "queue": 1 - before["queue"], "kchunk": 3}
        state["fail"] = True
        with pytest.raises(MemoryError):
            plan.set_tuning(**new)
        state["fail"] = False
        assert plan.get_tuning() == before
        num1, g21 = run()
        np.testing.assert_array_equal(num1, num0)
        np.testing.assert_array_equal(g21.view(np.uint32), g20.view(np.uint32))
        plan.set_tuning(**new)
        assert plan.get_tuning() == new
        num2, _ = run()
        np.testing.assert_array_equal(num2, num0)
        plan.close()
    finally:
        lib.dxg2_set_allocator(_lib.ALLOC_FN(), _lib.FREE_FN(), None)
        torch.cuda.synchronize()
