"""Serving loop on the GPU: every job's result in the mix equals the same job run alone
(exact integer numerators, identical g2), under grouped launches, several streams and
a tight memory budget.  Marked gpu."""
import ctypes
import sys

import numpy as np
import pytest

from oracle import g2 as og

pytestmark = pytest.mark.gpu

SHAPES = ((48, 40), (64, 64))
POOL_FRAMES = 600


def _announce(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def pool():
    from dynamix_amd.serving import InputPool
    return InputPool(SHAPES, POOL_FRAMES, seed=5, dense_mean=0.3, event_mean=0.05, tau_c=20.0)


def _jobs(n, seed):
    from dynamix_amd.serving import make_jobs
    return make_jobs(n, rate=2000.0, seed=seed, shapes=SHAPES, nframes=(100, POOL_FRAMES),
                     bins=(2, 7), dense_fraction=0.6, pool_frames=POOL_FRAMES)


def _alone(pool, spec):
    """The job run by itself through the single-plan entry points."""
    import torch
    from dynamix_amd import _lib
    from dynamix_amd.correlator.common import Plan
    from dynamix_amd.serving import job_qmask
    q = job_qmask(spec)
    plan = Plan(spec.shape, spec.nframes, q, np.uint8)
    nb, n = plan.n_bins, spec.nframes
    num = torch.empty((nb, n), dtype=torch.int64, device="cuda")
    sums = torch.empty_like(num)
    den = torch.empty((nb, n), dtype=torch.float64, device="cuda")
    g2 = torch.empty((nb, n), dtype=torch.float32, device="cuda")
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    outs = [_lib.ptr(x) for x in (num, sums, den, g2)]
    if spec.layout == "dense":
        f = pool.dense[spec.shape][spec.frame0:spec.frame0 + n]
        _lib.check(plan.lib.dxg2_correlate_dense(plan.handle, _lib.ptr(f), *outs, st))
        frames = f.cpu().numpy()
    else:
        t, d, o = pool.events[spec.shape]
        hw = spec.shape[0] * spec.shape[1]
        cap = max(int(t.shape[0]), 1)
        off = torch.empty(hw + 1, dtype=torch.int32, device="cuda")
        times = torch.empty(cap, dtype=torch.int32, device="cuda")
        data = torch.empty(cap, dtype=torch.uint8, device="cuda")
        scratch = torch.empty((hw + 1023) // 1024 + 1, dtype=torch.int32, device="cuda")
        _lib.check(plan.lib.dxg2_events_window(
            _lib.ptr(o), _lib.ptr(t), _lib.ptr(d), int(t.shape[0]), hw, _lib.DXG2_U8,
            spec.frame0, spec.frame0 + n, _lib.ptr(scratch), _lib.ptr(off), _lib.ptr(times),
            _lib.ptr(data), cap, st))
        _lib.check(plan.lib.dxg2_correlate_events_csr(plan.handle, _lib.ptr(times), _lib.ptr(data),
                                                      _lib.ptr(off), cap, *outs, st))
        frames = None
    torch.cuda.synchronize()
    res = {"num": num.cpu().numpy(), "sums": sums.cpu().numpy(), "g2": g2.cpu().numpy()}
    plan.close()
    return res, frames, q


def _check_mix_equals_alone(pool, done, check_oracle=1):
    checked = 0
    for j in done:
        alone, frames, q = _alone(pool, j.spec)
        np.testing.assert_array_equal(j.result["num"], alone["num"])
        np.testing.assert_array_equal(j.result["sums"], alone["sums"])
        np.testing.assert_array_equal(j.result["g2"].view(np.uint32), alone["g2"].view(np.uint32))
        if frames is not None and checked < check_oracle:
            ref = og.dense_g2(frames, q)
            np.testing.assert_array_equal(alone["num"], ref["num"])
            checked += 1


def test_mix_equals_alone(pool):
    from dynamix_amd.serving import Server
    _announce("serving mix vs alone")
    specs = _jobs(14, seed=3)
    assert {s.layout for s in specs} == {"dense", "events"}
    srv = Server(pool, max_batch=4, n_streams=2, depth=1, keep_results=True)
    try:
        stats = srv.run(specs)
    finally:
        srv.close()
    done = stats["records"]
    assert stats["jobs"] == len(specs)
    assert stats["mean_batch_size"] > 1.0, "no grouped launch happened"
    _check_mix_equals_alone(pool, done)


def test_tight_memory_budget(pool):
    """A budget of about two jobs forces admission to wait for retired plans."""
    from dynamix_amd.serving import Server
    _announce("serving tight budget")
    specs = _jobs(8, seed=11)
    srv = Server(pool, max_batch=3, n_streams=2, keep_results=True)
    fp = max(srv._prepare(s).footprint for s in specs)
    srv.budget = int(2.2 * fp)
    try:
        stats = srv.run(specs)
    finally:
        srv.close()
    assert stats["jobs"] == len(specs)
    _check_mix_equals_alone(pool, stats["records"], check_oracle=0)


def test_events_window_matches_numpy(pool):
    """dxg2_events_window equals the numpy window of the same pixel-major list."""
    import torch
    from dynamix_amd import _lib
    _announce("events window")
    shape = SHAPES[0]
    t, d, o = pool.events[shape]
    hw = shape[0] * shape[1]
    cap = max(int(t.shape[0]), 1)
    lib = _lib.load()
    tn, dn, on = t.cpu().numpy(), d.cpu().numpy(), o.cpu().numpy().astype(np.int64)
    for t0, t1 in [(0, POOL_FRAMES), (37, 250), (599, 600), (300, 300)]:
        off = torch.empty(hw + 1, dtype=torch.int32, device="cuda")
        times = torch.empty(cap, dtype=torch.int32, device="cuda")
        data = torch.empty(cap, dtype=torch.uint8, device="cuda")
        scratch = torch.empty((hw + 1023) // 1024 + 1, dtype=torch.int32, device="cuda")
        st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        _lib.check(lib.dxg2_events_window(_lib.ptr(o), _lib.ptr(t), _lib.ptr(d), int(t.shape[0]),
                                          hw, _lib.DXG2_U8, t0, t1, _lib.ptr(scratch),
                                          _lib.ptr(off), _lib.ptr(times), _lib.ptr(data), cap, st))
        torch.cuda.synchronize()
        pix = np.repeat(np.arange(hw), np.diff(on))
        keep = (tn >= t0) & (tn < t1)
        ref_off = np.zeros(hw + 1, np.int64)
        ref_off[1:] = np.cumsum(np.bincount(pix[keep], minlength=hw))
        got_off = off.cpu().numpy().astype(np.int64)
        np.testing.assert_array_equal(got_off, ref_off)
        m = int(ref_off[-1])
        np.testing.assert_array_equal(times.cpu().numpy()[:m], tn[keep] - t0)
        np.testing.assert_array_equal(data.cpu().numpy()[:m], dn[keep])
