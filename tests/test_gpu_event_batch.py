"""Grouped event launch (dxg2_batch_correlate_events): several jobs of different shapes,
frame counts (one and several lag windows) and bin counts in one pair launch give the
same numerators, sums, denominators and g2 bits as each job alone, match the event
oracle, and keep each job's input-error flags to its own plan."""
import ctypes
import sys

import numpy as np
import pytest

from oracle import events as ev
from oracle import g2 as og

pytestmark = pytest.mark.gpu


def _announce(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def _job(seed, shape, n, nbins, mean, dtype=np.uint8):
    import torch
    from dynamix_amd.synth import speckle_frames
    f = speckle_frames(shape, n, mean, 20.0, seed, dtype)
    data, times, offsets = ev.compress_all_stack(f)
    q = np.zeros(shape, np.int32)
    q.ravel()[:] = 1 + np.arange(q.size) % nbins
    q.ravel()[::7] = 0                      # some pixels unbinned
    dev = {"times": torch.from_numpy(times.astype(np.int32)).cuda(),
           "data": torch.from_numpy(data).cuda(),
           "offsets": torch.from_numpy(offsets.astype(np.int64).astype(np.uint32).view(np.int32)).cuda(),
           "nnz": int(data.size)}
    return {"q": q, "n": n, "host": (times, data, offsets), "dev": dev}


def _outs(nb, n):
    import torch
    return (torch.empty((nb, n), dtype=torch.int64, device="cuda"),
            torch.empty((nb, n), dtype=torch.int64, device="cuda"),
            torch.empty((nb, n), dtype=torch.float64, device="cuda"),
            torch.empty((nb, n), dtype=torch.float32, device="cuda"))


def _run(jobs, grouped, bad=None):
    import torch
    from dynamix_amd import _lib
    from dynamix_amd.correlator.common import Plan
    plans = [Plan(j["q"].shape, j["n"], j["q"], np.uint8) for j in jobs]
    outs = [_outs(p.n_bins, j["n"]) for p, j in zip(plans, jobs)]
    lib = plans[0].lib
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    times = [j["dev"]["times"] for j in jobs]
    if bad is not None:  # job `bad`: a frame index out of range in a binned pixel
        _t, _d, off = jobs[bad]["host"]
        qb = jobs[bad]["q"].ravel()
        p = next(i for i in range(qb.size) if qb[i] > 0 and off[i + 1] > off[i])
        times[bad] = times[bad].clone()
        times[bad][int(off[p])] = jobs[bad]["n"]
    if grouped:
        n = len(jobs)
        arr = lambda vals: (ctypes.c_void_p * n)(*vals)
        b = ctypes.c_void_p()
        _lib.check(lib.dxg2_batch_create(ctypes.byref(b)))
        try:
            _lib.check(lib.dxg2_batch_correlate_events(
                b, n, arr([p.handle.value for p in plans]), arr([_lib.ptr(t) for t in times]),
                arr([_lib.ptr(j["dev"]["data"]) for j in jobs]),
                arr([_lib.ptr(j["dev"]["offsets"]) for j in jobs]),
                (ctypes.c_int64 * n)(*[j["dev"]["nnz"] for j in jobs]),
                *[arr([_lib.ptr(o[i]) for o in outs]) for i in range(4)], st))
            torch.cuda.synchronize()
        finally:
            lib.dxg2_batch_destroy(b)
    else:
        for p, j, t, o in zip(plans, jobs, times, outs):
            _lib.check(lib.dxg2_correlate_events_csr(
                p.handle, _lib.ptr(t), _lib.ptr(j["dev"]["data"]), _lib.ptr(j["dev"]["offsets"]),
                j["dev"]["nnz"], *[_lib.ptr(x) for x in o], st))
        torch.cuda.synchronize()
    flags = []
    for p in plans:
        f = ctypes.c_int32()
        _lib.check(lib.dxg2_plan_event_status(p.handle, ctypes.byref(f), st))
        flags.append(f.value)
    res = [[x.cpu().numpy() for x in o] for o in outs]
    for p in plans:
        p.close()
    return res, flags


def _jobs():
    return [_job(1, (20, 24), 300, 4, 0.05),
            _job(2, (6, 7), 17000, 2, 0.02),     # several lag windows
            _job(3, (40, 50), 120, 9, 0.1),      # bins over many work items
            _job(4, (3, 3), 40, 1, 0.3)]


def test_grouped_events_equal_each_job_alone_and_the_oracle():
    _announce("grouped event launch vs alone vs oracle")
    jobs = _jobs()
    grouped, gflags = _run(jobs, True)
    alone, aflags = _run(jobs, False)
    assert gflags == aflags == [0] * len(jobs)
    for g, a in zip(grouped, alone):
        for x, y in zip(g, a):
            np.testing.assert_array_equal(x.view(np.uint8), y.view(np.uint8))
    for j, g in zip(jobs, grouped):
        t, d, o = j["host"]
        _, bins = og.qmask_bins(j["q"].shape, j["q"])
        s, num = ev.event_sums_numerators_fast(t, d, o.astype(np.uint32), j["q"], j["n"], bins)
        np.testing.assert_array_equal(g[0], num)
        np.testing.assert_array_equal(g[1], s)


def test_grouped_events_keep_errors_per_job():
    _announce("grouped event launch: a bad job flags only its plan")
    jobs = _jobs()
    grouped, flags = _run(jobs, True, bad=2)
    alone, _ = _run(jobs, False)
    assert flags[2] & 1 and [f for i, f in enumerate(flags) if i != 2] == [0, 0, 0]
    for i in (0, 1, 3):
        for x, y in zip(grouped[i], alone[i]):
            np.testing.assert_array_equal(x.view(np.uint8), y.view(np.uint8))
