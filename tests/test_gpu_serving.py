"""Serving loop on the GPU: every job's result in the mix equals the same job run alone
(exact integer numerators, identical g2), under grouped launches, several streams and
a tight memory budget; a job that cannot run fails alone.  Marked gpu."""
import ctypes
import sys

import numpy as np
import pytest

from oracle import events as ev
from oracle import g2 as og

pytestmark = pytest.mark.gpu

SHAPES = ((48, 40), (64, 64))


def _announce(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def _jobs(n, seed, shapes=SHAPES, nframes=(100, 600), bins=(2, 7), dense_fraction=0.6):
    from dynamix_amd.serving import make_jobs
    return make_jobs(n, rate=2000.0, seed=seed, shapes=shapes, nframes=nframes, bins=bins,
                     dense_fraction=dense_fraction)


def alone(spec):
    """The job by itself through the single-plan entry points, on its own input."""
    import torch
    from dynamix_amd import _lib
    from dynamix_amd.correlator.common import Plan
    from dynamix_amd.serving import job_input, job_qmask
    q = job_qmask(spec)
    plan = Plan(spec.shape, spec.nframes, q, np.uint8)
    nb, n = plan.n_bins, spec.nframes
    num = torch.empty((nb, n), dtype=torch.int64, device="cuda")
    sums = torch.empty_like(num)
    den = torch.empty((nb, n), dtype=torch.float64, device="cuda")
    g2 = torch.empty((nb, n), dtype=torch.float32, device="cuda")
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    outs = [_lib.ptr(x) for x in (num, sums, den, g2)]
    inp = job_input(spec)
    if spec.layout == "dense":
        _lib.check(plan.lib.dxg2_correlate_dense(plan.handle, _lib.ptr(inp), *outs, st))
    else:
        t, d, o = inp
        _lib.check(plan.lib.dxg2_correlate_events_csr(plan.handle, _lib.ptr(t), _lib.ptr(d),
                                                      _lib.ptr(o), int(t.shape[0]), *outs, st))
    torch.cuda.synchronize()
    res = {"num": num.cpu().numpy(), "sums": sums.cpu().numpy(), "g2": g2.cpu().numpy()}
    plan.close()
    return res, inp, q


def oracle_of(spec, inp, q):
    """(sums, num) of a job's input by the oracle: dense frames or events."""
    _, bins = og.qmask_bins(q.shape, q)
    if spec.layout == "dense":
        return og.sums_numerators_fast(inp.cpu().numpy(), q, bins)
    t, d, o = (x.cpu().numpy() for x in inp)
    return ev.event_sums_numerators_fast(t, d, o.view(np.uint32), q, spec.nframes, bins)


def check_mix_equals_alone(done, oracle_layouts=("dense", "events")):
    checked = set()
    for j in sorted(done, key=lambda j: j.npix * j.spec.nframes):
        res, inp, q = alone(j.spec)
        np.testing.assert_array_equal(j.result["num"], res["num"])
        np.testing.assert_array_equal(j.result["sums"], res["sums"])
        np.testing.assert_array_equal(j.result["g2"].view(np.uint32), res["g2"].view(np.uint32))
        if j.spec.layout in oracle_layouts and j.spec.layout not in checked:
            checked.add(j.spec.layout)     # the smallest job of each layout
            s, num = oracle_of(j.spec, inp, q)
            np.testing.assert_array_equal(res["num"], num)
            np.testing.assert_array_equal(res["sums"], s)
    return checked


def test_mix_equals_alone():
    from dynamix_amd.serving import Server
    _announce("serving mix vs alone")
    specs = _jobs(14, seed=3)
    assert {s.layout for s in specs} == {"dense", "events"}
    srv = Server(max_batch=4, n_streams=2, depth=1, keep_results=True)
    try:
        stats = srv.run(specs, realtime=False)   # all at once: dense jobs queue and group
    finally:
        srv.close()
    assert stats["jobs"] == len(specs) and stats["failed"] == 0, stats["failures"]
    assert stats["mean_batch_size"] > 1.0, "no grouped launch happened"
    assert any(getattr(j, "batch_size", 1) > 1 for j in stats["records"]
               if j.spec.layout == "events"), "no grouped event launch happened"
    assert check_mix_equals_alone(stats["records"]) == {"dense", "events"}


def test_each_job_has_its_own_correlation_time():
    """Two jobs that differ only in tau_c get different inputs, and each job's input is
    reproducible from its spec."""
    import dataclasses
    import torch
    from dynamix_amd.serving import job_input
    _announce("serving per-job tau_c")
    spec = _jobs(1, seed=21, dense_fraction=1.0)[0]
    a = job_input(spec)
    b = job_input(spec)
    c = job_input(dataclasses.replace(spec, tau_c=spec.tau_c * 3 + 7))
    torch.cuda.synchronize()
    assert torch.equal(a, b) and not torch.equal(a, c)


def test_tight_memory_budget_and_failing_job():
    """A budget of about two jobs forces admission to wait for retired jobs; a job over
    the budget fails alone and every other job completes."""
    from dynamix_amd.serving import Server
    _announce("serving tight budget + over-budget job")
    specs = _jobs(8, seed=11)
    big = _jobs(1, seed=12, shapes=((600, 700),), nframes=(5000, 5000), dense_fraction=1.0)[0]
    big.id, big.arrival = 99, specs[3].arrival
    srv = Server(keep_results=True, max_batch=3, n_streams=2)
    fp = max(srv.executor.prepare(s).footprint for s in specs)
    srv.budget = int(2.2 * fp)
    try:
        stats = srv.run(specs + [big])
    finally:
        srv.close()
    assert stats["jobs"] == len(specs)
    assert stats["failed"] == 1 and stats["failures"][0][0] == 99
    assert "MemoryError" in stats["failures"][0][1]
    check_mix_equals_alone(stats["records"], oracle_layouts=())
