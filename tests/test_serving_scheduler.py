"""The serving scheduler on the CPU, over a fake executor: jobs take simulated device
times, and chosen jobs fail in preparation, admission, dispatch or completion, or do not
fit the memory budget.  Each failure stays with its job; every other job completes,
memory and slots are accounted, and grouped launches happen."""
import threading
import time

import pytest

from dynamix_amd.serving import Job, Scheduler, make_jobs


class FakeExecutor(object):
    """Jobs run `ms_per_gop` of simulated time per 1e9 Gram ops on each stream."""

    def __init__(self, budget, fail=None, footprint=lambda s: 10 << 20, ms_per_gop=0.002):
        self.budget = budget
        self.fail = fail or {}
        self.footprint = footprint
        self.ms_per_gop = ms_per_gop
        self.lock = threading.Lock()
        self.resident = 0
        self.peak = 0
        self.retired, self.released = [], []
        self.dispatches = []
        self.stream_free_at = {}

    def _maybe_fail(self, stage, jobs):
        for j in jobs:
            if self.fail.get(j.spec.id) == stage:
                raise RuntimeError("injected %s failure of job %d" % (stage, j.spec.id))

    def prepare(self, spec, staged=None):
        job = Job(spec)
        job.npix = spec.shape[0] * spec.shape[1]
        job.ops = 2.0 * job.npix * spec.nframes * (spec.nframes + 1) / 2.0
        job.footprint = self.footprint(spec)
        self._maybe_fail("prepare", [job])
        return job

    def admit(self, job):
        self._maybe_fail("admit", [job])
        with self.lock:
            self.resident += job.footprint
            self.peak = max(self.peak, self.resident)
        job.admitted = True

    def dispatch(self, k, jobs):
        self._maybe_fail("dispatch", jobs)
        now = time.perf_counter()
        start = max(now, self.stream_free_at.get(k, now))
        end = start + sum(j.ops for j in jobs) / 1e9 * self.ms_per_gop / 1e3
        self.stream_free_at[k] = end
        self.dispatches.append([j.spec.id for j in jobs])
        return (end, [j.spec.id for j in jobs])

    def poll(self, handle):
        return time.perf_counter() >= handle[0]

    def complete(self, handle):
        return ["injected completion error" if self.fail.get(i) == "complete" else None
                for i in handle[1]]

    def _free(self, job):
        if getattr(job, "admitted", False):
            with self.lock:
                self.resident -= job.footprint
            job.admitted = False

    def retire(self, job):
        self._free(job)
        self.retired.append(job.spec.id)

    def release(self, job):
        self._free(job)
        self.released.append(job.spec.id)

    def finish(self):
        pass


def _specs(n=40, seed=1, rate=400.0):
    return make_jobs(n, rate, seed, shapes=((64, 64), (128, 96)), nframes=(200, 2000),
                     bins=(2, 20))


def test_all_jobs_complete_and_batch():
    """All jobs at once: the ready queue fills while both streams are busy, so later
    dense jobs go out in grouped dispatches."""
    ex = FakeExecutor(budget=1 << 40)
    sch = Scheduler(ex, max_batch=4, n_streams=2, depth=2)
    try:
        st = sch.run(_specs(), realtime=False)
    finally:
        sch.close()
    assert st["jobs"] == 40 and st["failed"] == 0
    assert sorted(ex.retired) == list(range(40))
    assert ex.resident == 0
    assert any(len(d) > 1 for d in ex.dispatches), "no grouped dispatch"
    layout = {s.id: s.layout for s in _specs()}
    for d in ex.dispatches:  # a launch holds one layout
        assert len({layout[i] for i in d}) == 1, d
    assert any(len(d) > 1 and layout[d[0]] == "events" for d in ex.dispatches), \
        "no grouped event dispatch"


def test_realtime_arrivals_latency():
    ex = FakeExecutor(budget=1 << 40)
    sch = Scheduler(ex, max_batch=4, n_streams=2, depth=2)
    try:
        st = sch.run(_specs(20))
    finally:
        sch.close()
    assert st["jobs"] == 20 and st["failed"] == 0
    assert st["latency_p99_ms"] >= st["latency_p50_ms"] > 0
    assert all(j.t_done >= j.spec.arrival for j in st["records"])


@pytest.mark.parametrize("stage", ["prepare", "admit", "dispatch", "complete"])
def test_a_failing_job_fails_alone(stage):
    bad = {5: stage, 17: stage}
    ex = FakeExecutor(budget=1 << 40, fail=bad)
    sch = Scheduler(ex, max_batch=4, n_streams=2, depth=2)
    try:
        st = sch.run(_specs())
    finally:
        sch.close()
    assert st["failed"] == 2, st["failures"]
    assert sorted(i for i, _ in st["failures"]) == [5, 17]
    assert all(stage in err or "injected" in err for _, err in st["failures"])
    assert st["jobs"] == 38
    assert sorted(ex.retired) == [i for i in range(40) if i not in bad]
    assert ex.resident == 0
    if stage == "dispatch":
        # the grouped launches that held a failing job were retried job by job
        assert all(5 not in d and 17 not in d for d in ex.dispatches)


def test_over_budget_job_fails_alone_and_budget_is_respected():
    big = {7}
    fp = lambda s: (3 << 30) if s.id in big else (300 << 20)
    ex = FakeExecutor(budget=1 << 30, footprint=fp)
    sch = Scheduler(ex, max_batch=4, n_streams=2, depth=2)
    try:
        st = sch.run(_specs(30))
    finally:
        sch.close()
    assert st["failed"] == 1 and st["failures"][0][0] == 7
    assert "MemoryError" in st["failures"][0][1]
    assert st["jobs"] == 29
    assert ex.peak <= ex.budget
    assert ex.resident == 0


def test_slot_limit_is_respected():
    ex = FakeExecutor(budget=1 << 40)
    seen = []
    orig = ex.admit

    def admit(job):
        orig(job)
        seen.append(ex.resident // (10 << 20))

    ex.admit = admit
    sch = Scheduler(ex, max_batch=4, n_streams=2, depth=1, max_resident=3)
    try:
        st = sch.run(_specs(25))
    finally:
        sch.close()
    assert st["jobs"] == 25
    assert max(seen) <= 3


class StagingExecutor(FakeExecutor):
    """Inputs staged ahead of admission: `held` counts the staged inputs of jobs not yet
    admitted, and `peak` the most device memory ever in use (staged + resident).  The
    reservation of a job is its footprint."""

    def __init__(self, budget, input_bytes, **kw):
        FakeExecutor.__init__(self, budget, **kw)
        self.inbytes = input_bytes
        self.held = 0
        self.staged_ids = []

    def input_bytes(self, spec):
        return self.inbytes(spec)

    def footprint_bound(self, spec):
        return self.footprint(spec)

    def stage(self, spec):
        with self.lock:
            self.held += self.inbytes(spec)
            self.peak = max(self.peak, self.held + self.resident)
            self.staged_ids.append(spec.id)
        return spec.id

    def prepare(self, spec, staged=None):
        job = FakeExecutor.prepare(self, spec, staged)
        job.staged = staged is not None
        return job

    def _unstage(self, job):
        if getattr(job, "staged", False):
            with self.lock:
                self.held -= self.inbytes(job.spec)
            job.staged = False

    def admit(self, job):
        FakeExecutor.admit(self, job)
        self._unstage(job)
        with self.lock:
            self.peak = max(self.peak, self.held + self.resident)

    def release(self, job):
        self._unstage(job)
        FakeExecutor.release(self, job)


def test_overload_keeps_staged_inputs_within_budget():
    """Every job arrives at once, far more than fits: inputs are staged only as memory
    frees up (staged + resident bytes never exceed the budget), the backlog waits on the
    host, every job completes, and a job whose input alone exceeds the budget fails
    alone without being staged."""
    big = 11
    inb = lambda s: (5 << 30) if s.id == big else (200 << 20)
    fp = lambda s: inb(s) + (50 << 20)
    ex = StagingExecutor(budget=1 << 30, input_bytes=inb, footprint=fp)
    sch = Scheduler(ex, max_batch=4, n_streams=2, depth=2, lookahead=8)
    try:
        st = sch.run(_specs(40), realtime=False)
    finally:
        sch.close()
    assert st["failed"] == 1 and st["failures"][0][0] == big, st["failures"]
    assert "MemoryError" in st["failures"][0][1]
    assert st["jobs"] == 39
    assert big not in ex.staged_ids and len(ex.staged_ids) == 39
    assert ex.peak <= ex.budget, (ex.peak, ex.budget)
    assert ex.held == 0 and ex.resident == 0
