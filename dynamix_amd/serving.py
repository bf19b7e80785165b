"""Serving loop: continuous batching of mixed dense / event g2 jobs on one GPU.

Jobs arrive by a seeded Poisson process.  Each job draws its frame shape, frame count,
bin count, layout (dense uint8 or pixel-major events) and AR(1) correlation time from
its seed (`make_jobs`), and its input is that job's own speckle data (`job_input`):
frames or events generated on the GPU from the job's seed and correlation time.

The loop is split in two:

  * `Scheduler` -- the policy, in plain Python: it stages the inputs of the next few
    jobs while their reserved footprint fits the memory budget, prepares arriving jobs
    on a small host thread pool, admits a prepared job when its device footprint fits
    the budget next to the jobs already resident and a slot is free, keeps a ready queue,
    and dispatches to `n_streams` compute streams with at most `depth` dispatches in
    flight each.  The dense jobs at the head of the queue (similar sizes, up to
    `max_batch`, under `batch_ops` Gram work) go out as ONE grouped launch, and so do
    up to `max_batch` event jobs (one pair launch over all of them).  Completions are
    polled and retired in stream order.  Jobs that arrive faster than they are served
    wait on the host, holding no device memory, so an overload raises latency rather
    than failing jobs.
    A job that fails -- in preparation, over the budget, at admission, at dispatch or
    at completion -- fails alone: it is released and recorded with its error, and
    every other job carries on.  A grouped dispatch that fails is retried job by job,
    so only the job that fails by itself is lost.
  * `GpuExecutor` -- the device side: plans (dxg2_plan_create), job inputs generated on
    an admission stream, grouped dense launches (dxg2_batch_correlate_dense: per-job
    gathers, one Gram launch over every job's tiles, per-job finalise), grouped event
    launches (dxg2_batch_correlate_events: one pair launch over every job's pixels,
    per-job finalise), completion events, and stream-ordered release of plans
    into torch's caching allocator (dxg2_set_allocator), so nothing waits on hipFree.

`Server` is the two together.  Tests drive the Scheduler with a CPU executor whose
jobs take simulated times and fail on demand.

Every job's numerators are exact integers, so a job's result does not depend on which
jobs share its launch: tests compare each job in the mix with the same job run alone.
"""
import collections
import ctypes
import functools
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .correlator.common import Plan

SERVE_SHAPES = ((256, 256), (514, 1030), (1030, 1065))
DENSE_MEAN = 0.02     # counts per pixel per frame of dense jobs
EVENT_MEAN = 1e-3     # counts per pixel per frame of event jobs


@dataclass
class JobSpec:
    id: int
    arrival: float          # seconds after the start of serving
    shape: tuple
    nframes: int
    n_bins: int
    layout: str             # "dense" or "events"
    tau_c: float            # AR(1) correlation time of the job's speckle, frames
    seed: int


def make_jobs(n_jobs, rate, seed, shapes=SERVE_SHAPES, nframes=(1000, 20000), bins=(10, 100),
              dense_fraction=0.7, tau_c=(10.0, 500.0)):
    """Seeded job stream: Poisson arrivals at `rate` jobs/s; every field of a job is
    drawn from its own seed."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=n_jobs)
    arrivals = np.cumsum(gaps) - gaps[0]
    jobs = []
    for i in range(n_jobs):
        js = int(rng.integers(0, 2 ** 63 - 1))
        r = np.random.default_rng(js)
        shape = tuple(shapes[int(r.integers(0, len(shapes)))])
        n = int(r.integers(nframes[0], nframes[1] + 1))
        nb = int(r.integers(bins[0], bins[1] + 1))
        layout = "dense" if r.random() < dense_fraction else "events"
        tc = float(r.uniform(tau_c[0], tau_c[1]))
        jobs.append(JobSpec(i, float(arrivals[i]), shape, n, nb, layout, tc, js))
    return jobs


@functools.lru_cache(maxsize=16)
def _radius_fraction(shape):
    """Distance of every pixel from the frame centre over the distance to the corner."""
    h, w = shape
    yy, xx = np.mgrid[0:h, 0:w]
    r = np.hypot(yy - (h - 1) / 2.0, xx - (w - 1) / 2.0)
    f = r / (float(r.max()) * (1 + 1e-9))
    f.flags.writeable = False
    return f


def job_qmask(spec):
    """n_bins equal-width annuli about the frame centre out to the corner: bin
    1 + floor(n_bins * r / r_corner)."""
    f = _radius_fraction(tuple(spec.shape))
    return (f * int(spec.n_bins)).astype(np.int32) + 1


def job_input(spec, stream=None, frames_out=None):
    """A job's own input, generated on the GPU from its seed and correlation time:
    dense uint8 frames (N, H, W) at DENSE_MEAN counts per pixel per frame (into
    `frames_out` if given; stream-ordered, no host wait), or pixel-major events
    (times, data, offsets) at EVENT_MEAN (waits for `stream` to size the lists).
    Deterministic: the same spec always gives the same input."""
    import torch
    lib = _lib.load()
    h, w = spec.shape
    mean = DENSE_MEAN if spec.layout == "dense" else EVENT_MEAN
    d_mean = _mean_map(tuple(spec.shape), mean)
    st = ctypes.c_void_p(stream.cuda_stream if stream is not None
                         else torch.cuda.current_stream().cuda_stream)
    seed = int(spec.seed) & 0xFFFFFFFFFFFFFFFF
    if spec.layout == "dense":
        frames = frames_out if frames_out is not None else torch.empty(
            (spec.nframes, h, w), dtype=torch.uint8, device="cuda")
        _lib.check(lib.dxg2_synth_speckle(_lib.ptr(frames), _lib.DXG2_U8, spec.nframes, h, w,
                                          _lib.ptr(d_mean), float(spec.tau_c), seed, st),
                   "dxg2_synth_speckle")
        return frames
    offsets = torch.empty(h * w + 1, dtype=torch.int32, device="cuda")
    nnz = ctypes.c_int64()
    args = (_lib.ptr(d_mean), h, w, spec.nframes, float(spec.tau_c), seed, _lib.DXG2_U8,
            _lib.ptr(offsets))
    _lib.check(lib.dxg2_synth_events(*args, None, None, 0, ctypes.byref(nnz), st),
               "dxg2_synth_events")
    n = max(int(nnz.value), 1)
    times = torch.empty(n, dtype=torch.int32, device="cuda")
    data = torch.empty(n, dtype=torch.uint8, device="cuda")
    _lib.check(lib.dxg2_synth_events(*args, _lib.ptr(times), _lib.ptr(data), n,
                                     ctypes.byref(nnz), st), "dxg2_synth_events")
    return times, data, offsets


_MEAN_MAPS = {}


def _mean_map(shape, mean):
    import torch
    key = (shape, mean, torch.cuda.current_device())
    t = _MEAN_MAPS.get(key)
    if t is None:
        t = torch.full(shape, float(mean), dtype=torch.float32, device="cuda")
        _MEAN_MAPS[key] = t
    return t


@dataclass
class Job:
    spec: JobSpec
    plan: object = None
    qmask: np.ndarray = None
    inputs: dict = field(default_factory=dict)
    outputs: tuple = ()
    footprint: int = 0
    t_admit: float = 0.0
    t_dispatch: float = 0.0
    t_done: float = 0.0
    npix: int = 0
    ops: float = 0.0
    prep_s: float = 0.0
    t_prepared: float = 0.0   # clock when preparation finished
    batch_size: int = 1
    solo: bool = False        # dispatch alone (after a grouped dispatch failed)
    error: str = None
    result: dict = None
    ready: object = None          # device event: outputs allocated, dense frames generated
    input_ready: object = None    # device event: event input generated


# ---------------------------------------------------------------------- scheduler
class Scheduler(object):
    """Admission, batching and dispatch policy over an executor (see module docstring).

    The executor provides prepare(spec, staged) -> Job (host work, may raise; `staged`
    is the future of stage(spec), an optional method that makes a job's input resident
    before the job arrives, or None), admit(job)
    (device memory and input, may raise), dispatch(k, jobs) -> handle (may raise),
    poll(handle) -> bool, complete(handle) -> [error or None per job], retire(job) and
    release(job) (after success / failure; never raise), and `budget` in bytes; with
    stage(), the optional footprint_bound(spec) bounds a job's footprint before it is
    prepared, and that much of the budget (at most all of it) is held from the moment
    its input is staged; a job whose input_bytes(spec) (optional) exceed the budget
    fails on arrival, unstaged.
    """

    def __init__(self, executor, max_batch=8, n_streams=2, depth=2, max_resident=32,
                 batch_ops=1e14, prep_threads=4, lookahead=6, clock=time.perf_counter,
                 sleep=time.sleep):
        self.ex = executor
        # Inputs of the next `lookahead` jobs are staged before they arrive (the job's
        # frames are resident when it is submitted) by the executor's optional stage(),
        # counted from the last job admitted, so that an overload queues on the host.
        self.lookahead = lookahead
        self.producer = ThreadPoolExecutor(max_workers=1)
        self.max_batch = max_batch
        self.n_streams = n_streams
        self.depth = depth
        self.max_resident = max_resident
        self.batch_ops = batch_ops   # Gram work (int8 ops) one grouped launch may collect
        self.prep = ThreadPoolExecutor(max_workers=prep_threads)
        self.clock = clock
        self.sleep = sleep
        self._t0 = 0.0

    def close(self):
        self.prep.shutdown(wait=True)
        self.producer.shutdown(wait=True)

    def _prepare(self, spec, staged):
        try:
            job = self.ex.prepare(spec, staged)
        except Exception as e:  # the job fails alone
            job = Job(spec)
            job.error = "prepare: %s: %s" % (type(e).__name__, e)
        job.t_prepared = self.clock() - self._t0
        return job

    def _group(self, ready):
        """Jobs for one grouped launch: the head of the queue, then later jobs of its
        layout.  Dense: no larger than twice the head, while the total Gram work stays
        under batch_ops (small jobs share a launch; a large job runs alone).  Events:
        up to max_batch jobs in one pair launch."""
        head = ready[0]
        dense = head.spec.layout == "dense"
        group, total = [head], head.ops
        if not head.solo:
            for j in list(ready)[1:]:
                if len(group) >= self.max_batch:
                    break
                if j.spec.layout != head.spec.layout or j.solo:
                    continue
                if dense and (j.ops > 2 * head.ops or total + j.ops > self.batch_ops):
                    continue
                group.append(j)
                total += j.ops
        for j in group:
            ready.remove(j)
        return group

    def run(self, specs, realtime=True):
        """Serve `specs` (arrival times in seconds from now if realtime, else all at
        once).  Returns per-job records and summary statistics."""
        ex = self.ex
        order = sorted(specs, key=lambda s: s.arrival)
        pending = collections.deque(order)
        stage = getattr(ex, "stage", None)
        footprint_bound = getattr(ex, "footprint_bound", None)
        input_bytes = getattr(ex, "input_bytes", None)
        staged = {}
        n_staged = 0
        # Inputs are staged, and jobs prepared, at most `lookahead` jobs beyond the last
        # one that left preparation (admitted or failed), and a job's reservation (the
        # executor's footprint_bound) counts against the memory budget from the moment
        # its input is staged: when jobs arrive faster than they are served, the backlog
        # waits on the host, holding no device memory.
        n_passed = 0
        staged_bytes = 0
        holds = {}                       # job id -> bytes reserved when it was staged
        too_big = set()                  # jobs whose input alone exceeds the budget
        arrived = collections.deque()    # arrived, not yet in preparation
        waiting = collections.deque()    # in preparation or prepared
        ready = collections.deque()      # admitted, not dispatched
        busy = [collections.deque() for _ in range(self.n_streams)]
        done, failed = [], []
        resident_bytes, n_resident = 0, 0
        t0 = self._t0 = self.clock()
        clock = lambda: self.clock() - t0

        def leave_waiting(j):
            nonlocal staged_bytes, n_passed
            staged_bytes -= holds.pop(j.spec.id, 0)
            n_passed += 1

        def fail(jobs, err, held):
            """held: the jobs were admitted (their footprint is resident); otherwise they
            fail on their way out of preparation."""
            nonlocal resident_bytes, n_resident
            t = clock()
            for j in jobs:
                j.error = j.error or err
                j.t_done = t
                ex.release(j)
                if held:
                    resident_bytes -= j.footprint
                    n_resident -= 1
                else:
                    leave_waiting(j)
                failed.append(j)

        while pending or arrived or waiting or ready or any(busy):
            now = clock()
            progressed = False
            while stage and n_staged < len(order) and n_staged - n_passed < self.lookahead:
                spec = order[n_staged]
                est = min(int(footprint_bound(spec)), ex.budget) if footprint_bound else 0
                if input_bytes and int(input_bytes(spec)) > ex.budget:
                    too_big.add(spec.id)
                elif resident_bytes + staged_bytes + est > ex.budget and \
                        (n_staged > n_passed or n_resident > 0):
                    break    # staged when memory frees up
                else:
                    staged[spec.id] = self.producer.submit(stage, spec)
                    holds[spec.id] = est
                    staged_bytes += est
                n_staged += 1
            while pending and (not realtime or pending[0].arrival <= now):
                arrived.append(pending.popleft())
            while arrived:
                spec = arrived[0]
                if spec.id in too_big:
                    arrived.popleft()
                    job = Job(spec)
                    job.t_prepared = clock()
                    fail([job], "MemoryError: input needs %d bytes, budget %d"
                         % (int(input_bytes(spec)), ex.budget), False)
                    progressed = True
                    continue
                if stage and spec.id not in staged:
                    break    # its input is not staged yet
                if not stage and len(waiting) >= self.lookahead:
                    break
                arrived.popleft()
                waiting.append(self.prep.submit(self._prepare, spec, staged.pop(spec.id, None)))
                progressed = True
            # completions, in stream order
            for q in busy:
                while q and ex.poll(q[0][0]):
                    handle, jobs = q.popleft()
                    try:
                        errs = ex.complete(handle)
                    except Exception as e:
                        errs = ["complete: %s: %s" % (type(e).__name__, e)] * len(jobs)
                    t = clock()
                    for j, err in zip(jobs, errs):
                        if err:
                            fail([j], err, True)
                            continue
                        j.t_done = t
                        ex.retire(j)
                        resident_bytes -= j.footprint
                        n_resident -= 1
                        done.append(j)
                    progressed = True
            # admission by footprint and slots
            while waiting and n_resident < self.max_resident and waiting[0].done():
                job = waiting[0].result()
                if job.error:
                    waiting.popleft()
                    fail([job], job.error, False)
                    progressed = True
                    continue
                if job.footprint > ex.budget:
                    waiting.popleft()
                    fail([job], "MemoryError: needs %d bytes, budget %d"
                         % (job.footprint, ex.budget), False)
                    progressed = True
                    continue
                # the job's reservation is replaced by its footprint
                others = staged_bytes - holds.get(job.spec.id, 0)
                if n_resident > 0 and resident_bytes + others + job.footprint > ex.budget:
                    break
                waiting.popleft()
                job.t_admit = clock()
                try:
                    ex.admit(job)
                except Exception as e:
                    fail([job], "admit: %s: %s" % (type(e).__name__, e), False)
                    progressed = True
                    continue
                leave_waiting(job)
                resident_bytes += job.footprint
                n_resident += 1
                ready.append(job)
                progressed = True
            # dispatch to the least loaded streams
            while ready:
                k = min(range(self.n_streams), key=lambda i: len(busy[i]))
                if len(busy[k]) >= self.depth:
                    break
                group = self._group(ready)
                t = clock()
                try:
                    handle = ex.dispatch(k, group)
                except Exception as e:
                    if len(group) > 1:    # retry each job alone, at the head of the queue
                        for j in reversed(group):
                            j.solo = True
                            ready.appendleft(j)
                    else:
                        fail(group, "dispatch: %s: %s" % (type(e).__name__, e), True)
                    progressed = True
                    continue
                for j in group:
                    j.t_dispatch = t
                    j.batch_size = len(group)
                busy[k].append((handle, group))
                progressed = True
            if not progressed:
                self.sleep(0.0002)
        ex.finish()
        st = summarize(done)
        st["failed"] = len(failed)
        st["failures"] = [(j.spec.id, j.error) for j in failed]
        st["failed_records"] = failed
        return st


# ------------------------------------------------------------------ GPU executor
_INSTALLED = []


class TorchAllocator(object):
    """Routes libdxg2's device allocations to torch's caching allocator while installed
    (dxg2_set_allocator): releasing a plan then returns its memory to torch's cache
    instead of calling hipFree, which waits for every kernel in flight.  Plans must be
    released only after their work has completed (the serving loop retires a job after
    its completion event)."""

    def __init__(self):
        import torch
        self.lib = _lib.load()
        dev = torch.cuda.current_device()

        def alloc(nbytes, stream, ctx):
            try:
                return torch.cuda.caching_allocator_alloc(int(nbytes), dev, int(stream or 0))
            except RuntimeError:  # out of memory: the library reports DXG2_ENOMEM
                return None

        def free(ptr, stream, ctx):
            torch.cuda.caching_allocator_delete(int(ptr))

        self._alloc = _lib.ALLOC_FN(alloc)  # keep the callbacks alive while installed
        self._free = _lib.FREE_FN(free)

    def __enter__(self):
        _lib.check(self.lib.dxg2_set_allocator(self._alloc, self._free, None), "dxg2_set_allocator")
        return self

    def __exit__(self, *exc):
        self.lib.dxg2_set_allocator(_lib.ALLOC_FN(), _lib.FREE_FN(), None)
        return False


class GpuExecutor(object):
    """Device side of the serving loop (see module docstring)."""

    def __init__(self, hbm_budget=None, reserve_bytes=8 << 30, n_streams=2, depth=2,
                 keep_results=False, persistent_gram=True):
        import threading
        import torch
        self.torch = torch
        self.lib = _lib.load()
        # Job memory comes from torch's caching allocator (TorchAllocator): what the
        # device has free plus what torch holds but is not using.
        free, _total = torch.cuda.mem_get_info()
        idle = torch.cuda.memory_reserved() - torch.cuda.memory_allocated()
        self.budget = int(hbm_budget) if hbm_budget else \
            max(0, int(free) + int(idle) - int(reserve_bytes))
        # A persistent Gram launch holds every CU until its queues drain, so kernels of
        # jobs on the other streams wait behind it; persistent_gram=False gives single
        # dense jobs one workgroup per tile instead, which the hardware interleaves with
        # other streams' work.
        self.plan_flags = 0 if persistent_gram else _lib.DXG2_FLAG_NO_QUEUE
        self.keep_results = keep_results
        self.streams = [torch.cuda.Stream() for _ in range(n_streams)]
        self.admit_stream = torch.cuda.Stream()   # dense job inputs
        # Event inputs are generated and sized while the job is prepared, on a stream of
        # their own (one call at a time): the host waits for that stream only.
        self.input_stream = torch.cuda.Stream()
        self.input_lock = threading.Lock()
        # Finished plans are freed on an otherwise idle stream: their work has completed,
        # so the frees complete at once and the pool can hand the memory to any stream
        # without making it wait for work queued behind a free on a busy stream.
        self.release_stream = torch.cuda.Stream()
        self.allocator = TorchAllocator().__enter__()  # before any plan or batch exists
        _INSTALLED.append(self.allocator)  # callbacks outlive every plan made with them
        # One batch object (job table + staged work list) per in-flight slot: a slot's
        # staging is reused only after its dispatch has completed, so no call waits.
        self.free_batches = [collections.deque() for _ in range(n_streams)]
        for k in range(n_streams):
            for _ in range(depth):
                h = ctypes.c_void_p()
                _lib.check(self.lib.dxg2_batch_create(ctypes.byref(h)), "dxg2_batch_create")
                self.free_batches[k].append(h)

    def close(self):
        for q in self.free_batches:
            while q:
                self.lib.dxg2_batch_destroy(q.popleft())
        self.allocator.__exit__(None, None, None)

    def stage(self, spec):
        """A job's input made resident before the job arrives (the frames or events a
        detector would have delivered): generated on the input stream."""
        with self.input_lock:
            with self.torch.cuda.stream(self.input_stream):
                inp = job_input(spec, self.input_stream)
                ev = self.torch.cuda.Event()
                ev.record(self.input_stream)
        return inp, ev

    def input_bytes(self, spec):
        """Device bytes of a job's input before it is made: the frames of a dense job;
        for an event job the offsets alone (its event lists are sized when made)."""
        h, w = spec.shape
        if spec.layout == "dense":
            return spec.nframes * h * w
        return (h * w + 1) * 4

    def footprint_bound(self, spec):
        """Device bytes a job may hold, bounded before its plan exists (the scheduler
        reserves them when the job's input is staged): the input (dense frames; event
        CSR lists at 1.25 times the expected number of events, as a pixel-frame holds an
        event with probability below EVENT_MEAN), the outputs and per-bin frame rows, the
        pixel lists, for dense jobs the slab (every pixel binned, each bin padded to whole
        64-pixel k-blocks, slab rows padded by up to 256) and the Gram work lists (one
        item per 256-frame tile pair per bin and k-chunk, in two lists)."""
        h, w = spec.shape
        hw, n, nb = h * w, spec.nframes, max(int(spec.n_bins), 1)
        rows = nb * (n + 2) * 8 * 4 + nb * n * (8 + 8 + 8 + 4) + hw * 4 + (1 << 16)
        if spec.layout == "dense":
            nkb = hw // 64 + nb
            nt = (n + 255) // 256
            items = nt * (nt + 1) // 2 * (nb + nkb // 2047 + 1)
            lists = 2 * 40 * items + 32 * nkb
            return n * hw + (hw + 64 * nb) * (n + 256) + rows + lists
        return int(1.25 * EVENT_MEAN * n * hw) * 5 + 4 * (hw + 1) * 4 + rows + hw

    def prepare(self, spec, staged=None):
        """Plan and footprint of a job; its input comes from stage() (waited for if it
        is late) or, without one, is generated here (events) or at admission (frames)."""
        t0 = time.perf_counter()
        job = Job(spec)
        inp = None
        if staged is not None:
            inp, job.input_ready = staged.result()
        elif spec.layout != "dense":
            inp, job.input_ready = self.stage(spec)
        if inp is not None and spec.layout == "dense":
            job.inputs["frames"] = inp
        elif inp is not None:
            t, d, o = inp
            job.inputs.update(times=t, data=d, offsets=o, nnz=int(t.shape[0]))
        job.qmask = job_qmask(spec)
        job.plan = Plan(spec.shape, spec.nframes, job.qmask, np.uint8, flags=self.plan_flags)
        nb, n = job.plan.n_bins, spec.nframes
        job.npix = int(job.plan.pixel_counts[:nb].sum())
        job.ops = 2.0 * job.npix * n * (n + 1) / 2.0
        out_bytes = nb * n * (8 + 8 + 8 + 4)
        h, w = spec.shape
        if spec.layout == "dense":
            job.footprint = job.plan.footprint(0) + out_bytes + n * h * w
        else:
            job.footprint = job.plan.footprint(3) + out_bytes + job.inputs["nnz"] * 5 + \
                (h * w + 1) * 4
        job.prep_s = time.perf_counter() - t0
        return job

    def admit(self, job):
        """Outputs of a job, and a dense job's frames generated on the admission stream."""
        torch = self.torch
        spec = job.spec
        nb, n = job.plan.n_bins, spec.nframes
        s = self.admit_stream
        with torch.cuda.stream(s):
            job.outputs = (torch.empty((nb, n), dtype=torch.int64, device="cuda"),
                           torch.empty((nb, n), dtype=torch.int64, device="cuda"),
                           torch.empty((nb, n), dtype=torch.float64, device="cuda"),
                           torch.empty((nb, n), dtype=torch.float32, device="cuda"))
            if spec.layout == "dense" and "frames" not in job.inputs:
                frames = torch.empty((n,) + tuple(spec.shape), dtype=torch.uint8,
                                     device="cuda")
                job_input(spec, s, frames_out=frames)
                job.inputs["frames"] = frames
        job.ready = torch.cuda.Event()
        job.ready.record(s)

    def dispatch(self, k, jobs):
        """One dispatch on compute stream k: a grouped dense launch or one event job."""
        torch = self.torch
        s = self.streams[k]
        st = ctypes.c_void_p(s.cuda_stream)
        batch = None
        for j in jobs:
            s.wait_event(j.ready)
            if getattr(j, "input_ready", None) is not None:
                s.wait_event(j.input_ready)   # events generated during preparation
        if jobs[0].spec.layout == "dense" and len(jobs) == 1:
            j = jobs[0]  # alone: the plan's resident work list, no per-call list building
            _lib.check(self.lib.dxg2_correlate_dense(
                j.plan.handle, _lib.ptr(j.inputs["frames"]), *[_lib.ptr(x) for x in j.outputs],
                st), "dxg2_correlate_dense")
        elif jobs[0].spec.layout == "dense":
            n = len(jobs)
            arr = lambda vals: (ctypes.c_void_p * n)(*vals)
            batch = self.free_batches[k].popleft()
            try:
                _lib.check(self.lib.dxg2_batch_correlate_dense(
                    batch, n, arr([j.plan.handle.value for j in jobs]),
                    arr([_lib.ptr(j.inputs["frames"]) for j in jobs]),
                    arr([_lib.ptr(j.outputs[0]) for j in jobs]),
                    arr([_lib.ptr(j.outputs[1]) for j in jobs]),
                    arr([_lib.ptr(j.outputs[2]) for j in jobs]),
                    arr([_lib.ptr(j.outputs[3]) for j in jobs]), st),
                    "dxg2_batch_correlate_dense")
            except Exception:
                self.free_batches[k].append(batch)   # the slot is not in flight
                raise
        elif len(jobs) == 1:
            j = jobs[0]
            _lib.check(self.lib.dxg2_correlate_events_csr(
                j.plan.handle, _lib.ptr(j.inputs["times"]), _lib.ptr(j.inputs["data"]),
                _lib.ptr(j.inputs["offsets"]), j.inputs["nnz"],
                *[_lib.ptr(x) for x in j.outputs], st), "dxg2_correlate_events_csr")
        else:
            n = len(jobs)
            arr = lambda vals: (ctypes.c_void_p * n)(*vals)
            batch = self.free_batches[k].popleft()
            try:
                _lib.check(self.lib.dxg2_batch_correlate_events(
                    batch, n, arr([j.plan.handle.value for j in jobs]),
                    arr([_lib.ptr(j.inputs["times"]) for j in jobs]),
                    arr([_lib.ptr(j.inputs["data"]) for j in jobs]),
                    arr([_lib.ptr(j.inputs["offsets"]) for j in jobs]),
                    (ctypes.c_int64 * n)(*[j.inputs["nnz"] for j in jobs]),
                    *[arr([_lib.ptr(j.outputs[i]) for j in jobs]) for i in range(4)], st),
                    "dxg2_batch_correlate_events")
            except Exception:
                self.free_batches[k].append(batch)   # the slot is not in flight
                raise
        for j in jobs:
            for t in list(j.outputs) + [v for v in j.inputs.values() if hasattr(v, "record_stream")]:
                t.record_stream(s)
        ev = torch.cuda.Event()
        ev.record(s)
        return (k, ev, batch, len(jobs))

    def poll(self, handle):
        return handle[1].query()

    def complete(self, handle):
        k, _ev, batch, n = handle
        if batch is not None:
            self.free_batches[k].append(batch)
        return [None] * n

    def _drop(self, job):
        job.outputs = ()
        job.inputs = {}
        if job.plan is not None:
            job.plan.close_async(self.release_stream)
            job.plan = None

    def retire(self, job):
        if self.keep_results:
            num, sums, den, g2 = job.outputs
            job.result = {"num": num.cpu().numpy(), "sums": sums.cpu().numpy(),
                          "g2": g2.cpu().numpy()}
        self._drop(job)

    def release(self, job):
        try:
            if job.plan is not None:
                # its work (if any was queued) must be done before the plan's memory goes
                self.torch.cuda.synchronize()
            self._drop(job)
        except Exception:
            job.plan = None

    def finish(self):
        self.torch.cuda.synchronize()


class Server(object):
    """Continuous-batching server for dense and event g2 jobs: a Scheduler over a
    GpuExecutor (see module docstring)."""

    def __init__(self, hbm_budget=None, reserve_bytes=8 << 30, max_batch=8, n_streams=2,
                 max_resident=32, keep_results=False, batch_ops=1e14, prep_threads=4,
                 depth=2, persistent_gram=True, lookahead=6):
        self.executor = GpuExecutor(hbm_budget, reserve_bytes, n_streams, depth, keep_results,
                                    persistent_gram)
        self.scheduler = Scheduler(self.executor, max_batch=max_batch, n_streams=n_streams,
                                   depth=depth, max_resident=max_resident, batch_ops=batch_ops,
                                   prep_threads=prep_threads, lookahead=lookahead)
        self.streams = self.executor.streams

    @property
    def budget(self):
        return self.executor.budget

    @budget.setter
    def budget(self, v):
        self.executor.budget = int(v)

    def run(self, specs, realtime=True):
        return self.scheduler.run(specs, realtime)

    def close(self):
        self.scheduler.close()
        self.executor.close()


def summarize(done):
    """Latency percentiles and throughput of served jobs."""
    if not done:
        return {"jobs": 0, "records": []}
    lat = np.array([j.t_done - j.spec.arrival for j in done])
    phases = {"prepare": np.array([max(j.t_prepared - j.spec.arrival, 0.0) for j in done]),
              "admission": np.array([j.t_admit - max(j.t_prepared, j.spec.arrival) for j in done]),
              "queue": np.array([j.t_dispatch - j.t_admit for j in done]),
              "execute": np.array([j.t_done - j.t_dispatch for j in done])}
    first = min(j.spec.arrival for j in done)
    last = max(j.t_done for j in done)
    pf = sum(j.npix * j.spec.nframes for j in done)
    dense = [j for j in done if j.spec.layout == "dense"]
    return {"jobs": len(done),
            "dense_jobs": len(dense),
            "event_jobs": len(done) - len(dense),
            "wall_s": last - first,
            "jobs_per_s": len(done) / (last - first) if last > first else None,
            "latency_p50_ms": float(np.percentile(lat, 50) * 1e3),
            "latency_p99_ms": float(np.percentile(lat, 99) * 1e3),
            "latency_max_ms": float(lat.max() * 1e3),
            "phase_p50_ms": {k: float(np.percentile(v, 50) * 1e3) for k, v in phases.items()},
            "phase_p99_ms": {k: float(np.percentile(v, 99) * 1e3) for k, v in phases.items()},
            "mean_batch_size": float(np.mean([j.batch_size for j in dense])) if dense else None,
            "pixel_frames_per_s": pf / (last - first) if last > first else None,
            "records": done}
