"""Serving loop: continuous batching of mixed dense / event g2 jobs on one GPU.

Jobs arrive by a seeded Poisson process.  Each job draws its frame shape, frame count,
bin count, layout (dense uint8 or pixel-major events) and AR(1) correlation time from
its seed (`make_jobs`).  The server:

  * prepares each arriving job on a small host thread pool (qmask, plan, footprint), so
    plan construction overlaps with dispatching;
  * admits a job when its device footprint (plan workspaces + input + outputs, from
    dxg2_plan_footprint) fits the HBM budget next to the jobs already resident, and a
    slot is free; inputs and partial sums then stay resident in HBM until the result is
    complete;
  * keeps a ready queue of admitted jobs and `n_streams` compute streams (2 by default:
    with the compute streams, the admission and release streams and the library's plan
    set-up stream within the process's hardware queues, no stream waits behind another's
    kernels), each with at
    most `depth` dispatches in flight.  When a stream has room, the dense jobs waiting
    at the head of the queue (similar sizes, up to `max_batch`) go out as ONE grouped
    launch (dxg2_batch_correlate_dense: per-job gathers, one Gram launch over every
    job's tiles, per-job finalise; a job alone uses its plan's resident work list); an
    event job at the head goes out alone.  Work of different streams overlaps on the
    device;
  * polls completion events, records latency = completion - arrival per job, and
    retires the job: its plan's workspaces go back to the device's stream-ordered pool
    (dxg2_plan_destroy_async, no device synchronisation) and its footprint leaves the
    budget at once.

Inputs come from a resident pool (`InputPool`), generated once before serving: per frame
shape, a dense uint8 stack and a pixel-major event list of `pool_frames` frames.  A dense
job uses frames [f0, f0 + N) of its shape's stack (a view, no copy); an event job gets
the events of frames [f0, f0 + N) (dxg2_events_window, on the device).  The pool's
correlation time is fixed per shape (the job's own tau_c is recorded but does not
change its data).

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

from . import _lib, synth
from .correlator.common import Plan

SERVE_SHAPES = ((256, 256), (514, 1030), (1030, 1065))


@dataclass
class JobSpec:
    id: int
    arrival: float          # seconds after the start of serving
    shape: tuple
    nframes: int
    n_bins: int
    layout: str             # "dense" or "events"
    tau_c: float
    seed: int
    frame0: int = 0         # first frame of the job in its shape's input pool


def make_jobs(n_jobs, rate, seed, shapes=SERVE_SHAPES, nframes=(1000, 20000), bins=(10, 100),
              dense_fraction=0.7, tau_c=(10.0, 500.0), pool_frames=20000):
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
        f0 = int(r.integers(0, pool_frames - n + 1))
        jobs.append(JobSpec(i, float(arrivals[i]), shape, n, nb, layout, tc, js, f0))
    return jobs


class InputPool(object):
    """Resident inputs per frame shape: dense uint8 frames and pixel-major events."""

    def __init__(self, shapes, pool_frames, seed, dense_mean=0.02, event_mean=1e-3,
                 tau_c=200.0, layouts=("dense", "events")):
        import torch
        self.pool_frames = pool_frames
        self.dense = {}
        self.events = {}
        for k, shape in enumerate(shapes):
            shape = tuple(shape)
            if "dense" in layouts:
                self.dense[shape] = synth.speckle_frames_device(
                    shape, pool_frames, dense_mean, tau_c, seed + 101 * k, np.uint8)
            if "events" in layouts:
                self.events[shape] = synth.speckle_events_device(
                    shape, pool_frames, event_mean, tau_c, seed + 101 * k + 7, np.uint8)
        torch.cuda.synchronize()

    def nbytes(self):
        n = sum(t.numel() for t in self.dense.values())
        for t, d, o in self.events.values():
            n += t.numel() * 4 + d.numel() + o.numel() * 4
        return n


@dataclass
class Job:
    spec: JobSpec
    plan: Plan = None
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
    t_prepared: float = 0.0   # absolute perf_counter when preparation finished
    batch_size: int = 1
    result: dict = None


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


class Server(object):
    """Continuous-batching server for dense and event g2 jobs (see module docstring)."""

    def __init__(self, pool, hbm_budget=None, reserve_bytes=8 << 30, max_batch=8, n_streams=2,
                 max_resident=32, keep_results=False, batch_ops=1e14, prep_threads=4,
                 depth=2, persistent_gram=True):
        import torch
        self.torch = torch
        self.pool = pool
        self.lib = _lib.load()
        # Job memory comes from torch's caching allocator (TorchAllocator): what the
        # device has free plus what torch holds but is not using.
        free, _total = torch.cuda.mem_get_info()
        idle = torch.cuda.memory_reserved() - torch.cuda.memory_allocated()
        self.budget = int(hbm_budget) if hbm_budget else \
            max(0, int(free) + int(idle) - int(reserve_bytes))
        self.max_batch = max_batch
        # A persistent Gram launch holds every CU until its queues drain, so kernels of
        # jobs on the other streams wait behind it; persistent_gram=False gives single
        # dense jobs one workgroup per tile instead, which the hardware interleaves with
        # other streams' work.  Measured on the 200-job mix: persistent p50 65 ms / p99
        # 310 ms, per-tile 72 ms / 431 ms.
        self.plan_flags = 0 if persistent_gram else _lib.DXG2_FLAG_NO_QUEUE
        self.batch_ops = batch_ops  # Gram work (int8 ops) one grouped launch may collect
        self.max_resident = max_resident
        self.keep_results = keep_results
        self.streams = [torch.cuda.Stream() for _ in range(n_streams)]
        self.admit_stream = torch.cuda.Stream()  # event windows of newly admitted jobs
        # Finished plans are freed on an otherwise idle stream: their work has completed,
        # so the frees complete at once and the pool can hand the memory to any stream
        # without making it wait for work queued behind a free on a busy stream.
        self.release_stream = torch.cuda.Stream()
        # One batch object (job table + staged work list) per in-flight slot: a slot's
        # staging is reused only after its dispatch has completed, so no call waits.
        self.allocator = TorchAllocator().__enter__()  # before any plan or batch exists
        _INSTALLED.append(self.allocator)  # callbacks outlive every plan made with them
        self.free_batches = [collections.deque() for _ in range(n_streams)]
        for k in range(n_streams):
            for _ in range(depth):
                h = ctypes.c_void_p()
                _lib.check(self.lib.dxg2_batch_create(ctypes.byref(h)), "dxg2_batch_create")
                self.free_batches[k].append(h)
        # dispatches in flight per stream, oldest first: (event, [jobs]); at most `depth`
        self.busy = [collections.deque() for _ in range(n_streams)]
        self.depth = depth
        # Host preparation (qmask, plan, footprint) runs beside the scheduler thread.
        self.prep = ThreadPoolExecutor(max_workers=prep_threads)
        self.resident = 0

    def close(self):
        self.prep.shutdown(wait=True)
        for q in self.free_batches:
            while q:
                self.lib.dxg2_batch_destroy(q.popleft())
        self.allocator.__exit__(None, None, None)

    # ------------------------------------------------------------------ jobs
    def _prepare(self, spec):
        """Plan and footprint of a job (host work; no device memory for inputs yet)."""
        t0 = time.perf_counter()
        job = Job(spec)
        job.qmask = job_qmask(spec)
        job.plan = Plan(spec.shape, spec.nframes, job.qmask, np.uint8, flags=self.plan_flags)
        nb, n = job.plan.n_bins, spec.nframes
        job.npix = int(job.plan.pixel_counts[:nb].sum())
        job.ops = 2.0 * job.npix * n * (n + 1) / 2.0
        out_bytes = nb * n * (8 + 8 + 8 + 4)
        if spec.layout == "dense":
            job.footprint = job.plan.footprint(0) + out_bytes
        else:
            t, d, o = self.pool.events[spec.shape]
            cap = max(int(t.shape[0]), 1)
            hw = spec.shape[0] * spec.shape[1]
            job.footprint = (job.plan.footprint(3) + out_bytes + cap * 5 + (hw + 1) * 4 +
                             ((hw + 1023) // 1024 + 1) * 4)
        job.t_prepared = time.perf_counter()
        job.prep_s = job.t_prepared - t0
        return job

    def _admit(self, job, stream):
        """Allocate a job's outputs (and event window) on `stream`; dense frames are a
        view of the pool."""
        torch = self.torch
        spec = job.spec
        nb, n = job.plan.n_bins, spec.nframes
        with torch.cuda.stream(stream):
            job.outputs = (torch.empty((nb, n), dtype=torch.int64, device="cuda"),
                           torch.empty((nb, n), dtype=torch.int64, device="cuda"),
                           torch.empty((nb, n), dtype=torch.float64, device="cuda"),
                           torch.empty((nb, n), dtype=torch.float32, device="cuda"))
            if spec.layout == "dense":
                job.inputs["frames"] = self.pool.dense[spec.shape][spec.frame0:spec.frame0 + n]
            else:
                t, d, o = self.pool.events[spec.shape]
                hw = spec.shape[0] * spec.shape[1]
                cap = max(int(t.shape[0]), 1)
                off = torch.empty(hw + 1, dtype=torch.int32, device="cuda")
                times = torch.empty(cap, dtype=torch.int32, device="cuda")
                data = torch.empty(cap, dtype=torch.uint8, device="cuda")
                scratch = torch.empty((hw + 1023) // 1024 + 1, dtype=torch.int32, device="cuda")
                _lib.check(self.lib.dxg2_events_window(
                    _lib.ptr(o), _lib.ptr(t), _lib.ptr(d), int(t.shape[0]), hw, _lib.DXG2_U8,
                    spec.frame0, spec.frame0 + n, _lib.ptr(scratch), _lib.ptr(off),
                    _lib.ptr(times), _lib.ptr(data), cap,
                    ctypes.c_void_p(stream.cuda_stream)), "dxg2_events_window")
                job.inputs.update(offsets=off, times=times, data=data, scratch=scratch,
                                  nnz=cap)
        job.ready = torch.cuda.Event()
        job.ready.record(stream)

    def _dispatch(self, k, jobs):
        """One dispatch on compute stream k: a grouped dense launch or one event job."""
        torch = self.torch
        s = self.streams[k]
        st = ctypes.c_void_p(s.cuda_stream)
        batch = None
        for j in jobs:
            s.wait_event(j.ready)
        if jobs[0].spec.layout == "dense" and len(jobs) == 1:
            j = jobs[0]  # alone: the plan's resident work list, no per-call list building
            _lib.check(self.lib.dxg2_correlate_dense(
                j.plan.handle, _lib.ptr(j.inputs["frames"]), *[_lib.ptr(x) for x in j.outputs],
                st), "dxg2_correlate_dense")
        elif jobs[0].spec.layout == "dense":
            n = len(jobs)
            arr = lambda vals: (ctypes.c_void_p * n)(*vals)
            batch = self.free_batches[k].popleft()
            _lib.check(self.lib.dxg2_batch_correlate_dense(
                batch, n, arr([j.plan.handle.value for j in jobs]),
                arr([_lib.ptr(j.inputs["frames"]) for j in jobs]),
                arr([_lib.ptr(j.outputs[0]) for j in jobs]),
                arr([_lib.ptr(j.outputs[1]) for j in jobs]),
                arr([_lib.ptr(j.outputs[2]) for j in jobs]),
                arr([_lib.ptr(j.outputs[3]) for j in jobs]), st), "dxg2_batch_correlate_dense")
        else:
            j = jobs[0]
            _lib.check(self.lib.dxg2_correlate_events_csr(
                j.plan.handle, _lib.ptr(j.inputs["times"]), _lib.ptr(j.inputs["data"]),
                _lib.ptr(j.inputs["offsets"]), j.inputs["nnz"],
                *[_lib.ptr(x) for x in j.outputs], st), "dxg2_correlate_events_csr")
        for j in jobs:
            for t in list(j.outputs) + [v for v in j.inputs.values() if hasattr(v, "record_stream")]:
                t.record_stream(s)
            j.batch_size = len(jobs)
        ev = torch.cuda.Event()
        ev.record(s)
        self.busy[k].append((ev, jobs, batch))

    def _group(self, ready):
        """Dense jobs for one grouped launch: the head of the queue, then later dense jobs
        no larger than twice the head, while the total Gram work stays under batch_ops
        (small jobs share a launch; a large job runs alone)."""
        head = ready[0]
        group, total = [head], head.ops
        for j in list(ready)[1:]:
            if len(group) >= self.max_batch:
                break
            if j.spec.layout != "dense" or j.ops > 2 * head.ops or total + j.ops > self.batch_ops:
                continue
            group.append(j)
            total += j.ops
        for j in group:
            ready.remove(j)
        return group

    def _retire(self, jobs, stream):
        for j in jobs:
            if self.keep_results:
                num, sums, den, g2 = j.outputs
                j.result = {"num": num.cpu().numpy(), "sums": sums.cpu().numpy(),
                            "g2": g2.cpu().numpy()}
            j.outputs = ()
            j.inputs = {}
            j.plan.close_async(stream)
            j.plan = None
            self.resident -= j.footprint

    # ------------------------------------------------------------------ loop
    def run(self, specs, realtime=True):
        """Serve `specs` (arrival times in seconds from now if realtime, else all at
        once).  Returns per-job records and summary statistics."""
        torch = self.torch
        pending = collections.deque(sorted(specs, key=lambda s: s.arrival))
        waiting = collections.deque()    # arrived, not admitted
        ready = collections.deque()      # admitted, not dispatched
        done = []
        n_resident = 0
        t0 = time.perf_counter()
        clock = lambda: time.perf_counter() - t0
        while pending or waiting or ready or any(len(b) for b in self.busy):
            now = clock()
            progressed = False
            while pending and (not realtime or pending[0].arrival <= now):
                spec = pending.popleft()
                waiting.append(self.prep.submit(self._prepare, spec))
            # completions
            for k, q in enumerate(self.busy):
                while q and q[0][0].query():  # a stream completes in order
                    ev, jobs, batch = q.popleft()
                    if batch is not None:
                        self.free_batches[k].append(batch)
                    t = clock()
                    for j in jobs:
                        j.t_done = t
                    self._retire(jobs, self.release_stream)
                    done.extend(jobs)
                    n_resident -= len(jobs)
                    progressed = True
            # admission by footprint and slots
            while waiting and n_resident < self.max_resident and waiting[0].done():
                job = waiting[0].result()
                if self.resident + job.footprint > self.budget:
                    if job.footprint > self.budget:
                        raise MemoryError("job %d needs %d bytes, budget %d"
                                          % (job.spec.id, job.footprint, self.budget))
                    break
                waiting.popleft()
                self.resident += job.footprint
                n_resident += 1
                job.t_admit = clock()
                self._admit(job, self.admit_stream)
                ready.append(job)
                progressed = True
            # dispatch to the least loaded streams: grouped dense launches, single event jobs
            while ready:
                k = min(range(len(self.streams)), key=lambda i: len(self.busy[i]))
                if len(self.busy[k]) >= self.depth:
                    break
                if ready[0].spec.layout == "dense":
                    group = self._group(ready)
                else:
                    group = [ready.popleft()]
                t = clock()
                for j in group:
                    j.t_dispatch = t
                self._dispatch(k, group)
                progressed = True
            if not progressed:
                time.sleep(0.0002)
        torch.cuda.synchronize()
        for j in done:
            j.t_prepared -= t0
        return summarize(done)


def summarize(done):
    """Latency percentiles and throughput of served jobs."""
    if not done:
        return {"jobs": 0}
    lat = np.array([j.t_done - j.spec.arrival for j in done])
    phases = {"prepare": np.array([max(j.t_prepared - j.spec.arrival, 0.0) for j in done]),
              "admission": np.array([j.t_admit - max(j.t_prepared, j.spec.arrival) for j in done]),
              "queue": np.array([j.t_dispatch - j.t_admit for j in done]),
              "execute": np.array([j.t_done - j.t_dispatch for j in done])}
    first = min(j.spec.arrival for j in done)
    last = max(j.t_done for j in done)
    pf = sum(j.npix * j.spec.nframes for j in done)
    return {"jobs": len(done),
            "dense_jobs": sum(j.spec.layout == "dense" for j in done),
            "event_jobs": sum(j.spec.layout == "events" for j in done),
            "wall_s": last - first,
            "jobs_per_s": len(done) / (last - first) if last > first else None,
            "latency_p50_ms": float(np.percentile(lat, 50) * 1e3),
            "latency_p99_ms": float(np.percentile(lat, 99) * 1e3),
            "latency_max_ms": float(lat.max() * 1e3),
            "phase_p50_ms": {k: float(np.percentile(v, 50) * 1e3) for k, v in phases.items()},
            "phase_p99_ms": {k: float(np.percentile(v, 99) * 1e3) for k, v in phases.items()},
            "mean_batch_size": float(np.mean([j.batch_size for j in done if j.spec.layout == "dense"]))
            if any(j.spec.layout == "dense" for j in done) else None,
            "pixel_frames_per_s": pf / (last - first) if last > first else None,
            "records": done}
