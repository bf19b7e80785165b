"""Benchmark: g2 pixel*frames/s of the dense batched i8-MFMA path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload headline]

A step is one dense correlate of the whole workload: qmask gather of every bin into
i8 slabs, one grouped i8-MFMA Gram launch over all bins with the fused tau
reduction, exact denominators and g2.  Frames are generated on the GPU (seeded AR(1)
speckle) before the timed region, so inputs are resident in HBM; g2, the int64
numerators, sums and FP64 denominators stay in HBM.

Printed as one JSON line.  Besides the metric it reports
  roofline      the Gram kernel's achieved int8 ops/s (algorithmic ops: the upper
                triangle with the diagonal, 2 * sum_b n_b * N (N+1)/2, per launch)
                over its average duration from HIP events recorded around each
                launch inside the timed region, against the 5.0 POP/s dense int8 peak
                (the persistent launch over per-XCD work queues, with the fill kernel
                that clears the queue heads);
  calc_std      the dense path with calc_std=True (dev(tau)) on the same stack;
  compress      the on-device frames -> events compaction of the same stack;
  ttcf          the two-time matrix of one bin of the same stack;
  cpu_baseline  the reference's numpy MatMul correlator (float32 sgemm + diagonal
                loop, restated in oracle/) on one bin of the same stack, all host cores;
  cpu_baseline_fft  the reference's FFT correlator (float32 real FFTs of zero-padded
                pixel series, restated in oracle/ with scipy.fft) on pixel chunks of the
                same stack, all host cores;
  plain_fp32    this library's one-bin-per-launch FP32 comparator on the same stack;
  fft           the FFT path (FP32 Stockham power spectra) on the same stack;
  uint16        the dense path on 2162 x 2068 x 5,000 uint16 frames (1 % hot pixels),
                50 linear-width bins: byte-split digit planes with the wide pixels
                (a count > 255) first, useful and executed ops;
  serving       the continuous-batching loop: 200 seeded jobs (shapes 256^2, 514x1030,
                1030x1065; 1k-20k frames; 10-100 bins; 70% dense uint8 / 30% events;
                AR(1) tau_c 10-500 frames, each job on its own generated input) at
                20 arrivals/s, and a sweep at 40 and 80/s: jobs/s, p50/p99 latency,
                mean dense batch size;
  events        the event path on 1030 x 1065 x 100,000 frames at 1e-3 counts/pixel/frame
                (generated as events on the GPU): pixel-major input, and frame-sorted
                (pixel, frame, count) input including its conversion.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("g2 pixel·frames/s at 514×1030×10k uint8, 20 bins; "
          "p50/p99 mixed-job latency")
INT8_PEAK_TOPS = 5000.0   # MI355X dense int8 MFMA, 5.0 POP/s
HBM_PEAK_GBS = 8000.0


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def read_traffic(name):
    """HBM bytes per launch measured with rocprofv3 --pmc, committed under profiles/."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


def cpu_baseline(frames_dev, qmask, target_s=15.0):
    """Reference-equivalent numpy MatMul correlator on whole bins of the same stack,
    bins taken in order of closeness to the mean bin size until target_s is spent."""
    import numpy as np
    import torch
    from oracle import g2 as og
    n = frames_dev.shape[0]
    counts = np.bincount(qmask.ravel())[1:]
    order = np.argsort(np.abs(counts - counts.mean())) + 1
    cores = len(os.sched_getaffinity(0))
    done, pf, spent = [], 0, 0.0
    for b in order:
        idx = torch.from_numpy(np.flatnonzero(qmask.ravel() == b)).cuda()
        a = frames_dev.view(n, -1).index_select(1, idx).cpu().numpy()   # (N, n_b) uint8
        t0 = time.perf_counter()
        og.py_dense_correlator_f32(a.reshape(n, 1, -1), np.ones((1, a.shape[1]), dtype=bool))
        spent += time.perf_counter() - t0
        done.append("%d (%d px)" % (b, a.shape[1]))
        pf += a.shape[1] * n
        if spent >= target_s:
            break
    return {"value": pf / spent, "unit": "pixel*frames/s", "cores": cores, "kind": "port",
            "sample": "numpy MatMulCorrelator (float32 sgemm + per-lag diagonal loop, "
                      "oracle/g2.py py_dense_correlator_f32) on whole bins %s x %d frames "
                      "of the same stack: %.1f s, OpenBLAS threads = %d"
                      % (", ".join(done), n, spent, cores)}


def gather_roofline(qmask, shape, n, gather_ms):
    """HBM roofline of the gather stage: every frame byte read once, and one byte per
    binned pixel (rounded up to whole 64-pixel k-blocks) per frame written to the slab."""
    import numpy as np
    counts = np.bincount(qmask.ravel())[1:]
    read = float(n) * shape[0] * shape[1]
    write = float(n) * int((64 * np.ceil(counts / 64.0)).sum())
    gbps = (read + write) / (gather_ms * 1e-3) / 1e9
    return {"ms": gather_ms, "bytes_read": read, "bytes_written": write, "GBps": gbps,
            "peak_GBps": HBM_PEAK_GBS, "frac_of_hbm": gbps / HBM_PEAK_GBS,
            "kernel": "gather_band_kernel"}


def cpu_baseline_fft(frames_dev, qmask, target_s=8.0, chunk=4096):
    """Reference-equivalent FFT correlator (oracle fft_correlator, scipy.fft with every
    host core) on chunks of `chunk` pixels of the bins, taken in order of closeness to
    the mean bin size, until target_s is spent.  Each chunk is transformed as one bin:
    the FFT work per pixel is the same as in the whole bin."""
    import numpy as np
    import torch
    from oracle import g2 as og
    n = frames_dev.shape[0]
    counts = np.bincount(qmask.ravel())[1:]
    order = np.argsort(np.abs(counts - counts.mean())) + 1
    cores = len(os.sched_getaffinity(0))
    done, pf, spent = [], 0, 0.0
    for b in order:
        pix = np.flatnonzero(qmask.ravel() == b)[:chunk]
        idx = torch.from_numpy(pix).cuda()
        a = frames_dev.view(n, -1).index_select(1, idx).cpu().numpy()   # (N, chunk) uint8
        t0 = time.perf_counter()
        og.fft_correlator(a.reshape(n, 1, -1), np.ones((1, a.shape[1]), dtype=np.int32),
                          workers=cores)
        spent += time.perf_counter() - t0
        done.append("%d" % b)
        pf += a.shape[1] * n
        if spent >= target_s:
            break
    return {"value": pf / spent, "unit": "pixel*frames/s", "cores": cores, "kind": "port",
            "sample": "FFT correlator (scipy.fft float32 rfft/irfft at Nf = %d, oracle/g2.py "
                      "fft_correlator) on the first %d pixels of bins %s x %d frames of the "
                      "same stack: %.1f s, scipy workers = %d"
                      % (og.next_fft_length(2 * n), chunk, ", ".join(done), n, spent, cores)}


def _timed(fn, reps, torch):
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / reps


def _stage_times(plan, fn, reps, torch):
    """The library's per-stage HIP-event times of `reps` calls of fn (ms)."""
    torch.cuda.synchronize()
    plan.set_profiling(True)
    for _ in range(reps):
        fn()
    t = plan.stage_times()
    plan.set_profiling(False)
    return t


def bench_fft(plan, frames, nb, n, pf):
    """FFT path on the same stack; its stage times come from the library's HIP events."""
    import torch
    from dynamix_amd import _lib
    num = torch.empty((nb, n), dtype=torch.float64, device="cuda")
    sums = torch.empty((nb, n), dtype=torch.int64, device="cuda")
    den = torch.empty((nb, n), dtype=torch.float64, device="cuda")
    g2 = torch.empty((nb, n), dtype=torch.float32, device="cuda")
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def step():
        _lib.check(plan.lib.dxg2_correlate_fft(plan.handle, _lib.ptr(frames), _lib.ptr(num),
                                               _lib.ptr(sums), _lib.ptr(den), _lib.ptr(g2), st))
    step()
    torch.cuda.synchronize()
    plan.set_profiling(True)
    ms = _timed(step, 2, torch)
    t = plan.stage_times()
    plan.set_profiling(False)
    from dynamix_amd.correlator.dense import fft_length
    L = fft_length(n) // 2
    npix = int(plan.pixel_counts[:nb].sum())
    flop = npix * 5.0 * L * math.log2(L)   # one complex FFT of L points per pixel
    return {"ms": ms, "value": pf / (ms * 1e-3), "unit": "pixel*frames/s",
            "stage_ms": {"gather_and_sums": t[0], "spectra_and_inverse": t[1], "finalise": t[2]},
            "fft_length": 2 * L, "spectra_gflops": flop / (t[1] * 1e-3) / 1e9}


def bench_std(plan, frames, nb, n, pf, base_ms, reps=2):
    """Dense path with calc_std on the same stack: dev(tau) from per-tile ratio passes."""
    import torch
    from dynamix_amd import _lib
    num = torch.empty((nb, n), dtype=torch.int64, device="cuda")
    sums = torch.empty_like(num)
    den = torch.empty((nb, n), dtype=torch.float64, device="cuda")
    g2 = torch.empty((nb, n), dtype=torch.float32, device="cuda")
    dev = torch.empty_like(g2)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def step():
        _lib.check(plan.lib.dxg2_correlate_dense_std(plan.handle, _lib.ptr(frames), _lib.ptr(num),
                                                     _lib.ptr(sums), _lib.ptr(den), _lib.ptr(g2),
                                                     _lib.ptr(dev), st))
    step()
    torch.cuda.synchronize()
    plan.set_profiling(True)
    ms = _timed(step, reps, torch)
    t = plan.stage_times()
    plan.set_profiling(False)
    size = ctypes.c_size_t()
    _lib.check(plan.lib.dxg2_plan_footprint(plan.handle, 4, ctypes.byref(size)))
    finite = torch.isfinite(dev).float().mean().item()
    out = {"ms": ms, "value": pf / (ms * 1e-3), "unit": "pixel*frames/s",
           "overhead_vs_res_only": ms / base_ms,
           "stage_ms": {"gather_and_sums": t[0], "gram_store_and_ratio_passes": t[1],
                        "finalise_and_dev": t[2]},
           "workspace_bytes": size.value, "finite_fraction": finite}
    del num, sums, den, g2, dev
    torch.cuda.empty_cache()
    return out


def bench_compress(frames, reps=2):
    """On-device frames -> events compaction (FramesCompressor) of the headline stack."""
    import torch
    from dynamix_amd.correlator.event import FramesCompressor
    FramesCompressor.compress_all_stack(frames)  # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        d, t, o = FramesCompressor.compress_all_stack(frames)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    nnz = int(d.shape[0])
    read = 2.0 * frames.numel() * frames.element_size()      # count and fill passes
    written = nnz * (4 + frames.element_size()) + (o.numel()) * 4
    return {"ms": ms, "events": nnz, "bytes_read": read, "bytes_written": written,
            "GBps": (read + written) / (ms * 1e-3) / 1e9,
            "frac_of_hbm": (read + written) / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "note": "includes the device-to-host read of the event count between the passes"}


def bench_ttcf(plan, frames, n, q=10, reps=2):
    """Two-time matrix (N x N float32) of one bin of the headline stack."""
    import torch
    from dynamix_amd import _lib
    trc = torch.empty((n, n), dtype=torch.float32, device="cuda")
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def step():
        _lib.check(plan.lib.dxg2_correlate_dense_ttcf(plan.handle, _lib.ptr(frames), q,
                                                      _lib.ptr(trc), st))
    step()
    torch.cuda.synchronize()
    ms = _timed(step, reps, torch)
    finite = torch.isfinite(trc).float().mean().item()
    out = {"ms": ms, "bin": q, "pixels": int(plan.pixel_counts[q - 1]),
           "matrix_bytes": n * n * 4, "finite_fraction": finite}
    del trc
    torch.cuda.empty_cache()
    return out


def bench_events(seed):
    """Event path on the events workload: pixel-major and frame-sorted COO input."""
    import numpy as np
    import torch
    from dynamix_amd import _lib, synth
    from dynamix_amd.correlator.common import Plan
    wl = synth.WORKLOADS["events"]
    shape, n = wl["shape"], wl["nframes"]
    q = synth.workload_qmask("events")
    t0 = time.perf_counter()
    times, data, offsets = synth.speckle_events_device(shape, n, wl["mean"], wl["tau_c"], seed)
    torch.cuda.synchronize()
    gen_s = time.perf_counter() - t0
    nnz = int(times.shape[0])
    counts = np.diff(offsets.cpu().numpy().astype(np.int64))
    binned = q.ravel() > 0
    pairs = int((counts[binned] * (counts[binned] + 1) // 2).sum())
    plan = Plan(shape, n, q, np.uint8, scale=np.bincount(q.ravel())[1:].astype(np.float64))
    nb = plan.n_bins
    num = torch.empty((nb, n), dtype=torch.int64, device="cuda")
    sums = torch.empty_like(num)
    den = torch.empty((nb, n), dtype=torch.float64, device="cuda")
    g2 = torch.empty((nb, n), dtype=torch.float32, device="cuda")
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    outs = [_lib.ptr(x) for x in (num, sums, den, g2)]

    def csr():
        _lib.check(plan.lib.dxg2_correlate_events_csr(plan.handle, _lib.ptr(times), _lib.ptr(data),
                                                      _lib.ptr(offsets), nnz, *outs, st))
    ms_csr = _timed(csr, 2, torch)
    stages_csr = _stage_times(plan, csr, 2, torch)
    num_csr = num.clone()
    flags = ctypes.c_int32()
    _lib.check(plan.lib.dxg2_plan_event_status(plan.handle, ctypes.byref(flags), st))
    pix, fr, cnt = synth.events_to_coo_device(times, data, offsets, shape, n)
    del times, data
    torch.cuda.empty_cache()

    def coo():
        _lib.check(plan.lib.dxg2_correlate_events_coo(plan.handle, _lib.ptr(pix), _lib.ptr(fr),
                                                      _lib.ptr(cnt), nnz, *outs, st))
    ms_coo = _timed(coo, 2, torch)
    stages_coo = _stage_times(plan, coo, 2, torch)
    same = bool((num == num_csr).all().item())
    # the i8-MFMA denominator's work: per bin 1 digit pass (max |S| <= 255) or 4 (two
    # digits, max |S| <= 16383); per pass and workgroup of 16 lag blocks c0.., 4 waves x
    # 24 MFMA 32x32x32 per 32 rows over K = rows - c0 rounded up to 96 rows
    smax = sums.abs().amax(dim=1).cpu().numpy()
    passes = np.where(smax <= 255, 1, np.where(smax <= 16383, 4, 0))
    rows = (n + 63) // 64
    steps = sum(-(-(rows - c0) // 96) * 3 for c0 in range(0, rows, 16))
    den_exec_ops = float(passes.sum()) * steps * 4 * 24 * 65536
    den_useful_ops = float(passes.sum()) * n * (n + 1)   # 2 ops per product, tau >= 0
    # LDS atomic-add ceiling of this GPU (random counters, the pair kernel's LDS
    # footprint): every pair costs at least one such add
    lds_rate = ctypes.c_double()
    _lib.check(plan.lib.dxg2_measure_lds_atomics(0, 4096, ctypes.byref(lds_rate), st))
    pairs_ms = stages_csr[1]
    hw = shape[0] * shape[1]
    conv_bytes = 9 * nnz + 2 * 5 * nnz + 4 * (hw + 1)
    # CSR call's algorithmic HBM traffic: each event's time and value read once, the
    # offsets and the binned pixel list once, the four (bin, tau) output planes written once.
    csr_bytes = 5 * nnz + 4 * (hw + 1) + 4 * int(binned.sum()) + nb * n * (8 + 8 + 8 + 4)
    plan.close()
    return {"workload": "events: %dx%d x %d frames, mean %g, %d bins, %d events, %d pairs"
                        % (shape[0], shape[1], n, wl["mean"], nb, nnz, pairs),
            "csr_ms": ms_csr, "coo_ms": ms_coo, "pairs_per_s": pairs / (ms_csr * 1e-3),
            "events_per_s": nnz / (ms_csr * 1e-3), "pixel_frames_per_s": int(binned.sum()) * n / (ms_csr * 1e-3),
            "csr_hbm_GBps": csr_bytes / (ms_csr * 1e-3) / 1e9,
            "coo_to_csr_GBps": conv_bytes / ((ms_coo - ms_csr) * 1e-3) / 1e9 if ms_coo > ms_csr else None,
            "coo_equals_csr": same, "input_flags": flags.value, "generate_s": gen_s,
            "stage_ms_csr": {"pairs_and_sums": stages_csr[1], "denominator_and_finalise": stages_csr[2]},
            "stage_ms_coo": {"coo_to_csr": stages_coo[0], "pairs_and_sums": stages_coo[1],
                             "denominator_and_finalise": stages_coo[2]},
            "den_kernel": {"kernel": "den_mfma_kernel", "bins_one_digit": int((passes == 1).sum()),
                           "bins_two_digits": int((passes == 4).sum()),
                           "bins_valu": int((passes == 0).sum()),
                           "executed_ops": den_exec_ops, "useful_ops": den_useful_ops,
                           "note": "the kernel's time is in the kernel trace (rocprofv3); the "
                                   "stage time above adds the prefix sums and the finaliser"},
            "pairs_kernel": {"kernel": "event_pairs_kernel", "pairs_per_s": pairs / (pairs_ms * 1e-3),
                             "lds_atomic_ceiling_per_s": lds_rate.value,
                             "frac_of_lds_atomic_ceiling": pairs / (pairs_ms * 1e-3) / lds_rate.value,
                             "note": "one LDS atomic add per pair at least; the ceiling is "
                                     "dxg2_measure_lds_atomics (random counters) on this GPU"}}


def bench_uint16(seed, reps=2, tune=True):
    """Dense path on the uint16 workload (byte-split digit planes)."""
    import numpy as np
    import torch
    from dynamix_amd import _lib, synth
    from dynamix_amd.correlator.common import Plan
    wl = synth.WORKLOADS["uint16"]
    shape, n = wl["shape"], wl["nframes"]
    q = synth.workload_qmask("uint16")
    frames = synth.workload_frames_device("uint16", seed)
    plan = Plan(shape, n, q, np.uint16)
    nb = plan.n_bins
    num = torch.empty((nb, n), dtype=torch.int64, device="cuda")
    sums = torch.empty_like(num)
    den = torch.empty((nb, n), dtype=torch.float64, device="cuda")
    g2 = torch.empty((nb, n), dtype=torch.float32, device="cuda")
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def step():
        _lib.check(plan.lib.dxg2_correlate_dense(plan.handle, _lib.ptr(frames), _lib.ptr(num),
                                                 _lib.ptr(sums), _lib.ptr(den), _lib.ptr(g2), st))
    tuning = None
    if tune:
        from dynamix_amd import autotune
        tuning = autotune.autotune(plan, frames, n, reps=1, record=False)
    step()
    torch.cuda.synchronize()
    plan.set_profiling(True)
    ms = _timed(step, reps, torch)
    gather_ms, gram_ms, fin_ms, _, _ = plan.stage_times()
    npix = int(plan.pixel_counts[:nb].sum())
    useful = 2.0 * npix * n * (n + 1) / 2.0
    # Executed digit-plane work: LL over every k-block of a bin, HH and the two cross
    # halves over the k-blocks holding its wide pixels (a count > 255 in some frame).
    wide = torch.zeros(shape, dtype=torch.bool, device="cuda")
    fi = frames.view(torch.int16).view(n, *shape)
    for t0 in range(0, n, 250):
        x = fi[t0:t0 + 250]
        wide |= ((x >= 256) | (x < 0)).any(dim=0)
    qd = torch.from_numpy(np.ascontiguousarray(q)).to("cuda", torch.int64)
    nwide = torch.bincount(qd[wide], minlength=nb + 1)[1:nb + 1].cpu().numpy()
    counts = np.asarray(plan.pixel_counts[:nb], dtype=np.int64)
    k_ll = int((64 * np.ceil(counts / 64.0)).sum())
    k_w = int((64 * np.ceil(nwide / 64.0)).sum())
    executed = 2.0 * n * (n + 1) / 2.0 * (k_ll + 3 * k_w)
    del wide, fi
    assert bool((den > 0).any().item()) and torch.isfinite(g2[den > 0]).all().item()
    out = {"workload": "uint16: %dx%d x %d frames uint16, mean %g, %g%% hot pixels at %g, "
                       "%d linear-width bins, %d binned px"
                       % (shape[0], shape[1], n, wl["mean"], 100 * wl["hot_fraction"],
                          wl["hot_mean"], nb, npix),
           "ms": ms, "value": npix * n / (ms * 1e-3), "unit": "pixel*frames/s",
           "stage_ms": {"gather": gather_ms, "gram": gram_ms, "finalise": fin_ms},
           "wide_pixels": int(nwide.sum()),
           "gram_useful_TOPs": useful / (gram_ms * 1e-3) / 1e12,
           "gram_useful_frac_of_peak": useful / (gram_ms * 1e-3) / 1e12 / INT8_PEAK_TOPS,
           "gram_executed_TOPs": executed / (gram_ms * 1e-3) / 1e12,
           "gram_executed_frac_of_peak": executed / (gram_ms * 1e-3) / 1e12 / INT8_PEAK_TOPS,
           "note": "digit planes: LL over all k-blocks, HH and cross only over the "
                   "k-blocks of each bin's wide pixels (ordered first)",
           "gram_tuning": plan.get_tuning(),
           "default_gram_ms": None if tuning is None else tuning["default_ms"]}
    plan.close()
    return out


def bench_serving(seed, n_jobs=200, rates=(20.0, 40.0, 80.0)):
    """Serving loop: seeded Poisson arrivals of mixed dense / event jobs, each on its own
    input (generated on the GPU from its seed and correlation time), at a sweep of
    arrival rates.  The first rate is the reported one; every rate reports jobs/s,
    p50/p99 latency and the mean dense batch size."""
    import torch
    from dynamix_amd import serving
    sweep = []
    for i, rate in enumerate(rates):
        specs = serving.make_jobs(n_jobs, rate, seed + 1 + i)
        srv = serving.Server()
        try:
            st = srv.run(specs)
        finally:
            srv.close()
        recs = st.pop("records")
        st.pop("failed_records", None)
        dense_ops = sum(r.ops for r in recs if r.spec.layout == "dense")
        st.update({"arrival_rate_per_s": rate, "streams": len(srv.streams),
                   "dense_gram_TOPs": dense_ops / st["wall_s"] / 1e12 if st["wall_s"] else None,
                   "jobs": "%d jobs: shapes %s, 1k-20k frames, 10-100 bins, 70%% dense uint8 / "
                           "30%% events, tau_c 10-500 frames per job"
                           % (n_jobs, list(serving.SERVE_SHAPES))})
        st["failures"] = [(i, str(e)[:200]) for i, e in st["failures"][:5]]
        log("serving at %.0f/s: %.1f jobs/s, p50 %.0f ms, p99 %.0f ms, batch %.2f, failed %d%s"
            % (rate, st["jobs_per_s"] or 0, st["latency_p50_ms"], st["latency_p99_ms"],
               st["mean_batch_size"] or 0, st["failed"],
               " (first: %s)" % st["failures"][0][1] if st["failures"] else ""))
        sweep.append(st)
        del srv, recs
        torch.cuda.empty_cache()
    out = dict(sweep[0])
    out["rate_sweep"] = [{k: st[k] for k in ("arrival_rate_per_s", "jobs_per_s", "latency_p50_ms",
                                             "latency_p99_ms", "mean_batch_size",
                                             "pixel_frames_per_s", "failed", "failures")}
                         for st in sweep]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="headline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-fft", action="store_true")
    ap.add_argument("--no-events", action="store_true")
    ap.add_argument("--no-uint16", action="store_true")
    ap.add_argument("--no-std", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the compaction and two-time sections")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default Gram tuning instead of timing the candidates")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--pmc-copy-gib", type=float, default=0.0,
                    help="after the run, copy this many GiB device to device (a kernel of "
                         "known traffic that scales rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build runs the whole workload on a "
              "single MI355X (multi-GPU sharding is a design note, not a measured path).",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from dynamix_amd import _lib, synth
    from dynamix_amd.correlator.common import Plan

    torch.cuda.set_device(0)
    wl = synth.WORKLOADS[args.workload]
    shape, n = wl["shape"], wl["nframes"]
    q = synth.workload_qmask(args.workload)
    log("generating %s frames %s x %d on the GPU" % (args.workload, shape, n))
    frames = synth.workload_frames_device(args.workload, args.seed)
    torch.cuda.synchronize()
    plan = Plan(shape, n, q, np.dtype(wl["dtype"]))
    nb = plan.n_bins
    num = torch.empty((nb, n), dtype=torch.int64, device="cuda")
    sums = torch.empty((nb, n), dtype=torch.int64, device="cuda")
    den = torch.empty((nb, n), dtype=torch.float64, device="cuda")
    g2 = torch.empty((nb, n), dtype=torch.float32, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    lib = plan.lib

    def step():
        _lib.check(lib.dxg2_correlate_dense(plan.handle, _lib.ptr(frames), _lib.ptr(num),
                                            _lib.ptr(sums), _lib.ptr(den), _lib.ptr(g2), stream))

    tuning = None
    if not args.no_autotune:
        from dynamix_amd import autotune
        log("autotune")
        tuning = autotune.autotune(plan, frames, n, record=False)
        log("tuning %s: gram %.3f ms (default %.3f ms)" % (
            {k: tuning[k] for k in ("tile", "band", "group", "queue")}, tuning["gram_ms"],
            tuning["default_ms"]))

    log("warmup %d" % args.warmup)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    plan.set_profiling(True)
    log("timing %d steps" % args.steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gather_ms, gram_ms, fin_ms, total_ms, ncalls = plan.stage_times()
    plan.set_profiling(False)
    ms = wall * 1e3 / args.steps
    npix = int(plan.pixel_counts[:nb].sum())
    pf = npix * n
    ops = 2.0 * npix * n * (n + 1) / 2.0
    achieved = ops / (gram_ms * 1e-3) / 1e12
    log("ms/step %.3f  gather %.3f gram %.3f finalize %.3f  (%d calls)" %
        (ms, gather_ms, gram_ms, fin_ms, ncalls))

    # sanity: g2 finite where denominators are positive
    assert torch.isfinite(g2[den > 0]).all().item()

    out = {
        "metric": METRIC,
        "value": pf / (ms * 1e-3),
        "unit": "pixel*frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "i8",
        "data": "synthetic",
        "config": {"workload": "%s: %dx%d x %d frames %s, %d annular bins, %d binned px"
                               % (args.workload, shape[0], shape[1], n, wl["dtype"], nb, npix),
                   "path": "dense batched i8-MFMA Gram (gather + gram + finalise)",
                   "stage_ms": {"gather": gather_ms, "gram": gram_ms, "finalise": fin_ms},
                   "gram_tuning": plan.get_tuning(),
                   "autotune": None if tuning is None else {
                       "gram_ms": tuning["gram_ms"], "default_gram_ms": tuning["default_ms"],
                       "candidates": len(tuning["trials"])}},
        "gather": gather_roofline(q, shape, n, gather_ms),
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": INT8_PEAK_TOPS,
                     "unit": "TOP/s", "frac": achieved / INT8_PEAK_TOPS,
                     "traffic": read_traffic("gram_i8_%s_traffic.txt" % args.workload),
                     "kernel": ("gram_i8_queue_kernel" if plan.get_tuning()["queue"]
                                else "gram_i8_kernel"), "ops_per_launch": ops},
    }

    if not args.no_plain:
        from dynamix_amd.correlator.dense import PlainMatMulCorrelator
        log("plain fp32 comparator")
        pc = PlainMatMulCorrelator(shape, n, q, extra_options={"return_device": True})
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pg2 = pc.correlate(frames)          # first call: plan workspaces allocated
        torch.cuda.synchronize()
        cold = (time.perf_counter() - t0) * 1e3
        reps = 2
        t0 = time.perf_counter()
        for _ in range(reps):               # warm: the same plan and workspaces
            pg2 = pc.correlate(frames)
        torch.cuda.synchronize()
        pms = (time.perf_counter() - t0) * 1e3 / reps
        fin = torch.isfinite(g2) & torch.isfinite(pg2) & (g2 != 0)
        rel = ((pg2[fin].double() - g2[fin].double()).abs() / g2[fin].double().abs()).max().item()
        out["plain_fp32"] = {"ms": pms, "cold_ms": cold, "warm_calls": reps,
                             "value": pf / (pms * 1e-3),
                             "speedup_of_i8": pms / ms, "max_rel_diff_vs_i8": rel}
        del pc, pg2
        torch.cuda.empty_cache()

    if not args.no_fft:
        log("fft path")
        out["fft"] = bench_fft(plan, frames, nb, n, pf)
    if not args.no_std:
        log("calc_std path")
        out["calc_std"] = bench_std(plan, frames, nb, n, pf, ms)
    if not args.no_extras:
        log("compaction path")
        out["compress"] = bench_compress(frames)
        log("two-time path")
        out["ttcf"] = bench_ttcf(plan, frames, n)
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(frames, q)
        if not args.no_fft:
            log("cpu baseline (FFT)")
            out["cpu_baseline_fft"] = cpu_baseline_fft(frames, q)
            if "fft" in out:
                out["fft"]["speedup_vs_cpu_fft"] = (out["fft"]["value"] /
                                                    out["cpu_baseline_fft"]["value"])
    if not (args.no_events and args.no_uint16 and args.no_serving):
        del frames, num, sums, den, g2
        plan.close()
        torch.cuda.empty_cache()
    if not args.no_uint16 and args.workload != "uint16":
        log("uint16 path")
        out["uint16"] = bench_uint16(args.seed, tune=not args.no_autotune)
        torch.cuda.empty_cache()
    if not args.no_events:
        log("event path")
        out["events"] = bench_events(args.seed)
        torch.cuda.empty_cache()
    if not args.no_serving:
        log("serving loop")
        out["serving"] = bench_serving(args.seed)
        s = out["serving"]
        log("serving: %.1f jobs/s, p50 %.0f ms, p99 %.0f ms" %
            (s["jobs_per_s"], s["latency_p50_ms"], s["latency_p99_ms"]))
    if args.pmc_copy_gib > 0:
        nbytes = int(args.pmc_copy_gib * (1 << 30))
        src = torch.ones(nbytes, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        # an elementwise kernel (a plain device copy may go to the DMA engines, which
        # the counters do not see): reads and writes nbytes once each
        torch.add(src, 1, out=dst)
        torch.cuda.synchronize()
        log("copied %d bytes device to device (elementwise add kernel)" % nbytes)
        del src, dst
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
