"""Autotuner of the grouped Gram launch.

The Gram's speed depends on the tile size (128 or 256 frames), on the raster band and
XCD chunk of its work list, on the launch form (persistent workgroups over per-XCD
work queues, or one workgroup per tile) and on the k-chunk (k-blocks per work item:
shorter items let the tiles that run together share a chunk of frame panel in L2).  The best choice depends on the number of
frames and on how the binned pixels split into bins, so results are cached under a key
built from (frame dtype, nframes, histogram of bin sizes in power-of-two k-block
buckets).  `autotune()` times candidates on real frames with the library's own HIP event
stage timers, keeps the fastest, applies it to the plan and records it; `apply_cached()`
applies a recorded result to a new plan of the same key.

The cache is a JSON file: $DXG2_TUNING_CACHE, or ~/.cache/dynamix_amd/gram_tuning.json.
"""
import ctypes
import json
import math
import os
import tempfile

import numpy as np

from . import _lib

KB = 64  # pixels per k-block of the slab

TILES = (256, 128)
QUEUE = (1, 0)
BANDS = (2, 4, 8)
GROUPS = (16, 32, 64)
KCHUNKS = (2047, 1024, 512, 256)


def cache_path(path=None):
    if path:
        return path
    env = os.environ.get("DXG2_TUNING_CACHE")
    if env:
        return env
    return os.path.join(os.path.expanduser("~"), ".cache", "dynamix_amd", "gram_tuning.json")


def tuning_key(dtype, nframes, pixel_counts):
    """Key of a Gram workload: dtype, frame count and the number of bins in each
    power-of-two bucket of k-blocks per bin (empty bins are left out)."""
    nkb = [int(math.ceil(int(n) / KB)) for n in pixel_counts if int(n) > 0]
    hist = {}
    for k in nkb:
        b = int(math.ceil(math.log2(k))) if k > 1 else 0
        hist[b] = hist.get(b, 0) + 1
    parts = ",".join("%d:%d" % (b, hist[b]) for b in sorted(hist))
    return "%s|N%d|kb%s" % (np.dtype(dtype).name, int(nframes), parts)


def plan_key(plan, nframes):
    return tuning_key(plan.dtype, nframes, plan.pixel_counts[:max(plan.n_bins, 0)])


def load_cache(path=None):
    p = cache_path(path)
    try:
        with open(p) as f:
            data = json.load(f)
        return data if isinstance(data, dict) else {}
    except (OSError, ValueError):
        return {}


def save_entry(key, entry, path=None):
    p = cache_path(path)
    data = load_cache(p)
    data[key] = entry
    d = os.path.dirname(p)
    if d:
        os.makedirs(d, exist_ok=True)
    fd, tmp = tempfile.mkstemp(dir=d or ".", prefix=".gram_tuning.")
    with os.fdopen(fd, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, p)


def apply_cached(plan, nframes, path=None):
    """Apply the recorded tuning of this plan's key; returns the entry or None."""
    entry = load_cache(path).get(plan_key(plan, nframes))
    if entry:
        plan.set_tuning(entry["tile"], entry["band"], entry["group"], entry["queue"],
                        entry.get("kchunk", 0))
    return entry


def _time_gram(plan, d_frames, bufs, stream, reps):
    lib = plan.lib

    def call():
        _lib.check(lib.dxg2_correlate_dense(plan.handle, _lib.ptr(d_frames),
                                            *[_lib.ptr(b) for b in bufs], stream),
                   "dxg2_correlate_dense")
    call()  # first call after a retune allocates nothing new but warms the caches
    plan.set_profiling(True)
    for _ in range(reps):
        call()
    t = plan.stage_times()
    plan.set_profiling(False)
    return float(t[1])


def autotune(plan, d_frames, nframes, reps=2, path=None, record=True):
    """Time the Gram over candidate tunings on `d_frames` (a GPU tensor of the plan's
    shape and dtype), apply the fastest to `plan` and (record=True) cache it.

    Three stages: tile size x launch form at the default band, chunk and k-chunk; then
    band x chunk for the best of those; then the k-chunk.  Returns the cache entry: the chosen tuning, its Gram
    time, the default's time and every timing measured.
    """
    import torch
    nb = max(plan.n_bins, 1)
    num = torch.empty((nb, nframes), dtype=torch.int64, device="cuda")
    sums = torch.empty_like(num)
    den = torch.empty((nb, nframes), dtype=torch.float64, device="cuda")
    g2 = torch.empty((nb, nframes), dtype=torch.float32, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    bufs = (num, sums, den, g2)
    default = plan.get_tuning()
    timings = []

    def trial(tile, band, group, queue, kchunk):
        plan.set_tuning(tile, band, group, queue, kchunk)
        ms = _time_gram(plan, d_frames, bufs, stream, reps)
        timings.append({"tile": tile, "band": band, "group": group, "queue": queue,
                        "kchunk": kchunk, "gram_ms": ms})
        return ms

    kc0 = default["kchunk"]
    default_ms = trial(default["tile"], default["band"], default["group"], default["queue"],
                       kc0)
    best = dict(default, gram_ms=default_ms)
    for tile in TILES:
        for queue in QUEUE:
            if (tile, queue) == (default["tile"], default["queue"]):
                continue
            ms = trial(tile, default["band"], default["group"], queue, kc0)
            if ms < best["gram_ms"]:
                best = {"tile": tile, "band": default["band"], "group": default["group"],
                        "queue": queue, "kchunk": kc0, "gram_ms": ms}
    tile, queue = best["tile"], best["queue"]
    for band in BANDS:
        for group in GROUPS:
            if (band, group) == (best["band"], best["group"]):
                continue
            ms = trial(tile, band, group, queue, kc0)
            if ms < best["gram_ms"]:
                best = {"tile": tile, "band": band, "group": group, "queue": queue,
                        "kchunk": kc0, "gram_ms": ms}
    for kchunk in KCHUNKS:
        if kchunk == best["kchunk"]:
            continue
        ms = trial(best["tile"], best["band"], best["group"], best["queue"], kchunk)
        if ms < best["gram_ms"]:
            best = dict(best, kchunk=kchunk, gram_ms=ms)
    plan.set_tuning(best["tile"], best["band"], best["group"], best["queue"], best["kchunk"])
    entry = dict(best, default_ms=default_ms, trials=timings)
    if record:
        save_entry(plan_key(plan, nframes), entry, path)
    return entry
