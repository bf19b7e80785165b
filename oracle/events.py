"""Event (sparse) correlator oracle (TEST INFRASTRUCTURE ONLY -- see oracle/__init__.py).

Event layout (the reference's, FramesCompressor docstring event.py:208-249): for every
pixel p in raster order, its non-zero samples sorted by frame, concatenated:
    data[nnz]     values,
    times[nnz]    int32 frame indices,
    offsets[H*W+1] uint32, events of pixel p are [offsets[p], offsets[p+1]).

Parity pinning:
  compress_all_stack     -- golden vectors from the reference's compress_all_stack
                            (event.py:267-284) + brute force on tiny stacks.
  compress_progressive   -- equality with compress_all_stack, the reference's own
                            self-consistency test (test_event.py:62-73).
  event_numerators       -- brute-force ordered pairs with Python integers; equality
                            with the dense oracle numerators on the same stack.
  event_sums_numerators_fast -- equality with event_sums_numerators, the brute force
                            and the dense oracle on the same stacks.
  event_g2               -- dense oracle g2 on the same stack (same formula, with the
                            EventCorrelator's scale factor semantics).
  coo_to_csr             -- brute force (stable sort by pixel) and the inverse
                            csr_to_coo round trip.
"""
import numpy as np

from .g2 import denominators, g2_fp64, qmask_bins, scale_factors


def compress_all_stack(frames):
    """Dense stack (N, H, W) -> (data, times, offsets).  event.py:267-284."""
    assert frames.ndim == 3
    n = frames.shape[0]
    per_pixel = np.moveaxis(frames, 0, -1).reshape(-1, n)   # (H*W, N), frame fastest
    pix, t = np.nonzero(per_pixel > 0)                         # pixel-major, t ascending
    data = per_pixel[pix, t]
    times = t.astype(np.int32)
    counts = np.bincount(pix, minlength=per_pixel.shape[0])
    offsets = np.zeros(per_pixel.shape[0] + 1, dtype=np.uint32)
    offsets[1:] = np.cumsum(counts)
    return data, times, offsets


class FramesCompressorOracle:
    """Frame-by-frame compaction (event.py:287-321) kept for the self-consistency check."""

    def __init__(self, shape, nframes, max_nnz, dtype=np.int8):
        self.npix = int(np.prod(shape))
        self.nframes = nframes
        self.counter = np.zeros(self.npix, dtype=np.int32)
        self.events = np.zeros((self.npix, max_nnz), dtype=dtype)
        self.times = np.zeros((self.npix, max_nnz), dtype=np.int32)
        self.frames_done = 0

    def process_frame(self, frame):
        f = frame.ravel()
        m = f > 0
        self.events[m, self.counter[m]] = f[m]
        self.times[m, self.counter[m]] = self.frames_done
        self.counter[m] += 1
        self.frames_done += 1

    def get_compacted_events(self):
        offsets = np.zeros(self.npix + 1, dtype=np.uint32)
        offsets[1:] = np.cumsum(self.counter)
        m = self.events.ravel() > 0
        return self.events.ravel()[m], self.times.ravel()[m], offsets


def event_sums_numerators(times, data, offsets, qmask, nframes, bins):
    """S_b(t) and Num_b(tau) from the event layout, exact int64.

    Follows the pair loop of evtcorrelator.cl:3-46: for every pixel of bin b and
    every ordered pair j <= i of its frame-sorted events, Num_b[t_i - t_j] +=
    d_i * d_j, and S_b[t_i] += d_i.  Vectorised per pixel with numpy.
    """
    q = np.asarray(qmask).ravel()
    nb = len(bins)
    s = np.zeros((nb, nframes), dtype=np.int64)
    num = np.zeros((nb, nframes), dtype=np.int64)
    bidx = {int(b): i for i, b in enumerate(bins)}
    offsets = np.asarray(offsets, dtype=np.int64)
    times = np.asarray(times, dtype=np.int64)
    data = np.asarray(data).astype(np.int64)
    for p in range(q.size):
        lo, hi = offsets[p], offsets[p + 1]
        if hi == lo or int(q[p]) not in bidx:
            continue
        i = bidx[int(q[p])]
        t = times[lo:hi]
        d = data[lo:hi]
        np.add.at(s[i], t, d)
        dt = t[:, None] - t[None, :]
        prod = d[:, None] * d[None, :]
        keep = dt >= 0
        np.add.at(num[i], dt[keep], prod[keep])
    return s, num


def event_sums_numerators_fast(times, data, offsets, qmask, nframes, bins):
    """S_b(t) and Num_b(tau) as event_sums_numerators, vectorised over events for
    detector-sized inputs (1e7 events in seconds).

    The same pair rule (evtcorrelator.cl:3-46): within a pixel, every pair of events
    j <= i adds d_i d_j at lag t_i - t_j.  Events are pixel-major and frame-sorted, so
    the pairs at distance k in the list are (e, e + k) with both in one pixel; pairs at
    distance k + 1 exist only where those at distance k do, so the candidate set
    shrinks with k.  Each distance is accumulated with np.bincount in float64, which is
    exact because every per-distance total stays below 2^53 (checked), then added into
    int64 rows.
    """
    q = np.asarray(qmask).ravel()
    nb = len(bins)
    offsets = np.asarray(offsets, dtype=np.int64)
    npix = offsets.size - 1
    bin_of = np.full(int(max(q.max(initial=0), 0)) + 1, -1, dtype=np.int64)
    bin_of[np.asarray(bins, dtype=np.int64)] = np.arange(nb)
    pix_bin = np.where(q > 0, bin_of[np.clip(q, 0, None)], -1)
    pix = np.repeat(np.arange(npix, dtype=np.int64), np.diff(offsets))
    keep = pix_bin[pix] >= 0
    pix = pix[keep]
    t = np.asarray(times, dtype=np.int64)[keep]
    d = np.asarray(data).astype(np.int64)[keep]
    b = pix_bin[pix]
    size = nb * nframes
    dmax = int(np.abs(d).max(initial=0))
    assert dmax * dmax * max(d.size, 1) < 2 ** 53, "per-distance totals exceed 2^53"
    s = np.bincount(b * nframes + t, weights=d.astype(np.float64), minlength=size)
    num = np.bincount(b * nframes, weights=(d * d).astype(np.float64), minlength=size)
    num = num.astype(np.int64)
    cand = np.arange(d.size - 1, dtype=np.int64)
    k = 1
    while cand.size:
        cand = cand[pix[cand + k] == pix[cand]]
        if not cand.size:
            break
        j = cand + k
        num += np.bincount(b[cand] * nframes + (t[j] - t[cand]),
                           weights=(d[cand] * d[j]).astype(np.float64),
                           minlength=size).astype(np.int64)
        k += 1
        cand = cand[cand + k < d.size]
    return s.astype(np.int64).reshape(nb, nframes), num.reshape(nb, nframes)


def event_brute_force(times, data, offsets, qmask, nframes):
    """Python-integer pair loop on tiny inputs (independent check)."""
    q = np.asarray(qmask).ravel()
    bins = np.arange(1, int(q.max()) + 1)
    num = np.zeros((len(bins), nframes), dtype=object)
    s = np.zeros((len(bins), nframes), dtype=object)
    num[:] = 0
    s[:] = 0
    for p in range(q.size):
        b = int(q[p])
        if b <= 0:
            continue
        ev = [(int(times[k]), int(data[k])) for k in range(int(offsets[p]), int(offsets[p + 1]))]
        for (ti, di) in ev:
            s[b - 1, ti] += di
            for (tj, dj) in ev:
                if ti >= tj:
                    num[b - 1, ti - tj] += di * dj
    return num, s


def event_g2(times, data, offsets, qmask, nframes, scale_factor=None):
    """EventCorrelator.correlate semantics (event.py:162-204, evtcorrelator.cl:58-76),
    exact numerators/denominators and FP64 g2 = scale_b * Num / Den."""
    shape = np.asarray(qmask).shape
    q, bins = qmask_bins(shape, qmask)
    s, num = event_sums_numerators(times, data, offsets, q, nframes, bins)
    den = denominators(s)
    sc = scale_factors(q, bins, scale_factor)
    return {"num": num, "sums": s, "den": den, "g2": g2_fp64(num, den, sc), "bins": bins}


def coo_to_csr(pix, frame, count, npix):
    """Frame-sorted COO (pixel, frame, count) -> pixel-major CSR (times, data, offsets).

    A stable sort by pixel keeps each pixel's events in frame order.
    """
    pix = np.asarray(pix, dtype=np.int64)
    order = np.argsort(pix, kind="stable")
    times = np.asarray(frame)[order].astype(np.int32)
    data = np.asarray(count)[order]
    offsets = np.zeros(npix + 1, dtype=np.uint32)
    offsets[1:] = np.cumsum(np.bincount(pix, minlength=npix))
    return times, data, offsets


def csr_to_coo(times, data, offsets):
    """Pixel-major CSR -> frame-sorted COO (pixel ascending within a frame)."""
    offsets = np.asarray(offsets, dtype=np.int64)
    npix = offsets.size - 1
    pix = np.repeat(np.arange(npix, dtype=np.uint32), np.diff(offsets))
    order = np.lexsort((pix, np.asarray(times)))
    return pix[order], np.asarray(times, dtype=np.uint32)[order], np.asarray(data)[order]
