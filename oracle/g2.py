"""Dense g2 oracle (TEST INFRASTRUCTURE ONLY -- see oracle/__init__.py).

For bin b with pixel set P_b (n_b pixels), frames t = 0..N-1 and lags tau = 0..N-1:

    S_b(t)    = sum_{p in P_b} I(t, p)
    Num_b(tau) = sum_{t=0}^{N-1-tau} sum_{p in P_b} I(t, p) I(t+tau, p)
    Den_b(tau) = sum_{t=0}^{N-1-tau} S_b(t) S_b(t+tau)
    g2_b(tau)  = scale_b * Num_b(tau) / Den_b(tau)        (scale_b = n_b by default)

This is what every reference correlator computes (README.md:11-15;
dynamix/correlator/dense.py:48-64 MatMul; dense.py:364-368 FFT;
dynamix/resources/opencl/densecorrelator.cl:54 and evtcorrelator.cl:70-75).

Parity pinning of each function (tests/test_oracle.py):
  qmask_bins, scale_factors   -- golden vectors of the reference's BaseCorrelator
                                 semantics + closed form (bins 1..max, None -> 1 bin).
  frame_sums, numerators      -- brute-force triple loops on tiny inputs; invariants
                                 Num(0) = sum I^2, permutation invariance, constant
                                 frames; equality with the event-list restatement.
  sums_numerators_fast        -- equality with frame_sums / numerators and the brute
                                 force on the same stacks.
  denominators                -- brute force with Python integers; np.correlate.
  g2_fp64                     -- golden vectors of the reference py_dense_correlator
                                 (within its float32 rounding), single-pixel bin == 1,
                                 constant frames closed form.
  matmul_correlator_f32       -- bit-exact against golden vectors written by running
                                 the reference py_dense_correlator (dense.py:32-70),
                                 res and (calc_std) dev.
  dense_dev_fp64              -- the golden dev vectors within the reference's float32
                                 rounding; brute-force loops on a tiny stack.
  two_time_fp64               -- a loop restatement of the reference's photon-pair
                                 count and TTCF normalisation (ecorrts_q,
                                 event_y.py:169-230; fecorrt.f:1-26) on tiny stacks;
                                 closed form for time-constant frames; symmetry;
                                 diagonal rule; the last diagonal corner.
  fft_correlator              -- against `numerators`/`denominators` within FP32
                                 tolerance; FFT length against the reference's
                                 get_next_power golden vectors (utils.py:50-60).
"""
from bisect import bisect_left
from itertools import product

import math

import numpy as np


# --------------------------------------------------------------------------- qmask
def qmask_bins(shape, qmask=None):
    """Bins of a qmask.  Restates BaseCorrelator._set_qmask (common.py:36-46).

    qmask None -> one bin holding every pixel; otherwise bins are 1..qmask.max()
    (0 and negative values are excluded pixels).  Returns (qmask_int32, bins).
    """
    if np.isscalar(shape):
        shape = (int(shape), int(shape))
    if qmask is None:
        return np.ones(shape, dtype=np.int32), np.array([1], dtype=np.int32)
    q = np.ascontiguousarray(qmask, dtype=np.int32)
    n_bins = int(q.max())
    return q, np.arange(1, n_bins + 1, dtype=np.int32)


def pixel_counts(qmask, bins):
    """n_b = number of pixels of each bin (common.py:70-72)."""
    q = np.asarray(qmask).ravel()
    cnt = np.bincount(np.clip(q, 0, None), minlength=int(bins[-1]) + 1 if len(bins) else 1)
    return cnt[np.asarray(bins, dtype=np.int64)].astype(np.int64)


def scale_factors(qmask, bins, scale_factor=None):
    """Per-bin scale factors as BaseCorrelator._set_scale_factor (common.py:56-72).

    Default: pixel count of each bin.  A scalar is broadcast only if there is one bin
    (the reference asserts len == n_bins, common.py:62-64).  Returned as float64 array.
    """
    if scale_factor is None:
        return pixel_counts(qmask, bins).astype(np.float64)
    if isinstance(scale_factor, dict):
        assert len(scale_factor) == len(bins)
        return np.array([scale_factor[b] for b in bins], dtype=np.float64)
    if not np.iterable(scale_factor):
        scale_factor = [scale_factor]
    assert len(scale_factor) == len(bins)
    return np.asarray(scale_factor, dtype=np.float64)


# ------------------------------------------------------------------- exact pieces
def gather_bin(frames, qmask, b):
    """A_b (N, n_b): the pixels of bin b in raster order (dense.py:48-49)."""
    n = frames.shape[0]
    return frames.reshape(n, -1)[:, np.asarray(qmask).ravel() == b]


def frame_sums(frames, qmask, bins):
    """S_b(t) as exact int64, shape (n_bins, N)."""
    n = frames.shape[0]
    flat = frames.reshape(n, -1)
    q = np.asarray(qmask).ravel()
    out = np.zeros((len(bins), n), dtype=np.int64)
    for i, b in enumerate(bins):
        out[i] = flat[:, q == b].astype(np.int64).sum(axis=1)
    return out


def _diag_sums(g):
    """sum_t g[t, t+tau] for tau = 0..N-1, exact for an integer-valued matrix."""
    n = g.shape[0]
    out = np.zeros(n, dtype=np.int64)
    for tau in range(n):
        out[tau] = int(np.trace(g, offset=tau))
    return out


def bin_numerator(a):
    """Num(tau) for one gathered bin a (N, n_b), exact int64.

    Uses a float64 BLAS Gram when every partial sum is an integer below 2**53 (then
    float64 arithmetic is exact), an int64 Gram otherwise.
    """
    a = np.asarray(a)
    n, k = a.shape
    if k == 0:
        return np.zeros(n, dtype=np.int64)
    amax = int(np.abs(a.astype(np.int64)).max()) if a.size else 0
    if amax * amax * k < 2 ** 53:
        af = a.astype(np.float64)
        g = af @ af.T
        g = g.astype(np.int64)
    else:
        ai = a.astype(np.int64)
        g = ai @ ai.T
    return _diag_sums(g)


def numerators(frames, qmask, bins):
    """Num_b(tau), exact int64, shape (n_bins, N)."""
    return np.stack([bin_numerator(gather_bin(frames, qmask, b)) for b in bins]) \
        if len(bins) else np.zeros((0, frames.shape[0]), np.int64)


def sums_numerators_fast(frames, qmask, bins):
    """(S_b(t), Num_b(tau)) as frame_sums / numerators, for detector-sized stacks.

    The stack is transposed once to pixel-major rows, each bin's rows are taken by
    index, and the bin's Gram A_b A_b^T is formed by one float64 BLAS product when every
    partial sum is an integer below 2^53 (exact), otherwise in int64.  Same quantities,
    one pass over the stack instead of one per bin.
    """
    n = frames.shape[0]
    q = np.asarray(qmask).ravel()
    xt = np.ascontiguousarray(frames.reshape(n, -1).T)      # (H*W, N)
    order = np.argsort(q, kind="stable")
    qs = q[order]
    s = np.zeros((len(bins), n), dtype=np.int64)
    num = np.zeros((len(bins), n), dtype=np.int64)
    for i, b in enumerate(bins):
        lo, hi = np.searchsorted(qs, b, "left"), np.searchsorted(qs, b, "right")
        if hi == lo:
            continue
        at = xt[order[lo:hi]]                                  # (n_b, N)
        s[i] = at.astype(np.int64).sum(axis=0)
        amax = int(np.abs(at).max())
        if amax * amax * (hi - lo) < 2 ** 53:
            af = at.astype(np.float64)
            g = (af.T @ af).astype(np.int64)
        else:
            ai = at.astype(np.int64)
            g = ai.T @ ai
        num[i] = _diag_sums(g)
    return s, num


def denominators(sums):
    """Den_b(tau) = sum_t S(t) S(t+tau), exact (Python integers when int64 could overflow).

    Same quantity the reference FFT path computes exactly with np.correlate on integer
    sums (dense.py:365-366).  Returned as an object array of Python ints when the bound
    exceeds int64, else int64.
    """
    sums = np.asarray(sums)
    n = sums.shape[-1]
    smax = int(np.abs(sums).max()) if sums.size else 0
    if smax * smax * max(n, 1) < 2 ** 62:
        out = np.zeros(sums.shape, dtype=np.int64)
        for i in range(sums.shape[0]):
            s = sums[i].astype(np.int64)
            out[i] = np.correlate(s, s, "full")[n - 1:]
        return out
    out = np.empty(sums.shape, dtype=object)
    for i in range(sums.shape[0]):
        s = [int(x) for x in sums[i]]
        for tau in range(n):
            out[i, tau] = sum(s[t] * s[t + tau] for t in range(n - tau))
    return out


def g2_fp64(num, den, scale):
    """g2 = scale * Num / Den in float64 (NaN where Den == 0, as 0/0 in the reference)."""
    num = np.asarray(num, dtype=np.float64)
    den = np.asarray(den, dtype=np.float64) if np.asarray(den).dtype != object else \
        np.vectorize(float, otypes=[np.float64])(den)
    scale = np.asarray(scale, dtype=np.float64).reshape(-1, 1)
    with np.errstate(divide="ignore", invalid="ignore"):
        return scale * num / den


def dense_g2(frames, qmask=None, scale_factor=None):
    """Full exact pipeline: returns dict(num, sums, den, g2) for a frame stack."""
    shape = frames.shape[1:]
    q, bins = qmask_bins(shape, qmask)
    s = frame_sums(frames, q, bins)
    num = numerators(frames, q, bins)
    den = denominators(s)
    sc = scale_factors(q, bins, scale_factor)
    return {"num": num, "sums": s, "den": den, "g2": g2_fp64(num, den, sc),
            "bins": bins, "qmask": q, "scale": sc}


# ----------------------------------------------------------- brute force (tiny)
def brute_force(frames, qmask):
    """Triple loops with Python integers; the independent check on tiny inputs."""
    n = frames.shape[0]
    q, bins = qmask_bins(frames.shape[1:], qmask)
    flat = frames.reshape(n, -1)
    qf = q.ravel()
    num = np.zeros((len(bins), n), dtype=object)
    s = np.zeros((len(bins), n), dtype=object)
    for i, b in enumerate(bins):
        pix = [p for p in range(flat.shape[1]) if qf[p] == b]
        for t in range(n):
            s[i, t] = sum(int(flat[t, p]) for p in pix)
        for tau in range(n):
            num[i, tau] = sum(int(flat[t, p]) * int(flat[t + tau, p])
                              for t in range(n - tau) for p in pix)
    den = np.zeros((len(bins), n), dtype=object)
    for i in range(len(bins)):
        for tau in range(n):
            den[i, tau] = sum(s[i, t] * s[i, t + tau] for t in range(n - tau))
    return num, s, den


# --------------------------------------------------- reference float32 algorithm
def py_dense_correlator_f32(frames, mask, calc_std=False):
    """The reference numpy correlator for one bin, in its own float32 arithmetic.

    Follows dense.py:32-70 step for step (gather cast to float32 :49, float32 mean :50,
    sgemm Gram :54, outer product of means :55, per-lag diagonal sums :61-64, and with
    calc_std the std of the per-t ratios along each diagonal :65-66), so that its
    float32 rounding matches the reference; it is the CPU baseline the GPU is timed
    beside.  Returns float64 res of length N (the reference stores it into float32), or
    (res, dev) with calc_std.
    """
    rows, cols = np.nonzero(mask > 0)
    a = np.asarray(frames[:, rows, cols], dtype=np.float32)
    nt, npx = a.shape
    m = a.mean(axis=1, dtype=np.float32).reshape(1, nt)
    gram = a @ a.T
    mm = m.T @ m
    res = np.zeros(nt)
    dev = np.zeros(nt)
    with np.errstate(divide="ignore", invalid="ignore"):
        for tau in range(nt):
            dn = np.diagonal(gram, offset=tau) / npx
            dd = np.diagonal(mm, offset=tau)
            res[tau] = np.sum(dn) / np.sum(dd)
            if calc_std:
                # a Python float divisor keeps the quotient in float32 (math.sqrt at :66)
                dev[tau] = np.std(dn / dd) / math.sqrt(len(dd))
    return (res, dev) if calc_std else res


def matmul_correlator_f32(frames, qmask=None, calc_std=False):
    """MatMulCorrelator.correlate (dense.py:128-142): one bin at a time, float32 out;
    with calc_std also the float32 dev array."""
    q, bins = qmask_bins(frames.shape[1:], qmask)
    out = np.zeros((len(bins), frames.shape[0]), dtype=np.float32)
    dev = np.zeros_like(out)
    for i, b in enumerate(bins):
        r = py_dense_correlator_f32(frames, q == b, calc_std)
        if calc_std:
            out[i], dev[i] = r
        else:
            out[i] = r
    return (out, dev) if calc_std else out


def dense_dev_fp64(frames, qmask=None):
    """dev(tau) of the reference's calc_std (dense.py:58-59,65-66) from exact integers:
    per bin and lag, the population std over t of r_t = n_b G(t, t+tau) / (S(t) S(t+tau))
    divided by sqrt(N - tau), in float64.  A zero frame sum on a diagonal gives inf/NaN
    ratios, hence NaN, as numpy does."""
    q, bins = qmask_bins(frames.shape[1:], qmask)
    n = frames.shape[0]
    dev = np.zeros((len(bins), n))
    with np.errstate(divide="ignore", invalid="ignore"):
        for i, b in enumerate(bins):
            a = gather_bin(frames, q, b).astype(np.float64)
            npx = a.shape[1]
            if npx == 0:
                dev[i] = np.nan
                continue
            g = a @ a.T               # exact: integer entries well below 2^53 in tests
            s = a.sum(axis=1)
            for tau in range(n):
                r = npx * np.diagonal(g, offset=tau) / (s[:n - tau] * s[tau:])
                dev[i, tau] = np.std(r) / np.sqrt(n - tau)
    return dev


# ------------------------------------------------------------------------ FFT path
def two_time_fp64(frames, qmask, q):
    """Two-time correlation matrix of bin q (1-based) as the reference's ecorrts_q builds
    `trc` for ttcf_par = q (event_y.py:169-230 with WXPCS/fecorrt.f:1-26), in FP64 from
    exact integers:

      cor(t, u) = sum_p I(t, p) I(u, p) for t != u: fecorrt counts every pair of photons
                  of one pixel in two frames and symmetrises (fecorrt.f:8-25);
      cor(t, t) = sum_p I(t, p) (I(t, p) - 1): the pairs of photons within one frame,
                  counted once each way by the symmetrisation;
      trc(t, u) = n_b cor(t, u) / (S(t) S(u)), n_b every pixel of the bin
                  (for_norm, event_y.py:189,208-209,223);
    then the diagonal entries 0 .. N-2 become the mean of the positive entries of the
    first off-diagonal (event_y.py:224-227); trc[N-1, N-1] keeps its value.  Where
    S(t) S(u) = 0 the entry is NaN (numpy gives 0/0 = NaN for photon counts, whose
    frame sum is 0 only with every count 0).
    """
    qq, bins = qmask_bins(frames.shape[1:], qmask)
    a = gather_bin(frames, qq, q).astype(np.float64)
    n = frames.shape[0]
    npx = a.shape[1]
    g = a @ a.T                      # exact: integer entries well below 2^53 in tests
    s = a.sum(axis=1)
    cor = g.copy()
    cor[np.diag_indices(n)] -= s
    ss = np.outer(s, s)
    with np.errstate(divide="ignore", invalid="ignore"):
        trc = np.where(ss != 0, npx * cor / np.where(ss != 0, ss, 1.0), np.nan)
    if n > 1:
        d1 = np.diagonal(trc, 1)
        pos = d1[d1 > 0]
        fill = pos.mean() if pos.size else np.nan
        idx = np.arange(n - 1)
        trc[idx, idx] = fill
    return trc


def two_time_loops(frames, qmask, q):
    """The same matrix by the reference's own steps, on small non-negative stacks:
    every photon becomes one (pixel, frame) entry, sorted by pixel then frame (the
    lexsort of event_y.py:176-178); fecorrt's loop counts each pair of entries of one
    pixel into cor(t_later, t_earlier) and then symmetrises cor(j, i) = cor(i, j) =
    cor(j, i) + cor(i, j) for j >= i (fecorrt.f:8-25); S(t) is the photon histogram
    over frames (event_y.py:195); trc = n_b cor / (S S^T) element by element; the
    diagonal rule last (event_y.py:224-227)."""
    qq, _ = qmask_bins(frames.shape[1:], qmask)
    n = frames.shape[0]
    flat = frames.reshape(n, -1)
    pix = [p for p in range(qq.size) if qq.flat[p] == q]
    entries = []                      # (pixel, frame) per photon
    for p in pix:
        for t in range(n):
            entries.extend([(p, t)] * int(flat[t, p]))
    entries.sort()
    cor = [[0] * n for _ in range(n)]
    for i in range(len(entries)):
        j = i + 1
        while j < len(entries) and entries[j][0] == entries[i][0]:
            cor[entries[j][1]][entries[i][1]] += 1
            j += 1
    for i in range(n):
        for j in range(i, n):
            cor[j][i] = cor[j][i] + cor[i][j]
            cor[i][j] = cor[j][i]
    hist = [0] * n
    for _, t in entries:
        hist[t] += 1
    trc = [[0.0] * n for _ in range(n)]
    for t in range(n):
        for u in range(n):
            den = hist[t] * hist[u]
            trc[t][u] = len(pix) * cor[t][u] / den if den else float("nan")
    pos = [trc[j][j + 1] for j in range(n - 1) if trc[j][j + 1] > 0]
    fill = sum(pos) / len(pos) if pos else float("nan")
    for j in range(n - 1):
        trc[j][j] = fill
    return np.array(trc)


def fft_powers():
    """All 2^a 3^b 5^c 7^d with a in [2,15], b<=9, c<=6, d<=5 (utils.py:29-47)."""
    vals = set()
    for a, b, c, d in product(range(2, 16), range(0, 10), range(0, 7), range(0, 6)):
        vals.add((2 ** a) * (3 ** b) * (5 ** c) * (7 ** d))
    return np.array(sorted(vals), dtype=np.int64)


_POWERS = None


def next_fft_length(n):
    """Smallest member of fft_powers() >= n (get_next_power, utils.py:50-60)."""
    global _POWERS
    if _POWERS is None:
        _POWERS = fft_powers()
    i = bisect_left(_POWERS.tolist(), int(n))
    return int(_POWERS[i])


def fft_correlator(frames, qmask=None, workers=-1):
    """FFTWCorrelator.correlate restated with scipy.fft (dense.py:314-322, 347-369).

    Per bin: float32 FFT of each pixel series zero-padded to Nf = next_fft_length(2N),
    product with the FFT of the time-reversed series, inverse FFT, sum over pixels,
    slice [N-1, 2N-1); denominator np.correlate of the integer frame sums / n_b.
    pyFFTW is absent here, so scipy.fft (single precision, `workers` threads) stands in.
    """
    import scipy.fft as sfft
    n = frames.shape[0]
    q, bins = qmask_bins(frames.shape[1:], qmask)
    nf = next_fft_length(2 * n)
    flat = frames.reshape(n, -1)
    out = np.zeros((len(bins), n), dtype=np.float32)
    for i, b in enumerate(bins):
        a = flat[:, q.ravel() == b]
        npix = a.shape[1]
        x = np.zeros((npix, nf), dtype=np.float32)
        x[:, :n] = a.T
        f1 = sfft.rfft(x, axis=-1, workers=workers)
        x[:, :n] = a.T[:, ::-1]
        f2 = sfft.rfft(x, axis=-1, workers=workers)
        f1 *= f2
        num = sfft.irfft(f1, n=nf, axis=-1, workers=workers)
        num = num.sum(axis=0)[n - 1:2 * n - 1]
        s = a.sum(axis=1)
        den = np.correlate(s, s, "full")[s.size - 1:] / npix
        with np.errstate(divide="ignore", invalid="ignore"):
            out[i] = num / den
    return out
