"""Exact denominators on the i8 matrix-core path: bins whose frame sums fit one i8 digit
(|S| <= 127, or 0 <= S <= 255 with unsigned frames) or two (|S| <= 16383) take
den_mfma_kernel, the others the packed 16-bit or FP64 kernels.  Every case compares Den
bit for bit with the oracle's exact integers and with a plan made with the matrix-core
path switched off (DXG2_DEN_MFMA=0, read at plan creation).  Marked gpu: run on an
MI355X."""
import os
import sys

import numpy as np
import pytest

from oracle import events as ev
from oracle import g2 as og

pytestmark = pytest.mark.gpu


def _announce(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def _qmask():
    """8 x 16 detector: bin 1 four pixels, bin 2 eight, bin 3 sixteen, bin 4 eighty."""
    q = np.zeros((8, 16), np.int32)
    q.flat[0:4] = 1
    q.flat[16:24] = 2
    q.flat[32:48] = 3
    q.flat[48:128] = 4
    return q


def _frames(seed, n, dtype):
    """Frame sums per bin: bin 1 |S| <= 124 (one signed digit), bin 2 up to 248 (one
    biased digit for unsigned frames, two for signed; one frame above 127), bin 3 up to
    1600 (two digits), bin 4 up to 20,400 (uint8; 80,000 uint16: above two digits, the
    FP64 kernel) or 10,240 (int8)."""
    rng = np.random.default_rng(seed)
    lo = -31 if dtype == np.int8 else 0
    f = np.zeros((n, 8, 16), np.int64)
    flat = f.reshape(n, -1)
    flat[:, 0:4] = rng.integers(lo, 32, size=(n, 4))
    flat[:, 16:24] = rng.integers(lo, 32, size=(n, 8))
    flat[:, 32:48] = rng.integers(lo, 101, size=(n, 16))
    hi4 = {"uint8": 255, "int8": 127, "uint16": 1000}[np.dtype(dtype).name]
    flat[:, 48:128] = rng.integers(-128 if dtype == np.int8 else 0, hi4 + 1, size=(n, 80))
    flat[n // 2, 16:24] = 31 if dtype != np.int8 else -31
    flat[n // 3, 48:128] = hi4
    return f.astype(dtype)


def _with_env(value, fn):
    old = os.environ.get("DXG2_DEN_MFMA")
    os.environ["DXG2_DEN_MFMA"] = value
    try:
        return fn()
    finally:
        if old is None:
            del os.environ["DXG2_DEN_MFMA"]
        else:
            os.environ["DXG2_DEN_MFMA"] = old


def _dense_den(frames, q):
    from dynamix_amd.correlator.dense import DenseCorrelator
    c = DenseCorrelator(q.shape, frames.shape[0], q, dtype=frames.dtype)
    c.correlate(frames)
    return c.last_denominators.cpu().numpy(), c.last_sums.cpu().numpy()


@pytest.mark.parametrize("dtype", [np.uint8, np.int8, np.uint16])
@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 1000, 1025, 4159])
def test_dense_den_mfma_exact(dtype, n):
    _announce("den mfma dense dtype=%s n=%d" % (np.dtype(dtype).name, n))
    q = _qmask()
    f = _frames(n, n, dtype)
    den, sums = _dense_den(f, q)
    ref_sums = og.frame_sums(f, q, [1, 2, 3, 4])
    np.testing.assert_array_equal(sums, ref_sums)
    np.testing.assert_array_equal(den, og.denominators(ref_sums).astype(np.float64))
    off, _ = _with_env("0", lambda: _dense_den(f, q))
    np.testing.assert_array_equal(den, off)


def _spot_check(den, sums, lags):
    for b in range(sums.shape[0]):
        s = sums[b].astype(np.int64)
        n = s.shape[0]
        for tau in lags:
            assert den[b, tau] == float(np.dot(s[:n - tau], s[tau:])), (b, tau)


def test_dense_den_mfma_long_series():
    """100,000 frames (1,563 rows of 64): every bin kind of the 8 x 16 detector (one
    signed digit, one biased digit, two digits, FP64) agrees with the VALU kernels on
    every lag and with the exact sum at spot lags."""
    _announce("den mfma dense n=100000")
    n = 100000
    q = _qmask()
    f = _frames(3, n, np.uint8)
    den, sums = _dense_den(f, q)
    np.testing.assert_array_equal(sums, og.frame_sums(f, q, [1, 2, 3, 4]))
    off, _ = _with_env("0", lambda: _dense_den(f, q))
    np.testing.assert_array_equal(den, off)
    _spot_check(den, sums, (0, 1, 63, 64, 1023, 4095, 65535, n - 1))


def test_events_den_mfma_long_series():
    """Sparse events over 100,000 frames: the matrix-core and 16-bit kernels agree on
    every lag (one-digit bins: a pixel with events in half the frames, counts < 64, and
    pixels with 3 % occupancy)."""
    _announce("den mfma events n=100000")
    from dynamix_amd.correlator.event import EventCorrelator
    n, shape = 100000, (4, 8)
    rng = np.random.default_rng(5)
    f = ((rng.random((n,) + shape) < 0.03) * rng.integers(1, 4, size=(n,) + shape))
    f[:, 0, 0] = rng.integers(0, 64, size=n) * (rng.random(n) < 0.5)
    f = f.astype(np.uint8)
    q = np.zeros(shape, np.int32)
    q[0, :] = 1     # up to 63 + 7 * 3
    q[1:, :] = 2    # 23 pixels of at most 3
    q[3, 7] = 0
    data, times, offsets = ev.compress_all_stack(f)

    def run():
        c = EventCorrelator(shape, n, max_events_count=int(np.diff(offsets).max()), qmask=q,
                            dtype=np.uint8)
        c.correlate(times, data, offsets)
        return c.last_denominators.cpu().numpy(), c.last_sums.cpu().numpy()

    den, sums = run()
    off, _ = _with_env("0", run)
    np.testing.assert_array_equal(den, off)
    _spot_check(den, sums, (0, 1, 63, 64, 4095, 65535, n - 1))


def test_fft_den_mfma_exact():
    _announce("den mfma fft")
    from dynamix_amd.correlator.dense import FFTCorrelator
    q = _qmask()
    f = _frames(7, 3000, np.uint8)

    def run():
        c = FFTCorrelator(q.shape, f.shape[0], q)
        c.correlate(f)
        return c.last_denominators.cpu().numpy()

    den = run()
    ref = og.denominators(og.frame_sums(f, q, [1, 2, 3, 4])).astype(np.float64)
    np.testing.assert_array_equal(den, ref)
    np.testing.assert_array_equal(den, _with_env("0", run))
