"""calc_std on the GPU: dev(tau) against the FP64 oracle (from exact integers) and the
reference's own float32 outputs; res unchanged by calc_std; bit-reproducible dev.
Marked gpu."""
import sys

import numpy as np
import pytest

from oracle import g2 as og

pytestmark = pytest.mark.gpu

# dev vs the FP64 oracle: the FP64 ratio statistics rounded once to float32, plus an
# absolute floor for spreads that are exactly zero in exact arithmetic (one ratio per
# diagonal at the last lag, or equal ratios whose FP64 roundings differ by an ulp).
DEV_RTOL = 1e-6
DEV_ATOL = 1e-12


def _announce(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def _corr(frames, qmask, **extra):
    from dynamix_amd.correlator.dense import MatMulCorrelator
    return MatMulCorrelator(frames.shape[1:], frames.shape[0], qmask, extra_options=extra)


def assert_dev_close(dev, ref):
    fin = np.isfinite(ref)
    assert np.array_equal(fin, np.isfinite(dev)), "NaN pattern differs"
    err = np.abs(dev[fin].astype(np.float64) - ref[fin])
    bound = DEV_RTOL * np.abs(ref[fin]) + DEV_ATOL
    assert (err <= bound).all(), (err / np.maximum(np.abs(ref[fin]), 1e-300)).max()


def _check(frames, qmask, **extra):
    c = _corr(frames, qmask, **extra)
    r = c.correlate(frames, calc_std=True)
    assert r.res.dtype == np.float32 and r.dev.dtype == np.float32
    ref = og.dense_g2(frames, qmask)
    np.testing.assert_array_equal(c.last_numerators.cpu().numpy(), ref["num"])
    np.testing.assert_array_equal(c.last_sums.cpu().numpy(), ref["sums"])
    assert_dev_close(r.dev, og.dense_dev_fp64(frames, qmask))
    return c, r


def _bins(rng, shape, nbins):
    q = rng.integers(0, nbins + 1, size=shape).astype(np.int32)
    for b in range(1, nbins + 1):
        q.flat[b] = b
    return q


@pytest.mark.parametrize("tile", [256, 128])
@pytest.mark.parametrize("n,shape,mean", [
    (7, (5, 6), 3.0),        # fewer frames than one tile
    (130, (9, 11), 2.0),     # straddles a 128 tile
    (300, (16, 20), 1.0),    # several tiles
    (520, (24, 30), 0.8),    # 2 * 256 + 8 frames
])
def test_std_u8_vs_fp64_oracle(n, shape, mean, tile):
    from dynamix_amd.synth import speckle_frames
    _announce("std u8 n=%d shape=%s tile=%d" % (n, shape, tile))
    f = speckle_frames(shape, n, mean, 20.0, n, np.uint8)
    _check(f, _bins(np.random.default_rng(n), shape, 3), tile=tile)


def test_std_zero_frames_give_nan_only_where_numpy_does():
    """One frame with no counts in bin 1: dev is NaN for every lag whose diagonal
    meets it (tau <= N-1-5 or tau <= 5), finite elsewhere; bin 3 is empty."""
    _announce("std zero frame / empty bin")
    rng = np.random.default_rng(4)
    n, shape = 300, (12, 14)
    f = rng.integers(1, 6, size=(n,) + shape).astype(np.uint8)
    q = rng.integers(1, 3, size=shape).astype(np.int32)
    f[5][q == 1] = 0
    c, r = _check(f, q)
    assert np.isnan(r.dev[0, : n - 5]).all() and np.isfinite(r.dev[0, n - 5:]).all()
    assert np.isfinite(r.dev[1]).all()
    q3 = q.copy()
    q3[0, 0] = 4  # bins 1, 2, 4 (one pixel) and an empty bin 3
    c3, r3 = _check(f, q3)
    assert np.isnan(r3.dev[2]).all()


@pytest.mark.parametrize("tile", [256, 128])
def test_std_u16_byte_split(tile):
    _announce("std u16 tile=%d" % tile)
    rng = np.random.default_rng(12)
    f = rng.integers(0, 65536, size=(300, 14, 15)).astype(np.uint16)
    f[rng.random(f.shape) < 0.3] = 0
    _check(f, _bins(rng, (14, 15), 3), tile=tile)


def test_std_i8_signed():
    _announce("std i8")
    rng = np.random.default_rng(13)
    f = rng.integers(-20, 60, size=(270, 10, 12)).astype(np.int8)
    _check(f, _bins(rng, (10, 12), 2))


@pytest.mark.parametrize("tile", [256, 128])
def test_std_k_split_and_chunks(monkeypatch, tile):
    """Work items over 3 k-blocks each (several items per tile) and a 1 MiB tile
    scratch (many chunks): the same dev."""
    _announce("std k-split + chunks tile=%d" % tile)
    monkeypatch.setenv("DXG2_GRAM_KCHUNK", "3")
    monkeypatch.setenv("DXG2_STD_SCRATCH_MB", "1")
    rng = np.random.default_rng(14)
    f = rng.integers(0, 256, size=(600, 30, 40)).astype(np.uint8)
    _check(f, _bins(rng, (30, 40), 2), tile=tile)
    g = rng.integers(0, 3000, size=(260, 20, 20)).astype(np.uint16)
    _check(g, _bins(rng, (20, 20), 2), tile=tile)


def test_std_reference_goldens(golden):
    """Against the reference py_dense_correlator(calc_std=True) outputs: dev within the
    float32 rounding the reference's ratios carry, res within its test tolerance."""
    for name in ["u8_small", "u8_big_values", "i8_sparse", "u16_mixed", "u8_onepix"]:
        _announce("std golden " + name)
        f = golden["dense_%s_frames" % name]
        q = golden["dense_%s_qmask" % name]
        res, dev = golden["dense_%s_res" % name], golden["dense_%s_dev" % name]
        c, r = _check(f, q)
        fin = np.isfinite(dev)
        assert np.array_equal(fin, np.isfinite(r.dev))
        assert np.abs(r.dev[fin] - dev[fin]).max(initial=0) < 1e-5 * max(1.0, np.abs(dev[fin]).max(initial=0))
        fin = np.isfinite(res)
        assert np.abs(r.res[fin] - res[fin]).max(initial=0) < 5e-3


def test_std_leaves_res_unchanged_and_is_reproducible():
    import torch
    from dynamix_amd.synth import speckle_frames
    _announce("std res unchanged / reproducible")
    f = speckle_frames((20, 24), 400, 0.5, 30.0, 9, np.uint8)
    q = _bins(np.random.default_rng(9), (20, 24), 4)
    c = _corr(f, q)
    g = c.correlate(f)
    num = c.last_numerators.clone()
    r1 = c.correlate(f, calc_std=True)
    assert torch.equal(num, c.last_numerators)
    np.testing.assert_array_equal(g.view(np.uint32), r1.res.view(np.uint32))
    r2 = c.correlate(torch.from_numpy(f).cuda(), calc_std=True)
    np.testing.assert_array_equal(r1.dev.view(np.uint32), r2.dev.view(np.uint32))
