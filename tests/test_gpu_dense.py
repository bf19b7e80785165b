"""GPU parity of the dense HIP path against the CPU oracle (exact numerators, sums,
denominators; g2 within 1e-6 relative of FP64).  Marked gpu: run on an MI355X."""
import sys

import numpy as np
import pytest

from oracle import g2 as og

pytestmark = pytest.mark.gpu

G2_RTOL = 1e-6   # g2 vs the FP64 oracle: float32 rounding of an FP64 ratio of exact integers


def _announce(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def _run(frames, qmask, cls=None, **kw):
    from dynamix_amd.correlator.dense import MatMulCorrelator
    cls = cls or MatMulCorrelator
    n = frames.shape[0]
    c = cls(frames.shape[1:], n, qmask, **kw)
    g2 = c.correlate(frames)
    return c, g2


def _check_exact(c, g2, ref):
    np.testing.assert_array_equal(c.last_sums.cpu().numpy(), ref["sums"])
    np.testing.assert_array_equal(c.last_numerators.cpu().numpy(), ref["num"])
    den_ref = np.asarray(ref["den"], dtype=np.float64) if ref["den"].dtype != object else \
        np.array([[float(x) for x in r] for r in ref["den"]])
    np.testing.assert_array_equal(c.last_denominators.cpu().numpy(), den_ref)
    assert_g2_close(g2, ref["g2"], G2_RTOL)


def assert_g2_close(g2, g, rtol):
    """Same NaN pattern; exact zeros where the FP64 value is 0; relative error <= rtol."""
    fin = np.isfinite(g)
    assert np.array_equal(fin, np.isfinite(g2)), "NaN pattern differs"
    zero = fin & (g == 0)
    assert (g2[zero] == 0).all()
    nz = fin & (g != 0)
    rel = np.abs(g2[nz].astype(np.float64) - g[nz]) / np.abs(g[nz])
    assert rel.max(initial=0.0) <= rtol, rel.max()


def _speckle(seed, n, shape, mean, tau_c=20.0, dtype=np.uint8):
    from dynamix_amd.synth import speckle_frames
    return speckle_frames(shape, n, mean, tau_c, seed, dtype)


@pytest.mark.parametrize("tile", [256, 128])
@pytest.mark.parametrize("n,shape,mean,nbins", [
    (7, (5, 6), 2.0, 2),          # fewer frames than one tile
    (130, (9, 11), 1.0, 3),       # frames straddle a 128 tile boundary
    (300, (16, 20), 0.3, 4),      # several tiles, several bins
    (257, (33, 40), 0.05, 5),     # 256 + 1 frames, sparse
    (600, (40, 50), 0.1, 3),      # several 256 tiles, odd k-block counts
])
def test_dense_u8_vs_oracle(n, shape, mean, nbins, tile):
    _announce("dense u8 n=%d shape=%s tile=%d" % (n, shape, tile))
    rng = np.random.default_rng(n)
    f = _speckle(n, n, shape, mean)
    q = rng.integers(0, nbins + 1, size=shape).astype(np.int32)
    for b in range(1, nbins + 1):
        q.flat[b] = b
    c, g2 = _run(f, q, extra_options={"tile": tile})
    _check_exact(c, g2, og.dense_g2(f, q))


def test_dense_u8_full_range_values():
    """Values up to 255 exercise the -128 bias and its correction terms."""
    _announce("dense u8 full range")
    rng = np.random.default_rng(1)
    f = rng.integers(0, 256, size=(150, 12, 13)).astype(np.uint8)
    q = rng.integers(0, 4, size=(12, 13)).astype(np.int32)
    c, g2 = _run(f, q)
    _check_exact(c, g2, og.dense_g2(f, q))


def test_dense_denominator_int128_path():
    """N max(S)^2 >= 2^52: the denominator takes the 128-bit accumulation path."""
    _announce("dense denominator int128 path")
    rng = np.random.default_rng(7)
    f = rng.integers(200, 256, size=(64, 256, 256)).astype(np.uint8)
    q = np.ones((256, 256), dtype=np.int32)
    s_max = float(f.reshape(64, -1).sum(axis=1, dtype=np.int64).max())
    assert 64 * s_max ** 2 >= 2.0 ** 52
    c, g2 = _run(f, q)
    _check_exact(c, g2, og.dense_g2(f, q))


@pytest.mark.parametrize("tile", [256, 128])
@pytest.mark.parametrize("n,shape,hi", [
    (7, (5, 6), 65535),       # fewer frames than one tile, full 16-bit range
    (300, (16, 20), 700),     # several tiles, values across the byte boundary
    (520, (24, 30), 65535),   # 256 * 2 + 8 frames, full range
])
def test_dense_u16_byte_split_vs_oracle(n, shape, hi, tile):
    """uint16 frames: two biased digit planes, HH + cross + LL Gram passes."""
    _announce("dense u16 n=%d shape=%s hi=%d tile=%d" % (n, shape, hi, tile))
    rng = np.random.default_rng(n + hi)
    f = rng.integers(0, hi + 1, size=(n,) + shape).astype(np.uint16)
    f[rng.random(f.shape) < 0.5] = 0  # sparse, as photon counts are
    q = rng.integers(0, 4, size=shape).astype(np.int32)
    for b in range(1, 4):
        q.flat[b] = b
    c, g2 = _run(f, q, extra_options={"tile": tile})
    _check_exact(c, g2, og.dense_g2(f, q))


@pytest.mark.parametrize("tile", [256, 128])
@pytest.mark.parametrize("n_wide", [0, 1, 37, 300])
def test_dense_u16_wide_pixels_first(n_wide, tile):
    """uint16 counts below 256 except on n_wide scattered pixels: the dense call orders
    each bin's wide pixels first and clips the high-byte Gram items to their k-blocks
    (none at all for n_wide = 0, one k-block for 1, part of a k-block for 37, several for
    300); numerators, sums and denominators stay exact, and calc_std (unclipped) agrees."""
    _announce("dense u16 wide-first n_wide=%d tile=%d" % (n_wide, tile))
    rng = np.random.default_rng(100 + n_wide)
    shape, n = (40, 60), 290
    f = rng.integers(0, 256, size=(n,) + shape).astype(np.uint16)
    f[rng.random(f.shape) < 0.6] = 0
    wide = rng.choice(shape[0] * shape[1], size=n_wide, replace=False)
    for p in wide:
        t = rng.integers(0, n, size=3)
        f[t, p // shape[1], p % shape[1]] = rng.integers(256, 65536, size=3)
    q = rng.integers(0, 6, size=shape).astype(np.int32)
    for b in range(1, 6):
        q.flat[b] = b
    c, g2 = _run(f, q, extra_options={"tile": tile})
    ref = og.dense_g2(f, q)
    _check_exact(c, g2, ref)
    from dynamix_amd.correlator.dense import MatMulCorrelator
    cs = MatMulCorrelator(shape, n, q, extra_options={"tile": tile})
    cs.correlate(f, calc_std=True)
    np.testing.assert_array_equal(cs.last_numerators.cpu().numpy(), ref["num"])


@pytest.mark.parametrize("unaligned", [False, True])
def test_dense_u16_wide_flags_odd_width(unaligned):
    """uint16 wide-pixel flags when H*W is not a multiple of 8 or the frames' device
    address is not 16-byte aligned (per-pixel flag pass), over several frame chunks:
    counts above 255 in the first and last frames and on both sides of a chunk edge."""
    import torch
    _announce("dense u16 wide flags odd width unaligned=%s" % unaligned)
    rng = np.random.default_rng(7)
    shape, n = (37, 41), 300
    f = rng.integers(0, 200, size=(n,) + shape).astype(np.uint16)
    f[rng.random(f.shape) < 0.6] = 0
    hw = shape[0] * shape[1]
    wide = rng.choice(hw, size=12, replace=False)
    for k, p in enumerate(wide):
        t = (0, n - 1, 149, 150, 151)[k % 5]
        f[t, p // shape[1], p % shape[1]] = 256 + 1000 * k
    q = rng.integers(0, 5, size=shape).astype(np.int32)
    ref = og.dense_g2(f, q)
    from dynamix_amd.correlator.dense import MatMulCorrelator
    c = MatMulCorrelator(shape, n, q)
    if unaligned:
        buf = torch.zeros(f.size + 1, dtype=torch.int16, device="cuda")
        buf[1:] = torch.from_numpy(f.ravel().view(np.int16)).cuda()
        view = buf[1:].view(torch.uint16).view(n, *shape)
        assert view.data_ptr() % 16 != 0
        g2 = c.correlate(view)
    else:
        g2 = c.correlate(f)
    _check_exact(c, g2, ref)


def test_dense_u16_speckle_with_hot_pixels():
    """Speckle counts with a hot-pixel tail (mean 40), 20 linear-width annuli."""
    _announce("dense u16 speckle hot pixels")
    from dynamix_amd import synth
    shape, n = (48, 52), 600
    mm = synth.mean_map(shape, 0.5, hot_fraction=0.01, hot_mean=40.0, seed=3)
    f = synth.speckle_frames(shape, n, 0.5, 50.0, 11, np.uint16, mean_map_=mm)
    q = synth.annular_qmask(shape, 20, widths="linear")
    c, g2 = _run(f, q)
    _check_exact(c, g2, og.dense_g2(f, q))


def test_dense_i8_signed_values():
    _announce("dense i8 signed")
    rng = np.random.default_rng(2)
    f = rng.integers(-128, 128, size=(140, 10, 10)).astype(np.int8)
    q = rng.integers(0, 3, size=(10, 10)).astype(np.int32)
    c, g2 = _run(f, q)
    ref = og.dense_g2(f, q)
    np.testing.assert_array_equal(c.last_numerators.cpu().numpy(), ref["num"])
    np.testing.assert_array_equal(c.last_sums.cpu().numpy(), ref["sums"])


def test_dense_qmask_none_and_empty_bins():
    _announce("dense qmask None / empty bin")
    f = _speckle(3, 64, (8, 9), 0.5)
    c, g2 = _run(f, None)
    _check_exact(c, g2, og.dense_g2(f, None))
    q = np.zeros((8, 9), np.int32)
    q[:4] = 1
    q[6:] = 3            # bin 2 has no pixels -> NaN row, as np.mean of an empty gather
    c, g2 = _run(f, q)
    ref = og.dense_g2(f, q)
    assert np.isnan(g2[1]).all()
    _check_exact(c, g2, ref)


def test_dense_single_pixel_bin_is_one():
    _announce("dense single pixel")
    f = _speckle(4, 50, (6, 6), 3.0)
    q = np.zeros((6, 6), np.int32)
    q[2, 3] = 1
    c, g2 = _run(f, q)
    fin = np.isfinite(g2)
    np.testing.assert_allclose(g2[fin], 1.0, rtol=1e-6)


@pytest.mark.parametrize("tile", [256, 128])
def test_dense_k_chunks_over_int32_limit(tile):
    """One bin of 160,000 pixels: the work list splits K into chunks of 2047 k-blocks
    so every int32 MFMA accumulator stays below 2^31 (values up to 255)."""
    _announce("dense k-chunk split tile=%d" % tile)
    rng = np.random.default_rng(5)
    f = rng.integers(0, 256, size=(24, 400, 400)).astype(np.uint8)
    c, g2 = _run(f, None, extra_options={"tile": tile})
    _check_exact(c, g2, og.dense_g2(f, None))


def test_dense_reference_goldens(golden):
    """Against the reference py_dense_correlator outputs (uint8 / int8 cases), within
    the reference's own tolerance (test_dense.py:86) and our FP64 bound."""
    for name in ["u8_small", "u8_big_values", "i8_sparse", "u8_onepix"]:
        _announce("golden " + name)
        f = golden["dense_%s_frames" % name]
        q = golden["dense_%s_qmask" % name]
        res = golden["dense_%s_res" % name]
        c, g2 = _run(f, q)
        fin = np.isfinite(res)
        assert np.array_equal(fin, np.isfinite(g2))
        assert np.abs(g2[fin] - res[fin]).max(initial=0) < 5e-3
        _check_exact(c, g2, og.dense_g2(f, q))


@pytest.mark.parametrize("dtype", [np.uint8, np.int8])
def test_dense_band_gather_and_uncovered_kblocks(dtype):
    """A 1030-wide detector makes 8-row bands; 40 interleaved bins give k-blocks that
    span one, two or more rows, so both the band workgroups and the general-gather
    workgroups of the launch run; 37 frames end inside an 8-frame stage."""
    _announce("dense band gather %s" % np.dtype(dtype).name)
    rng = np.random.default_rng(37)
    shape, n = (70, 1030), 37
    f = _speckle(37, n, shape, 0.8)
    if dtype == np.int8:
        f = (f.astype(np.int16) - 2).clip(-128, 127).astype(np.int8)
    q = rng.integers(0, 41, size=shape).astype(np.int32)
    for b in range(1, 41):
        q.flat[b] = b
    c, g2 = _run(f, q)
    _check_exact(c, g2, og.dense_g2(f, q))


def test_dense_unaligned_frames_and_wide_detector():
    """Frames whose device address is not 16-byte aligned take the general gather; a
    detector too wide for an 8-frame band stage has no band work.  Both match the
    oracle."""
    import torch
    _announce("dense unaligned / wide")
    from dynamix_amd.correlator.dense import MatMulCorrelator
    shape, n = (30, 1030), 21
    f = _speckle(21, n, shape, 0.5)
    q = np.random.default_rng(21).integers(0, 7, size=shape).astype(np.int32)
    buf = torch.zeros(f.size + 1, dtype=torch.uint8, device="cuda")
    buf[1:] = torch.from_numpy(f.ravel()).cuda()
    view = buf[1:].view(n, *shape)
    assert view.data_ptr() % 16 != 0
    c = MatMulCorrelator(shape, n, q)
    g2 = c.correlate(view)
    _check_exact(c, g2, og.dense_g2(f, q))
    wide, nw = (3, 6000), 19
    fw = _speckle(19, nw, wide, 0.5)
    qw = np.random.default_rng(19).integers(0, 4, size=wide).astype(np.int32)
    c, g2 = _run(fw, qw)
    _check_exact(c, g2, og.dense_g2(fw, qw))


def test_dense_device_input_and_repeat_calls():
    """A GPU tensor is borrowed; repeated calls give identical results (sums and
    numerators are re-zeroed each call, unlike the event reference's res_int)."""
    import torch
    _announce("dense device input / repeat")
    f = _speckle(6, 200, (20, 24), 0.2)
    q = np.random.default_rng(6).integers(0, 5, size=(20, 24)).astype(np.int32)
    from dynamix_amd.correlator.dense import MatMulCorrelator
    c = MatMulCorrelator((20, 24), 200, q)
    g_a = c.correlate(torch.from_numpy(f).cuda())
    n_a = c.last_numerators.clone()
    g_b = c.correlate(f)
    assert torch.equal(n_a, c.last_numerators)
    np.testing.assert_array_equal(g_a, g_b)


def test_dense_correlator_class_matches():
    _announce("DenseCorrelator class")
    from dynamix_amd.correlator.dense import DenseCorrelator
    f = _speckle(7, 90, (10, 12), 0.4)
    q = np.random.default_rng(7).integers(0, 3, size=(10, 12)).astype(np.int32)
    c = DenseCorrelator((10, 12), 90, q, dtype=np.uint8, profile=True)
    g2 = c.correlate(f)
    _check_exact(c, g2, og.dense_g2(f, q))
    st = c.stage_times()
    assert len(st) == 5 and st[4] == 1 and st[1] > 0
    with pytest.raises(AssertionError):
        c.correlate(f.astype(np.int8))


def test_plain_fp32_vs_oracle():
    _announce("plain fp32")
    from dynamix_amd.correlator.dense import PlainMatMulCorrelator
    f = _speckle(8, 300, (20, 30), 0.3)
    q = np.random.default_rng(8).integers(0, 4, size=(20, 30)).astype(np.int32)
    c = PlainMatMulCorrelator((20, 30), 300, q)
    g2 = c.correlate(f)
    ref = og.dense_g2(f, q)["g2"]
    fin = np.isfinite(ref)
    assert np.array_equal(fin, np.isfinite(g2))
    np.testing.assert_allclose(g2[fin], ref[fin], rtol=2e-4, atol=1e-6)


def test_device_speckle_generator_statistics():
    """On-device generator: mean count and g2(tau) ~ 1 + exp(-2 tau/tau_c) (Siegert)."""
    _announce("device speckle")
    import torch
    from dynamix_amd.synth import speckle_frames_device
    from dynamix_amd.correlator.dense import MatMulCorrelator
    f = speckle_frames_device((64, 64), 2000, 0.5, 20.0, seed=11)
    m = f.float().mean().item()
    assert 0.45 < m < 0.55, m
    c = MatMulCorrelator((64, 64), 2000, None)
    g2 = c.correlate(f)[0]
    for tau in (5, 10, 20):
        assert abs(g2[tau] - (1 + np.exp(-2 * tau / 20.0))) < 0.08, (tau, g2[tau])
    ref = og.dense_g2(f.cpu().numpy(), None)
    np.testing.assert_array_equal(c.last_numerators.cpu().numpy(), ref["num"])
