"""GPU FFT correlator vs the exact oracle, within the stated FP32 FFT tolerance, over
every transform length of the mixed-radix set and the uint8 / int8 / uint16 frames."""
import sys

import numpy as np
import pytest

from oracle import g2 as og

pytestmark = pytest.mark.gpu

NUM_TOL = 2e-7   # |Num_fft - Num_exact| / Num(0): FP32 FFT rounding of counting data
G2_RTOL = 1e-6   # g2 rounding on top of the numerator error carried through Den(tau)


def _announce(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def _qmask(n, shape, nbins):
    q = np.random.default_rng(n).integers(0, nbins + 1, size=shape).astype(np.int32)
    for b in range(1, nbins + 1):
        q.flat[b] = b
    return q


def _sparse_ref(f, q):
    """The oracle through the event pairs (long, sparse stacks): every non-zero sample,
    negative int8 values included, as a pixel-major event list."""
    from oracle import events as ev
    n = f.shape[0]
    per_pixel = np.moveaxis(f, 0, -1).reshape(-1, n)
    pix, t = np.nonzero(per_pixel != 0)
    d = per_pixel[pix, t]
    o = np.zeros(per_pixel.shape[0] + 1, np.int64)
    o[1:] = np.cumsum(np.bincount(pix, minlength=per_pixel.shape[0]))
    _, bins = og.qmask_bins(q.shape, q)
    s, num = ev.event_sums_numerators_fast(t, d, o, q, f.shape[0], bins)
    den = og.denominators(s)
    sc = og.scale_factors(q, bins)
    return {"sums": s, "num": num, "den": den, "scale": sc, "g2": og.g2_fp64(num, den, sc)}


def _check(c, g2, f, q, ref=None):
    """Exact sums and denominators; numerators within NUM_TOL * Num(0) per bin; g2 at
    every lag within the bound that numerator error gives through Den(tau)."""
    ref = ref or og.dense_g2(f, q)
    np.testing.assert_array_equal(c.last_sums.cpu().numpy(), ref["sums"])
    den = np.asarray(ref["den"], np.float64)
    np.testing.assert_array_equal(c.last_denominators.cpu().numpy(), den)
    num = c.last_numerators.cpu().numpy()
    num0 = np.maximum(np.abs(ref["num"][:, :1]).astype(np.float64), 1.0)
    err = np.abs(num - ref["num"]) / num0
    print("max |dNum|/Num(0) = %.3e" % err.max())
    assert err.max() <= NUM_TOL
    g = ref["g2"]
    fin = np.isfinite(g)
    assert np.array_equal(fin, np.isfinite(g2))
    with np.errstate(divide="ignore", invalid="ignore"):
        bound = ref["scale"][:, None] * NUM_TOL * num0 / np.abs(den) + G2_RTOL * np.abs(g)
    assert (np.abs(g2[fin] - g[fin]) <= bound[fin]).all()


# (frames, half-length L the plan uses)
LENGTHS = [(7, 1024), (1000, 1024), (1100, 1280), (1500, 2048), (2300, 2560), (4000, 4096),
           (5000, 5120), (8000, 8192), (10000, 10240), (12000, 16384)]


@pytest.mark.parametrize("n,L", LENGTHS)
def test_fft_u8_every_length(n, L):
    from dynamix_amd.correlator.dense import FFTCorrelator, fft_length
    from dynamix_amd.synth import speckle_frames
    _announce("fft u8 n=%d L=%d" % (n, L))
    assert fft_length(n) == 2 * L
    shape = (6, 7) if n > 3000 else (12, 10)
    f = speckle_frames(shape, n, 0.3, 30.0, n, np.uint8)
    q = _qmask(n, shape, 3)
    c = FFTCorrelator(shape, n, q)
    _check(c, c.correlate(f), f, q)


@pytest.mark.parametrize("n", [300, 3000, 10000])
def test_fft_int8_and_uint16(n):
    from dynamix_amd.correlator.dense import FFTCorrelator
    from dynamix_amd.synth import mean_map, speckle_frames
    _announce("fft i8/u16 n=%d" % n)
    shape = (8, 9)
    rng = np.random.default_rng(n)
    q = _qmask(n, shape, 2)
    i8 = rng.integers(-128, 128, size=(n,) + shape).astype(np.int8)
    i8[rng.random(i8.shape) < 0.6] = 0
    c = FFTCorrelator(shape, n, q)
    _check(c, c.correlate(i8), i8, q)
    full = rng.integers(0, 65536, size=(n,) + shape).astype(np.uint16)
    full[rng.random(full.shape) < 0.5] = 0
    _check(c, c.correlate(full), full, q)
    hot = speckle_frames(shape, n, 0.5, 40.0, n, np.uint16,
                         mean_map_=mean_map(shape, 0.5, 0.1, 300.0, n))
    assert hot.max() > 255
    _check(c, c.correlate(hot), hot, q)


@pytest.mark.parametrize("n,L,dtype", [(16385, 20480, "uint8"), (20000, 20480, "uint16"),
                                       (20000, 20480, "int8"), (45000, 65536, "uint8")])
def test_fft_long_series_four_step(n, L, dtype, monkeypatch):
    """N > 16384: the four-step transform through the HBM scratch (several chunks of
    k-blocks with a 1 MiB scratch cap), exact sums, numerators within tolerance."""
    from dynamix_amd.correlator.dense import FFTCorrelator, fft_length
    from dynamix_amd.synth import speckle_frames
    _announce("fft four-step n=%d L=%d %s" % (n, L, dtype))
    monkeypatch.setenv("DXG2_FFT_SCRATCH_MB", "1")   # one k-block per chunk
    assert fft_length(n) == 2 * L
    shape = (9, 16)
    if dtype == "int8":
        rng = np.random.default_rng(n)
        f = (rng.random((n,) + shape) < 0.03) * rng.integers(-5, 6, size=(n,) + shape)
        f = f.astype(np.int8)
    else:
        f = speckle_frames(shape, n, 0.04, 60.0, n, np.dtype(dtype))
        if dtype == "uint16":
            f[::97, 2, 3] = 40000
    q = _qmask(n, shape, 3)
    q[:, :8] = 1          # bin 1 spans several k-blocks
    c = FFTCorrelator(shape, n, q)
    _check(c, c.correlate(f), f, q, _sparse_ref(f, q))


@pytest.mark.parametrize("shape", [(7, 13), (5, 17), (16, 16)])
def test_fft_gather_quads_odd_widths(shape):
    """8-bit frames whose size is not a multiple of 4 bytes: the gather's dword loads of
    consecutive pixel quads sit at every byte alignment, quads break at row and bin
    boundaries, and the last quads of the stack take the byte loads; every bin's series
    must arrive intact (exact sums, numerators within the FFT tolerance)."""
    _announce("fft gather quads shape=%s" % (shape,))
    from dynamix_amd.correlator.dense import FFTCorrelator
    from dynamix_amd.synth import speckle_frames
    n = 300
    f = speckle_frames(shape, n, 2.0, 20.0, 5, np.uint8)
    q = np.ones(shape, np.int32)
    q.flat[::5] = 2                 # runs of four interrupted by another bin
    q.flat[3] = 0                   # a masked pixel inside the first quad
    c = FFTCorrelator(shape, n, q)
    g2 = c.correlate(f)
    _check(c, g2, f, q)


@pytest.mark.parametrize("dtype", [np.uint8, np.int8])
@pytest.mark.parametrize("offset", [0, 1])
def test_fft_aligned_quads_ring_bins(dtype, offset):
    """8-bit frames with 4 | H W and annular bins: the FFT path orders each bin's pixels as
    aligned detector quads first (one 4-byte load per quad and frame in the gather) and
    the rest after.  Sums and denominators stay exact and numerators within the FFT
    tolerance, with the frames 4-aligned (quad loads) and one byte past it (byte loads)."""
    import torch
    _announce("fft aligned quads dtype=%s offset=%d" % (np.dtype(dtype).name, offset))
    from dynamix_amd import synth
    from dynamix_amd.correlator.dense import FFTCorrelator
    shape, n = (40, 64), 700
    f = synth.speckle_frames(shape, n, 2.0, 20.0, 9, np.uint8)
    if dtype == np.int8:
        f = (f.astype(np.int16) - 3).clip(-128, 127).astype(np.int8)
    q = synth.annular_qmask(shape, 6)
    q[5, 7:30] = 0                    # a masked run breaks quads inside a ring
    buf = torch.zeros(f.size + offset, dtype=torch.uint8, device="cuda")
    buf[offset:] = torch.from_numpy(f.ravel().view(np.uint8)).cuda()
    view = buf[offset:].view(torch.int8 if dtype == np.int8 else torch.uint8).view(n, *shape)
    assert (view.data_ptr() % 4 == 0) == (offset == 0)
    c = FFTCorrelator(shape, n, q)
    g2 = c.correlate(view)
    _check(c, g2, f, q)
