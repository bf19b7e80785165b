"""Two-time correlation output on the GPU against the FP64 oracle (which a loop
restatement of the reference's event_y code pins).  Marked gpu."""
import sys

import numpy as np
import pytest

from oracle import g2 as og

pytestmark = pytest.mark.gpu

TRC_RTOL = 1e-6   # float32 rounding of an FP64 ratio of exact integers


def _announce(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def _check(trc, ref):
    assert trc.dtype == np.float32 and trc.shape == ref.shape
    assert np.array_equal(np.isnan(trc), np.isnan(ref)), "NaN pattern differs"
    m = ~np.isnan(ref)
    err = np.abs(trc[m].astype(np.float64) - ref[m])
    assert (err <= TRC_RTOL * np.abs(ref[m]) + 1e-30).all(), err.max()


def _bins(rng, shape, nbins):
    q = rng.integers(0, nbins + 1, size=shape).astype(np.int32)
    for b in range(1, nbins + 1):
        q.flat[b] = b
    return q


@pytest.mark.parametrize("tile", [256, 128])
@pytest.mark.parametrize("n,shape,mean", [(9, (5, 6), 3.0), (300, (16, 20), 1.0),
                                          (520, (24, 30), 0.8)])
def test_two_time_u8_vs_oracle(n, shape, mean, tile):
    from dynamix_amd.correlator.dense import MatMulCorrelator
    from dynamix_amd.synth import speckle_frames
    _announce("ttcf u8 n=%d tile=%d" % (n, tile))
    f = speckle_frames(shape, n, mean, 20.0, n, np.uint8)
    q = _bins(np.random.default_rng(n), shape, 3)
    c = MatMulCorrelator(shape, n, q, extra_options={"tile": tile})
    for b in (1, 3):
        _check(c.two_time(f, b), og.two_time_fp64(f, q, b))


def test_two_time_u16_i8_zero_frames_empty_bin(monkeypatch):
    from dynamix_amd.correlator.dense import MatMulCorrelator
    _announce("ttcf u16 / i8 / zero frames / empty bin / chunks")
    monkeypatch.setenv("DXG2_GRAM_KCHUNK", "2")
    monkeypatch.setenv("DXG2_STD_SCRATCH_MB", "1")
    rng = np.random.default_rng(5)
    g = rng.integers(0, 5000, size=(280, 12, 14)).astype(np.uint16)
    g[rng.random(g.shape) < 0.4] = 0
    q = _bins(rng, (12, 14), 2)
    q[0, 0] = 4  # bin 3 empty
    g[7][q == 1] = 0     # a frame with no counts in bin 1
    g[-1][q == 2] = 0    # an empty last frame in bin 2
    c = MatMulCorrelator((12, 14), 280, q)
    for b in (1, 2, 3, 4):
        _check(c.two_time(g, b), og.two_time_fp64(g, q, b))
    f = rng.integers(-10, 40, size=(150, 9, 9)).astype(np.int8)
    q8 = _bins(rng, (9, 9), 2)
    c8 = MatMulCorrelator((9, 9), 150, q8)
    _check(c8.two_time(f, 2), og.two_time_fp64(f, q8, 2))


def test_two_time_bad_bin():
    from dynamix_amd.correlator.dense import MatMulCorrelator
    f = np.ones((10, 4, 4), np.uint8)
    c = MatMulCorrelator((4, 4), 10, None)
    with pytest.raises(ValueError):
        c.two_time(f, 2)


@pytest.mark.parametrize("dtype", ["uint8", "uint16"])
def test_two_time_from_events_equals_dense_and_oracle(dtype):
    """EventCorrelator.two_time on the pixel-major events of a stack equals the dense
    two-time matrix of the same stack bit for bit, and the oracle within float32
    rounding; a bin with no pixels is all NaN."""
    from dynamix_amd.correlator.dense import MatMulCorrelator
    from dynamix_amd.correlator.event import EventCorrelator
    from dynamix_amd.synth import speckle_frames
    from oracle import events as ev
    _announce("ttcf from events %s" % dtype)
    shape, n = (14, 16), 300
    f = speckle_frames(shape, n, 0.4, 20.0, 11, np.dtype(dtype))
    if dtype == "uint16":
        f[5, 3, 4] = 4000
    q = _bins(np.random.default_rng(2), shape, 3)
    q[q == 2] = 4      # bin 2 has no pixels
    d, t, o = ev.compress_all_stack(f)
    ce = EventCorrelator(shape, n, max_events_count=1, qmask=q, dtype=np.dtype(dtype))
    cd = MatMulCorrelator(shape, n, q)
    for b in (1, 3, 4):
        te = ce.two_time(t, d, o, b)
        td = cd.two_time(f, b)
        np.testing.assert_array_equal(te.view(np.uint32), td.view(np.uint32))
        _check(te, og.two_time_fp64(f, q, b))
    assert np.isnan(ce.two_time(t, d, o, 2)).all()
    with pytest.raises(ValueError):
        ce.two_time(t, d, o, 5)
