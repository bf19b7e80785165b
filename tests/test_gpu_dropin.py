"""Drop-in surface on the GPU: the reference's default dtype "f" and wider frame types
run on the exact integer path with results identical to the native-type call; the
reference's CUDA class names work.  Marked gpu."""
import sys

import numpy as np
import pytest

from oracle import events as ev
from oracle import g2 as og

pytestmark = pytest.mark.gpu


def _announce(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def _stack(seed, n=130, shape=(12, 14), mean=0.8):
    from dynamix_amd.synth import speckle_frames
    return speckle_frames(shape, n, mean, 15.0, seed, np.uint8)


def _qmask(shape, nbins=3, seed=0):
    q = np.random.default_rng(seed).integers(0, nbins + 1, size=shape).astype(np.int32)
    for b in range(1, nbins + 1):
        q.flat[b] = b
    return q


@pytest.mark.parametrize("dt", ["float32", "float64", "int16", "int32", "uint32"])
def test_matmul_any_numeric_dtype_equals_native(dt):
    import torch
    from dynamix_amd.correlator.dense import MatMulCorrelator
    _announce("matmul dtype %s" % dt)
    f = _stack(3)
    q = _qmask(f.shape[1:])
    c = MatMulCorrelator(f.shape[1:], f.shape[0], q)
    ref = c.correlate(f)
    ref_num = c.last_numerators.cpu().numpy()
    for arr in (f.astype(dt), torch.from_numpy(f.astype(dt)).cuda()):
        g = c.correlate(arr)
        np.testing.assert_array_equal(g.view(np.uint32), ref.view(np.uint32))
        np.testing.assert_array_equal(c.last_numerators.cpu().numpy(), ref_num)


def test_wide_and_signed_values_take_uint16_and_int8():
    from dynamix_amd.correlator.dense import MatMulCorrelator
    _announce("matmul float32 wide / signed")
    f = _stack(4).astype(np.int32)
    q = _qmask(f.shape[1:])
    wide = f * 97                       # up to ~65535: the uint16 digit planes
    assert wide.max() > 255 and wide.max() <= 65535
    c = MatMulCorrelator(f.shape[1:], f.shape[0], q)
    c.correlate(wide.astype(np.float32))
    np.testing.assert_array_equal(c.last_numerators.cpu().numpy(), og.dense_g2(wide, q)["num"])
    signed = (f % 7) - 3                # in [-3, 3]: int8
    c.correlate(signed.astype(np.float32))
    np.testing.assert_array_equal(c.last_numerators.cpu().numpy(),
                                  og.dense_g2(signed.astype(np.int8), q)["num"])
    with pytest.raises(ValueError):
        c.correlate(f.astype(np.float32) + 0.5)
    with pytest.raises(ValueError):
        c.correlate(np.full(f.shape, 70000, np.int32))


def test_dense_correlator_default_dtype_f():
    from dynamix_amd.correlator.dense import DenseCorrelator
    _announce("DenseCorrelator dtype f")
    f = _stack(5)
    q = _qmask(f.shape[1:])
    c = DenseCorrelator(f.shape[1:], f.shape[0], q)
    assert c.dtype == np.float32
    g = c.correlate(f.astype(np.float32))
    ref = og.dense_g2(f, q)
    np.testing.assert_array_equal(c.last_numerators.cpu().numpy(), ref["num"])
    fin = np.isfinite(ref["g2"])
    assert (np.abs(g[fin] - ref["g2"][fin]) <= 1e-6 * np.abs(ref["g2"][fin])).all()
    with pytest.raises(AssertionError):
        c.correlate(f)                  # uint8 frames for a dtype "f" correlator


def test_event_correlator_default_dtype_f():
    from dynamix_amd.correlator.event import EventCorrelator
    _announce("EventCorrelator dtype f")
    f = _stack(6, mean=0.1)
    q = _qmask(f.shape[1:])
    d, t, o = ev.compress_all_stack(f)
    c = EventCorrelator(f.shape[1:], f.shape[0], max_events_count=int(np.diff(o).max()), qmask=q)
    g = c.correlate(t, d.astype(np.float32), o)
    ref = ev.event_g2(t, d, o, q, f.shape[0])
    np.testing.assert_array_equal(c.last_numerators.cpu().numpy(), ref["num"])
    c8 = EventCorrelator(f.shape[1:], f.shape[0], max_events_count=1, qmask=q, dtype=np.uint8)
    g8 = c8.correlate(t, d, o)
    np.testing.assert_array_equal(g.view(np.uint32), g8.view(np.uint32))
    pix, fr, cnt = ev.csr_to_coo(t, d, o)
    g_coo = c.correlate_coo(pix, fr, cnt.astype(np.float32))
    np.testing.assert_array_equal(g_coo.view(np.uint32), g8.view(np.uint32))


def test_reference_cuda_class_names():
    from dynamix_amd.correlator.cuda import CublasMatMulCorrelator, CUFFTCorrelator
    from dynamix_amd.correlator.dense import FFTCorrelator, MatMulCorrelator
    _announce("cuda class names")
    f = _stack(7, n=200)
    q = _qmask(f.shape[1:])
    a = CublasMatMulCorrelator(f.shape[1:], f.shape[0], q).correlate(f)
    b = MatMulCorrelator(f.shape[1:], f.shape[0], q).correlate(f)
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    c = CUFFTCorrelator(f.shape[1:], f.shape[0], q)
    x = c.correlate(f)
    y = FFTCorrelator(f.shape[1:], f.shape[0], q).correlate(f)
    np.testing.assert_array_equal(x.view(np.uint32), y.view(np.uint32))
