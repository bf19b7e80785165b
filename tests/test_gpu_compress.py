"""On-device frames -> events compaction (FramesCompressor) against the reference's own
outputs (golden vectors) and the oracle.  Marked gpu."""
import sys

import numpy as np
import pytest

from oracle import events as ev

pytestmark = pytest.mark.gpu


def _announce(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", ["evt_u8", "evt_i8"])
def test_compress_matches_reference_goldens(golden, name):
    from dynamix_amd.correlator.event import FramesCompressor
    _announce("compress golden " + name)
    frames = golden["%s_frames" % name]
    d, t, o = FramesCompressor.compress_all_stack(frames)
    np.testing.assert_array_equal(d, golden["%s_data" % name])
    np.testing.assert_array_equal(t, golden["%s_times" % name])
    np.testing.assert_array_equal(o, golden["%s_offsets" % name])
    assert d.dtype == frames.dtype and t.dtype == np.int32 and o.dtype == np.uint32
    fc = FramesCompressor(frames.shape[1:], frames.shape[0],
                          int((frames > 0).sum(axis=0).max()), frames.dtype)
    for f in frames:
        fc.process_frame(f)
    d2, t2, o2 = fc.get_compacted_events()
    np.testing.assert_array_equal(d2, golden["%s_prog_data" % name])
    np.testing.assert_array_equal(t2, golden["%s_prog_times" % name])
    np.testing.assert_array_equal(o2, golden["%s_prog_offsets" % name])


@pytest.mark.parametrize("shape,n,dtype", [
    ((5, 7), 3, np.uint8),        # odd frame size: byte-wide loads
    ((12, 10), 130, np.uint8),    # 4-byte words, 3 frame chunks
    ((16, 32), 700, np.int8),     # 16-byte words, signed values (negatives dropped)
    ((9, 13), 200, np.uint16),    # 2-byte elements, odd size
    ((8, 16), 4500, np.uint16),   # 64 frame chunks
])
def test_compress_all_stack_vs_oracle(shape, n, dtype):
    import torch
    from dynamix_amd.correlator.event import FramesCompressor
    _announce("compress %s n=%d %s" % (shape, n, np.dtype(dtype).name))
    rng = np.random.default_rng(n)
    if dtype == np.int8:
        f = rng.integers(-3, 4, size=(n,) + shape).astype(np.int8)
    else:
        f = (rng.random((n,) + shape) < 0.2) * rng.integers(1, 400 if dtype == np.uint16 else 250,
                                                           size=(n,) + shape)
        f = f.astype(dtype)
    ref = ev.compress_all_stack(f)
    got = FramesCompressor.compress_all_stack(f)
    for a, b in zip(got, ref):
        np.testing.assert_array_equal(a, b)
    # device in, device out; a frames view that starts at an odd byte offset
    big = torch.from_numpy(np.concatenate([f[:1], f]).view(
        np.int16 if dtype == np.uint16 else dtype)).cuda()
    dd, tt, oo = FramesCompressor.compress_all_stack(big[1:])
    np.testing.assert_array_equal(tt.cpu().numpy(), ref[1])
    np.testing.assert_array_equal(oo.cpu().numpy().view(np.uint32), ref[2])


def test_compressed_events_feed_the_event_correlator():
    """Device compaction straight into EventCorrelator equals the dense correlator."""
    import torch
    from dynamix_amd.correlator.dense import MatMulCorrelator
    from dynamix_amd.correlator.event import EventCorrelator, FramesCompressor
    from dynamix_amd.synth import speckle_frames
    _announce("compress -> events")
    f = speckle_frames((24, 30), 500, 0.05, 20.0, 3, np.uint8)
    q = np.random.default_rng(3).integers(0, 4, size=(24, 30)).astype(np.int32)
    d, t, o = FramesCompressor.compress_all_stack(torch.from_numpy(f).cuda())
    ec = EventCorrelator((24, 30), 500, int(torch.diff(o).max().item()) + 1, qmask=q,
                         dtype=np.uint8)
    g_ev = ec.correlate(t, d, o)
    dc = MatMulCorrelator((24, 30), 500, q)
    g_dn = dc.correlate(f)
    np.testing.assert_array_equal(np.isfinite(g_ev), np.isfinite(g_dn))
    fin = np.isfinite(g_dn)
    np.testing.assert_allclose(g_ev[fin], g_dn[fin], rtol=1e-6)


def test_streaming_overflow_raises():
    from dynamix_amd.correlator.event import FramesCompressor
    f = np.ones((5, 4, 4), np.uint8)
    fc = FramesCompressor((4, 4), 5, 3, np.uint8)
    for x in f:
        fc.process_frame(x)
    with pytest.raises(IndexError):
        fc.get_compacted_events()
