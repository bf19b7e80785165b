"""GPU parity of the event (sparse) HIP path against the CPU oracle and the dense path."""
import sys

import numpy as np
import pytest

from oracle import events as ev
from oracle import g2 as og

pytestmark = pytest.mark.gpu


def _announce(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def _sparse_stack(seed, n, shape, p=0.05, hi=4, dtype=np.uint8):
    from dynamix_amd.synth import speckle_frames
    f = speckle_frames(shape, n, p, 15.0, seed, dtype)
    return f


def _check(c, g2, ref):
    np.testing.assert_array_equal(c.last_numerators.cpu().numpy(), ref["num"])
    np.testing.assert_array_equal(c.last_sums.cpu().numpy(), ref["sums"])
    np.testing.assert_array_equal(c.last_denominators.cpu().numpy(),
                                  np.asarray(ref["den"], dtype=np.float64))
    g = ref["g2"]
    fin = np.isfinite(g)
    assert np.array_equal(fin, np.isfinite(g2))
    nz = fin & (g != 0)
    assert (g2[fin & (g == 0)] == 0).all()
    rel = np.abs(g2[nz] - g[nz]) / np.abs(g[nz])
    assert rel.max(initial=0) <= 1e-6


@pytest.mark.parametrize("n,shape,mean,nbins", [
    (50, (6, 7), 0.3, 3),
    (400, (20, 25), 0.05, 5),
    (17000, (4, 5), 0.02, 2),      # several lag windows; pixels past the LDS staging cap
    (300, (60, 80), 0.05, 2),      # bins over 1024 pixels: several work items per bin
])
def test_events_csr_vs_oracle_and_dense(n, shape, mean, nbins):
    _announce("events csr n=%d shape=%s" % (n, shape))
    from dynamix_amd.correlator.event import EventCorrelator
    f = _sparse_stack(n, n, shape, mean)
    q = np.random.default_rng(n).integers(0, nbins + 1, size=shape).astype(np.int32)
    for b in range(1, nbins + 1):
        q.flat[b] = b
    data, times, offsets = ev.compress_all_stack(f)
    c = EventCorrelator(shape, n, max_events_count=int(np.diff(offsets).max()), qmask=q,
                        dtype=np.uint8)
    g2 = c.correlate(times, data, offsets)
    ref = ev.event_g2(times, data, offsets, q, n)
    _check(c, g2, ref)
    dense = og.dense_g2(f, q)
    np.testing.assert_array_equal(c.last_numerators.cpu().numpy(), dense["num"])


def test_events_coo_matches_csr():
    _announce("events coo")
    from dynamix_amd.correlator.event import EventCorrelator
    f = _sparse_stack(3, 300, (12, 14), 0.1)
    q = np.random.default_rng(3).integers(0, 4, size=(12, 14)).astype(np.int32)
    data, times, offsets = ev.compress_all_stack(f)
    pix, fr, cnt = ev.csr_to_coo(times, data, offsets)
    c = EventCorrelator((12, 14), 300, 10, qmask=q, dtype=np.uint8)
    g2 = c.correlate_coo(pix, fr, cnt.astype(np.uint8))
    _check(c, g2, ev.event_g2(times, data, offsets, q, 300))


def test_events_coo_uint16():
    """uint16 counts take the 64-bit packed (frame, count) words through the sort."""
    _announce("events coo uint16")
    from dynamix_amd.correlator.event import EventCorrelator
    rng = np.random.default_rng(11)
    f = (rng.random((200, 9, 11)) < 0.15) * rng.integers(1, 3000, size=(200, 9, 11))
    f = f.astype(np.uint16)
    q = rng.integers(0, 5, size=(9, 11)).astype(np.int32)
    data, times, offsets = ev.compress_all_stack(f)
    pix, fr, cnt = ev.csr_to_coo(times, data, offsets)
    c = EventCorrelator((9, 11), 200, 10, qmask=q, dtype=np.uint16)
    g2 = c.correlate_coo(pix, fr, cnt.astype(np.uint16))
    _check(c, g2, ev.event_g2(times, data, offsets, q, 200))
    dense = og.dense_g2(f, q)
    np.testing.assert_array_equal(c.last_numerators.cpu().numpy(), dense["num"])


def test_events_coo_input_errors():
    """A pixel whose events are not in increasing frame order, or a frame past the
    series, is reported as invalid input (the sort keeps each pixel's input order)."""
    _announce("events coo errors")
    from dynamix_amd.correlator.event import EventCorrelator
    f = _sparse_stack(5, 100, (6, 7), 0.3)
    q = np.ones((6, 7), dtype=np.int32)
    data, times, offsets = ev.compress_all_stack(f)
    pix, fr, cnt = ev.csr_to_coo(times, data, offsets)
    c = EventCorrelator((6, 7), 100, 10, qmask=q, dtype=np.uint8)
    with pytest.raises(ValueError, match="strictly increasing"):
        c.correlate_coo(pix[::-1].copy(), fr[::-1].copy(), cnt[::-1].astype(np.uint8))
    bad = fr.copy()
    bad[len(bad) // 2] = 100
    with pytest.raises(ValueError, match="out of"):
        c.correlate_coo(pix, bad, cnt.astype(np.uint8))
    # the plan recovers: a valid list afterwards gives the reference result
    g2 = c.correlate_coo(pix, fr, cnt.astype(np.uint8))
    _check(c, g2, ev.event_g2(times, data, offsets, q, 100))


def test_events_scale_factor_and_int8_and_goldens(golden):
    _announce("events int8 + scale")
    from dynamix_amd.correlator.event import EventCorrelator
    for name in ["evt_u8", "evt_i8"]:
        f = golden["%s_frames" % name]
        data, times, offsets = golden["%s_data" % name], golden["%s_times" % name], \
            golden["%s_offsets" % name]
        n = f.shape[0]
        c = EventCorrelator(f.shape[1:], n, 10, dtype=f.dtype, scale_factor=1234.0)
        g2 = c.correlate(times, data, offsets)
        ref = ev.event_g2(times, data, offsets, np.ones(f.shape[1:], np.int32), n,
                          scale_factor=1234.0)
        _check(c, g2, ref)


@pytest.mark.parametrize("case", ["u8_staged", "u8_long", "i8_staged"])
def test_events_saturated_pixels_stay_exact(case):
    """Saturated counts on every pixel: 8-bit counts use 32-bit LDS lag counters while a
    wave's sum of d^2 stays within its bound, and send the pairs of later pixels straight
    to the 64-bit numerators.  Pixels of 252 events (the most a wave stages in LDS), 253
    and 256 (read from HBM) in the same waves (uint8 from a dense stack, against the dense
    oracle), pixels of 250, 251 and 252 staged events, so that the staged lists end at
    every alignment of their four-entry steps (int8 event lists of both signs, which the
    compressor would not produce, against the event oracle), and pixels with 4000
    unstaged events (uint8) pass that bound within one wave."""
    _announce("events saturated %s" % case)
    from dynamix_amd.correlator.event import EventCorrelator
    rng = np.random.default_rng(31)
    if case == "i8_staged":
        shape, n = (32, 32), 300
        npx = shape[0] * shape[1]
        counts = np.array([250, 251, 252])[np.arange(npx) % 3]
        times = np.concatenate([np.sort(rng.choice(n, size=k, replace=False))
                                for k in counts]).astype(np.int32)
        data = rng.choice(np.array([-128, 127], dtype=np.int8), size=times.size)
        offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32)
        q = np.ones(shape, np.int32)
        c = EventCorrelator(shape, n, 252, qmask=q, dtype=np.int8)
        g2 = c.correlate(times, data, offsets)
        _check(c, g2, ev.event_g2(times, data, offsets, q, n))
        return
    if case == "u8_long":
        shape, n = (4, 5), 4000
        f = np.full((n,) + shape, 255, dtype=np.uint8)
    else:
        shape, n = (32, 32), 300
        f = np.zeros((n,) + shape, dtype=np.uint8)
        for p in range(shape[0] * shape[1]):
            k = (252, 253, 256)[p % 3]
            f[rng.choice(n, size=k, replace=False), p // shape[1], p % shape[1]] = 255
    q = np.ones(shape, np.int32)
    data, times, offsets = ev.compress_all_stack(f)
    c = EventCorrelator(shape, n, int(np.diff(offsets).max()), qmask=q, dtype=np.uint8)
    g2 = c.correlate(times, data, offsets)
    _check(c, g2, og.dense_g2(f, q))


def test_events_repeat_calls_do_not_accumulate():
    _announce("events repeat")
    from dynamix_amd.correlator.event import EventCorrelator
    f = _sparse_stack(5, 120, (8, 8), 0.2)
    data, times, offsets = ev.compress_all_stack(f)
    c = EventCorrelator((8, 8), 120, 20, dtype=np.uint8)
    a = c.correlate(times, data, offsets)
    n1 = c.last_numerators.clone()
    b = c.correlate(times, data, offsets)
    assert (n1 == c.last_numerators).all().item()
    np.testing.assert_array_equal(a, b)


def test_events_invalid_input_raises():
    _announce("events invalid input")
    from dynamix_amd.correlator.event import EventCorrelator
    f = _sparse_stack(6, 60, (5, 5), 0.3)
    data, times, offsets = ev.compress_all_stack(f)
    c = EventCorrelator((5, 5), 60, 20, dtype=np.uint8, allow_reallocate=False,
                        total_events_count=data.size)
    bad = times.copy()
    k = int(offsets[1])  # first event of pixel 1: make pixel 0's last frame repeat
    if k >= 2:
        bad[k - 1] = bad[k - 2]
        with pytest.raises(ValueError, match="increasing"):
            c.correlate(bad, data, offsets)
    bad = times.copy()
    bad[0] = 60          # frame index out of range
    with pytest.raises(ValueError):
        c.correlate(bad, data, offsets)
    with pytest.raises(ValueError, match="allow_reallocate"):
        c.correlate(np.concatenate([times, times]), np.concatenate([data, data]), offsets)
    with pytest.raises(AssertionError):
        c.correlate(times, data, offsets[:-1])
    # a valid call after the errors is exact again
    g2 = c.correlate(times, data, offsets)
    _check(c, g2, ev.event_g2(times, data, offsets, np.ones((5, 5), np.int32), 60))


def test_device_event_generator_matches_dense_compaction():
    """dxg2_synth_events emits exactly the non-zero samples of dxg2_synth_speckle (same
    seed); the dense and event paths then give identical numerators."""
    _announce("device events vs dense")
    from dynamix_amd.correlator.dense import MatMulCorrelator
    from dynamix_amd.correlator.event import EventCorrelator
    from dynamix_amd.synth import (events_to_coo_device, speckle_events_device,
                                   speckle_frames_device)
    shape, n = (24, 20), 700
    f = speckle_frames_device(shape, n, 0.05, 40.0, seed=99)
    t, d, o = speckle_events_device(shape, n, 0.05, 40.0, seed=99)
    data, times, offsets = ev.compress_all_stack(f.cpu().numpy())
    np.testing.assert_array_equal(o.cpu().numpy().astype(np.uint32), offsets)
    np.testing.assert_array_equal(t.cpu().numpy(), times)
    np.testing.assert_array_equal(d.cpu().numpy(), data)
    q = np.random.default_rng(9).integers(0, 5, size=shape).astype(np.int32)
    cd = MatMulCorrelator(shape, n, q)
    cd.correlate(f)
    ce = EventCorrelator(shape, n, 50, qmask=q, dtype=np.uint8)
    ce.correlate(t, d, o)
    assert (cd.last_numerators == ce.last_numerators).all().item()
    pix, fr, cnt = events_to_coo_device(t, d, o, shape, n)
    assert (fr[1:] >= fr[:-1]).all().item()
    ce.correlate_coo(pix, fr, cnt)
    assert (cd.last_numerators == ce.last_numerators).all().item()


@pytest.mark.parametrize("dtype", [np.uint8, np.int8])
def test_events_lags_at_window_edges(dtype):
    """Pairs whose lags sit on and around the lag-window boundaries (multiples of the
    window length, for 4-, 8- and 16-wave windows of 2048 lags per wave less 64), one
    event apart in every window, over 70,000 frames: exact against the Python pair
    loop."""
    _announce("events window edges %s" % np.dtype(dtype).name)
    from dynamix_amd.correlator.event import EventCorrelator
    n, shape = 70000, (3, 4)
    edges = sorted({k * w + d for w in (8192 - 64, 16384 - 64, 32768 - 64, 8192, 32768)
                    for k in (1, 2) for d in (-2, -1, 0, 1, 2) if 0 < k * w + d < n - 1})
    rng = np.random.default_rng(9)
    lo = -3 if dtype == np.int8 else 1
    times, data, offsets = [], [], [0]
    for p in range(shape[0] * shape[1]):
        t0 = int(rng.integers(0, 200))
        ts = sorted({t0} | {t0 + e for e in edges if t0 + e < n} |
                    set(int(x) for x in rng.integers(0, n, size=40)))
        times += ts
        data += [int(v) if v != 0 else 1 for v in rng.integers(lo, 4, size=len(ts))]
        offsets.append(len(times))
    times = np.array(times, np.int32)
    data = np.array(data, dtype)
    offsets = np.array(offsets, np.uint32)
    q = np.array([[1, 1, 2, 2], [1, 2, 2, 0], [2, 1, 1, 2]], np.int32)
    c = EventCorrelator(shape, n, max_events_count=int(np.diff(offsets).max()), qmask=q,
                        dtype=dtype)
    c.correlate(times, data, offsets)
    num, sums = ev.event_brute_force(times, data, offsets, q, n)
    np.testing.assert_array_equal(c.last_numerators.cpu().numpy(), num.astype(np.int64))
    np.testing.assert_array_equal(c.last_sums.cpu().numpy(), sums.astype(np.int64))
