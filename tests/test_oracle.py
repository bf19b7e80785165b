"""Pins the CPU oracle: golden vectors of the reference, brute force, closed forms."""
import numpy as np
import pytest

from oracle import events as ev
from oracle import g2 as og

DENSE_CASES = ["u8_small", "u8_big_values", "i8_sparse", "u16_mixed", "u8_onepix"]


def _case(golden, name):
    return (golden["dense_%s_frames" % name], golden["dense_%s_qmask" % name],
            golden["dense_%s_res" % name], golden["dense_%s_dev" % name])


# ------------------------------------------------------------ reference goldens
@pytest.mark.parametrize("name", DENSE_CASES)
def test_matmul_f32_bit_exact_vs_reference(golden, name):
    frames, q, res, _ = _case(golden, name)
    with np.errstate(all="ignore"):
        mine = og.matmul_correlator_f32(frames, q)
    np.testing.assert_array_equal(mine, res)


@pytest.mark.parametrize("name", DENSE_CASES)
def test_matmul_f32_calc_std_bit_exact_vs_reference(golden, name):
    frames, q, res, dev = _case(golden, name)
    mine, mdev = og.matmul_correlator_f32(frames, q, calc_std=True)
    np.testing.assert_array_equal(mine, res)
    np.testing.assert_array_equal(mdev, dev)


@pytest.mark.parametrize("name", DENSE_CASES)
def test_exact_dev_matches_reference_within_f32(golden, name):
    """The FP64 dev from exact integers and the reference's float32 dev agree to the
    float32 rounding of the ratios they are the spread of."""
    frames, q, _, dev = _case(golden, name)
    d = og.dense_dev_fp64(frames, q)
    fin = np.isfinite(dev)
    assert np.array_equal(fin, np.isfinite(d))
    # the std of ratios r_t ~ 1 loses the float32 rounding of each r_t in absolute terms
    assert np.abs(d[fin] - dev[fin]).max() < 1e-5 * max(1.0, np.abs(dev[fin]).max())


def test_dev_brute_force_and_closed_forms():
    f, q = _tiny(21, n=9, hi=7)
    d = og.dense_dev_fp64(f, q)
    _, bins = og.qmask_bins(q.shape, q)
    for i, b in enumerate(bins):
        pix = [(y, x) for y in range(q.shape[0]) for x in range(q.shape[1]) if q[y, x] == b]
        n = f.shape[0]
        for tau in range(n):
            r = []
            for t in range(n - tau):
                g = sum(int(f[t, y, x]) * int(f[t + tau, y, x]) for y, x in pix)
                s1 = sum(int(f[t, y, x]) for y, x in pix)
                s2 = sum(int(f[t + tau, y, x]) for y, x in pix)
                r.append(len(pix) * g / (s1 * s2) if s1 * s2 else float("nan"))
            if any(v != v for v in r):
                assert np.isnan(d[i, tau])
                continue
            mu = sum(r) / len(r)
            var = sum((v - mu) ** 2 for v in r) / len(r)
            assert abs(d[i, tau] - (var ** 0.5) / len(r) ** 0.5) < 1e-12
    # last lag: one ratio, zero spread; frames identical up to a factor: r_t constant
    assert np.all(d[np.isfinite(d[:, -1]), -1] == 0.0)
    c = np.random.default_rng(2).integers(1, 9, size=(4, 5)).astype(np.uint8)
    fc = np.stack([c * k for k in (1, 2, 3, 1, 2, 3)]).astype(np.uint8)
    qq = np.ones((4, 5), np.int32)
    assert np.abs(og.dense_dev_fp64(fc, qq)).max() < 1e-15


@pytest.mark.parametrize("seed", range(3))
def test_two_time_vs_reference_loops(seed):
    rng = np.random.default_rng(seed)
    f = (rng.random((9, 4, 5)) < 0.5) * rng.integers(1, 6, size=(9, 4, 5))
    f = f.astype(np.uint8)
    q = rng.integers(0, 3, size=(4, 5)).astype(np.int32)
    q.flat[0], q.flat[1] = 1, 2
    for b in (1, 2):
        a = og.two_time_fp64(f, q, b)
        r = og.two_time_loops(f, q, b)
        assert np.array_equal(np.isnan(a), np.isnan(r))
        m = ~np.isnan(r)
        np.testing.assert_allclose(a[m], r[m], rtol=1e-13)


def test_two_time_closed_forms():
    rng = np.random.default_rng(4)
    c = rng.integers(1, 9, size=(5, 6)).astype(np.uint8)
    f = np.broadcast_to(c, (12, 5, 6)).copy()
    q = np.ones((5, 6), np.int32)
    trc = og.two_time_fp64(f, q, 1)
    cb = c.astype(np.float64).ravel()
    v = cb.size * (cb ** 2).sum() / cb.sum() ** 2
    off = ~np.eye(12, dtype=bool)
    np.testing.assert_allclose(trc[off], v, rtol=1e-14)
    np.testing.assert_allclose(np.diagonal(trc)[:-1], v, rtol=1e-14)
    # the corner keeps the photon pairs within the last frame: n_b sum c(c-1) / (sum c)^2
    np.testing.assert_allclose(trc[-1, -1], cb.size * (cb * (cb - 1)).sum() / cb.sum() ** 2,
                               rtol=1e-14)
    # a last frame with no counts: its row, column and corner are NaN (0/0); fecorrt
    # symmetrises the pair counts before the normalisation, so both triangles are NaN
    f0 = f.copy()
    f0[-1] = 0
    t0 = og.two_time_fp64(f0, q, 1)
    assert np.isnan(t0[-1]).all() and np.isnan(t0[:, -1]).all()
    assert np.isfinite(t0[:-1, :-1]).all()
    g = rng.integers(0, 4, size=(15, 5, 6)).astype(np.uint8)
    t2 = og.two_time_fp64(g, q, 1)
    assert np.array_equal(np.isnan(t2), np.isnan(t2.T))
    m = ~np.isnan(t2)
    np.testing.assert_array_equal(t2[m], t2.T[m])


@pytest.mark.parametrize("name", DENSE_CASES)
def test_exact_g2_matches_reference_within_f32(golden, name):
    """The exact FP64 g2 and the reference's float32 g2 agree to float32 rounding."""
    frames, q, res, _ = _case(golden, name)
    out = og.dense_g2(frames, q)
    g = out["g2"]
    fin = np.isfinite(res)
    assert np.array_equal(fin, np.isfinite(g))
    rel = np.abs(g[fin] - res[fin]) / np.maximum(np.abs(g[fin]), 1e-30)
    assert rel.max() < 2e-5
    # the reference's own test tolerance (test_dense.py:86)
    assert np.abs(g[fin] - res[fin]).max() < 5e-3


def test_qmask_and_scale_conventions(golden):
    q = golden["base_qmask"]
    qq, bins = og.qmask_bins(q.shape, q)
    np.testing.assert_array_equal(bins, golden["base_bins"])
    assert len(bins) == int(golden["base_n_bins"])
    np.testing.assert_array_equal(og.scale_factors(qq, bins), golden["base_scale"])
    np.testing.assert_array_equal(og.scale_factors(qq, bins, [10.0, 20.0, 30.0]),
                                  golden["base_scale_given"])
    qn, bn = og.qmask_bins((3, 4), None)
    np.testing.assert_array_equal(bn, golden["base_none_bins"])
    np.testing.assert_array_equal(qn, golden["base_none_qmask"])
    np.testing.assert_array_equal(og.scale_factors(qn, bn), golden["base_none_scale"])
    qs, bs = og.qmask_bins(4, None)
    assert qs.shape == tuple(golden["base_scalar_shape"])
    np.testing.assert_array_equal(og.scale_factors(qs, bs, 7.5), golden["base_scalar_scale"])


def test_fft_length_rule(golden):
    for n, v in zip(golden["nextpow_n"], golden["nextpow_val"]):
        assert og.next_fft_length(int(n)) == int(v)


def test_fft_length_closed_form():
    def ok(m):
        a = 0
        while m % 2 == 0:
            m //= 2
            a += 1
        for p in (3, 5, 7):
            while m % p == 0:
                m //= p
        return m == 1 and a >= 2
    for n in [1, 5, 37, 1001, 3001]:
        v = og.next_fft_length(n)
        assert v >= n and ok(v)
        assert not any(ok(m) for m in range(n, v))


@pytest.mark.parametrize("name", ["evt_u8", "evt_i8"])
def test_compaction_vs_reference(golden, name):
    frames = golden["%s_frames" % name]
    d, t, o = ev.compress_all_stack(frames)
    np.testing.assert_array_equal(d, golden["%s_data" % name])
    np.testing.assert_array_equal(t, golden["%s_times" % name])
    np.testing.assert_array_equal(o, golden["%s_offsets" % name])
    fc = ev.FramesCompressorOracle(frames.shape[1:], frames.shape[0],
                                   int((frames > 0).sum(axis=0).max()), frames.dtype)
    for f in frames:
        fc.process_frame(f)
    d2, t2, o2 = fc.get_compacted_events()
    np.testing.assert_array_equal(d2, golden["%s_prog_data" % name])
    np.testing.assert_array_equal(t2, golden["%s_prog_times" % name])
    np.testing.assert_array_equal(o2, golden["%s_prog_offsets" % name])


# ---------------------------------------------------------------- brute force
def _tiny(seed, n=7, shape=(3, 4), hi=6, dtype=np.uint8, nbins=2):
    rng = np.random.default_rng(seed)
    f = rng.integers(0, hi, size=(n,) + shape).astype(dtype)
    q = rng.integers(0, nbins + 1, size=shape).astype(np.int32)
    for b in range(1, nbins + 1):
        q.flat[b] = b
    return f, q


@pytest.mark.parametrize("seed", range(4))
def test_numerators_vs_brute_force(seed):
    f, q = _tiny(seed, hi=256 if seed % 2 else 4)
    num, s, den = og.brute_force(f, q)
    out = og.dense_g2(f, q)
    assert [[int(x) for x in r] for r in out["num"]] == [[int(x) for x in r] for r in num]
    assert [[int(x) for x in r] for r in out["sums"]] == [[int(x) for x in r] for r in s]
    assert [[int(x) for x in r] for r in out["den"]] == [[int(x) for x in r] for r in den]


def test_int8_negative_values_vs_brute_force():
    f, q = _tiny(9, hi=5)
    f = (f.astype(np.int16) - 2).astype(np.int8)
    num, s, den = og.brute_force(f, q)
    out = og.dense_g2(f, q)
    assert [[int(x) for x in r] for r in out["num"]] == [[int(x) for x in r] for r in num]
    assert [[int(x) for x in r] for r in out["den"]] == [[int(x) for x in r] for r in den]


def test_denominator_python_int_path():
    s = np.array([[2 ** 31 - 1, 2 ** 31 - 5, 7, 2 ** 30]], dtype=np.int64)
    d = og.denominators(s)
    n = s.shape[1]
    for tau in range(n):
        assert int(d[0, tau]) == sum(int(s[0, t]) * int(s[0, t + tau]) for t in range(n - tau))


def test_invariants():
    rng = np.random.default_rng(5)
    f = rng.integers(0, 9, size=(20, 5, 6)).astype(np.uint8)
    q = rng.integers(1, 4, size=(5, 6)).astype(np.int32)
    out = og.dense_g2(f, q)
    # Num(0) = sum I^2 over the bin
    for i, b in enumerate(out["bins"]):
        a = og.gather_bin(f, q, b).astype(np.int64)
        assert out["num"][i, 0] == (a * a).sum()
    # permutation of pixels leaves everything unchanged
    perm = rng.permutation(30)
    f2 = f.reshape(20, -1)[:, perm].reshape(f.shape)
    q2 = q.ravel()[perm].reshape(q.shape)
    out2 = og.dense_g2(f2, q2)
    np.testing.assert_array_equal(out["num"], out2["num"])
    np.testing.assert_array_equal(out["den"], out2["den"])
    # time-constant frames: g2 = n * sum c^2 / (sum c)^2 exactly
    c = rng.integers(1, 9, size=(5, 6)).astype(np.uint8)
    fc = np.broadcast_to(c, (11, 5, 6)).copy()
    o = og.dense_g2(fc, q)
    for i, b in enumerate(o["bins"]):
        cb = c[q == b].astype(np.float64)
        np.testing.assert_allclose(o["g2"][i], cb.size * (cb ** 2).sum() / cb.sum() ** 2,
                                   rtol=1e-14)
    # single-pixel bin: g2 == 1 wherever defined
    q1 = np.zeros((5, 6), np.int32)
    q1[2, 3] = 1
    o1 = og.dense_g2(f, q1)
    fin = np.isfinite(o1["g2"])
    np.testing.assert_allclose(o1["g2"][fin], 1.0, rtol=1e-14)


def test_fft_correlator_close_to_exact():
    rng = np.random.default_rng(3)
    f = rng.poisson(0.5, size=(64, 6, 7)).astype(np.uint8)
    q = rng.integers(0, 4, size=(6, 7)).astype(np.int32)
    out = og.dense_g2(f, q)
    r = og.fft_correlator(f, q)
    fin = np.isfinite(out["g2"])
    assert np.allclose(r[fin], out["g2"][fin], rtol=1e-3, atol=1e-4)


# ----------------------------------------------------------------- event path
@pytest.mark.parametrize("seed", range(3))
def test_event_numerators_equal_dense(seed):
    rng = np.random.default_rng(seed)
    f = (rng.random((30, 5, 6)) < 0.2).astype(np.uint8) * rng.integers(1, 5, size=(30, 5, 6)).astype(np.uint8)
    q = rng.integers(0, 4, size=(5, 6)).astype(np.int32)
    for b in range(1, 4):
        q.flat[b] = b
    d, t, o = ev.compress_all_stack(f)
    e = ev.event_g2(t, d, o, q, 30)
    dn = og.dense_g2(f, q)
    np.testing.assert_array_equal(e["num"], dn["num"])
    np.testing.assert_array_equal(e["sums"], dn["sums"])
    bnum, bs = ev.event_brute_force(t, d, o, q, 30)
    assert [[int(x) for x in r] for r in e["num"]] == [[int(x) for x in r] for r in bnum]
    assert [[int(x) for x in r] for r in e["sums"]] == [[int(x) for x in r] for r in bs]


def test_coo_csr_round_trip():
    rng = np.random.default_rng(11)
    f = (rng.random((25, 4, 5)) < 0.3).astype(np.uint8) * 3
    d, t, o = ev.compress_all_stack(f)
    pix, fr, cnt = ev.csr_to_coo(t, d, o)
    assert np.all(np.diff(fr.astype(np.int64)) >= 0)
    t2, d2, o2 = ev.coo_to_csr(pix, fr, cnt, 20)
    np.testing.assert_array_equal(t2, t)
    np.testing.assert_array_equal(d2, d)
    np.testing.assert_array_equal(o2, o)


# ------------------------------------------- fast restatements for detector sizes
@pytest.mark.parametrize("seed", range(4))
def test_fast_dense_oracle_equals_per_bin_and_brute_force(seed):
    """sums_numerators_fast (one transpose, float64 or int64 Gram) equals the per-bin
    restatement and the Python-integer brute force, int8 negatives included."""
    f, q = _tiny(seed, n=9, shape=(4, 5), hi=256 if seed % 2 else 5, nbins=3)
    if seed == 2:
        f = (f.astype(np.int16) - 2).astype(np.int8)
    _, bins = og.qmask_bins(f.shape[1:], q)
    s, num = og.sums_numerators_fast(f, q, bins)
    np.testing.assert_array_equal(s, og.frame_sums(f, q, bins))
    np.testing.assert_array_equal(num, og.numerators(f, q, bins))
    bnum, bs, _ = og.brute_force(f, q)
    assert [[int(x) for x in r] for r in num] == [[int(x) for x in r] for r in bnum]
    assert [[int(x) for x in r] for r in s] == [[int(x) for x in r] for r in bs]


def test_fast_dense_oracle_int64_path():
    """Values whose squares times the pixel count reach 2^53 take the int64 Gram."""
    rng = np.random.default_rng(4)
    f = rng.integers(60000, 65536, size=(6, 3, 3)).astype(np.uint16)
    q = np.ones((3, 3), np.int32)
    assert 65535 ** 2 * 9 < 2 ** 53   # float path here; force the int64 path below
    _, bins = og.qmask_bins((3, 3), q)
    s, num = og.sums_numerators_fast(f, q, bins)
    np.testing.assert_array_equal(num, og.numerators(f, q, bins))
    big = np.full((4, 1, 2100), 65535, dtype=np.uint16)   # 65535^2 * 2100 > 2^53
    qb = np.ones((1, 2100), np.int32)
    s2, num2 = og.sums_numerators_fast(big, qb, np.array([1]))
    assert [int(x) for x in num2[0]] == [65535 ** 2 * 2100 * (4 - tau) for tau in range(4)]


@pytest.mark.parametrize("seed", range(4))
def test_fast_event_oracle_equals_loops_and_dense(seed):
    """event_sums_numerators_fast equals the per-pixel restatement, the brute-force
    pair loop and the dense oracle on the same stack (pixels with many events, empty
    pixels, excluded pixels and a bin with no events)."""
    rng = np.random.default_rng(100 + seed)
    n = 40
    f = (rng.random((n, 6, 7)) < (0.05 + 0.2 * seed)).astype(np.uint8) * \
        rng.integers(1, 256 if seed % 2 else 4, size=(n, 6, 7)).astype(np.uint8)
    q = rng.integers(0, 5, size=(6, 7)).astype(np.int32)
    q[0, :] = 4
    f[:, 0, :] = 0           # bin 4 holds pixels without events only
    q.flat[8] = 1
    q.flat[9] = 2
    q.flat[10] = 3
    d, t, o = ev.compress_all_stack(f)
    _, bins = og.qmask_bins(q.shape, q)
    s, num = ev.event_sums_numerators_fast(t, d, o, q, n, bins)
    s0, num0 = ev.event_sums_numerators(t, d, o, q, n, bins)
    np.testing.assert_array_equal(s, s0)
    np.testing.assert_array_equal(num, num0)
    bnum, bs = ev.event_brute_force(t, d, o, q, n)
    assert [[int(x) for x in r] for r in num] == [[int(x) for x in r] for r in bnum]
    assert [[int(x) for x in r] for r in s] == [[int(x) for x in r] for r in bs]
    dn = og.dense_g2(f, q)
    np.testing.assert_array_equal(num, dn["num"])
    np.testing.assert_array_equal(s, dn["sums"])


def test_fast_event_oracle_single_long_pixel():
    """One pixel with an event in every frame: Num(tau) = sum_t d_t d_(t+tau)."""
    n = 64
    d = np.arange(1, n + 1, dtype=np.uint8)
    t = np.arange(n, dtype=np.int32)
    o = np.array([0, 0, n, n], dtype=np.uint32)
    q = np.array([[0, 1, 1]], np.int32)
    s, num = ev.event_sums_numerators_fast(t, d, o, q, n, np.array([1]))
    dd = d.astype(np.int64)
    np.testing.assert_array_equal(num[0], [int((dd[:n - k] * dd[k:]).sum()) for k in range(n)])
    np.testing.assert_array_equal(s[0], dd)
