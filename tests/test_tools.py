"""Host tools around the path (qmask construction, radial average, multitau, result
formatting) against loop restatements of dynamix/tools/tools.py and closed forms."""
from collections import namedtuple

import numpy as np
import pytest

from dynamix_amd import tools
from oracle import tools as ot

Result = namedtuple("Result", "res dev trc")


def _mask(shape, seed):
    rng = np.random.default_rng(seed)
    m = (rng.random(shape) < 0.05).astype(np.uint8)
    m[shape[0] // 2 - 1: shape[0] // 2 + 2, :] = 1   # a gap band
    return m


@pytest.mark.parametrize("seed", range(3))
def test_radial_average_vs_loops(seed):
    rng = np.random.default_rng(seed)
    shape = (23, 31)
    saxs = rng.random(shape) * 100
    m = _mask(shape, seed)
    cx, cy = 12.3 + seed, 9.7
    a = tools.radial_average(saxs, m, cx, cy)
    b = ot.radi_loops(saxs, m, cx, cy)
    np.testing.assert_array_equal(a[0][:, 0], b[0][:, 0])
    np.testing.assert_allclose(a[0][:, 1], b[0][:, 1], rtol=1e-12)
    np.testing.assert_array_equal(a[1], b[1])
    np.testing.assert_allclose(a[2], b[2], rtol=1e-6)


def test_radial_average_of_a_radial_pattern():
    """A pattern that depends only on the rounded radius is reproduced exactly."""
    shape, cx, cy = (40, 50), 20.0, 25.0
    yy, xx = np.indices(shape)
    ring = (np.sqrt((xx - cx) ** 2 + (yy - cy) ** 2).astype(np.float32) + 0.5).astype(np.int16)
    saxs = 3.0 * ring + 1.0
    rad, r, model = tools.radial_average(saxs, np.zeros(shape, np.uint8), cx, cy)
    np.testing.assert_allclose(rad[:-1, 1], 3.0 * rad[:-1, 0] + 1.0, rtol=1e-12)
    # np.histogram's last bin is closed: it holds the two largest radii (tools.py:36)
    top = ring.max()
    sel = ring >= top - 1
    np.testing.assert_allclose(rad[-1, 1], saxs[sel].mean(), rtol=1e-12)


@pytest.mark.parametrize("seed", range(2))
def test_make_qmask_vs_loops(seed):
    shape = (37, 45)
    m = _mask(shape, 10 + seed)
    # q ~ 9.4e-5 / A per pixel of radius here: 9 annuli over radii ~3 .. 30
    args = dict(cx=20.5, cy=17.0 + seed, mask=m, first_q=3e-4, width_q=4e-4,
                step_q=2.8e-4 + 2e-5 * seed, number_q=9, pix_size=75e-6, distance=5.0,
                wavelength=1.0)
    a = tools.make_qmask(shape, **args)
    b = ot.make_q_loops(shape, **args)
    assert a.dtype == np.uint16
    np.testing.assert_array_equal(a, b)
    assert a.max() > 3 and (a[m > 0] == 0).all()


def test_make_qmask_rings_are_annuli():
    shape, cx, cy = (60, 60), 30.0, 30.0
    a = tools.make_qmask(shape, cx, cy, np.zeros(shape, np.uint8), 5e-4, 4e-4, 6e-4, 5,
                         75e-6, 5.0, 1.0)
    yy, xx = np.indices(shape)
    r = np.sqrt((xx - cx) ** 2 + (yy - cy) ** 2)
    for n in range(1, 6):
        rr = r[a == n]
        assert rr.size and rr.max() - rr.min() < 12   # one ring each, radially thin
    assert np.all(np.diff([r[a == n].mean() for n in range(1, 6)]) > 0)


@pytest.mark.parametrize("n,par,dtype", [(100, 16, np.float32), (1000, 16, np.float32),
                                         (777, 8, np.float64), (15, 16, np.float32),
                                         (4096, 32, np.float32)])
def test_multitau_vs_loops(n, par, dtype):
    rng = np.random.default_rng(n)
    d = np.zeros((n, 3), dtype)
    d[:, 0] = np.arange(1, n + 1)
    d[:, 1] = 1 + np.exp(-np.arange(n) / 50.0) + rng.normal(0, 0.01, n)
    d[:, 2] = rng.random(n) * 0.01
    a = tools.multitau(d, par)
    b = ot.cftomt_loops(d, par)
    assert a.dtype == d.dtype
    np.testing.assert_array_equal(a, b)


def test_multitau_closed_forms():
    # a linear function stays on the line: each level averages neighbours
    n = 500
    d = np.zeros((n, 3))
    d[:, 0] = np.arange(1, n + 1)
    d[:, 1] = 2 * d[:, 0] + 3
    m = tools.multitau(d, 16)
    np.testing.assert_allclose(m[:, 1], 2 * m[:, 0] + 3, rtol=1e-12)
    assert np.all(np.diff(m[:, 0]) > 0)
    # 16 + 8 per level while the level holds >= 16 rows
    lengths, L = [], n - 16
    while L >= 16:
        lengths.append(8)
        L = len(range(16, L - 2, 2))
    assert len(m) == 16 + sum(lengths)


def test_format_result():
    rng = np.random.default_rng(7)
    shape, nf = (20, 24), 200
    q = rng.integers(0, 4, size=shape).astype(np.int32)
    for b in (1, 2, 3):
        q.flat[b] = b
    res = 1 + rng.random((3, nf)).astype(np.float32)
    dev = rng.random((3, nf)).astype(np.float32) * 0.01
    trc = rng.random((nf, nf)).astype(np.float32)
    ff = 1 + 0.1 * rng.random(shape)
    cd = 5 + rng.random(shape)
    per_q, table, t = tools.format_result(Result(res, dev, trc), q, ff, cd, 0.01, 2)
    assert t is trc
    assert len(per_q) == 3 and table.shape[1] == 1 + 2 * 3
    for i, m in enumerate(per_q, start=1):
        sel = q == i
        corr = (cd[sel] ** 2).mean() / cd[sel].mean() ** 2 * ((ff[sel] ** 2).mean() / ff[sel].mean() ** 2)
        cf = np.zeros((nf, 3), np.float32)
        cf[:, 0] = np.arange(1, nf + 1) * 0.01
        cf[:, 1] = res[i - 1] / corr
        cf[:, 2] = dev[i - 1] / corr
        np.testing.assert_array_equal(m, ot.cftomt_loops(cf))
        np.testing.assert_array_equal(table[:, 2 * i - 1:2 * i + 1], m[:, 1:])


# ------------------------------------------------ golden vectors of the reference tools
@pytest.mark.parametrize("name", ["a", "b"])
def test_radial_average_matches_reference(golden, name):
    """tools.radi (tools.py:14-50) run on the same inputs: radii, ring means, radius map
    and the interpolated pattern."""
    cx, cy = golden["radi_%s_center" % name]
    rad, r, model = tools.radial_average(golden["radi_%s_saxs" % name],
                                         golden["radi_%s_mask" % name], cx, cy)
    ref = golden["radi_%s_rad" % name]
    np.testing.assert_array_equal(rad[:, 0], ref[:, 0])
    # ring means: the same sums in another order (bincount vs np.histogram weights)
    np.testing.assert_allclose(rad[:, 1], ref[:, 1], rtol=1e-12, equal_nan=True)
    np.testing.assert_array_equal(r, golden["radi_%s_q" % name])
    np.testing.assert_array_equal(model, golden["radi_%s_model" % name])


@pytest.mark.parametrize("name", ["f32_16", "f64_8", "f32_32"])
def test_multitau_matches_reference(golden, name):
    """tools.cftomt (tools.py:146-172) run on the same functions, bit for bit."""
    d = golden["cftomt_%s_in" % name]
    par = int(golden["cftomt_%s_par" % name])
    out = tools.multitau(d, par)
    ref = golden["cftomt_%s_out" % name]
    assert out.shape == ref.shape and out.dtype == ref.dtype
    np.testing.assert_array_equal(out, ref)


def test_format_result_matches_reference(golden):
    """tools.format_result (tools.py:313-346) on the same result: the per-q multitau
    functions, the table for saving and the returned two-time matrix."""
    r = Result(golden["fmt_res"], golden["fmt_dev"], golden["fmt_trc"])
    per_q, table, trc = tools.format_result(r, golden["fmt_qmask"], golden["fmt_flat"],
                                            golden["fmt_cdata"], 0.01, 2)
    assert len(per_q) == 3
    for i, m in enumerate(per_q):
        np.testing.assert_array_equal(m, golden["fmt_q%d" % (i + 1)])
    np.testing.assert_array_equal(table, golden["fmt_table"])
    np.testing.assert_array_equal(trc, golden["fmt_trc_out"])
