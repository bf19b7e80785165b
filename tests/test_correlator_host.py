"""Host-side state of the correlator classes (no GPU): bins, pixel counts and scale
factors of BaseCorrelator against the golden vectors written by running the
reference's BaseCorrelator (common.py:10-80)."""
import numpy as np
import pytest

from dynamix_amd.correlator.common import BaseCorrelator, bin_table, frame_shape, scale_table


def _base(shape, nframes, qmask, scale_factor=None, extra_options=None):
    c = BaseCorrelator()
    c._set_parameters(shape, nframes, qmask, scale_factor, extra_options)
    return c


def test_bins_and_default_scale_match_reference(golden):
    q = golden["base_qmask"]
    c = _base(q.shape, 10, q)
    np.testing.assert_array_equal(c.bins, golden["base_bins"])
    assert c.n_bins == int(golden["base_n_bins"])
    assert c.output_shape == (c.n_bins, 10)
    np.testing.assert_array_equal([c.scale_factors[b] for b in c.bins], golden["base_scale"])
    np.testing.assert_array_equal(c.pixel_counts, [(q == b).sum() for b in c.bins])


def test_given_scale_factors_match_reference(golden):
    q = golden["base_qmask"]
    c = _base(q.shape, 10, q, [10.0, 20.0, 30.0])
    np.testing.assert_array_equal([c.scale_factors[b] for b in c.bins],
                                  golden["base_scale_given"])
    d = {1: 4.0, 2: 5.0, 3: 6.0}
    assert _base(q.shape, 10, q, d).scale_factors == d
    with pytest.raises(AssertionError):
        _base(q.shape, 10, q, [1.0, 2.0])


def test_no_qmask_and_scalar_shape_match_reference(golden):
    c = _base((3, 4), 5, None)
    np.testing.assert_array_equal(c.bins, golden["base_none_bins"])
    np.testing.assert_array_equal(c.qmask, golden["base_none_qmask"])
    np.testing.assert_array_equal([c.scale_factors[b] for b in c.bins], golden["base_none_scale"])
    s = _base(4, 5, None, 7.5)
    assert s.shape == tuple(golden["base_scalar_shape"])
    np.testing.assert_array_equal([s.scale_factors[b] for b in s.bins], golden["base_scalar_scale"])


def test_edge_cases():
    # an unused value below the maximum is an empty bin
    q = np.array([[0, 3], [3, 1]], np.int32)
    labels, n_bins, bins, counts = bin_table((2, 2), q)
    assert n_bins == 3 and list(bins) == [1, 2, 3] and list(counts) == [1, 0, 2]
    # a qmask without positive labels has no bins and one scale entry for bin 1
    z = _base((2, 3), 4, np.zeros((2, 3), np.int32))
    assert z.n_bins == 0 and len(z.bins) == 0 and z.scale_factors == {1: 6}
    assert scale_table((2, 3), 0, np.zeros(0, np.int32), np.zeros(0), 2.5) == {1: 2.5}
    assert frame_shape(7) == (7, 7) and frame_shape([3, 5]) == (3, 5)
    with pytest.raises(AssertionError):
        frame_shape((1, 2, 3))
    with pytest.raises(AssertionError):
        _base((2, 2), 4, np.zeros((3, 3), np.int32))
    c = _base((2, 2), 4, q, extra_options={"tile": 128})
    assert c.extra_options == {"tile": 128}
    with pytest.raises(ValueError):
        c._set_weights(np.ones((2, 2)))


@pytest.mark.parametrize("dt", [np.float32, np.float64, np.int16, np.int32, np.int64,
                                np.uint32, np.uint64])
def test_narrow_counts_numpy_and_torch(dt):
    """Counts of any numeric dtype go to the narrowest exact type holding their values,
    unchanged; numpy on the host and torch tensors alike."""
    import torch
    from dynamix_amd.correlator.common import narrow_counts
    info = np.finfo(dt) if np.dtype(dt).kind == "f" else np.iinfo(dt)
    cases = [([0, 3, 255], np.uint8), ([0, 256, 65535], np.uint16), ([-128, 0, 127], np.int8)]
    for vals, want in cases:
        if min(vals) < info.min or max(vals) > info.max:
            continue
        a = np.array(vals, dt)
        out, got = narrow_counts(a, torch)
        assert got == np.dtype(want) and out.dtype == np.dtype(want)
        np.testing.assert_array_equal(out.astype(np.int64), a.astype(np.int64))
        if np.dtype(dt) == np.uint64:
            continue   # torch has no general uint64 tensors from numpy
        t, got_t = narrow_counts(torch.from_numpy(a), torch)
        assert got_t == np.dtype(want)
        np.testing.assert_array_equal(t.numpy().astype(np.int64), a.astype(np.int64))


def test_narrow_counts_rejects_what_the_exact_path_cannot_hold():
    import torch
    from dynamix_amd.correlator.common import narrow_counts
    for bad in (np.array([0.5, 1.0], np.float32), np.array([np.nan], np.float32),
                np.array([-1, 200], np.int32), np.array([0, 70000], np.int32),
                np.array([1 + 2j])):
        with pytest.raises(ValueError):
            narrow_counts(bad, torch)
        if bad.dtype.kind != "c":
            with pytest.raises(ValueError):
                narrow_counts(torch.from_numpy(bad), torch)
    u8 = np.array([1, 2], np.uint8)
    out, dt = narrow_counts(u8, torch)
    assert out is u8 and dt == np.uint8


def test_reference_gpu_class_names_importable():
    from dynamix_amd.correlator import cuda, dense
    assert issubclass(cuda.CublasMatMulCorrelator, dense.MatMulCorrelator)
    assert issubclass(cuda.CUFFTCorrelator, dense.FFTCorrelator)
    assert dense.FFTWCorrelator is dense.FFTCorrelator
