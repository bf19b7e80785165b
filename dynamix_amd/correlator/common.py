"""Correlator base: the reference's BaseCorrelator conventions over a dxg2 plan.

Mirrors dynamix/correlator/common.py:10-80 (shape, qmask -> bins, scale factors,
extra options) and owns the device plan of the HIP library.  Inputs may be numpy
arrays (copied to the GPU) or torch tensors already on the GPU (borrowed, no copy);
results come back as host float32 (n_bins, nframes) like the reference, and the exact
numerators / sums / denominators of the last call stay on the device.
"""
import ctypes

import numpy as np

from .. import _lib


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("dynamix_amd correlators need a ROCm GPU (torch.cuda is unavailable)")
    return torch


def frame_shape(shape):
    """(H, W) of a frame: a scalar means a square detector of that side."""
    if np.ndim(shape) == 0:
        return (int(shape), int(shape))
    dims = tuple(int(v) for v in shape)
    assert len(dims) == 2, "a frame shape has two dimensions"
    return dims


def bin_table(shape, qmask):
    """Bins of a q-mask, as the reference defines them (common.py:36-46).

    Returns (qmask int32 (H, W), n_bins, bins, pixel_counts).  Without a qmask every
    pixel is in bin 1.  With one, the bins are 1 .. max(qmask) (0 and negative values
    exclude a pixel), so an unused value below the maximum is an empty bin.
    pixel_counts[i] is the number of pixels of bins[i].
    """
    if qmask is None:
        labels = np.ones(shape, dtype=np.int32)
    else:
        labels = np.ascontiguousarray(qmask, dtype=np.int32)
        assert labels.shape == tuple(shape), "qmask shape must equal frame shape"
    top = int(labels.max()) if labels.size else 0
    bins = np.arange(1, top + 1, dtype=np.int32)
    hist = np.bincount(labels[labels > 0].ravel(), minlength=top + 1) if top > 0 else \
        np.zeros(1, dtype=np.int64)
    return labels, top, bins, hist[1:top + 1].astype(np.int64)


def scale_table(shape, n_bins, bins, pixel_counts, scale_factor=None):
    """{bin: scale} as the reference builds it (common.py:56-72).

    Default: each bin's pixel count.  A given value is one scalar (only valid for one
    bin), a sequence with one value per bin, or a dict keyed by bin.  A qmask whose
    maximum is 0 has no bins; it keeps a single entry for bin 1, the given value or the
    frame's pixel count.
    """
    if n_bins == 0:
        return {1: scale_factor if scale_factor else int(np.prod(shape))}
    if scale_factor is None:
        return dict(zip(bins, pixel_counts))
    values = scale_factor if np.iterable(scale_factor) else [scale_factor]
    assert len(values) == n_bins, "one scale factor per bin"
    if isinstance(values, dict):
        return values
    return dict(zip(bins, values))


class BaseCorrelator(object):
    """State every correlator shares: frame geometry, bins and their scale factors.

    Attribute names and meanings follow the reference's BaseCorrelator (common.py:10-20):
    nframes, shape, qmask, bins, n_bins, output_shape = (n_bins, nframes), weights,
    scale_factors ({bin: value}) and extra_options.  pixel_counts (n_b per bin) is added.
    """

    nframes = shape = qmask = bins = n_bins = output_shape = None
    weights = scale_factors = pixel_counts = None

    def _set_parameters(self, shape, nframes, qmask, scale_factor, extra_options):
        self.nframes = int(nframes)
        self.shape = frame_shape(shape)
        self._qmask_given = qmask is not None
        self.qmask, self.n_bins, self.bins, self.pixel_counts = bin_table(self.shape, qmask)
        self.output_shape = (self.n_bins, self.nframes)
        self.scale_factors = scale_table(self.shape, self.n_bins, self.bins, self.pixel_counts,
                                         scale_factor)
        self.extra_options = dict(extra_options or {})

    def _set_weights(self, weights=None):
        """Per-pixel weights are not implemented by the reference either (common.py:48-54):
        anything but None raises ValueError."""
        if weights is not None:
            raise ValueError("Advanced weighting is not implemented yet")
        self.weights = None


class Plan(object):
    """Owns one dxg2 plan (device work lists and workspaces)."""

    def __init__(self, shape, nframes, qmask, dtype, scale=None, flags=0):
        # torch first: the plan allocates and uploads through HIP, and a HIP runtime
        # started by the library before torch is imported leaves torch without a device
        _torch()
        self.lib = _lib.load()
        self.dtype = np.dtype(dtype)
        if self.dtype not in _lib.DTYPE_CODES:
            raise NotImplementedError(
                "frames of dtype %s are not supported (uint8, int8, uint16)" % self.dtype)
        self._keep = []
        desc = _lib.PlanDesc()
        desc.height, desc.width = int(shape[0]), int(shape[1])
        desc.nframes = int(nframes)
        desc.dtype = _lib.DTYPE_CODES[self.dtype]
        desc.flags = int(flags)
        if qmask is not None:
            q = np.ascontiguousarray(qmask, dtype=np.int32)
            self._keep.append(q)
            desc.qmask = q.ctypes.data
        if scale is not None:
            sc = np.ascontiguousarray(scale, dtype=np.float64)
            self._keep.append(sc)
            desc.scale_factors = sc.ctypes.data
            desc.n_scale_factors = sc.size
        handle = ctypes.c_void_p()
        _lib.check(self.lib.dxg2_plan_create(ctypes.byref(desc), ctypes.byref(handle)),
                   "dxg2_plan_create")
        self.handle = handle
        nb = ctypes.c_int32()
        _lib.check(self.lib.dxg2_plan_query(self.handle, ctypes.byref(nb), None, None))
        self.n_bins = nb.value
        self.pixel_counts = np.zeros(max(self.n_bins, 1), dtype=np.int64)
        self.scale = np.zeros(max(self.n_bins, 1), dtype=np.float64)
        _lib.check(self.lib.dxg2_plan_query(self.handle, None, self.pixel_counts.ctypes.data,
                                            self.scale.ctypes.data))

    def get_tuning(self):
        """The plan's Gram launch tuning: dict(tile, band, group, queue, kchunk)."""
        t = _lib.Tuning()
        _lib.check(self.lib.dxg2_plan_get_tuning(self.handle, ctypes.byref(t)))
        return {"tile": t.tile, "band": t.band, "group": t.group, "queue": t.queue,
                "kchunk": t.kchunk}

    def set_tuning(self, tile, band, group, queue, kchunk=0):
        """Rebuild the Gram work lists for a tuning (waits for the device); kchunk 0 is
        the dtype's default."""
        t = _lib.Tuning(int(tile), int(band), int(group), int(queue), int(kchunk))
        _lib.check(self.lib.dxg2_plan_set_tuning(self.handle, ctypes.byref(t)),
                   "dxg2_plan_set_tuning")

    def footprint(self, path=0):
        out = ctypes.c_size_t()
        _lib.check(self.lib.dxg2_plan_footprint(self.handle, path, ctypes.byref(out)))
        return out.value

    def set_profiling(self, on=True):
        _lib.check(self.lib.dxg2_plan_set_profiling(self.handle, 1 if on else 0))

    def stage_times(self):
        """Average [gather, gram, finalise, total] ms and call count since last read."""
        t = (ctypes.c_float * 5)()
        _lib.check(self.lib.dxg2_plan_stage_times(self.handle, t, 5))
        return list(t)

    def close(self):
        """Release the plan (waits for the device)."""
        if getattr(self, "handle", None):
            self.lib.dxg2_plan_destroy(self.handle)
            self.handle = None

    def close_async(self, stream):
        """Release the plan's device memory in stream order on `stream` (a torch stream
        on which every call on the plan was issued or has completed)."""
        if getattr(self, "handle", None):
            self.lib.dxg2_plan_destroy_async(self.handle, ctypes.c_void_p(stream.cuda_stream))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


EXACT_DTYPES = (np.dtype(np.uint8), np.dtype(np.int8), np.dtype(np.uint16))


def _exact_target(lo, hi):
    """Narrowest frame type of the exact integer path holding [lo, hi], or None."""
    if lo >= 0 and hi <= 255:
        return np.dtype(np.uint8)
    if lo >= -128 and hi <= 127:
        return np.dtype(np.int8)
    if lo >= 0 and hi <= 65535:
        return np.dtype(np.uint16)
    return None


def _range_error(what, dt, lo, hi):
    return ValueError("%s of dtype %s hold values in [%s, %s]; the exact integer path takes "
                      "integer photon counts in [0, 65535] (uint8/uint16) or [-128, 127] "
                      "(int8)" % (what, dt, lo, hi))


def narrow_counts(x, torch, what="frames"):
    """Counts of any numeric dtype as the narrowest type of the exact integer path.

    The reference correlators cast frames to float32 (dense.py:49) and its OpenCL
    classes default to dtype "f" (common.py:155-162).  Photon counts are integers, so
    here any integer or float array whose values are all integers in [0, 65535] or
    [-128, 127] is converted, exactly, to uint8, int8 or uint16 (in that order of
    preference); anything else raises ValueError naming the allowed ranges.  numpy
    arrays are converted on the host, GPU tensors on the GPU.  Returns (array, dtype).
    """
    if isinstance(x, np.ndarray):
        dt = np.dtype(x.dtype)
        if dt in EXACT_DTYPES:
            return x, dt
        if x.size == 0:
            return x.astype(np.uint8), np.dtype(np.uint8)
        if dt.kind == "f" and not (np.isfinite(x).all() and (np.trunc(x) == x).all()):
            raise ValueError("%s of dtype %s hold non-integer values; the exact integer path "
                             "takes integer photon counts" % (what, dt))
        if dt.kind not in "biuf":
            raise ValueError("%s of dtype %s are not numeric counts" % (what, dt))
        lo, hi = x.min(), x.max()
        target = _exact_target(lo, hi)
        if target is None:
            raise _range_error(what, dt, lo, hi)
        return x.astype(target), target
    native = {torch.uint8: np.uint8, torch.int8: np.int8, torch.uint16: np.uint16}
    if x.dtype in native:
        return x, np.dtype(native[x.dtype])
    if x.numel() == 0:
        return x.to(torch.uint8), np.dtype(np.uint8)
    if x.dtype == torch.bool:
        return x.to(torch.uint8), np.dtype(np.uint8)
    if x.dtype.is_floating_point:
        if not bool(torch.isfinite(x).all()) or bool((torch.trunc(x) != x).any()):
            raise ValueError("%s of dtype %s hold non-integer values; the exact integer path "
                             "takes integer photon counts" % (what, x.dtype))
    elif x.dtype in (torch.uint32, torch.uint64):
        x = x.to(torch.int64) if x.dtype == torch.uint32 else x
    elif x.dtype.is_complex:
        raise ValueError("%s of dtype %s are not numeric counts" % (what, x.dtype))
    if x.dtype == torch.uint64:
        hi = int(x.max().item())
        lo = 0 if hi <= 65535 else int(x.min().item())
    else:
        lo_t, hi_t = torch.aminmax(x)
        lo, hi = lo_t.item(), hi_t.item()
    target = _exact_target(lo, hi)
    if target is None:
        raise _range_error(what, x.dtype, lo, hi)
    tdt = {np.dtype(np.uint8): torch.uint8, np.dtype(np.int8): torch.int8,
           np.dtype(np.uint16): torch.uint16}[target]
    if x.dtype == torch.uint64:
        x = x.to(torch.int64)
    return x.to(tdt), target


def to_device_frames(frames, shape, nframes, torch):
    """Frames on the GPU in a type of the exact integer path (narrow_counts): numpy
    arrays are copied; a GPU torch tensor of uint8/int8/uint16 is borrowed (made
    contiguous), one of another dtype is converted."""
    if isinstance(frames, np.ndarray):
        assert frames.shape == (nframes,) + tuple(shape), \
            "frames should have shape %s, got %s" % ((nframes,) + tuple(shape), frames.shape)
        frames, dt = narrow_counts(frames, torch)
        return torch.from_numpy(np.ascontiguousarray(frames)).to("cuda"), dt
    if isinstance(frames, torch.Tensor):
        assert tuple(frames.shape) == (nframes,) + tuple(shape), \
            "frames should have shape %s, got %s" % ((nframes,) + tuple(shape), tuple(frames.shape))
        if not frames.is_cuda:
            frames = frames.to("cuda")
        frames, dt = narrow_counts(frames, torch)
        return frames.contiguous(), dt
    raise ValueError("Unknown array type %s" % str(type(frames)))


def current_stream(torch):
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
