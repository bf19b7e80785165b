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


class BaseCorrelator(object):
    """Abstract base of all correlators (common.py:10-20)."""

    def __init__(self):
        self.nframes = None
        self.shape = None
        self.bins = None
        self.n_bins = None
        self.output_shape = None
        self.weights = None
        self.scale_factors = None

    def _set_parameters(self, shape, nframes, qmask, scale_factor, extra_options):
        self.nframes = int(nframes)
        self._set_shape(shape)
        self._set_qmask(qmask=qmask)
        self._set_scale_factor(scale_factor=scale_factor)
        self._configure_extra_options(extra_options)

    def _set_shape(self, shape):  # common.py:29-34
        if np.isscalar(shape):
            self.shape = (int(shape), int(shape))
        else:
            assert len(shape) == 2
            self.shape = tuple(int(s) for s in shape)

    def _set_qmask(self, qmask=None):  # common.py:36-46
        if qmask is None:
            self.bins = np.array([1], dtype=np.int32)
            self.n_bins = 1
            self.qmask = np.ones(self.shape, dtype=np.int32)
            self._qmask_given = False
        else:
            self.qmask = np.ascontiguousarray(qmask, dtype=np.int32)
            assert self.qmask.shape == tuple(self.shape), "qmask shape must equal frame shape"
            self.n_bins = int(self.qmask.max())
            self.bins = np.arange(1, self.n_bins + 1, dtype=np.int32)
            self._qmask_given = True
        self.output_shape = (self.n_bins, self.nframes)

    def _set_weights(self, weights=None):  # common.py:48-54
        if weights is None:
            self.weights = None
            return
        raise ValueError("Advanced weighting is not implemented yet")

    def _set_scale_factor(self, scale_factor=None):  # common.py:56-72
        if self.n_bins == 0:
            s = scale_factor or np.prod(self.shape)
            self.scale_factors = {1: s}
            return
        if scale_factor is not None:
            if not np.iterable(scale_factor):
                scale_factor = [scale_factor]
            assert len(scale_factor) == self.n_bins
            if isinstance(scale_factor, dict):
                self.scale_factors = dict(scale_factor)
            else:
                self.scale_factors = {k: v for k, v in zip(self.bins, scale_factor)}
        else:
            self.scale_factors = {}
            for bin_val in self.bins:
                self.scale_factors[bin_val] = np.sum(self.qmask == bin_val)

    def _configure_extra_options(self, extra_options):
        self.extra_options = {}
        if extra_options is not None:
            self.extra_options.update(extra_options)


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
        """The plan's Gram launch tuning: dict(tile, band, group, queue)."""
        t = _lib.Tuning()
        _lib.check(self.lib.dxg2_plan_get_tuning(self.handle, ctypes.byref(t)))
        return {"tile": t.tile, "band": t.band, "group": t.group, "queue": t.queue}

    def set_tuning(self, tile, band, group, queue):
        """Rebuild the Gram work lists for a tuning (waits for the device)."""
        t = _lib.Tuning(int(tile), int(band), int(group), int(queue))
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


def to_device_frames(frames, shape, nframes, torch):
    """numpy -> GPU copy; a GPU torch tensor is borrowed (made contiguous)."""
    if isinstance(frames, np.ndarray):
        assert frames.shape == (nframes,) + tuple(shape), \
            "frames should have shape %s, got %s" % ((nframes,) + tuple(shape), frames.shape)
        return torch.from_numpy(np.ascontiguousarray(frames)).to("cuda"), frames.dtype
    if isinstance(frames, torch.Tensor):
        assert tuple(frames.shape) == (nframes,) + tuple(shape), \
            "frames should have shape %s, got %s" % ((nframes,) + tuple(shape), tuple(frames.shape))
        if not frames.is_cuda:
            frames = frames.to("cuda")
        dt = {torch.uint8: np.uint8, torch.int8: np.int8, torch.uint16: np.uint16}.get(frames.dtype)
        if dt is None:
            raise NotImplementedError("frames of dtype %s are not supported" % frames.dtype)
        return frames.contiguous(), np.dtype(dt)
    raise ValueError("Unknown array type %s" % str(type(frames)))


def current_stream(torch):
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
