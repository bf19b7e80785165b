"""Event (sparse) correlator on the MI355X: drop-in for dynamix/correlator/event.py.

EventCorrelator   -- event.py:7-204.  Same constructor and correlate(vol_times,
                     vol_data, offsets) on the pixel-major event layout; also
                     correlate_coo(pixel, frame, count) for frame-sorted event lists.
FramesCompressor  -- event.py:207-321.  Producer of that layout, on the GPU
                     (k_compress.hip): whole stack or frame by frame.

Differences from the reference, all deliberate:
  * numerators are recomputed from zero on every call (the reference zeroes res_int
    only at allocation, event.py:98 vs 170-171, so repeated calls accumulate);
  * max_events_count is accepted but not a limit (the reference copies a pixel's
    events into private arrays of that size without checking, evtcorrelator.cl:29-34);
  * unsorted or out-of-range frames raise ValueError instead of indexing out of bounds;
  * numerators are exact 64-bit integers and the normalisation is exact with one FP64
    division (the reference accumulates it in float32, evtcorrelator.cl:70-73).
"""
import ctypes

import numpy as np

from .. import _lib
from .common import EXACT_DTYPES, BaseCorrelator, Plan, _torch, current_stream, narrow_counts

_FLAG_TEXT = {1: "frames out of [0, nframes) or not strictly increasing within a pixel",
              2: "offsets decreasing or past the number of events",
              4: "pixel index past the frame size"}


def _dtype_of(dtype):
    dt = np.dtype(dtype)
    if dt not in _lib.DTYPE_CODES:
        raise NotImplementedError("event counts of dtype %s are not supported "
                                  "(uint8, int8, uint16)" % dt)
    return dt


class EventCorrelator(BaseCorrelator):
    """Drop-in for EventCorrelator (event.py:15-204).

    ctx/devicetype/platformid/deviceid/block_size/memory/profile are OpenCL options and
    are accepted and ignored.  scale_factor is honoured as in the reference (a scalar
    for one bin, or one value per bin; default: pixel count of each bin).
    """

    def __init__(self, shape, nframes, max_events_count, total_events_count=None,
                 allow_reallocate=True, qmask=None, dtype="f", weights=None, scale_factor=None,
                 extra_options={}, ctx=None, devicetype="all", platformid=None, deviceid=None,
                 block_size=None, memory=None, profile=False):
        super().__init__()
        super()._set_parameters(shape, nframes, qmask, scale_factor, extra_options)
        self._set_weights(weights)
        self.dtype = np.dtype(dtype)
        self._set_events_size(max_events_count, total_events_count)
        self.allow_reallocate = allow_reallocate
        self._plans = {}
        self._ttcf = {}
        if self.dtype in EXACT_DTYPES:
            self._plan_for(self.dtype)   # built up front, as the reference allocates
        self.last_numerators = None
        self.last_sums = None
        self.last_denominators = None
        self.last_g2 = None

    def _plan_for(self, dtype):
        """The plan of one count type (counts of another dtype are narrowed first)."""
        p = self._plans.get(dtype)
        if p is None:
            q = self.qmask if self._qmask_given else None
            scale = np.array([self.scale_factors[b] for b in self.bins], dtype=np.float64) \
                if self.n_bins > 0 else None
            p = Plan(self.shape, self.nframes, q, dtype, scale=scale)
            self._plans[dtype] = p
        self._plan = p
        return p

    def _set_events_size(self, max_events_count, total_events_count):  # event.py:61-63
        self.max_events_count = max_events_count
        self.total_events_count = total_events_count or max_events_count * int(np.prod(self.shape))

    def _check_event_arrays(self, vol_times, vol_data, offsets):  # event.py:104-115
        for arr in [vol_times, vol_data, offsets]:
            assert arr.ndim == 1
        assert vol_times.shape[0] == vol_data.shape[0]
        assert offsets.shape[0] == int(np.prod(self.shape)) + 1
        numels = vol_data.shape[0]
        if numels > self.total_events_count:
            if self.allow_reallocate:
                self.total_events_count = numels
            else:
                raise ValueError("Too many events and allow_reallocate was set to False")

    def _outputs(self, torch):
        nb, n = max(self.n_bins, 0), self.nframes
        return (torch.empty((nb, n), dtype=torch.int64, device="cuda"),
                torch.empty((nb, n), dtype=torch.int64, device="cuda"),
                torch.empty((nb, n), dtype=torch.float64, device="cuda"),
                torch.empty((nb, n), dtype=torch.float32, device="cuda"))

    def _finish(self, torch, outs):
        num, sums, den, g2 = outs
        flags = ctypes.c_int32()
        _lib.check(self._plan.lib.dxg2_plan_event_status(self._plan.handle, ctypes.byref(flags),
                                                         current_stream(torch)))
        if flags.value:
            msg = "; ".join(t for b, t in _FLAG_TEXT.items() if flags.value & b)
            raise ValueError("invalid event input: " + msg)
        self.last_numerators, self.last_sums, self.last_denominators, self.last_g2 = outs
        if self.extra_options.get("return_device", False):
            return g2
        return g2.cpu().numpy()

    @staticmethod
    def _dev(torch, a, dtype):
        if isinstance(a, np.ndarray):
            return torch.from_numpy(np.ascontiguousarray(a, dtype=dtype)).to("cuda")
        if isinstance(a, torch.Tensor):
            return a.to(device="cuda").contiguous()
        raise ValueError("Unknown array type %s" % str(type(a)))

    def correlate(self, vol_times, vol_data, offsets):
        """g2 from pixel-major events (times int32, data, offsets uint32)."""
        torch = _torch()
        self._check_event_arrays(vol_times, vol_data, offsets)
        if isinstance(vol_data, np.ndarray):
            assert np.dtype(vol_data.dtype) == self.dtype, \
                "vol_data dtype %s != %s" % (vol_data.dtype, self.dtype)
        t = self._dev(torch, vol_times, np.int32)
        d, dt = narrow_counts(vol_data, torch, "event counts")
        d = self._dev(torch, d, dt)
        o = self._dev(torch, offsets, np.uint32)
        plan = self._plan_for(dt)
        outs = self._outputs(torch)
        if self.n_bins:
            _lib.check(plan.lib.dxg2_correlate_events_csr(
                plan.handle, _lib.ptr(t), _lib.ptr(d), _lib.ptr(o),
                ctypes.c_int64(int(vol_data.shape[0])), *[_lib.ptr(x) for x in outs],
                current_stream(torch)), "dxg2_correlate_events_csr")
        return self._finish(torch, outs)

    def correlate_coo(self, pixel, frame, count):
        """g2 from a frame-sorted event list (uint32 pixel, uint32 frame, uint8 count)."""
        torch = _torch()
        assert pixel.shape == frame.shape == count.shape and pixel.ndim == 1
        if pixel.shape[0] > self.total_events_count:
            if self.allow_reallocate:
                self.total_events_count = pixel.shape[0]
            else:
                raise ValueError("Too many events and allow_reallocate was set to False")
        px = self._dev(torch, pixel, np.uint32)
        fr = self._dev(torch, frame, np.uint32)
        ct, dt = narrow_counts(count, torch, "event counts")
        if dt == np.dtype(np.int8):
            raise ValueError("frame-sorted event counts are unsigned (uint8 or uint16)")
        ct = self._dev(torch, ct, dt)
        plan = self._plan_for(dt)
        outs = self._outputs(torch)
        if self.n_bins:
            _lib.check(plan.lib.dxg2_correlate_events_coo(
                plan.handle, _lib.ptr(px), _lib.ptr(fr), _lib.ptr(ct),
                ctypes.c_int64(int(pixel.shape[0])), *[_lib.ptr(x) for x in outs],
                current_stream(torch)), "dxg2_correlate_events_coo")
        return self._finish(torch, outs)

    def two_time(self, vol_times, vol_data, offsets, q):
        """Two-time correlation matrix of bin q (1-based) from pixel-major events: float32
        (nframes, nframes), the `trc` the reference's event correlator ecorrts_q returns
        for ttcf_par = q (event_y.py:169-230 with WXPCS/fecorrt.f:1-26).  The bin's events
        become a dense (nframes, n_b) stack on the GPU (dxg2_events_to_columns), whose
        two-time matrix comes from the exact integer Gram in FP64
        (dxg2_correlate_dense_ttcf) with n_b = every pixel of the bin."""
        from .dense import MatMulCorrelator
        torch = _torch()
        self._check_event_arrays(vol_times, vol_data, offsets)
        if not 1 <= int(q) <= max(self.n_bins, 0):
            raise ValueError("bin %d is not in 1..%d" % (int(q), max(self.n_bins, 0)))
        n = self.nframes
        cols = np.flatnonzero(self.qmask.ravel() == int(q)).astype(np.int32)
        if cols.size == 0:
            trc = torch.full((n, n), float("nan"), dtype=torch.float32, device="cuda")
        else:
            d, dt = narrow_counts(vol_data, torch, "event counts")
            t = self._dev(torch, vol_times, np.int32)
            d = self._dev(torch, d, dt)
            o = self._dev(torch, offsets, np.uint32)
            tdt = {np.dtype(np.uint8): torch.uint8, np.dtype(np.int8): torch.int8,
                   np.dtype(np.uint16): torch.uint16}[dt]
            dense = torch.empty((n, 1, cols.size), dtype=tdt, device="cuda")
            dcols = torch.from_numpy(cols).to("cuda")
            err = torch.zeros(1, dtype=torch.int32, device="cuda")
            lib = _lib.load()
            _lib.check(lib.dxg2_events_to_columns(
                _lib.ptr(o), _lib.ptr(t), _lib.ptr(d), int(np.prod(self.shape)), n,
                _lib.DTYPE_CODES[dt], _lib.ptr(dcols), int(cols.size), _lib.ptr(dense),
                _lib.ptr(err), current_stream(torch)), "dxg2_events_to_columns")
            if int(err.item()):
                raise ValueError("invalid event input: frames out of [0, nframes)")
            key = (int(q), dt)
            sub = self._ttcf.get(key)
            if sub is None:
                sub = MatMulCorrelator((1, int(cols.size)), n, None,
                                       extra_options={"return_device": True})
                self._ttcf[key] = sub
            trc = sub.two_time(dense, 1)
        if self.extra_options.get("return_device", False):
            return trc
        return trc.cpu().numpy()


def _np_dtype_of_tensor(t):
    """numpy dtype of a frames tensor (uint16 frames travel as int16 or uint16)."""
    name = str(t.dtype).replace("torch.", "")
    if name in ("int16", "uint16"):
        return np.dtype(np.uint16)
    return _dtype_of(name)


class FramesCompressor(object):
    """Dense frames -> (data, times, offsets) event layout (event.py:207-321), on the GPU.

    For every pixel in raster order, its samples with value > 0 (signed for int8), in
    frame order, are concatenated; offsets[p] .. offsets[p+1] index pixel p's events.
    numpy inputs give numpy outputs, like the reference; GPU tensors give GPU tensors
    that dxg2_correlate_events_csr takes directly.

    Differences from the reference, both deliberate: values are kept in the frames'
    dtype (the reference stores them in `dtype`, int8 by default, where uint8 counts
    above 127 turn negative and are then dropped from the data but not from the
    offsets); a pixel with more than max_nnz events raises IndexError from
    get_compacted_events, not from the process_frame call that overflowed.
    """

    def __init__(self, shape, nframes, max_nnz, dtype=np.int8):
        self.shape = tuple(int(x) for x in shape)
        self.nframes = int(nframes)
        self.dtype = dtype
        self.max_nnz = int(max_nnz)
        self.npix = int(np.prod(self.shape))
        self.frames_counter = 0
        self._state = None  # device slots, allocated by the first frame (its dtype)

    @staticmethod
    def compress_all_stack(frames):
        """Whole stack at once (event.py:267-284)."""
        torch = _torch()
        lib = _lib.load()
        on_device = not isinstance(frames, np.ndarray)
        if on_device:
            assert frames.dim() == 3
            d = frames.contiguous()
            dt = _np_dtype_of_tensor(d)
        else:
            assert frames.ndim == 3
            dt = np.dtype(frames.dtype)
            d = torch.from_numpy(np.ascontiguousarray(frames).view(
                np.int16 if dt == np.uint16 else dt)).cuda()
        dt = _dtype_of(dt)
        n, h, w = (int(x) for x in d.shape)
        nbytes = ctypes.c_size_t()
        _lib.check(lib.dxg2_compress_scratch_bytes(n, h, w, ctypes.byref(nbytes)))
        scratch = torch.empty(max(nbytes.value // 4, 1), dtype=torch.int32, device="cuda")
        offsets = torch.empty(h * w + 1, dtype=torch.int32, device="cuda")
        st = current_stream(torch)
        code = _lib.DTYPE_CODES[dt]
        _lib.check(lib.dxg2_compress_count(_lib.ptr(d), code, n, h, w, _lib.ptr(scratch),
                                           _lib.ptr(offsets), st), "dxg2_compress_count")
        nnz = int(offsets[h * w].item()) & 0xFFFFFFFF
        times = torch.empty(max(nnz, 1), dtype=torch.int32, device="cuda")
        data = torch.empty(max(nnz, 1), dtype=d.dtype, device="cuda")
        _lib.check(lib.dxg2_compress_fill(_lib.ptr(d), code, n, h, w, _lib.ptr(scratch),
                                          _lib.ptr(offsets), _lib.ptr(times), _lib.ptr(data), st),
                   "dxg2_compress_fill")
        data, times = data[:nnz], times[:nnz]
        if on_device:
            return data, times, offsets
        out = data.cpu().numpy()
        return (out.view(np.uint16) if dt == np.uint16 else out, times.cpu().numpy(),
                offsets.cpu().numpy().view(np.uint32))

    def _slots(self, torch, tdtype, code):
        if self._state is None:
            n = self.npix * self.max_nnz
            self._state = {
                "code": code,
                "counter": torch.zeros(self.npix, dtype=torch.int32, device="cuda"),
                "events": torch.zeros(max(n, 1), dtype=tdtype, device="cuda"),
                "times": torch.zeros(max(n, 1), dtype=torch.int32, device="cuda"),
                "err": torch.zeros(1, dtype=torch.int32, device="cuda"),
            }
        elif self._state["code"] != code:
            raise ValueError("frames of one compressor must share their dtype")
        return self._state

    def process_frame(self, frame):
        """Append one frame's events (event.py:287-297)."""
        torch = _torch()
        lib = _lib.load()
        if isinstance(frame, np.ndarray):
            dt = _dtype_of(frame.dtype)
            d = torch.from_numpy(np.ascontiguousarray(frame).view(
                np.int16 if dt == np.uint16 else dt)).cuda()
        else:
            d = frame.contiguous()
            dt = _np_dtype_of_tensor(d)
        assert d.numel() == self.npix, "frame size differs from the compressor's shape"
        st = self._slots(torch, d.dtype, _lib.DTYPE_CODES[dt])
        _lib.check(lib.dxg2_compress_push(
            _lib.ptr(d), st["code"], self.shape[0], self.shape[1], self.frames_counter,
            self.max_nnz, _lib.ptr(st["counter"]), _lib.ptr(st["events"]), _lib.ptr(st["times"]),
            _lib.ptr(st["err"]), current_stream(torch)), "dxg2_compress_push")
        self.frames_counter += 1

    def get_compacted_events(self, wait_for_all_frames=True, device=False):
        """(data, times, offsets) of the frames processed so far (event.py:300-321);
        GPU tensors with device=True."""
        if self.frames_counter < self.nframes - 1 and wait_for_all_frames:
            raise RuntimeError("Not all frames were compressed yet (%d/%d)"
                               % (self.frames_counter, self.nframes))
        torch = _torch()
        lib = _lib.load()
        if self._state is None:
            self._slots(torch, torch.int8, _lib.DXG2_I8)
        st = self._state
        if int(st["err"].item()):
            raise IndexError("a pixel has more than max_nnz=%d events" % self.max_nnz)
        scratch = torch.empty((self.npix + 1023) // 1024 + 2, dtype=torch.int32, device="cuda")
        offsets = torch.empty(self.npix + 1, dtype=torch.int32, device="cuda")
        total = int(st["counter"].sum().item())
        data = torch.empty(max(total, 1), dtype=st["events"].dtype, device="cuda")
        times = torch.empty(max(total, 1), dtype=torch.int32, device="cuda")
        _lib.check(lib.dxg2_compress_compact(
            _lib.ptr(st["counter"]), _lib.ptr(st["events"]), _lib.ptr(st["times"]), st["code"],
            self.shape[0], self.shape[1], self.max_nnz, _lib.ptr(scratch), _lib.ptr(offsets),
            _lib.ptr(data), _lib.ptr(times), current_stream(torch)), "dxg2_compress_compact")
        data, times = data[:total], times[:total]
        if device:
            return data, times, offsets
        out = data.cpu().numpy()
        return (out.view(np.uint16) if st["code"] == _lib.DXG2_U16 else out,
                times.cpu().numpy(), offsets.cpu().numpy().view(np.uint32))
