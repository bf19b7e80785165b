"""Dense correlators on the MI355X: drop-ins for dynamix/correlator/dense.py.

MatMulCorrelator  -- dense.py:118-142.  Same constructor and correlate(frames,
                     calc_std=False); the numerator is the grouped i8-MFMA Gram of
                     every bin in one launch instead of one float32 sgemm per bin.
DenseCorrelator   -- dense.py:145-247.  Same constructor (OpenCL device-selection
                     arguments are accepted and ignored); same HIP path.
PlainMatMulCorrelator -- the one-bin-per-launch FP32 form of
                     CublasMatMulCorrelator (cuda.py:19-125), kept as the comparator.
FFTCorrelator     -- dense.py:275-369 (FFTWCorrelator) and cuda.py:129-219
                     (CUFFTCorrelator): numerators from FP32 power spectra along the
                     time axis (FFTWCorrelator is the same class under the reference's
                     name).

extra_options understood by the dense classes:
  "tile": 128           -- 128-frame Gram tiles (default 256);
  "return_device": True -- return g2 as a GPU tensor;
  "autotune": True      -- tune the Gram launch for this workload (dynamix_amd.autotune):
                           a cached result is applied when the plan is built, otherwise
                           the first correlate() times the candidates on its frames;
  "graph": True         -- replay the dense chain as one HIP graph per call, over
                           device buffers that persist across calls.

Results: float32 (n_bins, nframes), tau = 0 in column 0, NaN where the denominator is
zero (an empty bin, or frames with no counts).  g2 = n_b * Num / Den with exact integer
Num and Den and one FP64 division, i.e. within float32 rounding of the FP64 value.
After each call `last_numerators`, `last_sums`, `last_denominators` (torch tensors on
the GPU, int64/int64/float64) hold the exact intermediate results.
"""
from collections import namedtuple

import numpy as np

from .. import _lib
from .common import BaseCorrelator, Plan, _torch, current_stream, to_device_frames

CorrelationResult = namedtuple("CorrelationResult", "res dev")


class _DenseHIP(BaseCorrelator):
    """Shared machinery: one dxg2 plan per frame dtype, device outputs, host result."""

    def _init_dense(self):
        self._plans = {}
        self._tuned = set()
        self._graph_bufs = {}
        self._last_plan = None
        self.last_numerators = None
        self.last_sums = None
        self.last_denominators = None
        self.last_g2 = None

    def _plan(self, dtype):
        dtype = np.dtype(dtype)
        p = self._plans.get(dtype)
        if p is None:
            q = self.qmask if self._qmask_given else None
            flags = _lib.DXG2_FLAG_TILE128 if self.extra_options.get("tile") == 128 else 0
            p = Plan(self.shape, self.nframes, q, dtype, scale=None, flags=flags)
            self._plans[dtype] = p
            if self.extra_options.get("autotune", False):
                from .. import autotune
                if autotune.apply_cached(p, self.nframes):
                    self._tuned.add(dtype)
        return p

    def _outputs(self, plan, dtype, torch, calc_std):
        """Output tensors of one call: fresh ones, or (graph mode) the plan's persistent
        set, whose addresses the captured graph holds."""
        nb, n = max(self.n_bins, 0), self.nframes
        graph = self.extra_options.get("graph", False) and not calc_std
        if graph and dtype in self._graph_bufs:
            return self._graph_bufs[dtype]
        out = (torch.empty((nb, n), dtype=torch.int64, device="cuda"),
               torch.empty((nb, n), dtype=torch.int64, device="cuda"),
               torch.empty((nb, n), dtype=torch.float64, device="cuda"),
               torch.empty((nb, n), dtype=torch.float32, device="cuda"),
               torch.empty((nb, n), dtype=torch.float32, device="cuda") if calc_std else None)
        if graph:
            self._graph_bufs[dtype] = out
        return out

    def _run_dense(self, frames, calc_std=False):
        torch = _torch()
        d_frames, dtype = to_device_frames(frames, self.shape, self.nframes, torch)
        plan = self._plan(dtype)
        self._last_plan = plan
        nb = max(self.n_bins, 0)
        if nb and self.extra_options.get("autotune", False) and dtype not in self._tuned:
            from .. import autotune
            autotune.autotune(plan, d_frames, self.nframes)
            self._tuned.add(dtype)
        graph = self.extra_options.get("graph", False) and not calc_std
        if graph:
            # the graph holds the frames' address too: keep a persistent input buffer
            key = ("frames", dtype)
            buf = self._graph_bufs.get(key)
            if buf is None or buf.shape != d_frames.shape:
                buf = torch.empty_like(d_frames)
                self._graph_bufs[key] = buf
            if buf.data_ptr() != d_frames.data_ptr():
                buf.copy_(d_frames)
            d_frames = buf
        num, sums, den, g2, dev = self._outputs(plan, dtype, torch, calc_std)
        if nb and graph:
            _lib.check(plan.lib.dxg2_correlate_dense_graph(
                plan.handle, _lib.ptr(d_frames), _lib.ptr(num), _lib.ptr(sums), _lib.ptr(den),
                _lib.ptr(g2), current_stream(torch)), "dxg2_correlate_dense_graph")
        elif nb and calc_std:
            _lib.check(plan.lib.dxg2_correlate_dense_std(
                plan.handle, _lib.ptr(d_frames), _lib.ptr(num), _lib.ptr(sums), _lib.ptr(den),
                _lib.ptr(g2), _lib.ptr(dev), current_stream(torch)), "dxg2_correlate_dense_std")
        elif nb:
            _lib.check(plan.lib.dxg2_correlate_dense(plan.handle, _lib.ptr(d_frames), _lib.ptr(num),
                                                     _lib.ptr(sums), _lib.ptr(den), _lib.ptr(g2),
                                                     current_stream(torch)),
                       "dxg2_correlate_dense")
        self.last_numerators, self.last_sums, self.last_denominators, self.last_g2 = num, sums, den, g2
        if graph and self.extra_options.get("return_device", False):
            g2 = g2.clone()  # the persistent buffer is rewritten by the next call
        if not self.extra_options.get("return_device", False):
            g2 = g2.cpu().numpy()
            dev = dev.cpu().numpy() if calc_std else None
        return CorrelationResult(g2, dev) if calc_std else g2


    def two_time(self, frames, q):
        """Two-time correlation matrix of bin q (1-based): float32 (nframes, nframes),
        the `trc` the reference's event correlator ecorrts_q returns for ttcf_par = q
        (event_y.py:169-230 with WXPCS/fecorrt.f:1-26: n_b every pixel of the bin, the
        pair counts symmetrised, the diagonal rule), from the exact integer Gram in FP64."""
        torch = _torch()
        d_frames, dtype = to_device_frames(frames, self.shape, self.nframes, torch)
        plan = self._plan(dtype)
        n = self.nframes
        trc = torch.empty((n, n), dtype=torch.float32, device="cuda")
        _lib.check(plan.lib.dxg2_correlate_dense_ttcf(plan.handle, _lib.ptr(d_frames), int(q),
                                                      _lib.ptr(trc), current_stream(torch)),
                   "dxg2_correlate_dense_ttcf")
        if self.extra_options.get("return_device", False):
            return trc
        return trc.cpu().numpy()


class MatMulCorrelator(_DenseHIP):
    """Drop-in for MatMulCorrelator (dense.py:118-142).  scale_factor is accepted and,
    as in the reference, not used: g2 is normalised by the pixel count of each bin."""

    def __init__(self, shape, nframes, qmask=None, scale_factor=None, extra_options={}):
        super().__init__()
        super()._set_parameters(shape, nframes, qmask, scale_factor, extra_options)
        self._init_dense()

    def correlate(self, frames, calc_std=False):
        """g2 (n_bins, nframes) float32; with calc_std a CorrelationResult(res, dev) whose
        dev(tau) is the population std over t of the ratios n_b G(t, t+tau) / (S(t)
        S(t+tau)) divided by sqrt(nframes - tau) (dense.py:58-59,65-66), computed in
        FP64 from the exact integer Gram."""
        return self._run_dense(frames, calc_std=bool(calc_std))


class DenseCorrelator(_DenseHIP):
    """Drop-in for the OpenCL DenseCorrelator (dense.py:145-247).

    `dtype` is the frames dtype (default "f", float32, as in the reference);
    correlate() asserts numpy frames match it, as the reference's _set_data does
    (common.py:178-180).  Frames of any numeric dtype run on the exact integer path in
    the narrowest of uint8 / int8 / uint16 that holds their values (narrow_counts:
    integer counts in [0, 65535] or [-128, 127], else ValueError).  ctx/devicetype/platformid/
    deviceid/block_size/memory are OpenCL device selection and are ignored; `profile`
    turns on per-stage HIP event timing (see `stage_times`).
    """

    def __init__(self, shape, nframes, qmask=None, dtype="f", weights=None, extra_options={},
                 ctx=None, devicetype="all", platformid=None, deviceid=None,
                 block_size=None, memory=None, profile=False):
        super().__init__()
        super()._set_parameters(shape, nframes, qmask, None, extra_options)
        self._set_weights(weights)
        self.dtype = np.dtype(dtype)
        self.profile = bool(profile)
        self._init_dense()

    def correlate(self, frames):
        dt = frames.dtype if isinstance(frames, np.ndarray) else None
        if dt is not None:
            assert np.dtype(dt) == self.dtype, "frames dtype %s != %s" % (dt, self.dtype)
        return self._run_dense(frames)

    def _plan(self, dtype):
        fresh = np.dtype(dtype) not in self._plans
        p = super()._plan(dtype)
        if fresh and self.profile:
            p.set_profiling(True)
        return p

    def stage_times(self):
        """Average [gather, gram, finalise, total] ms and the number of calls averaged,
        over the calls since the last read (profile=True), of the plan the last call used
        (frames of dtype "f" run on the narrowest exact integer type of their values)."""
        if self._last_plan is None:
            return [0.0] * 5
        return self._last_plan.stage_times()


class PlainMatMulCorrelator(_DenseHIP):
    """One bin per launch, FP32 Gram: the structure of CublasMatMulCorrelator
    (cuda.py:101-125) written as this library's own HIP kernels.  Used as the
    comparator of the batched integer path."""

    def __init__(self, shape, nframes, qmask=None, scale_factor=None, extra_options={}):
        super().__init__()
        super()._set_parameters(shape, nframes, qmask, scale_factor, extra_options)
        self._init_dense()

    def correlate(self, frames):
        torch = _torch()
        d_frames, dtype = to_device_frames(frames, self.shape, self.nframes, torch)
        plan = self._plan(dtype)
        g2 = torch.empty((max(self.n_bins, 0), self.nframes), dtype=torch.float32, device="cuda")
        if self.n_bins:
            _lib.check(plan.lib.dxg2_correlate_dense_plain(plan.handle, _lib.ptr(d_frames),
                                                           _lib.ptr(g2), current_stream(torch)),
                       "dxg2_correlate_dense_plain")
        self.last_g2 = g2
        if self.extra_options.get("return_device", False):
            return g2
        return g2.cpu().numpy()


# Half-lengths L (complex points) of the in-LDS FFT plans: 2^a 5^b, 1024 <= L <= 16384.
FFT_HALF_LENGTHS = (1024, 1280, 2048, 2560, 4096, 5120, 8192, 10240, 16384)


# Long series (N > 16384): L = L1 L2, a row length L1 and a column count L2.
FFT_ROW_LENGTHS = (8192, 5120, 4096, 2560, 2048, 1280, 1024)
FFT_COL_LENGTHS = (2, 4, 5, 8, 16)


def fft_length(nframes):
    """Nf = 2 L for the smallest L of FFT_HALF_LENGTHS with L >= nframes, or above 16384
    the smallest product L1 L2 >= nframes of FFT_ROW_LENGTHS x FFT_COL_LENGTHS: Nf >= 2 N - 1
    is all the linear autocorrelation needs (the reference uses get_next_power(2N),
    utils.py:50-60, another smooth length of the same kind).  0 when N > 131072."""
    for L in FFT_HALF_LENGTHS:
        if L >= nframes:
            return 2 * L
    long = [a * b for a in FFT_ROW_LENGTHS for b in FFT_COL_LENGTHS if a * b >= nframes]
    return 2 * min(long) if long else 0


class FFTCorrelator(_DenseHIP):
    """Drop-in for FFTWCorrelator / CUFFTCorrelator (dense.py:275-369, cuda.py:129-219).

    Frames of any dtype the exact paths take (uint8, int8, uint16, or wider types
    narrowed to them), nframes <= 131072 (above 16384 the transform is a four-step FFT
    through an HBM scratch of at most DXG2_FFT_SCRATCH_MB MiB).  `weights` and `precompute_fft_plans` are accepted;
    weights are ignored as in the reference FFT correlators (dense.py:277-285), and the
    FFT needs no plan.  scale_factor is ignored (the reference divides by the pixel count,
    dense.py:366).  Numerators carry FP32 FFT rounding: |Num - Num_exact| <= 2e-7 * Num(0)
    (asserted by the tests; measured <= 4.8e-8 for N up to 10,000); the denominator is
    exact.
    After each call `last_numerators` holds the FP64 numerators.
    """

    def __init__(self, shape, nframes, qmask=None, weights=None, scale_factor=None,
                 precompute_fft_plans=False, extra_options={}):
        super().__init__()
        super()._set_parameters(shape, nframes, qmask, scale_factor, extra_options)
        self._init_dense()
        self.precompute_fft_plans = precompute_fft_plans
        self.Nf = fft_length(self.nframes)
        self.n_pix = {int(b): int(n) for b, n in zip(self.bins, self.pixel_counts)}

    def correlate(self, frames):
        torch = _torch()
        d_frames, dtype = to_device_frames(frames, self.shape, self.nframes, torch)
        plan = self._plan(dtype)
        nb, n = max(self.n_bins, 0), self.nframes
        num = torch.empty((nb, n), dtype=torch.float64, device="cuda")
        sums = torch.empty((nb, n), dtype=torch.int64, device="cuda")
        den = torch.empty((nb, n), dtype=torch.float64, device="cuda")
        g2 = torch.empty((nb, n), dtype=torch.float32, device="cuda")
        if nb:
            _lib.check(plan.lib.dxg2_correlate_fft(plan.handle, _lib.ptr(d_frames), _lib.ptr(num),
                                                   _lib.ptr(sums), _lib.ptr(den), _lib.ptr(g2),
                                                   current_stream(torch)),
                       "dxg2_correlate_fft")
        self.last_numerators, self.last_sums, self.last_denominators, self.last_g2 = num, sums, den, g2
        if self.extra_options.get("return_device", False):
            return g2
        return g2.cpu().numpy()


FFTWCorrelator = FFTCorrelator
