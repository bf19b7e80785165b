"""The reference's GPU correlator names (dynamix/correlator/cuda.py) on the MI355X.

CublasMatMulCorrelator  -- cuda.py:19-125, "a CUDA-accelerated version of
                           MatMulCorrelator": same constructor, correlate(frames) ->
                           float32 (n_bins, nframes).  Here it is the batched exact
                           integer path of MatMulCorrelator (every bin in one i8-MFMA Gram
                           launch).  The reference's own structure -- one bin per launch,
                           FP32 sgemm Gram, FP32 diagonal sums and mean outer product -- is
                           PlainMatMulCorrelator, kept as the comparator.
CUFFTCorrelator         -- cuda.py:129-219: same constructor as FFTWCorrelator; the FP32
                           time-axis FFT path of FFTCorrelator.

The reference raises ImportError when pycuda / scikit-cuda are missing
(cuda.py:40-41,141-142); nothing here needs them.
"""
from .dense import FFTCorrelator, MatMulCorrelator, PlainMatMulCorrelator  # noqa: F401


class CublasMatMulCorrelator(MatMulCorrelator):
    """Drop-in for CublasMatMulCorrelator (cuda.py:19-125).  extra_options such as
    "cublas_handle" are accepted and ignored."""

    def __init__(self, shape, nframes, qmask=None, scale_factor=None, extra_options={}):
        super().__init__(shape, nframes, qmask=qmask, scale_factor=scale_factor,
                         extra_options=extra_options)

    def correlate(self, frames):
        return super().correlate(frames)


class CUFFTCorrelator(FFTCorrelator):
    """Drop-in for CUFFTCorrelator (cuda.py:129-219)."""

    def __init__(self, shape, nframes, qmask=None, weights=None, scale_factor=None,
                 precompute_fft_plans=False, extra_options={}):
        super().__init__(shape, nframes, qmask=qmask, weights=weights,
                         scale_factor=scale_factor, precompute_fft_plans=precompute_fft_plans,
                         extra_options=extra_options)
