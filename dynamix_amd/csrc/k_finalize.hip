// Exact denominator and g2 normalisation.
//
// Reference: Den_b(tau)/n_b from the per-frame bin sums -- float32 correlate_1D
// (resources/opencl/densecorrelator.cl:156-176) then _normalize_sums (correlator/dense.py:
// 215-221); np.correlate on the integer sums in the FFT path (dense.py:365-366); float32
// normalize_correlation in the event path (evtcorrelator.cl:58-76); the ratio
// res = Num / (Den / n_b) (dense.py:64, densecorrelator.cl:54).
//
// Here Den_b(tau) = sum_t S(t) S(t+tau) is accumulated in 128-bit integers (exact for
// every input size), Num_b(tau) is completed from the raw i8 Gram diagonals with the
// uint8 bias correction, and g2 = scale_b * Num / Den is formed once in FP64.
//
// Bias correction (uint8 frames are stored as s = v - 128):
//   G(t,t') = U(t,t') + 128 (S(t) + S(t')) - 2^14 K_pad
// summed over t = 0..N-1-tau:
//   Num(tau) = U(tau) + 128 (sum_{t<N-tau} S(t) + sum_{t>=tau} S(t)) - 2^14 K_pad (N-tau).
#include "dxg2_internal.h"

namespace dxg2 {

constexpr int FIN_CH = 1024;  // frames of S staged per pass

__device__ __forceinline__ double i128_to_double(__int128 x) {
    const long long hi = (long long)(x >> 64);
    const unsigned long long lo = (unsigned long long)x;
    return (double)hi * 18446744073709551616.0 + (double)lo;
}

// grid (ceil(N/256), n_bins), 256 threads: one thread per lag tau.  NumT is int64_t
// (exact numerators of the Gram and event paths) or double (FFT path, no bias).
template <typename NumT>
__global__ __launch_bounds__(256) void finalize_kernel(const int64_t* __restrict__ sums,
                                                       const NumT* __restrict__ ugram,
                                                       const BinDev* __restrict__ bins,
                                                       int nframes, int biased,
                                                       NumT* __restrict__ num,
                                                       double* __restrict__ den,
                                                       float* __restrict__ g2) {
    __shared__ int64_t sa[FIN_CH];
    __shared__ int64_t sb[FIN_CH + 256];
    const int b = blockIdx.y;
    const int tau0 = blockIdx.x * 256;
    const int tau = tau0 + threadIdx.x;
    const int64_t* S = sums + (int64_t)b * nframes;

    __int128 acc = 0;
    int64_t lo_sum = 0;  // sum_{t < N - tau} S(t)
    int64_t hi_sum = 0;  // sum_{t < N - tau} S(t + tau)
    const int tmax = nframes - tau0;  // frames t needed by the smallest lag of the block
    for (int t0 = 0; t0 < tmax; t0 += FIN_CH) {
        __syncthreads();
        for (int i = threadIdx.x; i < FIN_CH; i += 256) {
            const int t = t0 + i;
            sa[i] = t < nframes ? S[t] : 0;
        }
        for (int i = threadIdx.x; i < FIN_CH + 256; i += 256) {
            const int t = t0 + tau0 + i;
            sb[i] = t < nframes ? S[t] : 0;
        }
        __syncthreads();
        int lim = nframes - tau - t0;
        if (lim > FIN_CH) lim = FIN_CH;
        for (int k = 0; k < lim; ++k) {
            const int64_t x = sa[k];
            const int64_t y = sb[k + threadIdx.x];
            acc += (__int128)x * (__int128)y;
            lo_sum += x;
            hi_sum += y;
        }
    }
    if (tau >= nframes) return;
    const BinDev bd = bins[b];
    NumT n = ugram[(int64_t)b * nframes + tau];
    if (biased)
        n += (NumT)(128 * (lo_sum + hi_sum) -
                    (int64_t)16384 * ((int64_t)bd.nkb * KB) * (nframes - tau));
    const double d = i128_to_double(acc);
    const int64_t o = (int64_t)b * nframes + tau;
    if (num) num[o] = n;
    if (den) den[o] = d;
    g2[o] = (float)(bd.scale * ((double)n / d));
}

void launch_finalize(const int64_t* sums, const int64_t* ugram, const BinDev* bins,
                     int n_bins, int nframes, int biased, int64_t* num, double* den,
                     float* g2, hipStream_t s) {
    if (n_bins <= 0 || nframes <= 0) return;
    dim3 grid((unsigned)((nframes + 255) / 256), (unsigned)n_bins);
    finalize_kernel<int64_t><<<grid, dim3(256), 0, s>>>(sums, ugram, bins, nframes, biased, num,
                                                        den, g2);
}

void launch_finalize_f64(const int64_t* sums, const double* fnum, const BinDev* bins, int n_bins,
                         int nframes, double* num, double* den, float* g2, hipStream_t s) {
    if (n_bins <= 0 || nframes <= 0) return;
    dim3 grid((unsigned)((nframes + 255) / 256), (unsigned)n_bins);
    finalize_kernel<double><<<grid, dim3(256), 0, s>>>(sums, fnum, bins, nframes, 0, num, den, g2);
}

}  // namespace dxg2
