// FFT correlator: numerator from power spectra along the time axis.
//
// Reference: FFTWCorrelator._correlate_fft (correlator/dense.py:347-369) and
// CUFFTCorrelator (cuda.py:185-219): per bin, a float32 FFT of every pixel series
// zero-padded to Nf >= 2N (dense.py:289), times the FFT of the reversed series, one
// inverse FFT per pixel, a sum over pixels; the denominator is np.correlate of the
// integer frame sums (dense.py:365-366).
//
// Here, per bin:
//   Num(tau) = IFFT_Nf( sum_p |X_p(f)|^2 )[tau],   X_p = FFT_Nf(x_p zero-padded),
// (the product with the reversed series is the same autocorrelation; the inverse is
// linear, so one inverse per bin replaces one per pixel).
//  * gather_pm_kernel: frames -> pixel-major bytes xs[k][t] (k in the bin's raster
//    list order, t < nrows, zero past the last frame), 64 pixels x 256 frames per
//    workgroup through an LDS transpose.
//  * pm_sums_kernel: exact S_b(t) from xs.
//  * fft_power_kernel: one workgroup per 64 pixels of a bin; each pixel's series is
//    packed as z[n] = x[2n] + i x[2n+1] (L = Nf/2 complex points), transformed by an
//    FP32 Stockham FFT in LDS (Govindaraju et al., "High performance discrete Fourier
//    transforms on graphics processors", SC 2008: radix-4 passes, one radix-2 pass when
//    log2 L is odd), unpacked to X[k], k = 0..L, and |X[k]|^2 is accumulated in FP64.
//  * fft_inverse_kernel: Num(tau) = (P_0 + (-1)^tau P_L + 2 sum_{k=1}^{L-1} P_k
//    cos(2 pi k tau / Nf)) / Nf in FP64, the cosine by an FP64 rotation recurrence.
// Twiddle tables are computed on the host in FP64 and rounded to FP32.
#include "dxg2_internal.h"

namespace dxg2 {

// threads per FFT workgroup: L/4 butterflies per radix-4 pass, at most 1024 threads
template <int L>
struct FftCfg {
    static constexpr int THREADS = L / 4 < 1024 ? L / 4 : 1024;
};

// --------------------------------------------------------------- pixel-major gather
// grid (bin k-blocks, frame tiles of 256).  Lane = pixel j of the k-block; each wave
// stages 64 frames, the LDS tile is [64 pixels][256 frames + 4].
__global__ __launch_bounds__(256) void gather_pm_kernel(const uint8_t* __restrict__ frames,
                                                        int dtype, int64_t hw, int nframes,
                                                        int nrows, const int32_t* __restrict__ pix_idx,
                                                        const BinDev* __restrict__ bins,
                                                        const KbWork* __restrict__ work,
                                                        uint8_t* __restrict__ xs) {
    __shared__ uint8_t tile[64][256 + 4];
    const KbWork w = work[blockIdx.x];
    const BinDev bd = bins[w.bin];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int t0 = blockIdx.y * 256;
    const int p = pix_idx[bd.idx_off + (int64_t)w.kb * KB + lane];
    const int pc = p >= 0 ? p : 0;
    for (int f = wave; f < 256; f += 4) {
        const int t = t0 + f;
        uint8_t v = 0;
        if (t < nframes && p >= 0) {
            if (dtype == DXG2_U16) {
                const uint16_t x = reinterpret_cast<const uint16_t*>(frames)[(int64_t)t * hw + pc];
                v = x > 255 ? 255 : (uint8_t)x;   // not reached: the FFT path takes 8-bit input
            } else {
                v = frames[(int64_t)t * hw + pc];
            }
        }
        tile[lane][f] = v;
    }
    __syncthreads();
    // write: pixel row j gets 256 contiguous bytes; 64 threads per 1 KiB? -> thread
    // writes 16 bytes: pixel j = tid / 16, bytes (tid % 16) * 16 .. +16 of the tile row
    for (int c = threadIdx.x; c < 64 * 16; c += 256) {
        const int j = c >> 4;
        const int o = (c & 15) * 16;
        uint32_t wd[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            wd[q] = (uint32_t)tile[j][o + 4 * q] | ((uint32_t)tile[j][o + 4 * q + 1] << 8) |
                    ((uint32_t)tile[j][o + 4 * q + 2] << 16) |
                    ((uint32_t)tile[j][o + 4 * q + 3] << 24);
        uint8_t* dst = xs + bd.slab_off + ((int64_t)w.kb * KB + j) * nrows + t0 + o;
        *reinterpret_cast<uint4*>(dst) = make_uint4(wd[0], wd[1], wd[2], wd[3]);
    }
}

// S_b(t) = sum over the bin's pixels of xs[k][t].
// grid (frame tiles of 256, bins, pixel chunks of PM_CH).  Thread (r, g) = (tid / 16,
// tid % 16) reads 16 frames (16 bytes) of pixel rows chunk + r, r + 16, ... and adds
// them into 16-bit lanes (even and odd bytes of each word apart: at most PM_CH / 16 rows
// of 255, below 2^16); the 16 row groups are reduced in LDS and each frame's chunk total
// goes to sums (zeroed first) with one 64-bit atomic add.
constexpr int PM_CH = 1024;

__global__ __launch_bounds__(256) void pm_sums_kernel(const uint8_t* __restrict__ xs,
                                                      const BinDev* __restrict__ bins, int nframes,
                                                      int nrows, int64_t* __restrict__ sums) {
    __shared__ uint32_t part[16][256 + 1];
    const int b = blockIdx.y;
    const BinDev bd = bins[b];
    const int k0 = blockIdx.z * PM_CH;
    if (k0 >= bd.npix) return;
    const int k1 = min(bd.npix, k0 + PM_CH);
    const int t0 = blockIdx.x * 256;
    const int r = threadIdx.x >> 4, g = threadIdx.x & 15;
    uint32_t ev[4] = {0u, 0u, 0u, 0u}, od[4] = {0u, 0u, 0u, 0u};
    const uint8_t* base = xs + bd.slab_off + t0 + 16 * g;
    for (int k = k0 + r; k < k1; k += 16) {
        const uint4 v = *reinterpret_cast<const uint4*>(base + (int64_t)k * nrows);
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            ev[q] += w[q] & 0x00FF00FFu;
            od[q] += (w[q] >> 8) & 0x00FF00FFu;
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {  // frame 16 g + 4 q + {0, 1, 2, 3}
        part[r][16 * g + 4 * q + 0] = ev[q] & 0xFFFFu;
        part[r][16 * g + 4 * q + 1] = od[q] & 0xFFFFu;
        part[r][16 * g + 4 * q + 2] = ev[q] >> 16;
        part[r][16 * g + 4 * q + 3] = od[q] >> 16;
    }
    __syncthreads();
    const int t = t0 + threadIdx.x;
    if (t < nframes) {
        uint32_t tot = 0;
#pragma unroll
        for (int i = 0; i < 16; ++i) tot += part[i][threadIdx.x];
        if (tot) atomicAdd(reinterpret_cast<unsigned long long*>(sums + (int64_t)b * nframes + t),
                           (unsigned long long)tot);
    }
}

// --------------------------------------------------------------------- forward FFT
struct c32 {
    float x, y;
};
__device__ __forceinline__ c32 cmul(c32 a, c32 b) {
    return c32{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__device__ __forceinline__ c32 cadd(c32 a, c32 b) { return c32{a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ c32 csub(c32 a, c32 b) { return c32{a.x - b.x, a.y - b.y}; }
// -i * a
__device__ __forceinline__ c32 cmuli_neg(c32 a) { return c32{a.y, -a.x}; }

// grid: one workgroup per (bin, k-block) of 64 pixels.  LDS holds L complex points.
// tw[m] = exp(-2 pi i m / L), m < L, and tw[L + k] = exp(-pi i k / L), k <= L, the
// unpacking factor (host FP64, rounded to FP32).
template <int L>
__global__ __launch_bounds__(FftCfg<L>::THREADS) void fft_power_kernel(
    const uint8_t* __restrict__ xs, const BinDev* __restrict__ bins,
    const KbWork* __restrict__ work, int nframes, int nrows, const float2* __restrict__ tw,
    double* __restrict__ power) {
    __shared__ c32 z[L];
    constexpr int FFT_THREADS = FftCfg<L>::THREADS;
    constexpr int LOG2 = __builtin_ctz(L);
    constexpr int PER = L / FFT_THREADS;          // complex points per thread (load)
    constexpr int BF4 = L / 4 / FFT_THREADS;      // radix-4 butterflies per thread per pass
    constexpr int NK = (L + 1 + FFT_THREADS - 1) / FFT_THREADS;  // power bins per thread
    const KbWork w = work[blockIdx.x];
    const BinDev bd = bins[w.bin];
    const int tid = threadIdx.x;
    double acc[NK];
#pragma unroll
    for (int m = 0; m < NK; ++m) acc[m] = 0.0;

    const int kend = min(KB, bd.npix - w.kb * KB);
    for (int j = 0; j < kend; ++j) {
        const uint8_t* x = xs + bd.slab_off + ((int64_t)w.kb * KB + j) * nrows;
        // load z[n] = x[2n] + i x[2n+1] (zero past the last frame)
#pragma unroll
        for (int m = 0; m < PER; ++m) {
            const int n = tid + m * FFT_THREADS;
            const int t = 2 * n;
            float re = 0.f, im = 0.f;
            if (t + 1 < nrows) {
                const uint16_t two = *reinterpret_cast<const uint16_t*>(x + t);
                re = t < nframes ? (float)(two & 0xFF) : 0.f;
                im = t + 1 < nframes ? (float)(two >> 8) : 0.f;
            }
            z[n] = c32{re, im};
        }
        __syncthreads();
        int ns = 1;
        if (LOG2 & 1) {  // one radix-2 pass, Ns = 1
            c32 v[PER / 2][2];
#pragma unroll
            for (int m = 0; m < PER / 2; ++m) {
                const int jj = tid + m * FFT_THREADS;
                v[m][0] = z[jj];
                v[m][1] = z[jj + L / 2];
            }
            __syncthreads();
#pragma unroll
            for (int m = 0; m < PER / 2; ++m) {
                const int jj = tid + m * FFT_THREADS;
                z[2 * jj] = cadd(v[m][0], v[m][1]);
                z[2 * jj + 1] = csub(v[m][0], v[m][1]);
            }
            __syncthreads();
            ns = 2;
        }
        // kept rolled: unrolled, the twiddle loads of every pass would be hoisted above
        // the barriers and overflow the register file
#pragma unroll 1
        for (; ns < L; ns *= 4) {
            c32 v[BF4][4];
            const int tstep = L / (4 * ns);   // twiddle index step: exp(-2 pi i r k / (4 ns))
#pragma unroll
            for (int m = 0; m < BF4; ++m) {
                const int jj = tid + m * FFT_THREADS;
                const int k = jj % ns;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    c32 a = z[jj + r * (L / 4)];
                    if (r) {
                        const float2 t = tw[(r * k * tstep) & (L - 1)];
                        a = cmul(a, c32{t.x, t.y});
                    }
                    v[m][r] = a;
                }
            }
            __syncthreads();
#pragma unroll
            for (int m = 0; m < BF4; ++m) {
                const int jj = tid + m * FFT_THREADS;
                const int k = jj % ns;
                const c32 s02 = cadd(v[m][0], v[m][2]), d02 = csub(v[m][0], v[m][2]);
                const c32 s13 = cadd(v[m][1], v[m][3]), d13 = cmuli_neg(csub(v[m][1], v[m][3]));
                const int o = (jj / ns) * ns * 4 + k;
                z[o] = cadd(s02, s13);
                z[o + ns] = cadd(d02, d13);
                z[o + 2 * ns] = csub(s02, s13);
                z[o + 3 * ns] = csub(d02, d13);
            }
            __syncthreads();
        }
        // unpack X[k] = (Z[k] + conj Z[L-k]) / 2 - i W^k (Z[k] - conj Z[L-k]) / 2,
        // W = exp(-pi i / L); accumulate |X[k]|^2
#pragma unroll
        for (int m = 0; m < NK; ++m) {
            const int k = tid + m * FFT_THREADS;
            if (k <= L) {
                const c32 a = z[k & (L - 1)];
                const c32 b = z[(L - k) & (L - 1)];
                const c32 s = c32{a.x + b.x, a.y - b.y};   // Z[k] + conj Z[L-k]
                const c32 d = c32{a.x - b.x, a.y + b.y};   // Z[k] - conj Z[L-k]
                const float2 u = tw[L + k];                 // W^k = exp(-pi i k / L)
                const c32 wd = cmul(c32{u.x, u.y}, d);
                const float xr = 0.5f * (s.x + wd.y);      // s/2 - i wd/2
                const float xi = 0.5f * (s.y - wd.x);
                acc[m] += (double)xr * xr + (double)xi * xi;
            }
        }
        __syncthreads();
    }
    double* prow = power + (int64_t)w.bin * (L + 1);
#pragma unroll
    for (int m = 0; m < NK; ++m) {
        const int k = tid + m * FFT_THREADS;
        if (k <= L && acc[m] != 0.0) atomicAdd(prow + k, acc[m]);
    }
}

// L = 16384 (the largest series): radix-16 Stockham passes with the 16-point DFT in
// registers, 16384 = 16 * 16 * 16 * 4, so four passes and eight barriers per series
// instead of seven radix-4 passes.  Thread j of 1024 takes butterfly j of each radix-16
// pass (16 complex values, 32 VGPRs) and four butterflies of the final radix-4 pass.
// A pass with stride ns reads z[j + r L/R], multiplies by exp(-2 pi i r k / (R ns)),
// k = j mod ns, and writes output q to z[(j - k) R + k + q ns].  The first pass writes
// 16 apart, so LDS index i is stored at i + i/16: consecutive lanes then land on
// distinct banks.  DFT-16 = 4 x 4: DFT-4 over r1 for each r0, twiddles W16^(r0 q0),
// DFT-4 over r0 for each q0, output q0 + 4 q1.
constexpr int F16_L = 16384;
constexpr int F16_T = 1024;

__device__ __forceinline__ int fpad(int i) { return i + (i >> 4); }

__device__ __forceinline__ void dft4(c32& x0, c32& x1, c32& x2, c32& x3) {
    const c32 s02 = cadd(x0, x2), d02 = csub(x0, x2);
    const c32 s13 = cadd(x1, x3), d13 = cmuli_neg(csub(x1, x3));
    x0 = cadd(s02, s13);
    x1 = cadd(d02, d13);
    x2 = csub(s02, s13);
    x3 = csub(d02, d13);
}

// exp(-2 pi i m / 16) for m = r0 q0 in {0, 1, 2, 3, 4, 6, 9}
__device__ __forceinline__ c32 w16(int m) {
    constexpr float C1 = 0.92387953251128674f, S1 = 0.38268343236508978f;
    constexpr float H = 0.70710678118654752f;
    switch (m) {
        case 1: return c32{C1, -S1};
        case 2: return c32{H, -H};
        case 3: return c32{S1, -C1};
        case 4: return c32{0.f, -1.f};
        case 6: return c32{-H, -H};
        case 9: return c32{-C1, S1};
        default: return c32{1.f, 0.f};
    }
}

// in place, transposed: afterwards a[4 q0 + q1] = sum_r a[r] exp(-2 pi i r q / 16) for
// q = q0 + 4 q1 (the caller stores slot 4 q0 + q1 as output q0 + 4 q1)
__device__ __forceinline__ void dft16(c32 (&a)[16]) {
#pragma unroll
    for (int r0 = 0; r0 < 4; ++r0) dft4(a[r0], a[4 + r0], a[8 + r0], a[12 + r0]);
    // now a[4 q0 + r0] holds b_{r0}[q0]
#pragma unroll
    for (int r0 = 1; r0 < 4; ++r0)
#pragma unroll
        for (int q0 = 1; q0 < 4; ++q0) a[4 * q0 + r0] = cmul(a[4 * q0 + r0], w16(r0 * q0));
#pragma unroll
    for (int q0 = 0; q0 < 4; ++q0) dft4(a[4 * q0], a[4 * q0 + 1], a[4 * q0 + 2], a[4 * q0 + 3]);
}

// one radix-16 pass of stride NS (1, 16 or 256) for butterfly j (the thread index)
template <int NS>
__device__ __forceinline__ void pass16(c32* z, const float2* __restrict__ tw, int j) {
    const int k = j & (NS - 1);
    c32 a[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        c32 v = z[fpad(j + r * (F16_L / 16))];
        if (NS > 1 && r) {
            const float2 t = tw[(r * k * (F16_L / (16 * NS))) & (F16_L - 1)];
            v = cmul(v, c32{t.x, t.y});
        }
        a[r] = v;
    }
    dft16(a);
    __syncthreads();  // every read of this pass is done
    const int o = (j - k) * 16 + k;
#pragma unroll
    for (int q0 = 0; q0 < 4; ++q0)
#pragma unroll
        for (int q1 = 0; q1 < 4; ++q1) z[fpad(o + (q0 + 4 * q1) * NS)] = a[4 * q0 + q1];
    __syncthreads();
}

__global__ __launch_bounds__(F16_T) void fft16_power_kernel(
    const uint8_t* __restrict__ xs, const BinDev* __restrict__ bins,
    const KbWork* __restrict__ work, int nframes, int nrows, const float2* __restrict__ tw,
    double* __restrict__ power) {
    __shared__ c32 z[F16_L + F16_L / 16];
    constexpr int L = F16_L;
    constexpr int NK = (L + 1 + F16_T - 1) / F16_T;  // power bins per thread
    const KbWork w = work[blockIdx.x];
    const BinDev bd = bins[w.bin];
    const int tid = threadIdx.x;
    double acc[NK];
#pragma unroll
    for (int m = 0; m < NK; ++m) acc[m] = 0.0;

    const int kend = min(KB, bd.npix - w.kb * KB);
    for (int jp = 0; jp < kend; ++jp) {
        const uint8_t* x = xs + bd.slab_off + ((int64_t)w.kb * KB + jp) * nrows;
        // A thread's addresses, twiddles and unpack factors are the same for every
        // series; an opaque copy of the thread index per series keeps the compiler from
        // hoisting them all out of this loop into registers (which spilled to scratch):
        // they are recomputed, and the twiddles re-read from L1/L2, instead.
        int tj = tid;
        asm volatile("" : "+v"(tj));
        // z[n] = x[2n] + i x[2n+1] (zero past the last frame)
#pragma unroll
        for (int m = 0; m < L / F16_T; ++m) {
            const int n = tj + m * F16_T;
            const int t = 2 * n;
            float re = 0.f, im = 0.f;
            if (t + 1 < nrows) {
                const uint16_t two = *reinterpret_cast<const uint16_t*>(x + t);
                re = t < nframes ? (float)(two & 0xFF) : 0.f;
                im = t + 1 < nframes ? (float)(two >> 8) : 0.f;
            }
            z[fpad(n)] = c32{re, im};
        }
        __syncthreads();
        pass16<1>(z, tw, tj);
        pass16<16>(z, tw, tj);
        pass16<256>(z, tw, tj);
        // radix-4 pass of stride 4096: butterfly jj reads and writes z[jj + 4096 r] (jj <
        // 4096), so each thread works in place on its own four points
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const int jj = tj + m * F16_T;
            c32 v[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                c32 a = z[fpad(jj + r * (L / 4))];
                if (r) {
                    const float2 t = tw[(r * jj) & (L - 1)];
                    a = cmul(a, c32{t.x, t.y});
                }
                v[r] = a;
            }
            dft4(v[0], v[1], v[2], v[3]);
#pragma unroll
            for (int r = 0; r < 4; ++r) z[fpad(jj + r * (L / 4))] = v[r];
        }
        __syncthreads();
        // unpack X[k] = (Z[k] + conj Z[L-k]) / 2 - i W^k (Z[k] - conj Z[L-k]) / 2,
        // W = exp(-pi i / L); accumulate |X[k]|^2
#pragma unroll
        for (int m = 0; m < NK; ++m) {
            const int k = tj + m * F16_T;
            if (k <= L) {
                const c32 a = z[fpad(k & (L - 1))];
                const c32 b = z[fpad((L - k) & (L - 1))];
                const c32 s2 = c32{a.x + b.x, a.y - b.y};
                const c32 d = c32{a.x - b.x, a.y + b.y};
                const float2 u = tw[L + k];  // W^k = exp(-pi i k / L)
                const c32 wd = cmul(c32{u.x, u.y}, d);
                const float xr = 0.5f * (s2.x + wd.y);
                const float xi = 0.5f * (s2.y - wd.x);
                acc[m] += (double)xr * xr + (double)xi * xi;
            }
        }
        __syncthreads();
    }
    double* prow = power + (int64_t)w.bin * (L + 1);
#pragma unroll
    for (int m = 0; m < NK; ++m) {
        const int k = tid + m * F16_T;
        if (k <= L && acc[m] != 0.0) atomicAdd(prow + k, acc[m]);
    }
}

// Num(tau) from the power spectrum, FP64; grid (ceil(N / 256), n_bins).  Every
// thread reads the same ps[k] in step, so the spectrum is read as uniform loads.
__global__ __launch_bounds__(256) void fft_inverse_kernel(const double* __restrict__ power,
                                                          int L, int nframes,
                                                          double* __restrict__ num) {
    const int b = blockIdx.y;
    const double* ps = power + (int64_t)b * (L + 1);
    const int tau = blockIdx.x * 256 + threadIdx.x;
    if (tau >= nframes) return;
    const double nf = 2.0 * L;
    double sn, cs;
    sincospi(2.0 * tau / nf, &sn, &cs);      // rotation by 2 pi tau / Nf
    double c = cs, s = sn;                   // cos, sin of 2 pi k tau / Nf at k = 1
    double acc = 0.0;
    for (int k = 1; k < L; ++k) {
        acc += ps[k] * c;
        const double cn = c * cs - s * sn;
        s = s * cs + c * sn;
        c = cn;
        if ((k & 1023) == 0) {               // re-anchor the recurrence
            sincospi(2.0 * (double)(((int64_t)(k + 1) * tau) % (int64_t)(2 * L)) / nf, &s, &c);
        }
    }
    const double last = (tau & 1) ? -ps[L] : ps[L];
    num[(int64_t)b * nframes + tau] = (ps[0] + last + 2.0 * acc) / nf;
}

void launch_gather_pm(const uint8_t* frames, int dtype, int64_t hw, int nframes, int nrows,
                      const int32_t* pix_idx, const BinDev* bins, const KbWork* work, int n_work,
                      uint8_t* xs, hipStream_t s) {
    if (n_work <= 0) return;
    gather_pm_kernel<<<dim3((unsigned)n_work, (unsigned)(nrows / 256)), 256, 0, s>>>(
        frames, dtype, hw, nframes, nrows, pix_idx, bins, work, xs);
}

void launch_pm_sums(const uint8_t* xs, const BinDev* bins, int n_bins, int max_npix, int nframes,
                    int nrows, int64_t* sums, hipStream_t s) {
    if (n_bins <= 0 || max_npix <= 0) return;
    launch_fill_i64(sums, (int64_t)n_bins * nframes, s);
    pm_sums_kernel<<<dim3((unsigned)((nframes + 255) / 256), (unsigned)n_bins,
                          (unsigned)((max_npix + PM_CH - 1) / PM_CH)),
                     256, 0, s>>>(xs, bins, nframes, nrows, sums);
}

int launch_fft_power(int L, const uint8_t* xs, const BinDev* bins, const KbWork* work,
                     int n_work, int nframes, int nrows, const float2* tw, double* power,
                     hipStream_t s) {
    if (n_work <= 0) return 0;
#define DXG2_FFT_CASE(LL)                                                                     \
    case LL:                                                                                  \
        fft_power_kernel<LL><<<dim3((unsigned)n_work), FftCfg<LL>::THREADS, 0, s>>>(          \
            xs, bins, work, nframes, nrows, tw, power);                                       \
        return 0;
    switch (L) {
        DXG2_FFT_CASE(1024)
        DXG2_FFT_CASE(2048)
        DXG2_FFT_CASE(4096)
        DXG2_FFT_CASE(8192)
        case F16_L:
            fft16_power_kernel<<<dim3((unsigned)n_work), F16_T, 0, s>>>(xs, bins, work, nframes,
                                                                       nrows, tw, power);
            return 0;
        default:
            return -1;
    }
#undef DXG2_FFT_CASE
}

void launch_fft_inverse(const double* power, int L, int n_bins, int nframes, double* num,
                        hipStream_t s) {
    if (n_bins <= 0) return;
    fft_inverse_kernel<<<dim3((unsigned)((nframes + 255) / 256), (unsigned)n_bins), 256, 0, s>>>(
        power, L, nframes, num);
}

}  // namespace dxg2
