// Exact denominator and g2 normalisation.
//
// Reference: Den_b(tau)/n_b from the per-frame bin sums -- float32 correlate_1D
// (resources/opencl/densecorrelator.cl:156-176) then _normalize_sums (correlator/dense.py:
// 215-221); np.correlate on the integer sums in the FFT path (dense.py:365-366); float32
// normalize_correlation in the event path (evtcorrelator.cl:58-76); the ratio
// res = Num / (Den / n_b) (dense.py:64, densecorrelator.cl:54).
//
// Here Den_b(tau) = sum_t S(t) S(t+tau) is exact, Num_b(tau) is completed from the raw i8
// Gram diagonals with the uint8 bias correction, and g2 = scale_b * Num / Den is formed
// once in FP64.
//
// Exactness: a per-bin prep pass forms the prefix sums P(t) = sum_{u<t} S(u) and
// M = max_t |S(t)|.  When N M^2 < 2^52 every product and every partial sum of Den is an
// integer below 2^53, so FP64 fused multiply-adds are exact (den_fp64_kernel, register
// lag windows); otherwise the lag loop accumulates in 128-bit integers.  Both give the
// same Den.
//
// Bias correction (uint8 frames are stored as s = v - 128, padded pixels as -128):
//   G(t,t') = U(t,t') + 128 (S(t) + S(t')) - 2^14 K_pad
// summed over t = 0..N-1-tau:
//   Num(tau) = U(tau) + 128 (P(N-tau) + P(N) - P(tau)) - 2^14 K_pad (N-tau).
// uint16 frames v = 256 h + l are stored as two biased digit planes; with the plane
// sums S_H, S_L (prefix sums P_H, P_L) the same identity gives (mode 2; mode 3, below,
// when HH and X are clipped to the wide k-blocks)
//   HH = U_HH + 128 E_H - 2^14 K_pad (N-tau),  LL = U_LL + 128 E_L - 2^14 K_pad (N-tau),
//   X  = U_X + 128 (E_H + E_L) - 2^15 K_pad (N-tau),   E = P(N-tau) + P(N) - P(tau),
// and Num = 65536 HH + 256 X + LL (formed in 128 bits), S = 256 S_H + S_L.
#include "dxg2_internal.h"

namespace dxg2 {

constexpr int FIN_CH = 1024;       // frames of S staged per pass
constexpr int FIN_THREADS = 256;
constexpr int FIN_LAGS = 4;        // lags per thread: tau0 + tid + 256 j, j < FIN_LAGS
constexpr int FIN_SPAN = FIN_THREADS * FIN_LAGS;  // lags per block
constexpr int PREP_THREADS = 1024;

__device__ __forceinline__ double i128_to_double(__int128 x) {
    const long long hi = (long long)(x >> 64);
    const unsigned long long lo = (unsigned long long)x;
    return (double)hi * 18446744073709551616.0 + (double)lo;
}

__device__ __forceinline__ int64_t wave_sum64(int64_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Exclusive prefix scan of one value per thread over the block; returns the prefix and
// adds the block total to `carry`.  wave_tot: PREP_THREADS / 64 words of LDS.
__device__ __forceinline__ int64_t block_scan(int64_t v, int64_t* wave_tot, int64_t& carry) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int64_t x = v;  // inclusive scan within the wave
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int64_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) wave_tot[wave] = x;
    __syncthreads();
    int64_t before = carry, total = carry;
    for (int w = 0; w < PREP_THREADS / 64; ++w) {
        if (w < wave) before += wave_tot[w];
        total += wave_tot[w];
    }
    __syncthreads();
    carry = total;
    return before + x - v;
}

// Block (bin x, chunk y).  S(t) are the rows of plane_sums; PLANES >= 2 (uint16 digit
// planes): rows b (S_H) and n_bins + b (S_L), combined into sums[b][t] = 256 S_H + S_L;
// PLANES = 3 adds rows 2 n_bins + b (S_Lw, the low bytes over the wide k-blocks, which
// only the bias correction uses).  Writes the exclusive prefix sums of each plane,
// prefix[plane * n_bins + b][t] for t = 0..N, and smax[b] = max |S| (by atomic max into
// a row zeroed first).  A block scans PREP_CHUNK frames: its starting offsets are the sums
// of every earlier frame (a block reduction over them, so no pass waits for another
// block), then each pass takes PREP_ITEMS consecutive frames per thread -- fewer with more
// planes, to stay in registers -- with a serial prefix in registers and one block scan of
// the thread totals.
constexpr int PREP_CHUNK = 16384;
template <int PLANES>
__global__ __launch_bounds__(PREP_THREADS) void sums_prep_kernel(
    const int64_t* __restrict__ plane_sums, int64_t* __restrict__ sums, int nframes,
    int64_t* __restrict__ prefix, int64_t* __restrict__ smax) {
    __shared__ int64_t wave_tot[PREP_THREADS / 64];
    __shared__ int64_t wave_max[PREP_THREADS / 64];
    __shared__ int64_t red[3][PREP_THREADS / 64];
    constexpr bool WIDE = PLANES >= 2;
    constexpr int PREP_ITEMS = PLANES == 1 ? 8 : (PLANES == 2 ? 4 : 1);
    const int b = blockIdx.x, nb = gridDim.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t* S0 = plane_sums + (int64_t)b * nframes;
    const int64_t* S1 = plane_sums + (int64_t)(nb + b) * nframes;
    const int64_t* S2 = plane_sums + (int64_t)(2 * nb + b) * nframes;
    int64_t* P0 = prefix + (int64_t)b * (nframes + 1);
    int64_t* P1 = prefix + (int64_t)(nb + b) * (nframes + 1);
    int64_t* P2 = prefix + (int64_t)(2 * nb + b) * (nframes + 1);
    const int cb = blockIdx.y * PREP_CHUNK;
    const int ce = min(nframes, cb + PREP_CHUNK);
    // offsets: the planes' sums over frames [0, cb)
    int64_t c0 = 0, c1 = 0, c2 = 0, mx = 0;
    for (int t = tid; t < cb; t += PREP_THREADS) {
        c0 += S0[t];
        if (WIDE) c1 += S1[t];
        if (PLANES == 3) c2 += S2[t];
    }
    c0 = wave_sum64(c0);
    if (WIDE) c1 = wave_sum64(c1);
    if (PLANES == 3) c2 = wave_sum64(c2);
    if (lane == 0) {
        red[0][wave] = c0;
        red[1][wave] = c1;
        red[2][wave] = c2;
    }
    __syncthreads();
    c0 = c1 = c2 = 0;
    for (int w = 0; w < PREP_THREADS / 64; ++w) {
        c0 += red[0][w];
        c1 += red[1][w];
        c2 += red[2][w];
    }
    // exclusive prefixes of this thread's PREP_ITEMS frames of one plane, in place
    auto scan = [&](int64_t (&v)[PREP_ITEMS], int64_t& carry) {
        int64_t run = 0;
#pragma unroll
        for (int k = 0; k < PREP_ITEMS; ++k) {
            const int64_t x = v[k];
            v[k] = run;
            run += x;
        }
        const int64_t base = block_scan(run, wave_tot, carry);
#pragma unroll
        for (int k = 0; k < PREP_ITEMS; ++k) v[k] += base;
    };
    for (int t0 = cb; t0 < ce; t0 += PREP_THREADS * PREP_ITEMS) {
        const int tb = t0 + PREP_ITEMS * tid;
        int64_t v0[PREP_ITEMS], v1[PREP_ITEMS], v2[PREP_ITEMS];
#pragma unroll
        for (int k = 0; k < PREP_ITEMS; ++k) {
            const int t = tb + k;
            v0[k] = t < ce ? S0[t] : 0;
            if (WIDE) v1[k] = t < ce ? S1[t] : 0;
            if (PLANES == 3) v2[k] = t < ce ? S2[t] : 0;
            const int64_t v = WIDE ? 256 * v0[k] + v1[k] : v0[k];
            if (WIDE && t < ce) sums[(int64_t)b * nframes + t] = v;
            mx = max(mx, v < 0 ? -v : v);
        }
        scan(v0, c0);
        if (WIDE) scan(v1, c1);
        if (PLANES == 3) scan(v2, c2);
#pragma unroll
        for (int k = 0; k < PREP_ITEMS; ++k) {
            const int t = tb + k;
            if (t < ce) {
                P0[t] = v0[k];
                if (WIDE) P1[t] = v1[k];
                if (PLANES == 3) P2[t] = v2[k];
            }
        }
    }
    if (tid == 0 && ce == nframes) {
        P0[nframes] = c0;
        if (WIDE) P1[nframes] = c1;
        if (PLANES == 3) P2[nframes] = c2;
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) mx = max(mx, __shfl_xor(mx, d, 64));
    if (lane == 0) wave_max[wave] = mx;
    __syncthreads();
    if (tid == 0) {
        int64_t m = 0;
        for (int w = 0; w < PREP_THREADS / 64; ++w) m = max(m, wave_max[w]);
        __hip_atomic_fetch_max(reinterpret_cast<unsigned long long*>(smax + b),
                               (unsigned long long)m, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    }
}

// smax zeroed, then the chunks of every bin.
template <int PLANES>
static void launch_sums_prep(const int64_t* plane_sums, int64_t* sums, int n_bins, int nframes,
                             int64_t* prefix, int64_t* smax, hipStream_t s) {
    launch_fill_i64(smax, n_bins, s);
    const dim3 grid((unsigned)n_bins, (unsigned)((nframes + PREP_CHUNK - 1) / PREP_CHUNK));
    sums_prep_kernel<PLANES><<<grid, dim3(PREP_THREADS), 0, s>>>(plane_sums, sums, nframes,
                                                                 prefix, smax);
}

// Accumulate Den for the FIN_LAGS lags of this thread over the frames of the block.
// Stage element type E (double or int64) and accumulator A (double or __int128).
template <typename E, typename A>
__device__ __forceinline__ void den_lags(const int64_t* __restrict__ S, int nframes, int tau0,
                                         E* sa, E* sb, A (&acc)[FIN_LAGS]) {
    const int tid = threadIdx.x;
    const int tmax = nframes - tau0;  // frames t needed by the smallest lag of the block
    for (int t0 = 0; t0 < tmax; t0 += FIN_CH) {
        __syncthreads();
        for (int i = tid; i < FIN_CH; i += FIN_THREADS) {
            const int t = t0 + i;
            sa[i] = (E)(t < nframes ? S[t] : 0);
        }
        for (int i = tid; i < FIN_CH + FIN_SPAN; i += FIN_THREADS) {
            const int64_t t = (int64_t)t0 + tau0 + i;
            sb[i] = (E)(t < nframes ? S[t] : 0);
        }
        __syncthreads();
        // Terms past the end read the zero padding of sb, so one bound serves all lags.
        int lim = nframes - (tau0 + tid) - t0;
        if (lim > FIN_CH) lim = FIN_CH;
        for (int k = 0; k < lim; ++k) {
            const E x = sa[k];
#pragma unroll
            for (int j = 0; j < FIN_LAGS; ++j) {
                const E y = sb[k + tid + FIN_THREADS * j];
                if constexpr (sizeof(A) == 16)
                    acc[j] += (__int128)x * (__int128)y;
                else
                    acc[j] = __builtin_fma(x, y, acc[j]);
            }
        }
    }
}

// Den(tau) = sum_t S(t) S(t+tau) in FP64 for the bins where that is exact (N M^2 < 2^52,
// the others are left to finalize_kernel's 128-bit path).  Block (x, bin, z) covers the
// lags tau0 = DEN_SPAN x + [0, DEN_SPAN) over `cpb` passes of DEN_CH frames from
// t = DEN_CH cpb z; thread `tid` owns the DEN_W consecutive lags tau0 + DEN_W tid + j and
// keeps the window S(t + lag .. lag + DEN_W - 1) in registers: each step over t reads one
// new S value (plus the broadcast S(t)) for DEN_W FMAs, the window registers rotating with
// the unrolled step.  The staged lag-side values sit at index i + i / 16 in LDS, so the
// lanes' reads, DEN_W apart, fall on distinct banks.  Terms past the end of the series
// read the zero padding, so one bound serves all lags of a thread.  The blocks' partial
// sums go to den (zeroed first) by FP64 atomic adds: every product and partial sum is an
// integer below 2^53, so neither the FMA order nor the order of the adds changes Den.
constexpr int DEN_W = 16;                       // consecutive lags per thread
constexpr int DEN_SPAN = FIN_THREADS * DEN_W;   // lags per block
constexpr int DEN_CH = 1024;                    // frames staged per pass
constexpr int DEN_SB = DEN_CH + DEN_SPAN + DEN_W;  // lag-side values staged per pass
constexpr int DEN_BLOCKS = 2048;                // blocks aimed at per launch
// Bins whose frame sums all satisfy |S| <= DEN_I16_MAX take den_i16_kernel: 16 products
// of at most DEN_I16_MAX^2 stay below 2^31.
constexpr int64_t DEN_I16_MAX = 11585;

__device__ __forceinline__ int den_pad(int i) { return i + (i >> 4); }

typedef short den_s2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t pack_s16(int64_t a, int64_t b) {
    return (uint32_t)(uint16_t)(int16_t)a | ((uint32_t)(uint16_t)(int16_t)b << 16);
}

// Den(tau) for the bins with |S| <= DEN_I16_MAX (and N M^2 < 2^52): the same blocks,
// passes and register lag windows as den_fp64_kernel, on 16-bit integers two frames at a
// time.  LDS holds pairs (S(t), S(t+1)) packed in one word, so one v_dot2_i32_i16 adds
// S(t) S(t+lag) + S(t+1) S(t+1+lag) into a 32-bit accumulator; each thread's window
// advances by two packed words per step, and the 32-bit sums (16 products of at most
// DEN_I16_MAX^2 since the last flush) move to 64-bit accumulators every DEN_W frames.
// Exact integer arithmetic throughout; the blocks' partial sums are added into den
// (zeroed first) in FP64, exact below 2^53.
__global__ __launch_bounds__(FIN_THREADS) void den_i16_kernel(const int64_t* __restrict__ sums,
                                                              int nframes, int cpb,
                                                              const int64_t* __restrict__ smax,
                                                              int skip, double* __restrict__ den) {
    __shared__ uint32_t sa[DEN_CH / 2];
    __shared__ uint32_t sb[DEN_SB + DEN_SB / 16 + 1];
    const int b = blockIdx.y;
    const int tau0 = blockIdx.x * DEN_SPAN;
    const int tid = threadIdx.x;
    const double m = (double)smax[b];
    if (!((double)nframes * m * m < 4503599627370496.0)) return;  // 128-bit path
    if (smax[b] > DEN_I16_MAX) return;                            // den_fp64_kernel's bin
    if (smax[b] <= skip) return;                                  // den_mfma_kernel's bin
    const int tb = blockIdx.z * cpb * DEN_CH;
    const int te = min(nframes - tau0, tb + cpb * DEN_CH);
    if (tb >= te) return;
    const int64_t* S = sums + (int64_t)b * nframes;
    const int lag = tau0 + DEN_W * tid;  // smallest lag of this thread
    int64_t acc[DEN_W];
#pragma unroll
    for (int j = 0; j < DEN_W; ++j) acc[j] = 0;
    auto at = [&](int64_t t) -> int64_t { return t < nframes ? S[t] : 0; };
    for (int t0 = tb; t0 < te; t0 += DEN_CH) {
        __syncthreads();
        for (int i = tid; i < DEN_CH / 2; i += FIN_THREADS)
            sa[i] = pack_s16(at((int64_t)t0 + 2 * i), at((int64_t)t0 + 2 * i + 1));
        for (int i = tid; i < DEN_SB; i += FIN_THREADS) {
            const int64_t t = (int64_t)t0 + tau0 + i;
            sb[den_pad(i)] = pack_s16(at(t), at(t + 1));
        }
        __syncthreads();
        int lim = nframes - lag - t0;
        if (lim > DEN_CH) lim = DEN_CH;
        const int base = DEN_W * tid;
        uint32_t y[DEN_W];
#pragma unroll
        for (int j = 0; j < DEN_W - 2; ++j) y[j] = sb[den_pad(base + j)];
        y[DEN_W - 2] = y[DEN_W - 1] = 0u;
        for (int k0 = 0; k0 < lim; k0 += DEN_W) {
            int part[DEN_W];
#pragma unroll
            for (int j = 0; j < DEN_W; ++j) part[j] = 0;
#pragma unroll
            for (int st = 0; st < DEN_W / 2; ++st) {
                const int k = k0 + 2 * st;
                y[(2 * st + DEN_W - 2) % DEN_W] = sb[den_pad(k + base + DEN_W - 2)];
                y[(2 * st + DEN_W - 1) % DEN_W] = sb[den_pad(k + base + DEN_W - 1)];
                const den_s2 x = __builtin_bit_cast(den_s2, sa[k >> 1]);
#pragma unroll
                for (int j = 0; j < DEN_W; ++j)
                    part[j] = __builtin_amdgcn_sdot2(
                        x, __builtin_bit_cast(den_s2, y[(2 * st + j) % DEN_W]), part[j], false);
            }
#pragma unroll
            for (int j = 0; j < DEN_W; ++j) acc[j] += part[j];
        }
    }
#pragma unroll
    for (int j = 0; j < DEN_W; ++j)
        if (lag + j < nframes && acc[j] != 0)
            atomicAdd(den + (int64_t)b * nframes + lag + j, (double)acc[j]);
}

__global__ __launch_bounds__(FIN_THREADS) void den_fp64_kernel(const int64_t* __restrict__ sums,
                                                               int nframes, int cpb,
                                                               const int64_t* __restrict__ smax,
                                                               int skip,
                                                               double* __restrict__ den) {
    __shared__ double sa[DEN_CH];
    __shared__ double sb[DEN_SB + DEN_SB / 16 + 1];
    const int b = blockIdx.y;
    const int tau0 = blockIdx.x * DEN_SPAN;
    const int tid = threadIdx.x;
    const double m = (double)smax[b];
    if (!((double)nframes * m * m < 4503599627370496.0)) return;  // 128-bit path
    if (smax[b] <= DEN_I16_MAX) return;                           // den_i16_kernel's bin
    if (smax[b] <= skip) return;                                  // den_mfma_kernel's bin
    const int tb = blockIdx.z * cpb * DEN_CH;
    const int te = min(nframes - tau0, tb + cpb * DEN_CH);
    if (tb >= te) return;
    const int64_t* S = sums + (int64_t)b * nframes;
    const int lag = tau0 + DEN_W * tid;  // smallest lag of this thread
    double acc[DEN_W];
#pragma unroll
    for (int j = 0; j < DEN_W; ++j) acc[j] = 0.0;
    for (int t0 = tb; t0 < te; t0 += DEN_CH) {
        __syncthreads();
        for (int i = tid; i < DEN_CH; i += FIN_THREADS)
            sa[i] = t0 + i < nframes ? (double)S[t0 + i] : 0.0;
        for (int i = tid; i < DEN_SB; i += FIN_THREADS) {
            const int64_t t = (int64_t)t0 + tau0 + i;
            sb[den_pad(i)] = t < nframes ? (double)S[t] : 0.0;
        }
        __syncthreads();
        int lim = nframes - lag - t0;
        if (lim > DEN_CH) lim = DEN_CH;
        const int base = DEN_W * tid;
        double y[DEN_W];
#pragma unroll
        for (int j = 0; j < DEN_W - 1; ++j) y[j] = sb[den_pad(base + j)];
        y[DEN_W - 1] = 0.0;
        for (int k0 = 0; k0 < lim; k0 += DEN_W) {
#pragma unroll
            for (int st = 0; st < DEN_W; ++st) {
                y[(st + DEN_W - 1) % DEN_W] = sb[den_pad(k0 + st + base + DEN_W - 1)];
                const double x = sa[k0 + st];
#pragma unroll
                for (int j = 0; j < DEN_W; ++j)
                    acc[j] = __builtin_fma(x, y[(st + j) % DEN_W], acc[j]);
            }
        }
    }
#pragma unroll
    for (int j = 0; j < DEN_W; ++j)
        if (lag + j < nframes && acc[j] != 0.0)
            atomicAdd(den + (int64_t)b * nframes + lag + j, acc[j]);
}

// sum_{t < N - tau} X(t) + sum_{t >= tau} X(t) from the exclusive prefix sums P of X.
__device__ __forceinline__ int64_t edge_sums(const int64_t* P, int nframes, int tau) {
    return P[nframes - tau] + P[nframes] - P[tau];
}

// grid (ceil(N / FIN_SPAN), n_bins), FIN_THREADS threads.  NumT is int64_t (exact
// numerators of the Gram and event paths) or double (FFT path).  mode 0: numerators as
// accumulated; 1: uint8 bias correction; 2: uint16 digit planes (ugram rows b, n_bins + b,
// 2 n_bins + b hold the biased HH, LL and cross sums); 3: as 2, with the HH and cross
// sums taken over nwkb[b] k-blocks of the bin's wide pixels only (the high bytes of every
// other pixel are zero):
//   HH = U_HH + 128 E_H - 2^14 K_w (N-tau),   X = U_X + 128 (E_H + E_Lw) - 2^15 K_w (N-tau)
// with K_w = 64 nwkb[b] and E_Lw from the low-byte sums over those k-blocks (prefix
// plane 2).
template <typename NumT>
__global__ __launch_bounds__(FIN_THREADS) void finalize_kernel(
    const int64_t* __restrict__ sums, const NumT* __restrict__ ugram,
    const BinDev* __restrict__ bins, int nframes, int mode,
    const int64_t* __restrict__ prefix, const int64_t* __restrict__ smax,
    const double* __restrict__ dpre, const int* __restrict__ nwkb, NumT* __restrict__ num,
    double* __restrict__ den, float* __restrict__ g2) {
    __shared__ __attribute__((aligned(16))) unsigned char stage[(2 * FIN_CH + FIN_SPAN) * 8];
    const int b = blockIdx.y;
    const int tau0 = blockIdx.x * FIN_SPAN;
    const int tid = threadIdx.x;
    const int64_t* S = sums + (int64_t)b * nframes;
    const int64_t* P = prefix + (int64_t)b * (nframes + 1);

    double d[FIN_LAGS];
    const double m = (double)smax[b];
    if ((double)nframes * m * m < 4503599627370496.0) {  // 2^52: den_fp64_kernel's Den
#pragma unroll
        for (int j = 0; j < FIN_LAGS; ++j) {
            const int tau = tau0 + tid + FIN_THREADS * j;
            d[j] = tau < nframes ? dpre[(int64_t)b * nframes + tau] : 0.0;
        }
    } else {
        __int128 acc[FIN_LAGS] = {};
        den_lags<int64_t, __int128>(S, nframes, tau0, reinterpret_cast<int64_t*>(stage),
                                    reinterpret_cast<int64_t*>(stage) + FIN_CH, acc);
#pragma unroll
        for (int j = 0; j < FIN_LAGS; ++j) d[j] = i128_to_double(acc[j]);
    }

    const BinDev bd = bins[b];
    const int nb = gridDim.y;
    const int64_t kpad = (int64_t)bd.nkb * KB;  // pixels per digit plane, with padding
#pragma unroll
    for (int j = 0; j < FIN_LAGS; ++j) {
        const int tau = tau0 + tid + FIN_THREADS * j;
        if (tau >= nframes) continue;
        const int64_t o = (int64_t)b * nframes + tau;
        const int64_t pairs = kpad * (nframes - tau);  // padded pixel pairs on the diagonal
        double nd;
        NumT n;
        if (mode >= 2) {
            // biased digit products: (h-128)(h'-128) etc.; padding entries are -128 too
            const int64_t* P1 = prefix + (int64_t)(nb + b) * (nframes + 1);
            const int64_t eh = edge_sums(P, nframes, tau), el = edge_sums(P1, nframes, tau);
            int64_t elx = el, pairs_h = pairs;
            if (mode == 3) {
                const int64_t* P2 = prefix + (int64_t)(2 * nb + b) * (nframes + 1);
                elx = edge_sums(P2, nframes, tau);
                pairs_h = (int64_t)nwkb[b] * KB * (nframes - tau);
            }
            const int64_t stride = (int64_t)nb * nframes;
            const int64_t hh = (int64_t)ugram[o] + 128 * eh - 16384 * pairs_h;
            const int64_t ll = (int64_t)ugram[o + stride] + 128 * el - 16384 * pairs;
            const int64_t x = (int64_t)ugram[o + 2 * stride] + 128 * (eh + elx) - 32768 * pairs_h;
            const __int128 v = (__int128)hh * 65536 + (__int128)x * 256 + ll;
            n = (NumT)(int64_t)v;
            nd = i128_to_double(v);
        } else {
            n = ugram[o];
            if (mode == 1) n += (NumT)(128 * edge_sums(P, nframes, tau) - 16384 * pairs);
            nd = (double)n;
        }
        if (num) num[o] = n;
        if (den) den[o] = d[j];
        g2[o] = (float)(bd.scale * (nd / d[j]));
    }
}

// Den for every bin with N M^2 < 2^52: i8 MFMA for the bins whose |S| fits two digits
// (k_den.hip), packed 16-bit dot products for |S| <= DEN_I16_MAX, FP64 above.
static void launch_den_fp64(const int64_t* sums, int n_bins, int nframes, const int64_t* smax,
                            const int64_t* prefix, int planes, const DenMfma& dm, double* dpre,
                            hipStream_t s) {
    // bins with max |S| <= skip are den_mfma_kernel's
    const int skip = dm.v && dm.umax > 0 ? den_mfma_max() : -1;
    launch_fill_i64(reinterpret_cast<int64_t*>(dpre), (int64_t)n_bins * nframes, s);  // +0.0
    launch_den_mfma(sums, n_bins, nframes, smax, prefix, planes, dm, dpre, s);
    const int64_t lag_blocks = (nframes + DEN_SPAN - 1) / DEN_SPAN;
    const int64_t chunks = (nframes + DEN_CH - 1) / DEN_CH;
    // frame passes per block: about DEN_BLOCKS blocks in all (half of them, above the
    // last lag of their chunk, return at once)
    int64_t cpb = (lag_blocks * chunks * n_bins + DEN_BLOCKS - 1) / DEN_BLOCKS;
    if (cpb < 1) cpb = 1;
    const int64_t zb = (chunks + cpb - 1) / cpb;
    den_fp64_kernel<<<dim3((unsigned)lag_blocks, (unsigned)n_bins, (unsigned)zb),
                      dim3(FIN_THREADS), 0, s>>>(sums, nframes, (int)cpb, smax, skip, dpre);
    den_i16_kernel<<<dim3((unsigned)lag_blocks, (unsigned)n_bins, (unsigned)zb),
                     dim3(FIN_THREADS), 0, s>>>(sums, nframes, (int)cpb, smax, skip, dpre);
}

void launch_finalize(const int64_t* plane_sums, int64_t* sums, const int64_t* ugram,
                     const BinDev* bins, int n_bins, int nframes, int mode, int64_t* prefix,
                     int64_t* smax, double* dpre, const int* nwkb, int64_t* num, double* den,
                     float* g2, const DenMfma& dm, hipStream_t s) {
    if (n_bins <= 0 || nframes <= 0) return;
    if (mode == 3)
        launch_sums_prep<3>(plane_sums, sums, n_bins, nframes, prefix, smax, s);
    else if (mode == 2)
        launch_sums_prep<2>(plane_sums, sums, n_bins, nframes, prefix, smax, s);
    else
        launch_sums_prep<1>(sums, sums, n_bins, nframes, prefix, smax, s);
    launch_den_fp64(sums, n_bins, nframes, smax, prefix, mode >= 2 ? 2 : 1, dm, dpre, s);
    dim3 grid((unsigned)((nframes + FIN_SPAN - 1) / FIN_SPAN), (unsigned)n_bins);
    finalize_kernel<int64_t><<<grid, dim3(FIN_THREADS), 0, s>>>(
        sums, ugram, bins, nframes, mode, prefix, smax, dpre, nwkb, num, den, g2);
}

void launch_finalize_f64(const int64_t* sums, const double* fnum, const BinDev* bins, int n_bins,
                         int nframes, int64_t* prefix, int64_t* smax, double* dpre, double* num,
                         double* den, float* g2, const DenMfma& dm, hipStream_t s) {
    if (n_bins <= 0 || nframes <= 0) return;
    launch_sums_prep<1>(sums, const_cast<int64_t*>(sums), n_bins, nframes, prefix, smax, s);
    launch_den_fp64(sums, n_bins, nframes, smax, prefix, 1, dm, dpre, s);
    dim3 grid((unsigned)((nframes + FIN_SPAN - 1) / FIN_SPAN), (unsigned)n_bins);
    finalize_kernel<double><<<grid, dim3(FIN_THREADS), 0, s>>>(
        sums, fnum, bins, nframes, 0, prefix, smax, dpre, nullptr, num, den, g2);
}

}  // namespace dxg2
