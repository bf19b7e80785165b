// Exact denominators Den_b(tau) = sum_t S(t) S(t+tau) on i8 matrix cores, for the bins
// whose frame sums fit two i8 digits.
//
// Reference: the denominator is correlate_1D over the frame sums
// (resources/opencl/densecorrelator.cl:156-176), np.correlate in the FFT path
// (correlator/dense.py:365-366) and the event path's sums (evtcorrelator.cl:58-76).
//
// Reshape a series x as rows of 64 frames, T[r][s] = x(64 r + s) (R = ceil(N/64) rows,
// x = 0 past the end).  For tau = 64 c + j (0 <= j < 64)
//   Den(64 c + j) = sum_s M_c[s][s + j],
//   M_c[s][s'] = sum_r T[r][s] x(64 (r + c) + s'),     s' in [0, 128),
// i.e. M_c = T^T W_c with W_c[r][s'] = x(64 (r + c) + s'): one 64 x 128 product with the
// long dimension r (R - c rows) as K, then 64 diagonal sums.  Of the 2 x 4 blocks of 32 x
// 32 in M_c, blocks (s-half 0, s'-block 3) and (1, 0) hold no diagonal j in [0, 64) and
// are skipped; the six others take one v_mfma_i32_32x32x32_i8 each per 32 rows of K.
//
// Operands: V[s][u] = x(64 u + s), the series transposed (64 rows of `up` bytes).  The A
// operand of rows s is V[s][r ..] (aligned); the B operand of lag block c and column s' is
// V[s' mod 64][r + c (+1 for s' >= 64) ..], 16 bytes at any byte offset.  A workgroup
// takes 16 consecutive lag blocks c0 .. c0 + 15, four per wave (c0 + 4w + i): a wave's B
// operands start at c0 + 4w + o (o = 0..4), so per row a 16-byte and a 4-byte load at
// c0 + 4w + r (4-byte aligned) hold every window, cut out with v_alignbyte at shifts o
// that are compile-time constants (the wave index is a template argument).  Per 32 rows
// of K a wave issues 6 loads for 24 MFMAs into 12 accumulators.
//
// Digits, by the bin's M = max |S|:
//   M <= 127: one pass over x = S;
//   127 < M <= 255 with unsigned frames: one pass over x = S - 128, then
//     Den(tau) = U(tau) + 128 E(tau) - 16384 (N - tau),  E = P(N-tau) + P(N) - P(tau),
//     with P the prefix sums of S (sums_prep_kernel);
//   M <= 16383: S = 128 h + l (h = S >> 7 in [-128, 127], l = S & 127), four passes
//     Den = 16384 HH + 128 (HL + LH) + LL, HL(tau) = sum_t h(t) l(t+tau).
// Accumulators: two blocks of |x x'| <= 2^14 over at most R <= 32768 rows stay below
// 2^30 in int32; the diagonal sums and the digit weights go to 64-bit LDS accumulators.  Den is exact
// (an integer below 2^53 stored as FP64), so these bins agree with den_i16_kernel and
// den_fp64_kernel to the last bit.
#include "dxg2_internal.h"

namespace dxg2 {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef __attribute__((address_space(1))) const v4i global_v4i;

// 16-byte load through a global-address-space pointer (a generic pointer would compile
// to a flat load).
__device__ __forceinline__ v4i load16(const int8_t* p) {
    return *reinterpret_cast<global_v4i*>(reinterpret_cast<uintptr_t>(p));
}

constexpr int DEN_MFMA_MAX = 16383;  // largest max |S| of a bin on this path
constexpr int PREP_QUADS = 16;        // 4-column groups of V per prep workgroup

// Passes for bin with max |S| = m: 1 (one digit, `bias` 0 or 128), 4 (two digits), 0
// (not this path's).
__device__ __forceinline__ int den_passes(int64_t m, int umax, int& bias) {
    bias = 0;
    if (m <= 127) return 1;
    if (m <= umax) {
        bias = 128;
        return 1;
    }
    return m <= DEN_MFMA_MAX ? 4 : 0;
}

// The digit planes of bin blockIdx.y, columns [u0, u0 + 4 PREP_QUADS): plane 0 is x (one
// digit) or h, plane 1 is l.  Thread (s, quad) reads S(64 (u0 + 4 quad + i) + s) for
// i < 4 -- the 64 lanes of a wave read 64 consecutive frames -- and writes 4 bytes.
__global__ __launch_bounds__(256) void den_mfma_prep_kernel(const int64_t* __restrict__ sums,
                                                            int nframes, int up,
                                                            const int64_t* __restrict__ smax,
                                                            int umax, int8_t* __restrict__ v) {
    const int b = blockIdx.y;
    int bias;
    const int passes = den_passes(smax[b], umax, bias);
    if (!passes) return;
    const int s = threadIdx.x & 63;
    const int64_t* S = sums + (int64_t)b * nframes;
    int8_t* v0 = v + (int64_t)b * 2 * 64 * up;
    for (int qd = threadIdx.x >> 6; qd < PREP_QUADS; qd += 4) {
        const int u = (blockIdx.x * PREP_QUADS + qd) * 4;
        if (u >= up) break;
        uint32_t w0 = 0, w1 = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int64_t t = 64 * (int64_t)(u + i) + s;
            const int64_t x = t < nframes ? S[t] : 0;
            uint32_t d0, d1 = 0;
            if (passes == 1) {
                d0 = (uint32_t)(uint8_t)(int8_t)(t < nframes ? x - bias : 0);
            } else {
                d0 = (uint32_t)(uint8_t)(int8_t)(x >> 7);
                d1 = (uint32_t)(x & 127);
            }
            w0 |= d0 << (8 * i);
            w1 |= d1 << (8 * i);
        }
        *reinterpret_cast<uint32_t*>(v0 + (int64_t)s * up + u) = w0;
        if (passes == 4) *reinterpret_cast<uint32_t*>(v0 + (int64_t)(64 + s) * up + u) = w1;
    }
}

// 16 bytes of the 20 in (lo, hi) from byte OFF (0 <= OFF <= 4).
template <int OFF>
__device__ __forceinline__ v4i window(const v4i& lo, uint32_t hi) {
    const uint32_t w[5] = {(uint32_t)lo.x, (uint32_t)lo.y, (uint32_t)lo.z, (uint32_t)lo.w, hi};
    constexpr int A = OFF >> 2, B = OFF & 3;
    v4i r;
    if constexpr (B == 0) {
        r.x = (int)w[A];
        r.y = (int)w[A + 1];
        r.z = (int)w[A + 2];
        r.w = (int)w[A + 3];
    } else {
        r.x = (int)__builtin_amdgcn_alignbyte(w[A + 1], w[A], B);
        r.y = (int)__builtin_amdgcn_alignbyte(w[A + 2], w[A + 1], B);
        r.z = (int)__builtin_amdgcn_alignbyte(w[A + 3], w[A + 2], B);
        r.w = (int)__builtin_amdgcn_alignbyte(w[A + 4], w[A + 3], B);
    }
    return r;
}

// 16 and 4 bytes at a 4-byte aligned global address.
typedef int v4i_a4 __attribute__((ext_vector_type(4), aligned(4)));
typedef __attribute__((address_space(1))) const v4i_a4 global_v4i_a4;
typedef __attribute__((address_space(1))) const uint32_t global_u32;
__device__ __forceinline__ v4i load16_a4(const int8_t* p) {
    const v4i_a4 v = *reinterpret_cast<global_v4i_a4*>(reinterpret_cast<uintptr_t>(p));
    return v4i{v.x, v.y, v.z, v.w};
}
__device__ __forceinline__ uint32_t load4(const int8_t* p) {
    return *reinterpret_cast<global_u32*>(reinterpret_cast<uintptr_t>(p));
}

#define DEN_MFMA(a, b, c) __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c, 0, 0, 0)

// Lag blocks c = c0 + 4 W + i (i < 4) over K = rows - c0 rows (later rows meet zeros):
// M_c for plane pair (X, Y), diagonal sums times `weight` into acc[4 W + i][j].
template <int W>
__device__ __forceinline__ void den_wave(const int8_t* __restrict__ vx,
                                         const int8_t* __restrict__ vy, int up, int c0, int K,
                                         int rows, int64_t weight,
                                         unsigned long long (*acc)[64]) {
    const int lane = threadIdx.x & 63, l32 = lane & 31, hl = lane >> 5;
    const int8_t* pa0 = vx + (int64_t)l32 * up + 16 * hl;
    const int8_t* pa1 = pa0 + 32 * (int64_t)up;
    // B rows from byte 4 W past c0: the windows of this wave start at offsets 0 .. 4
    const int8_t* pb0 = vy + (int64_t)l32 * up + c0 + 16 * hl + 4 * W;
    const int8_t* pb1 = pb0 + 32 * (int64_t)up;
    // m[i][d]: diagonal group d (j = 32 d + column - row) of lag block c0 + 4 W + i; the
    // two blocks of a group, (0, d) and (1, d + 1), share one accumulator
    v16i m[4][3];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int e = 0; e < 16; ++e) m[i][k][e] = 0;
    // Three register sets, each the operands of 32 rows, rotate through an unrolled
    // loop: a set is reloaded (two steps ahead) right after its MFMAs, so every load has
    // two steps of MFMAs (1,536 cycles) to arrive, and no register is copied.  The last
    // iteration may run up to two steps past K and load two more ahead; those rows meet
    // zero B columns (r + c >= rows) and stay inside V (c0 + K + 96 + 63 + 32 + 15 <
    // rows + 192 <= up).
    struct Ops {
        v4i a0, a1, b0l, b1l;
        uint32_t b0h, b1h;
    };
    auto load = [&](Ops& o, int r) {
        o.a0 = load16(pa0 + r);
        o.a1 = load16(pa1 + r);
        o.b0l = load16_a4(pb0 + r);
        o.b0h = load4(pb0 + r + 16);
        o.b1l = load16_a4(pb1 + r);
        o.b1h = load4(pb1 + r + 16);
    };
    // B(c0 + 4W + i): blocks 0, 1 (s' < 64) at offset 4W + i of rows l32 and l32 + 32,
    // blocks 2, 3 (s' >= 64) at offset 4W + i + 1 (window offsets i and i + 1)
    auto step = [&](const Ops& o) {
#define DEN_STEP(I)                                                                          \
        {                                                                                    \
            const v4i f0 = window<I>(o.b0l, o.b0h), f1 = window<I>(o.b1l, o.b1h);            \
            const v4i g0 = window<I + 1>(o.b0l, o.b0h), g1 = window<I + 1>(o.b1l, o.b1h);    \
            m[I][0] = DEN_MFMA(o.a0, f0, m[I][0]);                                           \
            m[I][1] = DEN_MFMA(o.a0, f1, m[I][1]);                                           \
            m[I][2] = DEN_MFMA(o.a0, g0, m[I][2]);                                           \
            m[I][0] = DEN_MFMA(o.a1, f1, m[I][0]);                                           \
            m[I][1] = DEN_MFMA(o.a1, g0, m[I][1]);                                           \
            m[I][2] = DEN_MFMA(o.a1, g1, m[I][2]);                                           \
        }
        DEN_STEP(0)
        DEN_STEP(1)
        DEN_STEP(2)
        DEN_STEP(3)
#undef DEN_STEP
    };
    Ops s0, s1, s2;
    load(s0, 0);
    load(s1, 32);
    // the scheduler would sink each load group to its first use, exposing its latency:
    // fences keep every group above the MFMAs it overlaps
    for (int r0 = 0; r0 < K; r0 += 96) {
        load(s2, r0 + 64);
        __builtin_amdgcn_sched_barrier(0);
        step(s0);
        __builtin_amdgcn_sched_barrier(0);
        load(s0, r0 + 96);
        __builtin_amdgcn_sched_barrier(0);
        step(s1);
        __builtin_amdgcn_sched_barrier(0);
        load(s1, r0 + 128);
        __builtin_amdgcn_sched_barrier(0);
        step(s2);
        __builtin_amdgcn_sched_barrier(0);
    }
    // 32 x 32 accumulator layout: column l32, rows 8 (e / 4) + 4 hl + e % 4.  Block
    // (sh, sb) holds s = 32 sh + row, s' = 32 sb + l32, diagonal j = 32 (sb - sh) + l32 -
    // row, so group d sums into j = 32 d + l32 - row where 0 <= j < 64.
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (c0 + 4 * W + i >= rows) break;
        unsigned long long* a = acc[4 * W + i];
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const int d = l32 - (8 * (e >> 2) + 4 * hl + (e & 3));  // -31 .. 31
            if (d >= 0) atomicAdd(a + d, (unsigned long long)((int64_t)m[i][0][e] * weight));
            atomicAdd(a + 32 + d, (unsigned long long)((int64_t)m[i][1][e] * weight));
            if (d < 0) atomicAdd(a + 64 + d, (unsigned long long)((int64_t)m[i][2][e] * weight));
        }
    }
}

// Workgroup: bin blockIdx.y, lag blocks c0 = 16 blockIdx.x .. c0 + 15; then Den(64 c + j)
// for those blocks into den (plain stores; each lag is written once).
__global__ __launch_bounds__(256) void den_mfma_kernel(
    const int8_t* __restrict__ v, int nframes, int rows, int up,
    const int64_t* __restrict__ smax, int umax, const int64_t* __restrict__ prefix,
    int planes, double* __restrict__ den) {
    __shared__ unsigned long long acc[16][64];
    const int b = blockIdx.y;
    int bias;
    const int passes = den_passes(smax[b], umax, bias);
    if (!passes) return;  // the same for the whole workgroup
    const int tid = threadIdx.x;
    for (int i = tid; i < 16 * 64; i += 256) acc[i >> 6][i & 63] = 0ull;
    __syncthreads();
    const int c0 = 16 * blockIdx.x;
    const int K = rows - c0;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int8_t* vh = v + (int64_t)b * 2 * 64 * up;  // x, or the high digit
    const int8_t* vl = vh + 64 * (int64_t)up;          // the low digit
    for (int p = 0; p < passes; ++p) {
        // passes: x x | h h (2^14), h l (2^7), l h (2^7), l l (1)
        const int8_t* vx = p < 2 ? vh : vl;
        const int8_t* vy = (passes == 1 || p == 0 || p == 2) ? vh : vl;
        const int64_t weight = passes == 1 ? 1 : (p == 0 ? 16384 : (p == 3 ? 1 : 128));
        switch (wave) {
            case 0: den_wave<0>(vx, vy, up, c0, K, rows, weight, acc); break;
            case 1: den_wave<1>(vx, vy, up, c0, K, rows, weight, acc); break;
            case 2: den_wave<2>(vx, vy, up, c0, K, rows, weight, acc); break;
            default: den_wave<3>(vx, vy, up, c0, K, rows, weight, acc); break;
        }
    }
    __syncthreads();
    for (int i = tid; i < 16 * 64; i += 256) {
        const int64_t tau = 64 * (int64_t)(c0 + (i >> 6)) + (i & 63);
        if (tau >= nframes) continue;
        int64_t d = (int64_t)acc[i >> 6][i & 63];
        if (bias) {
            const int64_t* P0 = prefix + (int64_t)b * (nframes + 1);
            const int64_t* P1 = prefix + ((int64_t)gridDim.y + b) * (nframes + 1);
            // prefix of S: one plane, or 256 S_H + S_L from the two byte planes
            auto P = [&](int64_t t) -> int64_t {
                return planes == 1 ? P0[t] : 256 * P0[t] + P1[t];
            };
            d += 128 * (P(nframes - tau) + P(nframes) - P(tau)) -
                 16384 * ((int64_t)nframes - tau);
        }
        den[(int64_t)b * nframes + tau] = (double)d;
    }
}

}  // namespace

int den_mfma_max() { return DEN_MFMA_MAX; }

int64_t den_mfma_columns(int nframes) {
    const int64_t rows = ((int64_t)nframes + 63) / 64;
    if (nframes <= 0 || rows > 32768) return 0;
    return (rows + 15) / 16 * 16 + 192;
}

void launch_den_mfma(const int64_t* sums, int n_bins, int nframes, const int64_t* smax,
                     const int64_t* prefix, int planes, const DenMfma& dm, double* den,
                     hipStream_t s) {
    if (!dm.v || dm.umax <= 0 || n_bins <= 0 || nframes <= 0) return;
    const int rows = (nframes + 63) / 64;
    const int up = (int)dm.up;
    den_mfma_prep_kernel<<<dim3((unsigned)((up + 4 * PREP_QUADS - 1) / (4 * PREP_QUADS)),
                                (unsigned)n_bins),
                           dim3(256), 0, s>>>(sums, nframes, up, smax, dm.umax, dm.v);
    // lag groups along x, smallest lags (the longest workgroups) first: consecutive
    // workgroups share a bin, so each XCD's L2 holds the planes of a few bins at a time
    den_mfma_kernel<<<dim3((unsigned)((rows + 15) / 16), (unsigned)n_bins), dim3(256), 0, s>>>(
        dm.v, nframes, rows, up, smax, dm.umax, prefix, planes, den);
}

}  // namespace dxg2
