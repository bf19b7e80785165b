// calc_std: the spread of the per-frame-pair ratios along each diagonal.
//
// Reference (correlator/dense.py:58-59,65-66): for one bin with n_b pixels,
//   dev(tau) = std_t( (G(t, t+tau) / n_b) / (m_t m_{t+tau}) ) / sqrt(N - tau),
// the population std over t = 0 .. N-1-tau, with m_t = S(t) / n_b the frame mean.  The
// ratio is r_t = n_b G(t, t+tau) / (S(t) S(t+tau)).  A zero frame sum on a diagonal
// makes numpy's ratio inf or nan there, and the std nan.
//
// The ratio is not linear in G, so it needs every element of the Gram, not only the
// diagonal sums.  The calc_std Gram launch writes each work item's int32 tile into a
// scratch slot (work items of one tile: k-ranges, and for uint16 the HH, LL and cross
// products), and one workgroup per tile here
//   * forms the exact int64 G(t, u) = sum_w w U_w + c(t) + c(u) - Kc (bias corrections
//     of the i8 digit planes, below) and r in FP64, a band of rows at a time in LDS;
//   * walks each of the tile's 2T-1 diagonals and keeps (count, mean, M2) of its
//     ratios, shifted by the segment's first ratio so that M2 does not cancel;
// and the finaliser combines the segments of each diagonal (Chan, Golub and LeVeque's
// pairwise update) in a fixed order, so dev is reproducible bit for bit.
//
// Bias corrections (bytes stored as v - 128, K = 64 nkb slots per plane, padded slots
// hold v = 0):
//   uint8:  G = U + 128 (S_t + S_u) - 2^14 K
//   uint16: HH = U_HH + 128 (H_t + H_u) - 2^14 K,  LL likewise with L,
//           X  = U_X + 128 (E_t + E_u) - 2^15 K with E = H + L,
//           G  = 65536 HH + 256 X + LL
//              = 65536 U_HH + 256 U_X + U_LL + c_t + c_u - 2^14 K 66049,
//           c_t = 128 (65792 H_t + 257 L_t).
#include "dxg2_internal.h"

namespace dxg2 {

typedef int v4i __attribute__((ext_vector_type(4)));

// Per-frame bin sums straight from the slab (the calc_std launch needs them before its
// first tile is reduced).  Block = 256 frames of one gather work item (<= 16 k-blocks).
template <bool WIDE, bool BIAS>
__global__ __launch_bounds__(256) void slab_sums_kernel(const int8_t* __restrict__ slab,
                                                        const BinDev* __restrict__ bins,
                                                        const GatherWork* __restrict__ work,
                                                        int n_bins, int nframes, int nrows,
                                                        int64_t* __restrict__ sums) {
    const GatherWork w = work[blockIdx.x];
    const int t = blockIdx.y * 256 + threadIdx.x;
    if (t >= nframes) return;
    const BinDev bd = bins[w.bin];
    const int64_t plane = (int64_t)bd.nkb * nrows * KB;
    int acc = 0, acc_l = 0;  // |sum| <= 16 * 64 * 128
    for (int kb = w.kb0; kb < w.kb0 + w.nkb; ++kb) {
        const int8_t* row = slab + bd.slab_off;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const v4i x = *reinterpret_cast<const v4i*>(row + slab_offset(kb, t, 16 * q, nrows));
            acc = __builtin_amdgcn_sdot4(x.x, 0x01010101, acc, false);
            acc = __builtin_amdgcn_sdot4(x.y, 0x01010101, acc, false);
            acc = __builtin_amdgcn_sdot4(x.z, 0x01010101, acc, false);
            acc = __builtin_amdgcn_sdot4(x.w, 0x01010101, acc, false);
        }
        if (WIDE) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const v4i x = *reinterpret_cast<const v4i*>(row + plane +
                                                            slab_offset(kb, t, 16 * q, nrows));
                acc_l = __builtin_amdgcn_sdot4(x.x, 0x01010101, acc_l, false);
                acc_l = __builtin_amdgcn_sdot4(x.y, 0x01010101, acc_l, false);
                acc_l = __builtin_amdgcn_sdot4(x.z, 0x01010101, acc_l, false);
                acc_l = __builtin_amdgcn_sdot4(x.w, 0x01010101, acc_l, false);
            }
        }
    }
    const int64_t bias = BIAS ? (int64_t)128 * KB * w.nkb : 0;
    const int64_t h = acc + bias;
    if (h != 0)
        atomicAdd(reinterpret_cast<unsigned long long*>(sums + (int64_t)w.bin * nframes + t),
                  (unsigned long long)h);
    if (WIDE) {
        const int64_t l = acc_l + bias;
        if (l != 0)
            atomicAdd(reinterpret_cast<unsigned long long*>(
                          sums + ((int64_t)n_bins + w.bin) * nframes + t),
                      (unsigned long long)l);
    }
}

void launch_slab_sums(const int8_t* slab, int dtype, const BinDev* bins, const GatherWork* work,
                      int n_work, int n_bins, int nframes, int nrows, int64_t* sums,
                      hipStream_t s) {
    if (n_work <= 0 || nframes <= 0) return;
    const dim3 grid((unsigned)n_work, (unsigned)((nframes + 255) / 256));
    if (dtype == DXG2_U16)
        slab_sums_kernel<true, true><<<grid, dim3(256), 0, s>>>(slab, bins, work, n_bins, nframes,
                                                               nrows, sums);
    else if (dtype == DXG2_U8)
        slab_sums_kernel<false, true><<<grid, dim3(256), 0, s>>>(slab, bins, work, n_bins,
                                                                nframes, nrows, sums);
    else
        slab_sums_kernel<false, false><<<grid, dim3(256), 0, s>>>(slab, bins, work, n_bins,
                                                                 nframes, nrows, sums);
}

// Frame sum S(t) and bias term c(t) of frame t of bin b.  mode 0: int8 (no bias),
// 1: uint8, 2: uint16 digit planes (sums holds S_H rows, then S_L rows).
__device__ __forceinline__ void frame_terms(const int64_t* __restrict__ sums, int mode,
                                            int n_bins, int nframes, int b, int t, int64_t& S,
                                            int64_t& c) {
    if (mode == 2) {
        const int64_t h = sums[(int64_t)b * nframes + t];
        const int64_t l = sums[((int64_t)n_bins + b) * nframes + t];
        S = 256 * h + l;
        c = 128 * (65792 * h + 257 * l);
    } else {
        S = sums[(int64_t)b * nframes + t];
        c = mode == 1 ? 128 * S : 0;
    }
}

template <int T>
__global__ __launch_bounds__(256) void std_tile_kernel(const StdTile* __restrict__ tiles,
                                                       const BinDev* __restrict__ bins,
                                                       const int* __restrict__ scratch,
                                                       const int64_t* __restrict__ sums,
                                                       int mode, int n_bins, int nframes,
                                                       StdPart* __restrict__ part) {
    constexpr int BR = 8192 / T;        // rows per LDS band (64 KiB of doubles)
    constexpr int RPP = 256 / T;        // rows per load pass
    constexpr int NDIAG = 2 * T - 1;
    constexpr int SPT = (NDIAG + 255) / 256;  // diagonals per thread
    __shared__ double x[BR][T];         // ratio, NaN where the element is not counted
    __shared__ int64_t rowS[BR], rowC[BR];  // S(t), c(t) of the band's rows

    const StdTile tl = tiles[blockIdx.x];
    const BinDev bd = bins[tl.bin];
    const int tid = threadIdx.x;
    const int c = tid % T;
    const int rsub = tid / T;
    const int t0 = tl.ti * T;
    const int u = tl.tj * T + c;
    const bool diag = tl.ti == tl.tj;
    const int64_t K = (int64_t)KB * bd.nkb;
    const int64_t Kc = mode == 0 ? 0 : (mode == 1 ? 16384 * K : 16384 * K * 66049);
    const int w_hh = mode == 2 ? 65536 : 1;
    const double npix = (double)bd.npix;
    int64_t S_u = 0, c_u = 0;
    if (u < nframes) frame_terms(sums, mode, n_bins, nframes, tl.bin, u, S_u, c_u);
    const int64_t TT = (int64_t)T * T;
    const int* base = scratch + (int64_t)tl.slot0 * TT;
    const int nk = tl.n_hh + tl.n_ll + tl.n_x;

    int cnt[SPT];
    double k0[SPT], s1[SPT], s2[SPT];
#pragma unroll
    for (int j = 0; j < SPT; ++j) {
        cnt[j] = 0;
        k0[j] = s1[j] = s2[j] = 0.0;
    }

    for (int r0 = 0; r0 < T && t0 + r0 < nframes; r0 += BR) {
        __syncthreads();
        if (tid < BR) {
            int64_t S_t = 0, c_t = 0;
            if (t0 + r0 + tid < nframes)
                frame_terms(sums, mode, n_bins, nframes, tl.bin, t0 + r0 + tid, S_t, c_t);
            rowS[tid] = S_t;
            rowC[tid] = c_t;
        }
        __syncthreads();
        // the band's rows in batches of 8, so that their tile loads are in flight together
#pragma unroll 8
        for (int rr = rsub; rr < BR; rr += RPP) {
            const int r = r0 + rr;
            const int t = t0 + r;
            double val = __builtin_nan("");
            const int64_t S_t = rowS[rr];
            if (t < nframes && u < nframes && (!diag || c >= r) && S_t != 0 && S_u != 0) {
                const int* e = base + (int64_t)r * T + c;
                int64_t g = rowC[rr] + c_u - Kc;
                if (nk == 1) {
                    g += e[0];  // uint8 / int8 with one k-chunk: the common case
                } else {
                    int k = 0;
                    for (; k < tl.n_hh; ++k) g += (int64_t)w_hh * e[k * TT];
                    for (; k < tl.n_hh + tl.n_ll; ++k) g += e[k * TT];
                    for (; k < nk; ++k) g += (int64_t)256 * e[k * TT];
                }
                val = npix * (double)g / ((double)S_t * (double)S_u);
            }
            x[rr][c] = val;
        }
        __syncthreads();
        // Each thread's diagonals walk the band's rows in order (the statistics depend on
        // that order, which is fixed); the SPT diagonals of a thread advance together.
        int lo[SPT], hi[SPT], dd[SPT];
        int rmin = BR, rmax = 0;
#pragma unroll
        for (int j = 0; j < SPT; ++j) {
            const int s = tid + 256 * j;
            dd[j] = s - (T - 1);
            lo[j] = s < NDIAG ? max(0, -dd[j] - r0) : BR;
            hi[j] = s < NDIAG ? min(BR, T - dd[j] - r0) : 0;
            rmin = min(rmin, lo[j]);
            rmax = max(rmax, hi[j]);
        }
        for (int rr = rmin; rr < rmax; ++rr) {
#pragma unroll
            for (int j = 0; j < SPT; ++j) {
                if (rr >= lo[j] && rr < hi[j]) {
                    const double v = x[rr][r0 + rr + dd[j]];
                    if (v == v) {
                        if (cnt[j] == 0) k0[j] = v;
                        const double dv = v - k0[j];
                        s1[j] += dv;
                        s2[j] = fma(dv, dv, s2[j]);
                        ++cnt[j];
                    }
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < SPT; ++j) {
        const int s = tid + 256 * j;
        if (s < NDIAG) {
            StdPart p{0.0, 0.0, 0.0};
            if (cnt[j] > 0) {
                const double n = (double)cnt[j];
                p.n = n;
                p.mean = k0[j] + s1[j] / n;
                p.m2 = fmax(s2[j] - s1[j] * s1[j] / n, 0.0);
            }
            part[(int64_t)tl.idx * NDIAG + s] = p;
        }
    }
}

void launch_std_tiles(int tile, const StdTile* tiles, int n_tiles, const BinDev* bins,
                      const int* scratch, const int64_t* sums, int mode, int n_bins, int nframes,
                      StdPart* part, hipStream_t s) {
    if (n_tiles <= 0) return;
    if (tile == 128)
        std_tile_kernel<128><<<dim3((unsigned)n_tiles), dim3(256), 0, s>>>(
            tiles, bins, scratch, sums, mode, n_bins, nframes, part);
    else
        std_tile_kernel<256><<<dim3((unsigned)n_tiles), dim3(256), 0, s>>>(
            tiles, bins, scratch, sums, mode, n_bins, nframes, part);
}

// Two-time correlation of one bin (event_y.py:137-157): trc(t, u) = n_b G(t, u) /
// (S(t) S(u)) for u > t, mirrored to (u, t); NaN where S(t) S(u) = 0.  One workgroup per
// tile, a band of BR rows at a time: the row-major upper entries are written straight
// from the lanes (consecutive u), the mirrored ones from an LDS transpose, 32 lanes per
// output row (128-byte runs).
template <int T>
__global__ __launch_bounds__(256) void ttcf_tile_kernel(const StdTile* __restrict__ tiles,
                                                        const BinDev* __restrict__ bins,
                                                        const int* __restrict__ scratch,
                                                        const int64_t* __restrict__ sums,
                                                        int mode, int n_bins, int nframes,
                                                        float* __restrict__ trc) {
    constexpr int BR = 32;
    constexpr int RPP = 256 / T;
    __shared__ float x[BR][T + 1];

    const StdTile tl = tiles[blockIdx.x];
    const BinDev bd = bins[tl.bin];
    const int tid = threadIdx.x;
    const int c = tid % T;
    const int rsub = tid / T;
    const int t0 = tl.ti * T;
    const int u0 = tl.tj * T;
    const int u = u0 + c;
    const int64_t K = (int64_t)KB * bd.nkb;
    const int64_t Kc = mode == 0 ? 0 : (mode == 1 ? 16384 * K : 16384 * K * 66049);
    const int w_hh = mode == 2 ? 65536 : 1;
    const double npix = (double)bd.npix;
    const int64_t N = nframes;
    int64_t S_u = 0, c_u = 0;
    if (u < nframes) frame_terms(sums, mode, n_bins, nframes, tl.bin, u, S_u, c_u);
    const int64_t TT = (int64_t)T * T;
    const int* base = scratch + (int64_t)tl.slot0 * TT;

    for (int r0 = 0; r0 < T && t0 + r0 < nframes; r0 += BR) {
        __syncthreads();
        for (int rr = rsub; rr < BR; rr += RPP) {
            const int r = r0 + rr;
            const int t = t0 + r;
            float val = 0.0f;
            // the last diagonal entry keeps the pairs within frame N-1: G(t, t) - S(t)
            const bool corner = t == nframes - 1 && u == t;
            if (t < nframes && u < nframes && (u > t || corner)) {
                int64_t S_t, c_t;
                frame_terms(sums, mode, n_bins, nframes, tl.bin, t, S_t, c_t);
                if (S_t != 0 && S_u != 0) {
                    const int* e = base + (int64_t)r * T + c;
                    int64_t g = c_t + c_u - Kc;
                    int k = 0;
                    for (; k < tl.n_hh; ++k) g += (int64_t)w_hh * e[k * TT];
                    for (; k < tl.n_hh + tl.n_ll; ++k) g += e[k * TT];
                    for (; k < tl.n_hh + tl.n_ll + tl.n_x; ++k) g += (int64_t)256 * e[k * TT];
                    if (corner) g -= S_t;
                    val = (float)(npix * (double)g / ((double)S_t * (double)S_u));
                } else {
                    val = __builtin_nanf("");
                }
                trc[(int64_t)t * N + u] = val;
            }
            x[rr][c] = val;
        }
        __syncthreads();
        for (int k = tid; k < T * BR; k += 256) {
            const int cu = k / BR, rr = k % BR;
            const int uu = u0 + cu, t = t0 + r0 + rr;
            if (t < nframes && uu < nframes && uu > t) trc[(int64_t)uu * N + t] = x[rr][cu];
        }
    }
}

// The diagonal of the two-time matrix: entries 0 .. N-2 take the mean of the positive
// entries of the first off-diagonal (event_y.py:224-227); entry (N-1, N-1), the pairs
// within the last frame, is written by ttcf_tile_kernel.  One workgroup.
__global__ __launch_bounds__(1024) void ttcf_diag_kernel(float* __restrict__ trc, int nframes,
                                                          const int64_t* __restrict__ sums,
                                                          int mode, int n_bins, int b) {
    __shared__ double ssum[1024];
    __shared__ int scnt[1024];
    const int tid = threadIdx.x;
    const int64_t N = nframes;
    double s = 0.0;
    int n = 0;
    for (int j = tid; j < nframes - 1; j += 1024) {
        const float v = trc[(int64_t)j * N + j + 1];
        if (v > 0.0f) {
            s += (double)v;
            ++n;
        }
    }
    ssum[tid] = s;
    scnt[tid] = n;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if (tid < off) {
            ssum[tid] += ssum[tid + off];
            scnt[tid] += scnt[tid + off];
        }
        __syncthreads();
    }
    const float fill = scnt[0] > 0 ? (float)(ssum[0] / scnt[0]) : __builtin_nanf("");
    for (int j = tid; j < nframes - 1; j += 1024) trc[(int64_t)j * N + j] = fill;
}

__global__ void fill_f32_kernel(float* __restrict__ p, int64_t n, float v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

void launch_ttcf_tiles(int tile, const StdTile* tiles, int n_tiles, const BinDev* bins,
                       const int* scratch, const int64_t* sums, int mode, int n_bins,
                       int nframes, float* trc, hipStream_t s) {
    if (n_tiles <= 0) return;
    if (tile == 128)
        ttcf_tile_kernel<128><<<dim3((unsigned)n_tiles), dim3(256), 0, s>>>(
            tiles, bins, scratch, sums, mode, n_bins, nframes, trc);
    else
        ttcf_tile_kernel<256><<<dim3((unsigned)n_tiles), dim3(256), 0, s>>>(
            tiles, bins, scratch, sums, mode, n_bins, nframes, trc);
}

void launch_ttcf_diag(float* trc, int nframes, const int64_t* sums, int mode, int n_bins, int b,
                      hipStream_t s) {
    if (nframes <= 0) return;
    ttcf_diag_kernel<<<dim3(1), dim3(1024), 0, s>>>(trc, nframes, sums, mode, n_bins, b);
}

void launch_fill_f32(float* p, int64_t n, float v, hipStream_t s) {
    if (n <= 0) return;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    fill_f32_kernel<<<dim3((unsigned)blocks), dim3(256), 0, s>>>(p, n, v);
}

// dev(tau) of every bin: the segments of diagonal tau lie in tiles (ti, ti + D) for
// D = tau / T (slot tau - D T + T - 1) and D + 1 (slot tau - D T - 1, if >= 0).
__global__ __launch_bounds__(256) void dev_finalize_kernel(const StdPart* __restrict__ part,
                                                           const BinDev* __restrict__ bins,
                                                           int nframes, int T, float* __restrict__ dev) {
    const int b = blockIdx.x;
    const int nT = (nframes + T - 1) / T;
    const int64_t ntri = (int64_t)nT * (nT + 1) / 2;
    const int ndiag = 2 * T - 1;
    const bool empty = bins[b].npix == 0;
    for (int tau = threadIdx.x; tau < nframes; tau += blockDim.x) {
        double n = 0.0, mean = 0.0, m2 = 0.0;
        const int D0 = tau / T;
        for (int D = D0; D <= D0 + 1 && D < nT && !empty; ++D) {
            const int s = tau - D * T + T - 1;
            if (s < 0) continue;
            for (int ti = 0; ti + D < nT; ++ti) {
                const int64_t idx = (int64_t)b * ntri + (int64_t)ti * nT -
                                    (int64_t)ti * (ti - 1) / 2 + D;
                const StdPart p = part[idx * ndiag + s];
                if (p.n == 0.0) continue;
                const double nn = n + p.n;
                const double delta = p.mean - mean;
                mean += delta * (p.n / nn);
                m2 += p.m2 + delta * delta * (n * p.n / nn);
                n = nn;
            }
        }
        float out = __builtin_nanf("");
        if (!empty && n == (double)(nframes - tau))
            out = (float)(sqrt(m2 / n) / sqrt(n));
        dev[(int64_t)b * nframes + tau] = out;
    }
}

void launch_dev_finalize(const StdPart* part, const BinDev* bins, int n_bins, int nframes,
                         int tile, float* dev, hipStream_t s) {
    if (n_bins <= 0 || nframes <= 0) return;
    dev_finalize_kernel<<<dim3((unsigned)n_bins), dim3(256), 0, s>>>(part, bins, nframes, tile,
                                                                     dev);
}

}  // namespace dxg2
