// Grouped i8-MFMA Gram over every q-bin with the diagonal (tau) reduction fused.
//
// Reference numerator: per bin, a float32 sgemm of the gathered (N, n_b) matrix with
// itself followed by a host loop over the N diagonals (correlator/dense.py:54,61-64),
// or the OpenCL loop of densecorrelator.cl:6-57.  Here one launch covers every bin:
// a work item is one T x T output tile (rows ti, columns tj >= ti) of one bin's Gram
// G = A A^T over a range of k-blocks.  v_mfma_i32_32x32x32_i8 accumulates exact int32
// dot products (the work list caps a range at 2047 k-blocks so |sum| < 2^31), and the
// epilogue folds the tile into its 2T-1 diagonals tau = (tj-ti)*T + c - r in LDS
// before one 64-bit global atomic per diagonal.  Nothing of size N x N is stored
// (except by the calc_std variant, which also writes its tile to a scratch slot).
//
// Data movement: the slab is stored fragment-contiguous (slab_offset()): the 32 frames
// x 32 bytes an MFMA operand takes from one k-block are one 1 KiB run in lane order.
// Each wave (2 x 2 waves per tile, WM x WN blocks of 32 x 32 each) loads its operand
// fragments straight from global memory into VGPRs with one global_load_dwordx4 per
// fragment -- no LDS staging, no barriers in the loop.  Two register slots form a
// software pipeline: the loads of k-block kb + 2 are issued right after the MFMAs of
// kb have consumed the slot, so a full k-block of MFMAs (WM*WN*2 instructions) covers
// every load's latency.  Two compiler behaviours must be kept out of the loop:
//   * the machine scheduler sinks loads towards their first use (it would undo the
//     pipeline), so sched_barrier fences pin the order MFMA group / load group;
//   * the waitcnt pass takes the minimum over paths, so a conditionally issued load
//     makes it wait for every load in flight: every path issues the same loads (the
//     k-block index is clamped to the item's last k-block near the end).
// Both operands of a wave are re-read by the wave beside it; those reads hit L1/L2.
//
// Diagonal tiles (ti == tj): the wave below the diagonal has nothing to compute and
// skips the loop; the other waves compute whole 32 x 32 blocks and the epilogue drops
// lags below the diagonal.  The two waves on the diagonal also sum the bytes of their
// A fragments (v_dot4), which gives the per-frame bin sums S_b(t) (densecorrelator.cl:
// 65-150) with no extra pass over the data.
//
// uint16 frames arrive as two biased digit planes H, L (k-blocks [0, nkb) and
// [nkb, 2 nkb) of the bin).  Work items then come in three kinds, each adding into its
// own output row: H H^T over [0, nkb), L L^T over [nkb, 2 nkb) (both in the plain
// kernel), and the cross product [H|L] [L|H]^T = H L^T + L H^T over [0, 2 nkb) in the
// CROSS kernel, whose B operand reads k-block kb +- nkb.  The cross Gram is symmetric,
// so the same upper-triangle tiles cover it; its diagonal tiles run as full tiles.
//
// LDS diagonal sums stay in 32-bit words: every int32 value v is split as
// v = hi * 2^16 + lo with lo in [0, 2^16) and hi in [-2^15, 2^15); a diagonal of a
// T <= 256 tile collects at most 256 values, so sum(lo) < 2^24 and |sum(hi)| < 2^23
// are exact, and the 64-bit total hi_sum * 2^16 + lo_sum is formed once per slot.
#include <algorithm>

#include "dxg2_internal.h"

namespace dxg2 {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef __attribute__((address_space(1))) const v4i global_v4i;

// 16-byte load through a global-address-space pointer: a generic pointer (one read from
// a job descriptor) would give flat loads, which the waitcnt pass also counts as LDS
// traffic and drains before every MFMA group.
__device__ __forceinline__ v4i gload16(const int8_t* p) {
    return *reinterpret_cast<global_v4i*>(reinterpret_cast<uintptr_t>(p));
}

// 64-bit atomic add through a global-address-space pointer (see gload16).
__device__ __forceinline__ void gatomic_add(int64_t* p, int64_t v) {
    __hip_atomic_fetch_add(reinterpret_cast<__attribute__((address_space(1))) unsigned long long*>(
                               reinterpret_cast<uintptr_t>(p)),
                           (unsigned long long)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A bin descriptor read through a global-address-space pointer (two 16-byte loads).
static_assert(sizeof(BinDev) == 32, "BinDev is loaded as two 16-byte words");
__device__ __forceinline__ BinDev gload_bin(const BinDev* p) {
    union {
        v4i v[2];
        BinDev b;
    } u;
    u.v[0] = gload16(reinterpret_cast<const int8_t*>(p));
    u.v[1] = gload16(reinterpret_cast<const int8_t*>(p) + 16);
    return u.b;
}

template <int WM, int WN>
struct GramShape {
    static constexpr int TM = 2 * WM * 32;  // 2 x 2 waves of WM x WN MFMA blocks
    static constexpr int TN = 2 * WN * 32;
    static constexpr int NDIAG = TM + TN - 1;
};

template <int WM, int WN, bool DIAG, bool CROSS, bool STORE = false>
struct GramTile {
    using S = GramShape<WM, WN>;
    const int8_t* __restrict__ slab_bin;
    int kb0, kb1, row0, col0, nframes, nrows;
    int tid, lane, wr, wc;
    int nkbp;  // CROSS: k-blocks per digit plane (B reads kb + nkbp or kb - nkbp)
    unsigned int* sDiag;

    v16i acc[WM][WN];
    v4i fa[2][2][WM];  // [slot][ks][block]
    v4i fb[2][2][WN];
    int rs[WM];        // DIAG: partial frame sums of this lane's A rows

    __device__ __forceinline__ void load(int s, int kb, const int8_t* pa, const int8_t* pb,
                                         int64_t kstride) {
        const int8_t* a = pa + kb * kstride;
        const int kbb = CROSS ? (kb < nkbp ? kb + nkbp : kb - nkbp) : kb;
        const int8_t* b = pb + kbb * kstride;
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
#pragma unroll
            for (int m = 0; m < WM; ++m) fa[s][ks][m] = gload16(a + m * 2048 + ks * 1024);
#pragma unroll
            for (int n = 0; n < WN; ++n) fb[s][ks][n] = gload16(b + n * 2048 + ks * 1024);
        }
    }

    __device__ __forceinline__ void mma(int s) {
#pragma unroll
        for (int ks = 0; ks < 2; ++ks)
#pragma unroll
            for (int m = 0; m < WM; ++m) {
#pragma unroll
                for (int n = 0; n < WN; ++n)
                    acc[m][n] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[s][ks][m], fb[s][ks][n],
                                                                      acc[m][n], 0, 0, 0);
                if (DIAG) {
                    const v4i x = fa[s][ks][m];
                    rs[m] = __builtin_amdgcn_sdot4(x.x, 0x01010101, rs[m], false);
                    rs[m] = __builtin_amdgcn_sdot4(x.y, 0x01010101, rs[m], false);
                    rs[m] = __builtin_amdgcn_sdot4(x.z, 0x01010101, rs[m], false);
                    rs[m] = __builtin_amdgcn_sdot4(x.w, 0x01010101, rs[m], false);
                }
            }
    }

    // STORE: also write the int32 tile to `tile` (row-major T x T) and leave the frame
    // sums to a separate pass (`sums` unused).
    __device__ __forceinline__ void run(int biased, int64_t* __restrict__ out,
                                        int64_t* __restrict__ sums,
                                        int* __restrict__ tile = nullptr) {
        for (int i = tid; i < 2 * S::NDIAG; i += 256) sDiag[i] = 0u;
#pragma unroll
        for (int m = 0; m < WM; ++m) {
            rs[m] = 0;
#pragma unroll
            for (int n = 0; n < WN; ++n)
#pragma unroll
                for (int v = 0; v < 16; ++v) acc[m][n][v] = 0;
        }

        // Fragment (k-block kb, 32-frame group g, half ks) starts at
        // slab_bin + (kb * ngroups + g) * 2048 + ks * 1024; lane l reads 16 bytes at +16 l.
        const int64_t kstride = (int64_t)(nrows >> 5) * 2048;
        const int8_t* pa = slab_bin + (int64_t)((row0 >> 5) + wr * WM) * 2048 + lane * 16;
        const int8_t* pb = slab_bin + (int64_t)((col0 >> 5) + wc * WN) * 2048 + lane * 16;
        // idle: every block of this wave is below the diagonal, or lies wholly in the
        // padding frames past the last one (a STORE tile's readers skip both parts)
        const bool idle = (DIAG && wr > wc) || row0 + wr * WM * 32 >= nframes ||
                          col0 + wc * WN * 32 >= nframes;
        if (!idle && kb0 < kb1) {
            const int last = kb1 - 1;
            load(0, kb0, pa, pb, kstride);
            load(1, min(kb0 + 1, last), pa, pb, kstride);
            for (int kb = kb0; kb < kb1; kb += 2) {
                mma(0);
                __builtin_amdgcn_sched_barrier(0);
                load(0, min(kb + 2, last), pa, pb, kstride);
                __builtin_amdgcn_sched_barrier(0);
                // An odd count ends with slot 1 holding a repeat of the last k-block:
                // its A fragments are zeroed so those MFMAs add nothing (no tail code
                // after the loop, whose live fragments would spill).
                const bool repeat = kb + 1 > last;
#pragma unroll
                for (int ks = 0; ks < 2; ++ks)
#pragma unroll
                    for (int m = 0; m < WM; ++m) {
                        const v4i x = fa[1][ks][m];
                        fa[1][ks][m] = v4i{repeat ? 0 : x.x, repeat ? 0 : x.y, repeat ? 0 : x.z,
                                           repeat ? 0 : x.w};
                    }
                mma(1);
                __builtin_amdgcn_sched_barrier(0);
                load(1, min(kb + 3, last), pa, pb, kstride);
                __builtin_amdgcn_sched_barrier(0);
            }
        }

        if (DIAG && !STORE && wr == wc) {
            // lanes l and l + 32 hold the two 32-byte halves of frame row (l & 31)
            const int64_t bias = biased ? (int64_t)128 * (kb1 - kb0) * KB : 0;
#pragma unroll
            for (int m = 0; m < WM; ++m) {
                const int v = rs[m] + __shfl_xor(rs[m], 32);
                const int t = row0 + wr * WM * 32 + m * 32 + lane;
                const int64_t s64 = (int64_t)v + bias;
                if (lane < 32 && t < nframes && s64 != 0) gatomic_add(sums + t, s64);
            }
        }

        // Epilogue.  32x32 i32 accumulator layout: lane holds column (lane & 31) and
        // rows 8*(v/4) + 4*(lane/32) + (v%4), v = 0..15.  Diagonal slot c - r + TM - 1.
        __syncthreads();
        if (!idle) {
            // slot = lane part + (n - m) 32 - 8 (v / 4) - v % 4; the second part is a
            // compile-time constant >= MINOFF, so every atomic is one base VGPR plus an
            // immediate offset (no per-element address registers).
            constexpr int MINOFF = -(WM - 1) * 32 - 15;
            unsigned int* sd = sDiag + (wc * WN * 32 + (lane & 31)) -
                               (wr * WM * 32 + 4 * (lane >> 5)) + S::TM - 1 + MINOFF;
#pragma unroll
            for (int m = 0; m < WM; ++m)
#pragma unroll
                for (int n = 0; n < WN; ++n) {
#pragma unroll
                    for (int v = 0; v < 16; ++v) {
                        const int k = (n - m) * 32 - 8 * (v >> 2) - (v & 3) - MINOFF;
                        const int val = acc[m][n][v];
                        atomicAdd(sd + k, (unsigned)(val & 0xFFFF));
                        atomicAdd(sd + S::NDIAG + k, (unsigned)(val >> 16));
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
            if (STORE) {
                // Lanes 0-31 and 32-63 of one store write two 128-byte row segments.
#pragma unroll
                for (int m = 0; m < WM; ++m)
#pragma unroll
                    for (int n = 0; n < WN; ++n) {
                        const int c = wc * WN * 32 + n * 32 + (lane & 31);
                        const int rbase = wr * WM * 32 + m * 32 + 4 * (lane >> 5);
#pragma unroll
                        for (int v = 0; v < 16; ++v)
                            tile[(rbase + 8 * (v >> 2) + (v & 3)) * S::TN + c] = acc[m][n][v];
                    }
            }
        }
        __syncthreads();
        for (int i = tid; i < S::NDIAG; i += 256) {
            const long long v =
                (long long)(int)sDiag[S::NDIAG + i] * 65536 + (long long)sDiag[i];
            const int tau = col0 - row0 + i - (S::TM - 1);
            if (v != 0 && tau >= 0 && tau < nframes) gatomic_add(out + tau, v);
        }
    }
};

// K-range of item `w` and the CROSS kernel's plane distance.  nwkb (uint16 plans whose
// wide pixels -- those with a count above 255 -- have k-blocks of their own; null: every
// item whole, distance nkb): per bin, `clip` k-blocks at the start of the high plane
// hold the wide pixels' high bytes, and
//  * HH items (plain, k-blocks below nkb) stop at clip; LL items are whole;
//  * cross items: the first half [0, nkb) stops at clip and reads its B operand `off`
//    k-blocks further; the second half [nkb, 2 nkb) moves to [off, off + clip) and reads
//    B `off` k-blocks back.  off = clip when the wide pixels' low bytes follow their high
//    bytes in the high plane (2 clip <= nkb), else nkb (clip = nkb: every pixel, the high
//    and low planes in list order).
struct ItemRange {
    int kb0, kb1, nkbp;
};
template <bool CROSS>
__device__ __forceinline__ ItemRange item_range(const GramWork& w, int nkb, const int* nwkb) {
    if (!nwkb) return ItemRange{w.kb0, w.kb1, nkb};
    const int clip = nwkb[w.bin];
    const int off = 2 * clip <= nkb ? clip : nkb;
    if (w.kb0 < nkb) return ItemRange{w.kb0, min(w.kb1, clip), off};
    if (!CROSS) return ItemRange{w.kb0, w.kb1, nkb};
    return ItemRange{off + (w.kb0 - nkb), off + min(w.kb1 - nkb, clip), off};
}

// One work item: the tile of `w` in a bin whose slab starts at slab_bin, over k-blocks
// [r.kb0, r.kb1) (a CROSS item's B operand r.nkbp k-blocks away).
template <int WM, int WN, bool CROSS, bool STORE>
__device__ __forceinline__ void gram_item(const GramWork& w, ItemRange r, const int8_t* slab_bin,
                                          int nframes, int nrows, int biased,
                                          int64_t* __restrict__ out, int64_t* __restrict__ sums,
                                          int* __restrict__ tile, unsigned int* sDiag) {
    using S = GramShape<WM, WN>;
    const int tid = threadIdx.x;
    const int wave = tid >> 6;
    if (!CROSS && w.ti == w.tj) {
        GramTile<WM, WN, true, false, STORE> t{slab_bin, r.kb0, r.kb1, w.ti * S::TM,
                                               w.tj * S::TN, nframes, nrows, tid, tid & 63,
                                               wave >> 1, wave & 1, 0, sDiag};
        t.run(biased, out, sums, tile);
    } else {
        GramTile<WM, WN, false, CROSS, STORE> t{slab_bin, r.kb0, r.kb1, w.ti * S::TM,
                                                w.tj * S::TN, nframes, nrows, tid, tid & 63,
                                                wave >> 1, wave & 1, r.nkbp, sDiag};
        t.run(biased, out, sums, tile);
    }
}

template <int WM, int WN, bool CROSS>
__global__ __launch_bounds__(256) void gram_i8_kernel(const int8_t* __restrict__ slab,
                                                      const BinDev* __restrict__ bins,
                                                      const GramWork* __restrict__ work,
                                                      int nframes, int nrows, int biased,
                                                      const int* __restrict__ nwkb,
                                                      int64_t* __restrict__ ugram,
                                                      int64_t* __restrict__ sums) {
    using S = GramShape<WM, WN>;
    static_assert(S::TM == S::TN, "square tiles");
    __shared__ unsigned int sDiag[2 * S::NDIAG];
    const GramWork w = work[blockIdx.x];
    if (w.bin < 0) return;  // padding item
    const BinDev bd = bins[w.bin];
    const ItemRange r = item_range<CROSS>(w, bd.nkb, nwkb);
    if (r.kb1 <= r.kb0) return;  // nothing of this item is left
    gram_item<WM, WN, CROSS, false>(w, r, slab + bd.slab_off, nframes, nrows, biased,
                                    ugram + (int64_t)w.out * nframes,
                                    sums + (int64_t)w.out * nframes, nullptr, sDiag);
}

// XCD (accelerator complex die) this wave runs on: HW_REG_XCC_ID bits [3:0].
__device__ __forceinline__ int xcc_id() {
    return (int)(__builtin_amdgcn_s_getreg((20) | (0 << 6) | (15 << 11)) & 0xF);
}

// Next work item for a workgroup on XCD `home`: the head of its own XCD's queue, then
// (once that is drained) the other XCDs' queues in turn.  Queue x holds items
// [qoff[x], qoff[x+1]) of the work list; heads[x] counts the items taken from it.
__device__ __forceinline__ int next_item(int home, const int* __restrict__ qoff, int* heads) {
    for (int k = 0; k < 8; ++k) {
        const int x = (home + k) & 7;
        const int n = qoff[x + 1] - qoff[x];
        if (__hip_atomic_load(heads + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= n)
            continue;
        const int i = atomicAdd(heads + x, 1);
        if (i < n) return qoff[x] + i;
    }
    return -1;
}

// Persistent grouped launch: one workgroup per resident slot, each taking tiles from
// the device work queues until all are empty.  Tiles of one queue share frame panels
// in their XCD's L2; stealing evens out the XCDs' finishing times.  heads[] must be
// zero at launch.
template <int WM, int WN, bool CROSS>
__global__ __launch_bounds__(256) void gram_i8_queue_kernel(
    const int8_t* __restrict__ slab, const BinDev* __restrict__ bins,
    const GramWork* __restrict__ work, const int* __restrict__ qoff, int* heads, int nframes,
    int nrows, int biased, const int* __restrict__ nwkb, int64_t* __restrict__ ugram,
    int64_t* __restrict__ sums) {
    using S = GramShape<WM, WN>;
    static_assert(S::TM == S::TN, "square tiles");
    __shared__ unsigned int sDiag[2 * S::NDIAG];
    __shared__ int s_item;
    const int home = xcc_id();
    for (;;) {
        if (threadIdx.x == 0) s_item = next_item(home, qoff, heads);
        __syncthreads();
        const int it = s_item;
        if (it < 0) break;
        const GramWork w = work[it];
        const BinDev bd = bins[w.bin];
        const ItemRange r = item_range<CROSS>(w, bd.nkb, nwkb);
        if (r.kb1 > r.kb0)  // the same for every thread of the workgroup
            gram_item<WM, WN, CROSS, false>(w, r, slab + bd.slab_off, nframes, nrows, biased,
                                            ugram + (int64_t)w.out * nframes,
                                            sums + (int64_t)w.out * nframes, nullptr, sDiag);
        __syncthreads();  // s_item and sDiag are reused by the next item
    }
}

template <int WM, int WN, bool CROSS>
int queue_grid() {
    static int grid[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (!grid[dev]) {
        int cus = 0, per = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            cus = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per, reinterpret_cast<const void*>(gram_i8_queue_kernel<WM, WN, CROSS>), 256, 0) !=
                hipSuccess ||
            per < 1)
            per = 1;
        grid[dev] = cus * per;
    }
    return grid[dev];
}

void gram_queue_prepare(int tile) {
    if (tile == 128) {
        (void)queue_grid<2, 2, false>();
        (void)queue_grid<2, 2, true>();
    } else {
        (void)queue_grid<4, 4, false>();
        (void)queue_grid<4, 4, true>();
    }
}

void launch_gram_i8_queue(int tile, bool cross, const int8_t* slab, const BinDev* bins,
                          const GramWork* work, int n_work, const int* qoff, int* heads,
                          int nframes, int nrows, int biased, const int* nwkb, int64_t* ugram,
                          int64_t* sums, hipStream_t s) {
    if (n_work <= 0) return;
    const dim3 block(256);
#define DXG2_QUEUE_LAUNCH(WM, CR)                                                             \
    gram_i8_queue_kernel<WM, WM, CR><<<dim3((unsigned)std::min(n_work, queue_grid<WM, WM, CR>())), \
                                       block, 0, s>>>(slab, bins, work, qoff, heads, nframes,    \
                                                      nrows, biased, nwkb, ugram, sums)
    if (tile == 128) {
        if (cross)
            DXG2_QUEUE_LAUNCH(2, true);
        else
            DXG2_QUEUE_LAUNCH(2, false);
    } else {
        if (cross)
            DXG2_QUEUE_LAUNCH(4, true);
        else
            DXG2_QUEUE_LAUNCH(4, false);
    }
#undef DXG2_QUEUE_LAUNCH
}

// calc_std variant: the same tiles (work items carry a scratch slot), which also write
// their int32 tile to tiles + slot * T * T for the per-element ratio pass.  Frame sums
// come from a separate pass over the slab, so that a chunk of tiles can be reduced as
// soon as its Gram is done.
template <int WM, int WN, bool CROSS>
__global__ __launch_bounds__(256) void gram_i8_store_kernel(
    const int8_t* __restrict__ slab, const BinDev* __restrict__ bins,
    const GramWork* __restrict__ work, int nframes, int nrows, int biased,
    int64_t* __restrict__ ugram, int* __restrict__ tiles) {
    using S = GramShape<WM, WN>;
    static_assert(S::TM == S::TN, "square tiles");
    __shared__ unsigned int sDiag[2 * S::NDIAG];
    const GramWork w = work[blockIdx.x];
    if (w.bin < 0) return;  // padding item
    const BinDev bd = bins[w.bin];
    gram_item<WM, WN, CROSS, true>(w, ItemRange{w.kb0, w.kb1, bd.nkb}, slab + bd.slab_off,
                                   nframes, nrows, biased, ugram + (int64_t)w.out * nframes,
                                   nullptr, tiles + (int64_t)w.slot * S::TM * S::TN, sDiag);
}

// Grouped launch over several plans (continuous batching of dense jobs): the same tile
// code, with the slab, bins, frame counts and outputs of each item's job.
template <int WM, int WN, bool CROSS>
__global__ __launch_bounds__(256) void gram_i8_jobs_kernel(const GramJob* __restrict__ jobs,
                                                           const GramWork* __restrict__ work) {
    using S = GramShape<WM, WN>;
    static_assert(S::TM == S::TN, "square tiles");
    __shared__ unsigned int sDiag[2 * S::NDIAG];
    const GramWork w = work[blockIdx.x];
    if (w.bin < 0) return;  // padding item
    const GramJob j = jobs[w.job];
    const BinDev bd = gload_bin(j.bins + w.bin);
    gram_item<WM, WN, CROSS, false>(w, ItemRange{w.kb0, w.kb1, bd.nkb}, j.slab + bd.slab_off,
                                    j.nframes, j.nrows, j.biased,
                                    j.ugram + (int64_t)w.out * j.nframes,
                                    j.sums + (int64_t)w.out * j.nframes, nullptr, sDiag);
}

void launch_gram_i8_jobs(int tile, bool cross, const GramJob* jobs, const GramWork* work,
                         int n_work, hipStream_t s) {
    if (n_work <= 0) return;
    const dim3 grid((unsigned)n_work), block(256);
    if (tile == 128) {
        if (cross)
            gram_i8_jobs_kernel<2, 2, true><<<grid, block, 0, s>>>(jobs, work);
        else
            gram_i8_jobs_kernel<2, 2, false><<<grid, block, 0, s>>>(jobs, work);
    } else {
        if (cross)
            gram_i8_jobs_kernel<4, 4, true><<<grid, block, 0, s>>>(jobs, work);
        else
            gram_i8_jobs_kernel<4, 4, false><<<grid, block, 0, s>>>(jobs, work);
    }
}

int gram_tile_size(int variant) { return variant == 1 ? 128 : 256; }

void launch_gram_i8_store(int tile, bool cross, const int8_t* slab, const BinDev* bins,
                          const GramWork* work, int n_work, int nframes, int nrows, int biased,
                          int64_t* ugram, int* tiles, hipStream_t s) {
    if (n_work <= 0) return;
    const dim3 grid((unsigned)n_work), block(256);
    if (tile == 128) {
        if (cross)
            gram_i8_store_kernel<2, 2, true><<<grid, block, 0, s>>>(slab, bins, work, nframes,
                                                                    nrows, biased, ugram, tiles);
        else
            gram_i8_store_kernel<2, 2, false><<<grid, block, 0, s>>>(slab, bins, work, nframes,
                                                                     nrows, biased, ugram, tiles);
    } else {
        if (cross)
            gram_i8_store_kernel<4, 4, true><<<grid, block, 0, s>>>(slab, bins, work, nframes,
                                                                    nrows, biased, ugram, tiles);
        else
            gram_i8_store_kernel<4, 4, false><<<grid, block, 0, s>>>(slab, bins, work, nframes,
                                                                     nrows, biased, ugram, tiles);
    }
}

void launch_gram_i8(int tile, bool cross, const int8_t* slab, const BinDev* bins,
                    const GramWork* work, int n_work, int nframes, int nrows, int biased,
                    const int* nwkb, int64_t* ugram, int64_t* sums, hipStream_t s) {
    if (n_work <= 0) return;
    const dim3 grid((unsigned)n_work), block(256);
    if (tile == 128) {
        if (cross)
            gram_i8_kernel<2, 2, true><<<grid, block, 0, s>>>(slab, bins, work, nframes, nrows,
                                                              biased, nwkb, ugram, sums);
        else
            gram_i8_kernel<2, 2, false><<<grid, block, 0, s>>>(slab, bins, work, nframes, nrows,
                                                               biased, nwkb, ugram, sums);
    } else {
        if (cross)
            gram_i8_kernel<4, 4, true><<<grid, block, 0, s>>>(slab, bins, work, nframes, nrows,
                                                              biased, nwkb, ugram, sums);
        else
            gram_i8_kernel<4, 4, false><<<grid, block, 0, s>>>(slab, bins, work, nframes, nrows,
                                                               biased, nwkb, ugram, sums);
    }
}

}  // namespace dxg2
