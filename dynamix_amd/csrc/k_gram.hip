// Grouped i8-MFMA Gram over every q-bin with the diagonal (tau) reduction fused.
//
// Reference numerator: per bin, a float32 sgemm of the gathered (N, n_b) matrix with
// itself followed by a host loop over the N diagonals (correlator/dense.py:54,61-64),
// or the OpenCL loop of densecorrelator.cl:6-57.  Here one launch covers every bin:
// a work item is one T x T output tile (rows ti, columns tj >= ti) of one bin's Gram
// G = A A^T over a range of k-blocks.  Both operands are frame panels of the same i8
// slab; v_mfma_i32_32x32x32_i8 accumulates exact int32 dot products (the work list
// caps a range at 2047 k-blocks so |sum| < 2^31), and the epilogue folds the tile
// into its 2T-1 diagonals tau = (tj-ti)*T + c - r in LDS before one 64-bit global
// atomic per diagonal.  Nothing of size N x N is ever stored.
//
// Data movement: frame panels go from HBM straight into LDS with
// global_load_lds_dwordx4 (no register staging) through a ring of 4 one-k-block
// stages, so three stages are in flight while the MFMAs consume the fourth.  Slab rows
// are padded with zero rows to a multiple of 256 frames, so every load is in bounds
// and rows past the last frame add nothing.  In LDS a panel row is 64 bytes = four
// 16-byte quads, stored at quad (q ^ ((row >> 2) & 3)): the MFMA fragment reads (16
// consecutive rows, one quad) and the DMA writes (contiguous 1 KiB) are both free of
// bank conflicts.
//
// Diagonal tiles (ti == tj) read each frame row of their panel exactly once per
// k-range, so they also produce the per-frame bin sums S_b(t) (densecorrelator.cl:
// 65-150) from the staged bytes with v_dot4 -- no extra pass over the data.
//
// uint16 frames arrive as two biased digit planes H, L (k-blocks [0, nkb) and
// [nkb, 2 nkb) of the bin).  Work items then come in three kinds, each adding into its
// own output row: H H^T over [0, nkb), L L^T over [nkb, 2 nkb) (both in the plain
// kernel), and the cross product [H|L] [L|H]^T = H L^T + L H^T over [0, 2 nkb) in the
// CROSS kernel, whose B operand reads k-block kb +- nkb.  The cross Gram is symmetric,
// so the same upper-triangle tiles cover it; its diagonal tiles run as full tiles
// (A != B) and the epilogue drops their lags below the diagonal.
//
// The loop is kept free of stores and of extra control flow: the compiler's tracking of
// which LDS stage each DMA writes is easily lost (it then waits for every DMA in flight
// before an LDS read), so the cross-plane variant is a separate kernel.
//
// LDS diagonal sums stay in 32-bit words: every int32 value v is split as
// v = hi * 2^16 + lo with lo in [0, 2^16) and hi in [-2^15, 2^15); a diagonal of a
// T <= 256 tile collects at most 256 values, so sum(lo) < 2^24 and |sum(hi)| < 2^23
// are exact, and the 64-bit total hi_sum * 2^16 + lo_sum is formed once per slot.
#include "dxg2_internal.h"

namespace dxg2 {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef __attribute__((address_space(3))) void lds_void;

template <int WM, int WN>
struct GramShape {
    static constexpr int TM = 2 * WM * 32;          // 2 x 2 waves of WM x WN MFMA blocks
    static constexpr int TN = 2 * WN * 32;
    static constexpr int PANEL = TM * KB;           // bytes of one staged panel (1 k-block)
    static constexpr int NDIAG = TM + TN - 1;
    static constexpr int LOADS = TM / 64;           // 1 KiB DMA instructions per wave per panel
};

// Physical LDS offset of (row, logical 16-byte quad q) in a staged panel.
__device__ __forceinline__ int swz(int row, int q) { return row * KB + ((q ^ (row >> 2)) & 3) * 16; }

// Issue the DMA of one panel (rows row0 .. row0+TM of k-block kb) into LDS `dst`.
// Instruction j of wave w fills LDS rows (w*LOADS + j)*16 .. +16; lane i writes the
// 16 bytes at dst + 1 KiB*(w*LOADS+j) + 16 i, i.e. row +i/4, physical quad i%4, so it
// fetches the logical quad (i%4) ^ ((row >> 2) & 3) of that row.
template <int WM, int WN>
__device__ __forceinline__ void dma_panel(const int8_t* __restrict__ slab_bin, int kb, int row0,
                                          int nrows, int8_t* dst, int tid) {
    using S = GramShape<WM, WN>;
    const int w = tid >> 6;
    const int i = tid & 63;
#pragma unroll
    for (int j = 0; j < S::LOADS; ++j) {
        const int g = w * S::LOADS + j;
        const int row = g * 16 + (i >> 2);
        const int q = ((i & 3) ^ (row >> 2)) & 3;
        const int8_t* src = slab_bin + ((int64_t)kb * nrows + row0 + row) * KB + q * 16;
        __builtin_amdgcn_global_load_lds((const void*)src, (lds_void*)(dst + g * 1024), 16, 0, 0);
    }
}

// s_waitcnt vmcnt(N), other counters left at their maximum.
template <int N>
__device__ __forceinline__ void wait_vm() {
    __builtin_amdgcn_s_waitcnt((N & 0xF) | ((N >> 4) << 14) | (0x7 << 4) | (0xF << 8));
}

// LDS of one workgroup: a ring of 4 stages for each operand, as separate arrays so
// every access names its stage statically (the compiler then waits only for the DMA
// that wrote the stage being read).
template <int WM, int WN>
struct GramLds {
    using S = GramShape<WM, WN>;
    int8_t* a[4];
    int8_t* b[4];
    unsigned int* diag;
};

template <int WM, int WN, bool DIAG, bool CROSS, bool STORE = false>
struct GramTile {
    using S = GramShape<WM, WN>;
    static constexpr int L = DIAG ? S::LOADS : 2 * S::LOADS;  // DMA per thread per stage
    const int8_t* __restrict__ slab_bin;
    int kb0, kb1, row0, col0, nframes, nrows;
    int tid, lane, wr, wc;
    int nkbp;  // CROSS: k-blocks per digit plane (B reads kb + nkbp or kb - nkbp)
    GramLds<WM, WN> lds;
    v16i acc[WM][WN];
    int rowsum;

    __device__ __forceinline__ static bool live(int wr, int wc, int mi, int ni) {
        return !DIAG || (wc * WN * 32 + ni * 32 + 31 >= wr * WM * 32 + mi * 32);
    }

    template <int SLOT>
    __device__ __forceinline__ void issue(int kb) {
        dma_panel<WM, WN>(slab_bin, kb, row0, nrows, lds.a[SLOT], tid);
        if (!DIAG)
            dma_panel<WM, WN>(slab_bin, CROSS ? (kb < nkbp ? kb + nkbp : kb - nkbp) : kb, col0,
                              nrows, lds.b[SLOT], tid);
    }

    // Consume stage kb from ring slot SLOT; refill slot (SLOT + 3) % 4 with kb + 3.
    template <int SLOT>
    __device__ __forceinline__ void stage(int kb) {
        const int later = kb1 - 1 - kb;  // stages issued after kb that may still be in flight
        if (later >= 2)
            wait_vm<2 * L>();
        else if (later == 1)
            wait_vm<L>();
        else
            wait_vm<0>();
        // The barrier orders (a) every wave's completed DMA of this stage before any
        // read of it and (b) every read of the previous stage before its slot is refilled.
        __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
        __builtin_amdgcn_s_barrier();
        if (kb + 3 < kb1) issue<(SLOT + 3) & 3>(kb + 3);
        const int8_t* a_s = lds.a[SLOT];
        const int8_t* b_s = DIAG ? lds.a[SLOT] : lds.b[SLOT];
        if (DIAG && tid < S::TM) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const v4i x = *reinterpret_cast<const v4i*>(a_s + swz(tid, q));
                rowsum = __builtin_amdgcn_sdot4(x.x, 0x01010101, rowsum, false);
                rowsum = __builtin_amdgcn_sdot4(x.y, 0x01010101, rowsum, false);
                rowsum = __builtin_amdgcn_sdot4(x.z, 0x01010101, rowsum, false);
                rowsum = __builtin_amdgcn_sdot4(x.w, 0x01010101, rowsum, false);
            }
        }
        const int frag_row = lane & 31;
        const int frag_q = lane >> 5;
#pragma unroll
        for (int ks = 0; ks < KB / 32; ++ks) {
            v4i a[WM], b[WN];
#pragma unroll
            for (int mi = 0; mi < WM; ++mi)
                a[mi] = *reinterpret_cast<const v4i*>(
                    a_s + swz(wr * WM * 32 + mi * 32 + frag_row, ks * 2 + frag_q));
#pragma unroll
            for (int ni = 0; ni < WN; ++ni)
                b[ni] = *reinterpret_cast<const v4i*>(
                    b_s + swz(wc * WN * 32 + ni * 32 + frag_row, ks * 2 + frag_q));
#pragma unroll
            for (int mi = 0; mi < WM; ++mi)
#pragma unroll
                for (int ni = 0; ni < WN; ++ni)
                    if (live(wr, wc, mi, ni))
                        acc[mi][ni] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[mi], b[ni],
                                                                            acc[mi][ni], 0, 0, 0);
        }
    }

    // STORE: also write the int32 tile to `tile` (row-major T x T) and leave the frame
    // sums to a separate pass (`sums` unused).
    __device__ __forceinline__ void run(int biased, int64_t* __restrict__ out,
                                        int64_t* __restrict__ sums,
                                        int* __restrict__ tile = nullptr) {
        unsigned int* sDiag = lds.diag;
        for (int i = tid; i < 2 * S::NDIAG; i += 256) sDiag[i] = 0u;
#pragma unroll
        for (int mi = 0; mi < WM; ++mi)
#pragma unroll
            for (int ni = 0; ni < WN; ++ni)
#pragma unroll
                for (int v = 0; v < 16; ++v) acc[mi][ni][v] = 0;
        rowsum = 0;

        if (kb0 < kb1) issue<0>(kb0);
        if (kb0 + 1 < kb1) issue<1>(kb0 + 1);
        if (kb0 + 2 < kb1) issue<2>(kb0 + 2);
        for (int kb = kb0; kb < kb1; kb += 4) {
            stage<0>(kb);
            if (kb + 1 < kb1) stage<1>(kb + 1);
            if (kb + 2 < kb1) stage<2>(kb + 2);
            if (kb + 3 < kb1) stage<3>(kb + 3);
        }

        if (DIAG && !STORE && tid < S::TM) {
            const int t = row0 + tid;
            const int64_t v = (int64_t)rowsum + (biased ? (int64_t)128 * (kb1 - kb0) * KB : 0);
            if (t < nframes && v != 0)
                atomicAdd(reinterpret_cast<unsigned long long*>(sums + t), (unsigned long long)v);
        }

        // Epilogue.  32x32 i32 accumulator layout: lane holds column (lane & 31) and
        // rows 8*(v/4) + 4*(lane/32) + (v%4), v = 0..15.  Diagonal slot c - r + TM - 1.
        __syncthreads();
#pragma unroll
        for (int mi = 0; mi < WM; ++mi)
#pragma unroll
            for (int ni = 0; ni < WN; ++ni) {
                if (live(wr, wc, mi, ni)) {
                    const int cbase = wc * WN * 32 + ni * 32 + (lane & 31);
                    const int rbase = wr * WM * 32 + mi * 32 + 4 * (lane >> 5);
#pragma unroll
                    for (int v = 0; v < 16; ++v) {
                        const int r = rbase + 8 * (v >> 2) + (v & 3);
                        const int val = acc[mi][ni][v];
                        const int slot = cbase - r + S::TM - 1;
                        atomicAdd(&sDiag[slot], (unsigned)(val & 0xFFFF));
                        atomicAdd(&sDiag[S::NDIAG + slot], (unsigned)(val >> 16));
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        if (STORE) {
            // Lanes 0-31 and 32-63 of one store write two 128-byte row segments.
#pragma unroll
            for (int mi = 0; mi < WM; ++mi)
#pragma unroll
                for (int ni = 0; ni < WN; ++ni) {
                    if (live(wr, wc, mi, ni)) {
                        const int c = wc * WN * 32 + ni * 32 + (lane & 31);
                        const int rbase = wr * WM * 32 + mi * 32 + 4 * (lane >> 5);
#pragma unroll
                        for (int v = 0; v < 16; ++v)
                            tile[(rbase + 8 * (v >> 2) + (v & 3)) * S::TN + c] = acc[mi][ni][v];
                    }
                }
        }
        __syncthreads();
        for (int i = tid; i < S::NDIAG; i += 256) {
            const long long v =
                (long long)(int)sDiag[S::NDIAG + i] * 65536 + (long long)sDiag[i];
            const int tau = col0 - row0 + i - (S::TM - 1);
            if (v != 0 && tau >= 0 && tau < nframes)
                atomicAdd(reinterpret_cast<unsigned long long*>(out + tau), (unsigned long long)v);
        }
    }
};

template <int WM, int WN, bool CROSS>
__global__ __launch_bounds__(256) void gram_i8_kernel(const int8_t* __restrict__ slab,
                                                      const BinDev* __restrict__ bins,
                                                      const GramWork* __restrict__ work,
                                                      int nframes, int nrows, int biased,
                                                      int64_t* __restrict__ ugram,
                                                      int64_t* __restrict__ sums) {
    using S = GramShape<WM, WN>;
    static_assert(S::TM == S::TN, "square tiles");
    __shared__ __attribute__((aligned(1024))) int8_t a0[S::PANEL], a1[S::PANEL], a2[S::PANEL],
        a3[S::PANEL];
    __shared__ __attribute__((aligned(1024))) int8_t b0[S::PANEL], b1[S::PANEL], b2[S::PANEL],
        b3[S::PANEL];
    __shared__ unsigned int sDiag[2 * S::NDIAG];

    const GramWork w = work[blockIdx.x];
    if (w.bin < 0) return;  // padding item
    const BinDev bd = bins[w.bin];
    const GramLds<WM, WN> lds{{a0, a1, a2, a3}, {b0, b1, b2, b3}, sDiag};
    const int tid = threadIdx.x;
    const int wave = tid >> 6;
    int64_t* out = ugram + (int64_t)w.out * nframes;
    int64_t* srow = sums + (int64_t)w.out * nframes;
    if (!CROSS && w.ti == w.tj) {
        GramTile<WM, WN, true, false> t{slab + bd.slab_off, w.kb0, w.kb1, w.ti * S::TM,
                                        w.tj * S::TN, nframes, nrows, tid, tid & 63,
                                        wave >> 1, wave & 1, 0, lds};
        t.run(biased, out, srow);
    } else {
        GramTile<WM, WN, false, CROSS> t{slab + bd.slab_off, w.kb0, w.kb1, w.ti * S::TM,
                                         w.tj * S::TN, nframes, nrows, tid, tid & 63,
                                         wave >> 1, wave & 1, bd.nkb, lds};
        t.run(biased, out, srow);
    }
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
    __shared__ __attribute__((aligned(1024))) int8_t a0[S::PANEL], a1[S::PANEL], a2[S::PANEL],
        a3[S::PANEL];
    __shared__ __attribute__((aligned(1024))) int8_t b0[S::PANEL], b1[S::PANEL], b2[S::PANEL],
        b3[S::PANEL];
    __shared__ unsigned int sDiag[2 * S::NDIAG];

    const GramWork w = work[blockIdx.x];
    if (w.bin < 0) return;  // padding item
    const BinDev bd = bins[w.bin];
    const GramLds<WM, WN> lds{{a0, a1, a2, a3}, {b0, b1, b2, b3}, sDiag};
    const int tid = threadIdx.x;
    const int wave = tid >> 6;
    int64_t* out = ugram + (int64_t)w.out * nframes;
    int* tile = tiles + (int64_t)w.slot * S::TM * S::TN;
    if (!CROSS && w.ti == w.tj) {
        GramTile<WM, WN, true, false, true> t{slab + bd.slab_off, w.kb0, w.kb1, w.ti * S::TM,
                                              w.tj * S::TN, nframes, nrows, tid, tid & 63,
                                              wave >> 1, wave & 1, 0, lds};
        t.run(biased, out, nullptr, tile);
    } else {
        GramTile<WM, WN, false, CROSS, true> t{slab + bd.slab_off, w.kb0, w.kb1, w.ti * S::TM,
                                               w.tj * S::TN, nframes, nrows, tid, tid & 63,
                                               wave >> 1, wave & 1, bd.nkb, lds};
        t.run(biased, out, nullptr, tile);
    }
}

// Grouped launch over several plans (continuous batching of dense jobs): the same tile
// code, with the slab, bins, frame counts and outputs of each item's job.
template <int WM, int WN, bool CROSS>
__global__ __launch_bounds__(256) void gram_i8_jobs_kernel(const GramJob* __restrict__ jobs,
                                                           const GramWork* __restrict__ work) {
    using S = GramShape<WM, WN>;
    static_assert(S::TM == S::TN, "square tiles");
    __shared__ __attribute__((aligned(1024))) int8_t a0[S::PANEL], a1[S::PANEL], a2[S::PANEL],
        a3[S::PANEL];
    __shared__ __attribute__((aligned(1024))) int8_t b0[S::PANEL], b1[S::PANEL], b2[S::PANEL],
        b3[S::PANEL];
    __shared__ unsigned int sDiag[2 * S::NDIAG];

    const GramWork w = work[blockIdx.x];
    if (w.bin < 0) return;  // padding item
    const GramJob j = jobs[w.job];
    const BinDev bd = j.bins[w.bin];
    const GramLds<WM, WN> lds{{a0, a1, a2, a3}, {b0, b1, b2, b3}, sDiag};
    const int tid = threadIdx.x;
    const int wave = tid >> 6;
    int64_t* out = j.ugram + (int64_t)w.out * j.nframes;
    int64_t* srow = j.sums + (int64_t)w.out * j.nframes;
    if (!CROSS && w.ti == w.tj) {
        GramTile<WM, WN, true, false> t{j.slab + bd.slab_off, w.kb0, w.kb1, w.ti * S::TM,
                                        w.tj * S::TN, j.nframes, j.nrows, tid, tid & 63,
                                        wave >> 1, wave & 1, 0, lds};
        t.run(j.biased, out, srow);
    } else {
        GramTile<WM, WN, false, CROSS> t{j.slab + bd.slab_off, w.kb0, w.kb1, w.ti * S::TM,
                                         w.tj * S::TN, j.nframes, j.nrows, tid, tid & 63,
                                         wave >> 1, wave & 1, bd.nkb, lds};
        t.run(j.biased, out, srow);
    }
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
                    int64_t* ugram, int64_t* sums, hipStream_t s) {
    if (n_work <= 0) return;
    const dim3 grid((unsigned)n_work), block(256);
    if (tile == 128) {
        if (cross)
            gram_i8_kernel<2, 2, true><<<grid, block, 0, s>>>(slab, bins, work, nframes, nrows,
                                                              biased, ugram, sums);
        else
            gram_i8_kernel<2, 2, false><<<grid, block, 0, s>>>(slab, bins, work, nframes, nrows,
                                                               biased, ugram, sums);
    } else {
        if (cross)
            gram_i8_kernel<4, 4, true><<<grid, block, 0, s>>>(slab, bins, work, nframes, nrows,
                                                              biased, ugram, sums);
        else
            gram_i8_kernel<4, 4, false><<<grid, block, 0, s>>>(slab, bins, work, nframes, nrows,
                                                               biased, ugram, sums);
    }
}

}  // namespace dxg2
