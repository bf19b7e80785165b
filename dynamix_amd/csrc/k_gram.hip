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
// LDS diagonal sums stay in 32-bit words: every int32 value v is split as
// v = hi * 2^16 + lo with lo in [0, 2^16) and hi in [-2^15, 2^15); a diagonal of a
// T <= 256 tile collects at most 256 values, so sum(lo) < 2^24 and |sum(hi)| < 2^23
// are exact, and the 64-bit total hi_sum * 2^16 + lo_sum is formed once per slot.
//
// Diagonal tiles (ti == tj) read each frame row of their panel exactly once per
// k-range, so they also produce the per-frame bin sums S_b(t) (densecorrelator.cl:
// 65-150) from the staged bytes with v_dot4 -- no extra pass over the data.
//
// Pipeline: a stage is 2 k-blocks (128 bytes of every panel row).  Global loads for
// stage s+1 are issued into registers before the MFMAs of stage s, written to the
// other LDS buffer after them, and one barrier per stage orders the two buffers.
#include "dxg2_internal.h"

namespace dxg2 {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

template <int WM, int WN>
struct GramShape {
    static constexpr int TM = 2 * WM * 32;          // 2 x 2 waves of WM x WN MFMA blocks
    static constexpr int TN = 2 * WN * 32;
    static constexpr int SKB = 2;                   // k-blocks per stage
    static constexpr int SK = SKB * KB;             // bytes of a panel row per stage
    static constexpr int STRIDE = SK + 16;          // padded LDS row (144 B)
    static constexpr int PANEL = TM * STRIDE;       // bytes of one staged panel
    static constexpr int NDIAG = TM + TN - 1;
    static constexpr int ROUNDS = TM * SK / (256 * 16);  // 16-byte loads per thread per panel
    static constexpr int ROWS_PER_ROUND = 256 * 16 / KB; // 64 rows of one k-block
};

// Global -> registers for one panel and one stage.  Round r covers k-block part
// (r & 1) of rows (r >> 1) * 64 .. +64; thread tid loads 16 bytes at row tid/4,
// chunk tid%4, so a round is one contiguous 4 KiB run of the slab.  Rows past the
// last frame and k-blocks past the item's range are clamped to valid addresses and
// zeroed by value (an i8 zero adds nothing to the Gram).
template <int WM, int WN, int R>
__device__ __forceinline__ void load_stage(int4 (&r)[R],
                                           const int8_t* __restrict__ slab_bin, int kb, int kb1,
                                           int row0, int nframes, int tid) {
    using S = GramShape<WM, WN>;
#pragma unroll
    for (int q = 0; q < S::ROUNDS; ++q) {
        const int part = q & 1;
        const int row = (q >> 1) * S::ROWS_PER_ROUND + (tid >> 2);
        const int t = row0 + row;
        const int kbp = kb + part;
        const bool ok = (t < nframes) && (kbp < kb1);
        const int tc = t < nframes ? t : nframes - 1;
        const int kc = kbp < kb1 ? kbp : kb;
        int4 v = *reinterpret_cast<const int4*>(slab_bin + ((int64_t)kc * nframes + tc) * KB +
                                                (tid & 3) * 16);
        if (!ok) v = make_int4(0, 0, 0, 0);
        r[q] = v;
    }
}

template <int WM, int WN, int R>
__device__ __forceinline__ void store_stage(const int4 (&r)[R],
                                            int8_t* s, int tid) {
    using S = GramShape<WM, WN>;
#pragma unroll
    for (int q = 0; q < S::ROUNDS; ++q) {
        const int row = (q >> 1) * S::ROWS_PER_ROUND + (tid >> 2);
        *reinterpret_cast<int4*>(s + row * S::STRIDE + (q & 1) * KB + (tid & 3) * 16) = r[q];
    }
}

// Per-frame sums of the staged rows this thread loaded (diagonal tiles only):
// rowsum[j] accumulates row (tid/4) + 64 j over this thread's 16-byte chunks.
template <int WM, int WN, int R>
__device__ __forceinline__ void sum_stage(const int4 (&r)[R], int (&rowsum)[R / 2]) {
    using S = GramShape<WM, WN>;
#pragma unroll
    for (int q = 0; q < S::ROUNDS; ++q) {
        int s = rowsum[q >> 1];
        s = __builtin_amdgcn_sdot4(r[q].x, 0x01010101, s, false);
        s = __builtin_amdgcn_sdot4(r[q].y, 0x01010101, s, false);
        s = __builtin_amdgcn_sdot4(r[q].z, 0x01010101, s, false);
        s = __builtin_amdgcn_sdot4(r[q].w, 0x01010101, s, false);
        rowsum[q >> 1] = s;
    }
}

// One tile.  DIAG: ti == tj, a single panel feeds both operands, the MFMA blocks
// strictly below the main diagonal (which only feed tau < 0) are skipped, and the
// per-frame sums of the panel rows are accumulated.
template <int WM, int WN, bool DIAG>
__device__ __forceinline__ void gram_tile(const int8_t* __restrict__ slab_bin, int kb0, int kb1,
                                          int row0, int col0, int nframes, int biased,
                                          int8_t* lds, unsigned int* sDiag,
                                          int64_t* __restrict__ out, int64_t* __restrict__ sums) {
    using S = GramShape<WM, WN>;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wr = wave >> 1;
    const int wc = wave & 1;
    // LDS: buffer b holds A at lds + b * 2 * PANEL, B right after it.
    auto sA = [&](int b) { return lds + b * 2 * S::PANEL; };
    auto sB = [&](int b) { return DIAG ? sA(b) : lds + b * 2 * S::PANEL + S::PANEL; };

    for (int i = tid; i < 2 * S::NDIAG; i += 256) sDiag[i] = 0u;

    v16i acc[WM][WN];
#pragma unroll
    for (int mi = 0; mi < WM; ++mi)
#pragma unroll
        for (int ni = 0; ni < WN; ++ni)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[mi][ni][v] = 0;

    int rowsum[S::ROUNDS / 2];
#pragma unroll
    for (int j = 0; j < S::ROUNDS / 2; ++j) rowsum[j] = 0;

    const int frag_row = lane & 31;
    const int frag_k = (lane >> 5) * 16;

    int4 ra[S::ROUNDS], rb[S::ROUNDS];
    load_stage<WM, WN>(ra, slab_bin, kb0, kb1, row0, nframes, tid);
    if (!DIAG) load_stage<WM, WN>(rb, slab_bin, kb0, kb1, col0, nframes, tid);
    if (DIAG) sum_stage<WM, WN>(ra, rowsum);
    store_stage<WM, WN>(ra, sA(0), tid);
    if (!DIAG) store_stage<WM, WN>(rb, sB(0), tid);
    __syncthreads();

    int buf = 0;
    for (int kb = kb0; kb < kb1; kb += S::SKB) {
        const bool more = kb + S::SKB < kb1;
        if (more) {
            load_stage<WM, WN>(ra, slab_bin, kb + S::SKB, kb1, row0, nframes, tid);
            if (!DIAG) load_stage<WM, WN>(rb, slab_bin, kb + S::SKB, kb1, col0, nframes, tid);
        }
        const int8_t* a_s = sA(buf);
        const int8_t* b_s = sB(buf);
        // Fragments of k-step ks+1 are read while the MFMAs of k-step ks run; the
        // scheduling barrier keeps the compiler from hoisting further LDS reads.
        v4i fa[2][WM], fb[2][WN];
#pragma unroll
        for (int mi = 0; mi < WM; ++mi)
            fa[0][mi] = *reinterpret_cast<const v4i*>(
                a_s + (wr * WM * 32 + mi * 32 + frag_row) * S::STRIDE + frag_k);
#pragma unroll
        for (int ni = 0; ni < WN; ++ni)
            fb[0][ni] = *reinterpret_cast<const v4i*>(
                b_s + (wc * WN * 32 + ni * 32 + frag_row) * S::STRIDE + frag_k);
#pragma unroll
        for (int ks = 0; ks < S::SK / 32; ++ks) {
            const int cur = ks & 1;
            if (ks + 1 < S::SK / 32) {
#pragma unroll
                for (int mi = 0; mi < WM; ++mi)
                    fa[cur ^ 1][mi] = *reinterpret_cast<const v4i*>(
                        a_s + (wr * WM * 32 + mi * 32 + frag_row) * S::STRIDE + (ks + 1) * 32 +
                        frag_k);
#pragma unroll
                for (int ni = 0; ni < WN; ++ni)
                    fb[cur ^ 1][ni] = *reinterpret_cast<const v4i*>(
                        b_s + (wc * WN * 32 + ni * 32 + frag_row) * S::STRIDE + (ks + 1) * 32 +
                        frag_k);
            }
#pragma unroll
            for (int mi = 0; mi < WM; ++mi)
#pragma unroll
                for (int ni = 0; ni < WN; ++ni) {
                    const bool live =
                        !DIAG || (wc * WN * 32 + ni * 32 + 31 >= wr * WM * 32 + mi * 32);
                    if (live)
                        acc[mi][ni] = __builtin_amdgcn_mfma_i32_32x32x32_i8(
                            fa[cur][mi], fb[cur][ni], acc[mi][ni], 0, 0, 0);
                }
            __builtin_amdgcn_sched_barrier(0);
        }
        if (more) {
            if (DIAG) sum_stage<WM, WN>(ra, rowsum);
            store_stage<WM, WN>(ra, sA(buf ^ 1), tid);
            if (!DIAG) store_stage<WM, WN>(rb, sB(buf ^ 1), tid);
        }
        __syncthreads();
        buf ^= 1;
    }

    if (DIAG) {
        // rows (tid/4) + 64 j: combine the 4 chunks of a row, add the bias back.
        const int64_t kitems = (int64_t)(kb1 - kb0) * KB;
#pragma unroll
        for (int j = 0; j < S::ROUNDS / 2; ++j) {
            int s = rowsum[j];
            s += __shfl_xor(s, 1);
            s += __shfl_xor(s, 2);
            const int t = row0 + j * S::ROWS_PER_ROUND + (tid >> 2);
            if ((tid & 3) == 0 && t < nframes) {
                const int64_t v = (int64_t)s + (biased ? 128 * kitems : 0);
                if (v != 0)
                    atomicAdd(reinterpret_cast<unsigned long long*>(sums + t),
                              (unsigned long long)v);
            }
        }
    }

    // Epilogue.  32x32 i32 accumulator layout: lane holds column (lane & 31) and rows
    // 8*(v/4) + 4*(lane/32) + (v%4), v = 0..15.  Diagonal slot c - r + TM - 1.
    // One block at a time keeps the accumulator reads from being hoisted together.
#pragma unroll
    for (int mi = 0; mi < WM; ++mi)
#pragma unroll
        for (int ni = 0; ni < WN; ++ni) {
            const bool live = !DIAG || (wc * WN * 32 + ni * 32 + 31 >= wr * WM * 32 + mi * 32);
            if (live) {
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
    __syncthreads();
    for (int i = tid; i < S::NDIAG; i += 256) {
        const long long v = (long long)(int)sDiag[S::NDIAG + i] * 65536 + (long long)sDiag[i];
        const int tau = col0 - row0 + i - (S::TM - 1);
        if (v != 0 && tau >= 0 && tau < nframes)
            atomicAdd(reinterpret_cast<unsigned long long*>(out + tau), (unsigned long long)v);
    }
}

template <int WM, int WN>
__global__ __launch_bounds__(256) void gram_i8_kernel(const int8_t* __restrict__ slab,
                                                      const BinDev* __restrict__ bins,
                                                      const GramWork* __restrict__ work,
                                                      int nframes, int biased,
                                                      int64_t* __restrict__ ugram,
                                                      int64_t* __restrict__ sums) {
    using S = GramShape<WM, WN>;
    static_assert(S::TM == S::TN, "square tiles");
    __shared__ __attribute__((aligned(16))) int8_t lds[4 * S::PANEL];
    __shared__ unsigned int sDiag[2 * S::NDIAG];

    const GramWork w = work[blockIdx.x];
    if (w.bin < 0) return;  // padding item
    const BinDev bd = bins[w.bin];
    const int8_t* slab_bin = slab + bd.slab_off;
    int64_t* out = ugram + (int64_t)w.bin * nframes;
    int64_t* srow = sums + (int64_t)w.bin * nframes;
    if (w.ti == w.tj)
        gram_tile<WM, WN, true>(slab_bin, w.kb0, w.kb1, w.ti * S::TM, w.tj * S::TN, nframes,
                                biased, lds, sDiag, out, srow);
    else
        gram_tile<WM, WN, false>(slab_bin, w.kb0, w.kb1, w.ti * S::TM, w.tj * S::TN, nframes,
                                 biased, lds, sDiag, out, srow);
}

int gram_tile_size(int variant) { return variant == 1 ? 128 : 256; }

void launch_gram_i8(int tile, const int8_t* slab, const BinDev* bins, const GramWork* work,
                    int n_work, int nframes, int biased, int64_t* ugram, int64_t* sums,
                    hipStream_t s) {
    if (n_work <= 0) return;
    if (tile == 128)
        gram_i8_kernel<2, 2><<<dim3((unsigned)n_work), dim3(256), 0, s>>>(
            slab, bins, work, nframes, biased, ugram, sums);
    else
        gram_i8_kernel<4, 4><<<dim3((unsigned)n_work), dim3(256), 0, s>>>(
            slab, bins, work, nframes, biased, ugram, sums);
}

}  // namespace dxg2
