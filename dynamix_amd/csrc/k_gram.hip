// Grouped i8-MFMA Gram over every q-bin with the diagonal (tau) reduction fused.
//
// Reference numerator: per bin, a float32 sgemm of the gathered (N, n_b) matrix with
// itself followed by a host loop over the N diagonals (correlator/dense.py:54,61-64),
// or the OpenCL loop of densecorrelator.cl:6-57.  Here one launch covers every bin:
// a work item is one TxT output tile (rows ti, columns tj >= ti) of one bin's Gram
// G = A A^T over a range of k-blocks.  Both operands are frame panels of the same i8
// slab; v_mfma_i32_32x32x32_i8 accumulates exact int32 dot products (the work list
// caps a range at 2047 k-blocks so |sum| < 2^31), and the epilogue folds the tile
// into its 2T-1 diagonals tau = (tj-ti)*T + c - r in LDS (64-bit) before one 64-bit
// global atomic per diagonal.  Nothing of size N x N is ever stored.
#include "dxg2_internal.h"

namespace dxg2 {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

template <int WM, int WN>
struct GramShape {
    static constexpr int TM = 2 * WM * 32;              // 2 x 2 waves
    static constexpr int TN = 2 * WN * 32;
    static constexpr int LDS_STRIDE = KB + 16;         // 80-byte rows: conflict-free b128 reads
    static constexpr int NDIAG = TM + TN - 1;
    static constexpr int CA = TM * (KB / 16) / 256;     // 16-byte chunks per thread per panel
};

// Loads the CA 16-byte chunks this thread moves for a TM-row panel starting at frame
// row0 of k-block pk.  Rows past the last frame are clamped to a valid address and
// zeroed by value (the i8 zero contributes nothing to the Gram).
template <int CA>
__device__ __forceinline__ void load_panel(int4 (&r)[CA], const int8_t* __restrict__ pk,
                                           int row0, int nframes, int tid) {
#pragma unroll
    for (int c = 0; c < CA; ++c) {
        const int chunk = tid + c * 256;
        const int t = row0 + (chunk >> 2);
        const int tc = t < nframes ? t : nframes - 1;
        int4 v = *reinterpret_cast<const int4*>(pk + (int64_t)tc * KB + (chunk & 3) * 16);
        if (t >= nframes) v = make_int4(0, 0, 0, 0);
        r[c] = v;
    }
}

template <int CA, int STRIDE>
__device__ __forceinline__ void store_panel(const int4 (&r)[CA], int8_t* s, int tid) {
#pragma unroll
    for (int c = 0; c < CA; ++c) {
        const int chunk = tid + c * 256;
        *reinterpret_cast<int4*>(s + (chunk >> 2) * STRIDE + (chunk & 3) * 16) = r[c];
    }
}

// One tile.  DIAG: ti == tj, a single panel feeds both operands and the MFMA blocks
// strictly below the main diagonal (which only feed tau < 0) are skipped.
template <int WM, int WN, bool DIAG>
__device__ __forceinline__ void gram_tile(const int8_t* __restrict__ slab_bin, int kb0, int kb1,
                                          int row0, int col0, int nframes, int8_t* sA,
                                          int8_t* sB, unsigned long long* sDiag,
                                          int64_t* __restrict__ out) {
    using S = GramShape<WM, WN>;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wr = wave >> 1;
    const int wc = wave & 1;
    const int8_t* sBp = DIAG ? sA : sB;

    for (int i = tid; i < S::NDIAG; i += 256) sDiag[i] = 0ull;

    v16i acc[WM][WN];
#pragma unroll
    for (int mi = 0; mi < WM; ++mi)
#pragma unroll
        for (int ni = 0; ni < WN; ++ni)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[mi][ni][v] = 0;

    const int frag_row = lane & 31;
    const int frag_k = (lane >> 5) * 16;
    const int64_t kstride = (int64_t)nframes * KB;

    int4 ra[S::CA], rb[S::CA];
    load_panel<S::CA>(ra, slab_bin + kb0 * kstride, row0, nframes, tid);
    if (!DIAG) load_panel<S::CA>(rb, slab_bin + kb0 * kstride, col0, nframes, tid);

    for (int kb = kb0; kb < kb1; ++kb) {
        __syncthreads();
        store_panel<S::CA, S::LDS_STRIDE>(ra, sA, tid);
        if (!DIAG) store_panel<S::CA, S::LDS_STRIDE>(rb, sB, tid);
        __syncthreads();
        if (kb + 1 < kb1) {
            load_panel<S::CA>(ra, slab_bin + (kb + 1) * kstride, row0, nframes, tid);
            if (!DIAG) load_panel<S::CA>(rb, slab_bin + (kb + 1) * kstride, col0, nframes, tid);
        }
#pragma unroll
        for (int ks = 0; ks < KB / 32; ++ks) {
            v4i a[WM], b[WN];
#pragma unroll
            for (int mi = 0; mi < WM; ++mi)
                a[mi] = *reinterpret_cast<const v4i*>(
                    sA + (wr * WM * 32 + mi * 32 + frag_row) * S::LDS_STRIDE + ks * 32 + frag_k);
#pragma unroll
            for (int ni = 0; ni < WN; ++ni)
                b[ni] = *reinterpret_cast<const v4i*>(
                    sBp + (wc * WN * 32 + ni * 32 + frag_row) * S::LDS_STRIDE + ks * 32 + frag_k);
#pragma unroll
            for (int mi = 0; mi < WM; ++mi)
#pragma unroll
                for (int ni = 0; ni < WN; ++ni) {
                    const bool live = !DIAG || (wc * WN * 32 + ni * 32 + 31 >= wr * WM * 32 + mi * 32);
                    if (live)
                        acc[mi][ni] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[mi], b[ni],
                                                                            acc[mi][ni], 0, 0, 0);
                }
        }
    }

    // Epilogue.  32x32 i32 accumulator layout: lane holds column (lane & 31) and rows
    // 8*(v/4) + 4*(lane/32) + (v%4), v = 0..15.  Diagonal slot c - r + TM - 1.
    __syncthreads();
#pragma unroll
    for (int mi = 0; mi < WM; ++mi)
#pragma unroll
        for (int ni = 0; ni < WN; ++ni) {
            const bool live = !DIAG || (wc * WN * 32 + ni * 32 + 31 >= wr * WM * 32 + mi * 32);
            if (!live) continue;
            const int cbase = wc * WN * 32 + ni * 32 + (lane & 31);
            const int rbase = wr * WM * 32 + mi * 32 + 4 * (lane >> 5);
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int r = rbase + 8 * (v >> 2) + (v & 3);
                const int val = acc[mi][ni][v];
                if (val != 0)
                    atomicAdd(&sDiag[cbase - r + S::TM - 1], (unsigned long long)(long long)val);
            }
        }
    __syncthreads();
    for (int i = tid; i < S::NDIAG; i += 256) {
        const long long v = (long long)sDiag[i];
        const int tau = col0 - row0 + i - (S::TM - 1);
        if (v != 0 && tau >= 0 && tau < nframes)
            atomicAdd(reinterpret_cast<unsigned long long*>(out + tau), (unsigned long long)v);
    }
}

template <int WM, int WN>
__global__ __launch_bounds__(256) void gram_i8_kernel(const int8_t* __restrict__ slab,
                                                      const BinDev* __restrict__ bins,
                                                      const GramWork* __restrict__ work,
                                                      int nframes,
                                                      int64_t* __restrict__ ugram) {
    using S = GramShape<WM, WN>;
    static_assert(S::TM == S::TN, "square tiles");
    __shared__ __attribute__((aligned(16))) int8_t sA[S::TM * S::LDS_STRIDE];
    __shared__ __attribute__((aligned(16))) int8_t sB[S::TN * S::LDS_STRIDE];
    __shared__ unsigned long long sDiag[S::NDIAG];

    const GramWork w = work[blockIdx.x];
    if (w.bin < 0) return;  // padding item
    const BinDev bd = bins[w.bin];
    const int8_t* slab_bin = slab + bd.slab_off;
    int64_t* out = ugram + (int64_t)w.bin * nframes;
    if (w.ti == w.tj)
        gram_tile<WM, WN, true>(slab_bin, w.kb0, w.kb1, w.ti * S::TM, w.tj * S::TN, nframes,
                                sA, sB, sDiag, out);
    else
        gram_tile<WM, WN, false>(slab_bin, w.kb0, w.kb1, w.ti * S::TM, w.tj * S::TN, nframes,
                                 sA, sB, sDiag, out);
}

int gram_tile_size(int variant) { return variant == 1 ? 128 : 256; }

void launch_gram_i8(int tile, const int8_t* slab, const BinDev* bins, const GramWork* work,
                    int n_work, int nframes, int64_t* ugram, hipStream_t s) {
    if (n_work <= 0) return;
    if (tile == 128)
        gram_i8_kernel<2, 2><<<dim3((unsigned)n_work), dim3(256), 0, s>>>(slab, bins, work,
                                                                           nframes, ugram);
    else
        gram_i8_kernel<4, 4><<<dim3((unsigned)n_work), dim3(256), 0, s>>>(slab, bins, work,
                                                                           nframes, ugram);
}

}  // namespace dxg2
