// uint16 frames: wide pixels apart.
//
// A uint16 count v = 256 h + l is correlated as two biased i8 digit planes (k_gram.hip):
// the Gram of the high bytes (HH) and the cross Gram of high and low bytes cost three
// quarters of the work, yet h is non-zero only for pixels that ever count above 255.
// Per call:
//  * the general gather (gather_slots_kernel<FLAG>) writes every pixel's low bytes to the
//    low plane in list order and flags the pixels with h != 0 in some frame, in the
//    same pass over the stack;
//  * wide_partition_kernel reorders each bin's pixel list stably, wide pixels first, and
//    sets clip = nwkb[b] = ceil(n_wide / 64), or nkb when 2 clip > nkb (then every
//    pixel goes through HH and the cross Gram, in list order);
//  * wide_gather_kernel writes the first clip k-blocks of the wide-first list to the
//    high plane: high bytes to k-blocks [0, clip), low bytes to [clip, 2 clip) (with
//    clip = nkb: every k-block's high bytes, list order);
//  * HH items stop after clip k-blocks and the cross items read [0, clip) against its
//    low copy (item_range in k_gram.hip); LL runs over the whole low plane;
//  * wide_low_sums_kernel gives the low-byte frame sums over the k-blocks the cross Gram
//    reads, which the bias correction of the clipped terms needs (k_finalize.hip,
//    mode 3).
// The order of pixels within a bin, and which copy of a low byte a term reads, change no
// numerator or frame sum.
#include "dxg2_internal.h"

namespace dxg2 {

// Exclusive scan of one value per thread over a 1024-thread block; adds the block total
// to `carry` and returns the value's exclusive prefix.
__device__ __forceinline__ int wide_block_scan(int v, int* wave_tot, int& carry) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) wave_tot[wave] = x;
    __syncthreads();
    int before = carry, total = carry;
    for (int w = 0; w < 16; ++w) {
        if (w < wave) before += wave_tot[w];
        total += wave_tot[w];
    }
    __syncthreads();
    carry = total;
    return before + x - v;
}

// One 1024-thread block per bin: the bin's pixel list, stably reordered wide-first, into
// pix_w (same layout as pix; the padding entries past n_b are left as they are), and
// nwkb[b] = ceil(n_wide / 64), or nkb when twice that exceeds nkb; nwkb[n_bins] gets the
// largest of them.  Two passes over the list: count, then place.
__global__ __launch_bounds__(1024) void wide_partition_kernel(const int32_t* __restrict__ pix,
                                                              const BinDev* __restrict__ bins,
                                                              const uint8_t* __restrict__ flag,
                                                              int32_t* __restrict__ pix_w,
                                                              int* __restrict__ nwkb) {
    __shared__ int wave_tot[16];
    const BinDev bd = bins[blockIdx.x];
    const int32_t* src = pix + bd.idx_off;
    int32_t* dst = pix_w + bd.idx_off;
    const int n = bd.npix;
    int nw = 0;
    for (int k0 = 0; k0 < n; k0 += 1024) {
        const int k = k0 + threadIdx.x;
        const int w = k < n ? flag[src[k]] : 0;
        (void)wide_block_scan(w, wave_tot, nw);
    }
    int cw = 0, cn = 0;
    for (int k0 = 0; k0 < n; k0 += 1024) {
        const int k = k0 + threadIdx.x;
        const int p = k < n ? src[k] : -1;
        const int w = k < n ? flag[p] : 0;
        const int rw = wide_block_scan(w, wave_tot, cw);
        const int rn = wide_block_scan(k < n ? 1 - w : 0, wave_tot, cn);
        if (k < n) dst[w ? rw : nw + rn] = p;
    }
    if (threadIdx.x == 0) {
        const int clip = (nw + KB - 1) / KB;
        const int c = 2 * clip <= bd.nkb ? clip : bd.nkb;
        nwkb[blockIdx.x] = c;
        atomicMax(nwkb + gridDim.x, c);  // the largest over the bins (zeroed beforehand)
    }
}

// S_Lw(t): the low bytes of frame t summed over the clip = nwkb[b] k-blocks the cross
// Gram reads them from (k-blocks clip .. 2 clip - 1 of the bin, or the whole low plane
// nkb .. 2 nkb - 1 when clip = nkb; stored biased), into lsums[b][t].
__global__ __launch_bounds__(256) void wide_low_sums_kernel(const int8_t* __restrict__ slab,
                                                            const BinDev* __restrict__ bins,
                                                            const int* __restrict__ nwkb,
                                                            int nframes, int nrows,
                                                            int64_t* __restrict__ lsums) {
    const int b = blockIdx.y;
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= nframes) return;
    const BinDev bd = bins[b];
    const int nw = nwkb[b];
    const int off = 2 * nw <= bd.nkb ? nw : bd.nkb;
    const int8_t* base = slab + bd.slab_off;
    int64_t total = 0;
    for (int kb = off; kb < off + nw; ++kb) {
        int acc = 0;  // |acc| <= 64 * 128
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int4 x = *reinterpret_cast<const int4*>(base + slab_offset(kb, t, 16 * q, nrows));
            acc = __builtin_amdgcn_sdot4(x.x, 0x01010101, acc, false);
            acc = __builtin_amdgcn_sdot4(x.y, 0x01010101, acc, false);
            acc = __builtin_amdgcn_sdot4(x.z, 0x01010101, acc, false);
            acc = __builtin_amdgcn_sdot4(x.w, 0x01010101, acc, false);
        }
        total += acc;
    }
    lsums[(int64_t)b * nframes + t] = total + (int64_t)128 * KB * nw;
}

void launch_wide_partition(const int32_t* pix, const BinDev* bins, int n_bins,
                           const uint8_t* flag, int32_t* pix_w, int* nwkb, hipStream_t s) {
    if (n_bins <= 0) return;
    wide_partition_kernel<<<dim3((unsigned)n_bins), 1024, 0, s>>>(pix, bins, flag, pix_w, nwkb);
}

void launch_wide_low_sums(const int8_t* slab, const BinDev* bins, const int* nwkb, int n_bins,
                          int nframes, int nrows, int64_t* lsums, hipStream_t s) {
    if (n_bins <= 0 || nframes <= 0) return;
    wide_low_sums_kernel<<<dim3((unsigned)((nframes + 255) / 256), (unsigned)n_bins), 256, 0, s>>>(
        slab, bins, nwkb, nframes, nrows, lsums);
}

}  // namespace dxg2
