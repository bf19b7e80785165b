// uint16 frames: wide pixels apart.
//
// A uint16 count v = 256 h + l is correlated as two biased i8 digit planes (k_gram.hip):
// the Gram of the high bytes (HH) and the cross Gram of high and low bytes cost three
// quarters of the work, yet h is non-zero only for pixels that ever count above 255.
// Per call:
//  * the general gather (gather_slots_kernel<FLAG>) writes every pixel's low bytes to the
//    low plane in list order and flags the pixels with h != 0 in some frame, in the
//    same pass over the stack;
//  * wide_compact_kernel lists each bin's wide pixels (n_wide of them) and
//    wide_clip_kernel sets clip = nwkb[b] = ceil(n_wide / 64), or nkb when 2 clip > nkb
//    (then every pixel goes through HH and the cross Gram, in list order);
//  * wide_gather_kernel writes the first clip k-blocks of the wide list (padded past
//    n_wide) to the high plane: high bytes to k-blocks [0, clip), low bytes to
//    [clip, 2 clip) (with clip = nkb: every k-block's high bytes, list order);
//  * HH items stop after clip k-blocks and the cross items read [0, clip) against its
//    low copy (item_range in k_gram.hip); LL runs over the whole low plane;
//  * wide_low_sums_kernel gives the low-byte frame sums over the k-blocks the cross Gram
//    reads, which the bias correction of the clipped terms needs (k_finalize.hip,
//    mode 3).
// The order of pixels within a bin, and which copy of a low byte a term reads, change no
// numerator or frame sum.
#include "dxg2_internal.h"

namespace dxg2 {

// The wide pixels of each bin, compacted: one thread per list entry; a wide pixel takes
// the next free entry of its bin's wide list in pix_w (in no particular order: no sum
// depends on it), counted in count[b] (zeroed beforehand).  Grid (entry chunks of 256,
// bins).
__global__ __launch_bounds__(256) void wide_compact_kernel(const int32_t* __restrict__ pix,
                                                           const BinDev* __restrict__ bins,
                                                           const uint8_t* __restrict__ flag,
                                                           int32_t* __restrict__ pix_w,
                                                           int* __restrict__ count) {
    const int b = blockIdx.y;
    const BinDev bd = bins[b];
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= bd.npix) return;
    const int p = pix[bd.idx_off + k];
    if (flag[p]) {
        const int pos = atomicAdd(count + b, 1);
        pix_w[bd.idx_off + pos] = p;
    }
}

// clip = nwkb[b] = ceil(count[b] / 64), or nkb when twice that exceeds nkb; nwkb[n_bins]
// gets the largest clip (zeroed beforehand).
__global__ __launch_bounds__(256) void wide_clip_kernel(const BinDev* __restrict__ bins,
                                                        int n_bins, const int* __restrict__ count,
                                                        int* __restrict__ nwkb) {
    for (int b = threadIdx.x; b < n_bins; b += 256) {
        const int nkb = bins[b].nkb;
        const int c = (count[b] + KB - 1) / KB;
        const int clip = 2 * c <= nkb ? c : nkb;
        nwkb[b] = clip;
        atomicMax(nwkb + n_bins, clip);
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

void launch_wide_partition(const int32_t* pix, const BinDev* bins, int n_bins, int max_npix,
                           const uint8_t* flag, int32_t* pix_w, int* nwkb, hipStream_t s) {
    if (n_bins <= 0) return;
    int* count = nwkb + n_bins + 1;
    launch_fill_i32(nwkb + n_bins, n_bins + 1, s);  // the maximum and the counts
    if (max_npix > 0)
        wide_compact_kernel<<<dim3((unsigned)((max_npix + 255) / 256), (unsigned)n_bins), 256, 0,
                              s>>>(pix, bins, flag, pix_w, count);
    wide_clip_kernel<<<1, 256, 0, s>>>(bins, n_bins, count, nwkb);
}

void launch_wide_low_sums(const int8_t* slab, const BinDev* bins, const int* nwkb, int n_bins,
                          int nframes, int nrows, int64_t* lsums, hipStream_t s) {
    if (n_bins <= 0 || nframes <= 0) return;
    wide_low_sums_kernel<<<dim3((unsigned)((nframes + 255) / 256), (unsigned)n_bins), 256, 0, s>>>(
        slab, bins, nwkb, nframes, nrows, lsums);
}

}  // namespace dxg2
