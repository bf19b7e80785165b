// uint16 frames: wide pixels first.
//
// A uint16 count v = 256 h + l is correlated as two biased i8 digit planes (k_gram.hip):
// the Gram of the high bytes (HH) and the cross Gram of high and low bytes cost three
// quarters of the work, yet h is non-zero only for pixels that ever count above 255.
// Per call, before the gather:
//  * wide_flag_kernel streams the stack once and marks the pixels with h != 0 in any
//    frame;
//  * wide_partition_kernel reorders each bin's pixel list stably, wide pixels first, and
//    records the bin's wide k-blocks nwkb = ceil(n_wide / 64).
// The gather then fills the slab in that order, so the high plane of k-blocks >= nwkb is
// all zero (biased: all -128); the HH and cross work items stop after nwkb k-blocks, and
// wide_low_sums_kernel supplies the low-byte frame sums over those k-blocks, which the
// bias correction of the clipped cross Gram needs (k_finalize.hip, mode 3).  The order of
// pixels within a bin does not change any numerator or frame sum.
#include "dxg2_internal.h"

namespace dxg2 {

// Grid (pixel groups, frame chunks).  A thread ORs the counts of 8 pixels (16 bytes of a
// frame row) when the stack is 16-byte aligned and H*W is a multiple of 8, else of one
// pixel, over frames [y*fch, (y+1)*fch), and sets the flag byte of each pixel with a high
// byte by a vector atomicOr on the flag words (pixel 4j+k is byte k of word j), which
// start at zero.  Splitting the frame axis keeps many more loads in flight than one
// thread per pixel group over all frames; a flag word is written only when a chunk holds
// a count > 255, which is rare.
template <bool VEC>
__global__ __launch_bounds__(256) void wide_flag_kernel(const uint16_t* __restrict__ frames,
                                                        int64_t hw, int nframes, int fch,
                                                        uint32_t* __restrict__ flagw) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int t0 = blockIdx.y * fch;
    const int t1 = min(nframes, t0 + fch);
    if (VEC) {
        if (8 * i >= hw) return;
        const uint16_t* src = frames + 8 * i;
        uint32_t acc[4] = {0u, 0u, 0u, 0u};
#pragma unroll 8
        for (int t = t0; t < t1; ++t) {
            const uint4 v = *reinterpret_cast<const uint4*>(src + (int64_t)t * hw);
            acc[0] |= v.x;
            acc[1] |= v.y;
            acc[2] |= v.z;
            acc[3] |= v.w;
        }
        uint32_t f[2] = {0u, 0u};
#pragma unroll
        for (int q = 0; q < 4; ++q) {  // pixel 2q in the low half of word q, 2q+1 high
            const int sh = 8 * (2 * q & 3);
            f[q >> 1] |= ((acc[q] & 0xFF00u) ? 1u : 0u) << sh;
            f[q >> 1] |= ((acc[q] & 0xFF000000u) ? 1u : 0u) << (sh + 8);
        }
        if (f[0]) atomicOr(flagw + 2 * i, f[0]);
        if (f[1]) atomicOr(flagw + 2 * i + 1, f[1]);
    } else {
        if (i >= hw) return;
        uint32_t acc = 0u;
        for (int t = t0; t < t1; ++t) acc |= frames[(int64_t)t * hw + i];
        if (acc & 0xFF00u) atomicOr(flagw + (i >> 2), 1u << (8 * (int)(i & 3)));
    }
}

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
// nwkb[b] = ceil(n_wide / 64).  Two passes over the list: count, then place.
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
    if (threadIdx.x == 0) nwkb[blockIdx.x] = (nw + KB - 1) / KB;
}

// S_Lw(t): the low bytes of frame t summed over the first nwkb[b] k-blocks of bin b's low
// plane (k-blocks nkb .. nkb + nwkb - 1, stored biased), into lsums[b][t].
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
    const int8_t* base = slab + bd.slab_off;
    int acc = 0;
    for (int kb = bd.nkb; kb < bd.nkb + nw; ++kb) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int4 x = *reinterpret_cast<const int4*>(base + slab_offset(kb, t, 16 * q, nrows));
            acc = __builtin_amdgcn_sdot4(x.x, 0x01010101, acc, false);
            acc = __builtin_amdgcn_sdot4(x.y, 0x01010101, acc, false);
            acc = __builtin_amdgcn_sdot4(x.z, 0x01010101, acc, false);
            acc = __builtin_amdgcn_sdot4(x.w, 0x01010101, acc, false);
        }
    }
    lsums[(int64_t)b * nframes + t] = (int64_t)acc + (int64_t)128 * KB * nw;
}

void launch_wide_first(const uint16_t* frames, int64_t hw, int nframes, const int32_t* pix,
                       const BinDev* bins, int n_bins, uint8_t* flag, int32_t* pix_w, int* nwkb,
                       hipStream_t s) {
    if (n_bins <= 0 || nframes <= 0) return;
    uint32_t* flagw = reinterpret_cast<uint32_t*>(flag);  // 256-byte aligned, >= hw + 3 bytes
    launch_fill_u32(flagw, (hw + 3) / 4, s);
    const bool vec = reinterpret_cast<uintptr_t>(frames) % 16 == 0 && hw % 8 == 0;
    const int64_t bx = ((vec ? hw / 8 : hw) + 255) / 256;
    // about 16k workgroups in all, each thread over at least 128 frames
    const int64_t want = std::max<int64_t>(1, 16384 / std::max<int64_t>(bx, 1));
    const int ny = (int)std::max<int64_t>(1, std::min<int64_t>(want, nframes / 128));
    const int fch = (nframes + ny - 1) / ny;
    const dim3 grid((unsigned)bx, (unsigned)((nframes + fch - 1) / fch));
    if (vec)
        wide_flag_kernel<true><<<grid, 256, 0, s>>>(frames, hw, nframes, fch, flagw);
    else
        wide_flag_kernel<false><<<grid, 256, 0, s>>>(frames, hw, nframes, fch, flagw);
    wide_partition_kernel<<<dim3((unsigned)n_bins), 1024, 0, s>>>(pix, bins, flag, pix_w, nwkb);
}

void launch_wide_low_sums(const int8_t* slab, const BinDev* bins, const int* nwkb, int n_bins,
                          int nframes, int nrows, int64_t* lsums, hipStream_t s) {
    if (n_bins <= 0 || nframes <= 0) return;
    wide_low_sums_kernel<<<dim3((unsigned)((nframes + 255) / 256), (unsigned)n_bins), 256, 0, s>>>(
        slab, bins, nwkb, nframes, nrows, lsums);
}

}  // namespace dxg2
