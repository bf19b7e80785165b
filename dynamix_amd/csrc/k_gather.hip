// qmask gather: frames (N, H, W) -> per-bin i8 slabs + exact per-frame bin sums S_b(t).
//
// The reference gathers one bin at a time on the host (correlator/dense.py:48-49,
// dense.py:317-319) and sums frames per bin in a separate pass
// (resources/opencl/densecorrelator.cl:65-150).  Here one launch covers every bin:
// each workgroup owns (bin, up to 16 k-blocks of 64 pixels, 64 frames), reads the
// pixels through the bin's raster-ordered index list, writes 1 KiB contiguous slab
// runs per wave and folds the per-frame sums into one 64-bit atomic per frame.
#include "dxg2_internal.h"

namespace dxg2 {

__global__ void fill_i64_kernel(int64_t* __restrict__ p, int64_t n) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = 0;
}

void launch_fill_i64(int64_t* p, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    fill_i64_kernel<<<dim3((unsigned)blocks), dim3(256), 0, s>>>(p, n);
}

// Work layout: blockIdx.x -> GatherWork (bin, first k-block, k-block count <= 16),
// blockIdx.y -> 64-frame tile.  Wave w takes k-blocks kb0+w, kb0+w+4, ...; in a
// k-block, lane l takes pixels (l&3)*16..+16 and frames t0 + 16*it + (l>>2), it=0..3,
// and stores its 16 bytes at slab[kb][t][(l&3)*16] -- a wave writes 1 KiB runs.
// BIAS (uint8 input): store v ^ 0x80 = v - 128 as i8, so every byte fits the signed
// MFMA operand; padded pixels (index -1) read as 0 and become -128 the same way.
template <bool BIAS>
__global__ __launch_bounds__(256) void gather_i8_kernel(
    const uint8_t* __restrict__ frames, int64_t hw, int nframes,
    const int32_t* __restrict__ pix_idx, const BinDev* __restrict__ bins,
    const GatherWork* __restrict__ work, int8_t* __restrict__ slab,
    int64_t* __restrict__ sums) {
    __shared__ int s_part[64];
    const GatherWork w = work[blockIdx.x];
    const BinDev bd = bins[w.bin];
    const int t0 = blockIdx.y * 64;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int g = lane & 3;
    const int f = lane >> 2;
    if (threadIdx.x < 64) s_part[threadIdx.x] = 0;
    __syncthreads();

    int acc[4] = {0, 0, 0, 0};
    for (int kb = w.kb0 + wave; kb < w.kb0 + w.nkb; kb += 4) {
        const int4* idx4 =
            reinterpret_cast<const int4*>(pix_idx + bd.idx_off + (int64_t)kb * KB + g * 16);
        int ix[16];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int4 v = idx4[q];
            ix[4 * q + 0] = v.x;
            ix[4 * q + 1] = v.y;
            ix[4 * q + 2] = v.z;
            ix[4 * q + 3] = v.w;
        }
        int8_t* dst = slab + bd.slab_off + (int64_t)kb * nframes * KB + g * 16;
#pragma unroll
        for (int it = 0; it < 4; ++it) {
            const int t = t0 + it * 16 + f;
            if (t < nframes) {
                const uint8_t* fr = frames + (int64_t)t * hw;
                uint32_t word[4];
                int s = 0;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    uint32_t wd = 0;
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        const int p = ix[4 * q + b];
                        const uint32_t v = p >= 0 ? (uint32_t)fr[p] : 0u;
                        s += BIAS ? (int)v : (int)(int8_t)v;
                        wd |= v << (8 * b);
                    }
                    word[q] = BIAS ? (wd ^ 0x80808080u) : wd;
                }
                *reinterpret_cast<uint4*>(dst + (int64_t)t * KB) =
                    make_uint4(word[0], word[1], word[2], word[3]);
                acc[it] += s;
            }
        }
    }
#pragma unroll
    for (int it = 0; it < 4; ++it) {
        acc[it] += __shfl_xor(acc[it], 1);
        acc[it] += __shfl_xor(acc[it], 2);
    }
    if (g == 0) {
#pragma unroll
        for (int it = 0; it < 4; ++it)
            if (acc[it] != 0) atomicAdd(&s_part[it * 16 + f], acc[it]);
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        const int t = t0 + threadIdx.x;
        const int v = s_part[threadIdx.x];
        if (t < nframes && v != 0)
            atomicAdd(reinterpret_cast<unsigned long long*>(sums + (int64_t)w.bin * nframes + t),
                      (unsigned long long)(int64_t)v);
    }
}

void launch_gather_i8(const uint8_t* frames, int dtype, int64_t hw, int nframes,
                      const int32_t* pix_idx, const BinDev* bins, const GatherWork* work,
                      int n_work, int8_t* slab, int64_t* sums, hipStream_t s) {
    if (n_work <= 0 || nframes <= 0) return;
    dim3 grid((unsigned)n_work, (unsigned)((nframes + 63) / 64));
    if (dtype == DXG2_U8)
        gather_i8_kernel<true><<<grid, dim3(256), 0, s>>>(frames, hw, nframes, pix_idx, bins,
                                                         work, slab, sums);
    else
        gather_i8_kernel<false><<<grid, dim3(256), 0, s>>>(frames, hw, nframes, pix_idx, bins,
                                                          work, slab, sums);
}

}  // namespace dxg2
