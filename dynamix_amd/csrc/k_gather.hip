// qmask gather: frames (N, H, W) -> per-bin i8 slabs slab[kb][t][64] (uint16 frames:
// two digit planes).
//
// The reference gathers one bin at a time on the host (correlator/dense.py:48-49,
// dense.py:317-319).  Here one launch covers every bin.  A wave owns one k-block (64
// consecutive entries of a bin's raster-ordered pixel list) and a 64-frame tile:
//   * each load instruction reads entry `lane` of one frame, so the 64 lanes read 64
//     consecutive list entries -- runs of neighbouring pixels, few cache lines;
//   * after four frames each lane holds 4 frames x 1 pixel; a 4x4 byte transpose
//     inside every lane quad (DPP quad broadcast + v_perm) gives lane 4q+i the dword
//     of pixels 4q..4q+3 for frame t+i, so one store instruction writes rows t..t+3
//     of the k-block: four 64-byte runs of the fragment-contiguous slab layout.
// The per-frame bin sums are formed later by the Gram kernel's diagonal tiles.
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

__global__ void fill_i32_kernel(int* __restrict__ p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

void launch_fill_i32(int* p, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    fill_i32_kernel<<<dim3((unsigned)blocks), dim3(256), 0, s>>>(p, n);
}

constexpr int GATHER_TF = 64;  // frames per workgroup tile

// Broadcast lane j of every quad to the whole quad (DPP quad_perm [j, j, j, j]).
template <int J>
__device__ __forceinline__ uint32_t quad_bcast(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, J * 0x55, 0xF, 0xF, false);
}

// 4x4 byte transpose inside a lane quad: lane 4q+i holds bytes of frames t..t+3 of pixel
// 4q+i in x; returns the dword of pixels 4q..4q+3 for frame t+i (i = lane & 3).
__device__ __forceinline__ uint32_t quad_transpose(uint32_t x, uint32_t sel_lo, uint32_t sel_hi) {
    const uint32_t y0 = quad_bcast<0>(x);
    const uint32_t y1 = quad_bcast<1>(x);
    const uint32_t y2 = quad_bcast<2>(x);
    const uint32_t y3 = quad_bcast<3>(x);
    return __builtin_amdgcn_perm(y1, y0, sel_lo) | __builtin_amdgcn_perm(y3, y2, sel_hi);
}

// Work layout: blockIdx.x -> GatherWork (bin, first k-block, k-block count <= 16),
// blockIdx.y -> 64-frame tile.  Wave w handles k-blocks kb0 + w, kb0 + w + 4, ...
// BIAS (uint8 and uint16 input): bytes are stored as v ^ 0x80 = v - 128 (i8 MFMA
// operands); padded list entries (index -1) read as 0 and become -128 the same way.
// T = uint16_t writes the high bytes to k-block kb and the low bytes to k-block
// kb + nkb of the bin (two digit planes).
template <typename T, bool BIAS>
__global__ __launch_bounds__(256) void gather_i8_kernel(
    const T* __restrict__ frames, int64_t hw, int nframes, int nrows,
    const int32_t* __restrict__ pix_idx, const BinDev* __restrict__ bins,
    const GatherWork* __restrict__ work, int8_t* __restrict__ slab) {
    constexpr bool WIDE = sizeof(T) == 2;
    const GatherWork w = work[blockIdx.x];
    const BinDev bd = bins[w.bin];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int qi = lane & 3;        // frame offset this lane stores after the transpose
    const int q4 = (lane >> 2) * 4; // first pixel of this lane's stored dword
    // byte selectors for v_perm: lo = (y0[qi], y1[qi], 0, 0), hi = (0, 0, y2[qi], y3[qi])
    const uint32_t sel_lo = (uint32_t)qi | ((uint32_t)(4 + qi) << 8) | 0x0c0c0000u;
    const uint32_t sel_hi = 0x00000c0cu | ((uint32_t)qi << 16) | ((uint32_t)(4 + qi) << 24);
    const int t0 = blockIdx.y * GATHER_TF;
    const int tend = min(nframes, t0 + GATHER_TF);
    const int64_t plane = (int64_t)bd.nkb * nrows * KB;  // bytes between digit planes

    for (int kb = w.kb0 + wave; kb < w.kb0 + w.nkb; kb += 4) {
        const int p = pix_idx[bd.idx_off + (int64_t)kb * KB + lane];
        const int pc = p >= 0 ? p : 0;
        const uint32_t keep = p >= 0 ? 0xFFFFu : 0u;
        int8_t* dst = slab + bd.slab_off;
        for (int t = t0; t < tend; t += 4) {
            uint32_t xl = 0, xh = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int tj = t + j < nframes ? t + j : nframes - 1;
                const uint32_t v = (uint32_t)frames[(int64_t)tj * hw + pc] & keep;
                xl |= (v & 0xFFu) << (8 * j);
                if (WIDE) xh |= (v >> 8) << (8 * j);
            }
            uint32_t lo = quad_transpose(xl, sel_lo, sel_hi);
            if (BIAS) lo ^= 0x80808080u;
            const int64_t o = slab_offset(kb, t + qi, q4, nrows);
            if (WIDE) {
                const uint32_t hi = quad_transpose(xh, sel_lo, sel_hi) ^ 0x80808080u;
                if (t + qi < nframes) {
                    *reinterpret_cast<uint32_t*>(dst + o) = hi;
                    *reinterpret_cast<uint32_t*>(dst + plane + o) = lo;
                }
            } else if (t + qi < nframes) {
                *reinterpret_cast<uint32_t*>(dst + o) = lo;
            }
        }
    }
}

void launch_gather_i8(const uint8_t* frames, int dtype, int64_t hw, int nframes, int nrows,
                      const int32_t* pix_idx, const BinDev* bins, const GatherWork* work,
                      int n_work, int8_t* slab, hipStream_t s) {
    if (n_work <= 0 || nframes <= 0) return;
    dim3 grid((unsigned)n_work, (unsigned)((nframes + GATHER_TF - 1) / GATHER_TF));
    if (dtype == DXG2_U16)
        gather_i8_kernel<uint16_t, true><<<grid, dim3(256), 0, s>>>(
            reinterpret_cast<const uint16_t*>(frames), hw, nframes, nrows, pix_idx, bins, work,
            slab);
    else if (dtype == DXG2_U8)
        gather_i8_kernel<uint8_t, true><<<grid, dim3(256), 0, s>>>(frames, hw, nframes, nrows,
                                                                  pix_idx, bins, work, slab);
    else
        gather_i8_kernel<uint8_t, false><<<grid, dim3(256), 0, s>>>(frames, hw, nframes, nrows,
                                                                   pix_idx, bins, work, slab);
}

}  // namespace dxg2
