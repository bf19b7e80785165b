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
//
// For uint8/int8 frames the band gather below does most k-blocks: it reads the frames
// with 16-byte loads through LDS instead of one byte per lane and load.
#include "dxg2_internal.h"

namespace dxg2 {

typedef __attribute__((address_space(3))) void lds_void;

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

// Band gather (uint8 / int8 frames).  A workgroup takes one band of detector rows and
// up to BAND_KMAX k-blocks whose pixels lie in those rows, for BAND_FRAMES frames.  Each
// wave loads its k-blocks' pixel quads (relative to the band) and slab addresses once,
// into registers; then, 4 frames at a time, the workgroup stages the band's rows of the
// 4 frames in LDS -- one contiguous byte range per frame, read with 16-byte loads from
// its 16-byte aligned start -- and lane 4 q + i assembles list entries 4q .. 4q+3 of
// frame t + i from LDS and stores them as one dword (4 x 64-byte runs per store, as in
// gather_i8_kernel).  A slab address splits into a k-block part and a frame part
// (slab_offset(kb, t, kk) = slab_offset(kb, 0, kk) + (t/32) 2048 + (t%32) 16).  The
// work list gives each k-block to the band of its first row, with 2 rows of overlap
// staged below the band, so k-blocks spanning up to 3 rows are covered; the others go
// to gather_i8_kernel.
constexpr int BAND_KMAX = 32;       // k-blocks per wave
constexpr int BAND_FRAMES = 64;     // frames per workgroup

template <bool BIAS>
__global__ __launch_bounds__(256) void gather_band_kernel(
    const uint8_t* __restrict__ frames, int64_t hw, int width, int nframes, int nrows, int srow,
    const int32_t* __restrict__ pix_idx, const BinDev* __restrict__ bins,
    const BandWork* __restrict__ bands, const int2* __restrict__ band_kb,
    int8_t* __restrict__ slab) {
    extern __shared__ __attribute__((aligned(16))) uint8_t stage[];  // 4 x srow bytes
    __shared__ int shift[4];
    const BandWork bw = bands[blockIdx.x];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int qq = lane >> 2;  // quad of list entries 4qq .. 4qq+3
    const int i = lane & 3;    // frame t + i
    const int64_t band0 = (int64_t)bw.y0 * width;
    const int nk = (bw.k1 - bw.k0 - wave + 3) / 4;  // this wave's k-blocks: k0 + wave + 4j

    int4 pq[BAND_KMAX];
    int64_t sbase[BAND_KMAX];
#pragma unroll
    for (int j = 0; j < BAND_KMAX; ++j) {
        if (j < nk) {
            const int2 bk = band_kb[bw.k0 + wave + 4 * j];
            const BinDev bd = bins[bk.x];
            const int4 p = *reinterpret_cast<const int4*>(pix_idx + bd.idx_off +
                                                          (int64_t)bk.y * KB + 4 * qq);
            pq[j] = make_int4(p.x >= 0 ? (int)(p.x - band0) : -1, p.y >= 0 ? (int)(p.y - band0) : -1,
                              p.z >= 0 ? (int)(p.z - band0) : -1, p.w >= 0 ? (int)(p.w - band0) : -1);
            sbase[j] = bd.slab_off + slab_offset(bk.y, 0, 4 * qq, nrows);
        }
    }

    const int64_t total = (int64_t)nframes * hw;
    const int64_t span = (int64_t)bw.rows * width;
    const int tb = blockIdx.y * BAND_FRAMES;
    const int te = min(nframes, tb + BAND_FRAMES);
    for (int t0 = tb; t0 < te; t0 += 4) {
        __syncthreads();  // the previous frames' stage has been read
        // DMA the 16-byte chunks of each frame's rows straight into LDS: the chunks of a
        // 64-lane group land contiguously (lane l at base + 16 l), as the stage is laid
        // out.  A chunk that would run past the end of the frame buffer is left to the
        // byte loads below.
        int64_t tail_a = -1;
        int tail_f = 0, tail_c = 0;
        for (int f = 0; f < 4 && t0 + f < nframes; ++f) {
            const int64_t a0 = (int64_t)(t0 + f) * hw + band0;
            const int64_t A = a0 & ~(int64_t)15;
            if (threadIdx.x == 0) shift[f] = (int)(a0 - A);
            const int nchunk = (int)((a0 + span - A + 15) >> 4);
            for (int g = wave; g * 64 < nchunk; g += 4) {
                const int c = g * 64 + lane;
                const int64_t a = A + 16 * (int64_t)c;
                if (c < nchunk && a + 16 <= total)
                    __builtin_amdgcn_global_load_lds(
                        (const void*)(frames + a),
                        (lds_void*)(stage + f * srow + 16 * g * 64), 16, 0, 0);
                else if (c < nchunk) {
                    tail_a = a;
                    tail_f = f;
                    tail_c = c;
                }
            }
        }
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): every DMA of this thread landed
        if (tail_a >= 0) {  // bytes inside the buffer only
            uint8_t* d = stage + tail_f * srow + 16 * tail_c;
            for (int e = 0; e < 16; ++e) d[e] = tail_a + e < total ? frames[tail_a + e] : 0;
        }
        __syncthreads();
        const int t = t0 + i;
        if (t < nframes) {
            const uint8_t* st = stage + i * srow + shift[i];
            const int64_t fo = (int64_t)(t >> 5) * 2048 + (t & 31) * 16;
#pragma unroll
            for (int j = 0; j < BAND_KMAX; ++j) {
                if (j < nk) {
                    const int4 p = pq[j];
                    uint32_t v = (p.x >= 0 ? (uint32_t)st[p.x] : 0u) |
                                 ((p.y >= 0 ? (uint32_t)st[p.y] : 0u) << 8) |
                                 ((p.z >= 0 ? (uint32_t)st[p.z] : 0u) << 16) |
                                 ((p.w >= 0 ? (uint32_t)st[p.w] : 0u) << 24);
                    if (BIAS) v ^= 0x80808080u;
                    *reinterpret_cast<uint32_t*>(slab + sbase[j] + fo) = v;
                }
            }
        }
    }
}

void launch_gather_band(const uint8_t* frames, int dtype, int64_t hw, int width, int nframes,
                        int nrows, int srow, const int32_t* pix_idx, const BinDev* bins,
                        const BandWork* bands, int n_bands, const int2* band_kb, int8_t* slab,
                        hipStream_t s) {
    if (n_bands <= 0 || nframes <= 0) return;
    const dim3 grid((unsigned)n_bands, (unsigned)((nframes + BAND_FRAMES - 1) / BAND_FRAMES));
    const size_t lds = (size_t)4 * srow;
    if (dtype == DXG2_U8)
        gather_band_kernel<true><<<grid, dim3(256), lds, s>>>(frames, hw, width, nframes, nrows,
                                                              srow, pix_idx, bins, bands, band_kb,
                                                              slab);
    else
        gather_band_kernel<false><<<grid, dim3(256), lds, s>>>(frames, hw, width, nframes, nrows,
                                                               srow, pix_idx, bins, bands,
                                                               band_kb, slab);
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
