// qmask gather: frames (N, H, W) -> per-bin i8 slabs slab[kb][t][64] (uint16 frames:
// two digit planes).
//
// The reference gathers one bin at a time on the host (correlator/dense.py:48-49,
// dense.py:317-319).  Here one launch covers every bin.  A wave owns four k-blocks (64
// consecutive entries each of a bin's raster-ordered pixel list) and a frame tile:
//   * each load instruction reads entry `lane` of one frame, so the 64 lanes read 64
//     consecutive list entries -- runs of neighbouring pixels, few cache lines;
//   * after eight frames each lane holds 8 frames x 1 pixel per k-block; a 4x4 byte
//     transpose inside every lane quad (DPP quad broadcast + v_perm) gives lane 4q+i
//     the dword of pixels 4q..4q+3 for frame t+i (and t+4+i), so one store instruction
//     writes rows t..t+3 of the k-block: four 64-byte runs of the fragment-contiguous
//     slab layout.
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
// blockIdx.y -> 64-frame tile.  Wave w handles k-blocks kb0 + 4 w .. kb0 + 4 w + 3.
// BIAS (uint8 and uint16 input): bytes are stored as v ^ 0x80 = v - 128 (i8 MFMA
// operands); padded list entries (index -1) read as 0 and become -128 the same way.
// T = uint16_t writes the high bytes to k-block kb and the low bytes to k-block
// kb + nkb of the bin (two digit planes); with nwkb (pixel lists ordered wide-first) the
// high bytes of k-blocks past nwkb[bin], all zero and never read by the clipped Gram, are
// not written.
template <typename T, bool BIAS>
__device__ __forceinline__ void gather_general(const T* __restrict__ frames, int64_t hw,
                                               int nframes, int nrows,
                                               const int32_t* __restrict__ pix_idx,
                                               const BinDev* __restrict__ bins, const GatherWork w,
                                               int t0, int tend, int8_t* __restrict__ slab,
                                               const int* __restrict__ nwkb = nullptr) {
    constexpr bool WIDE = sizeof(T) == 2;
    constexpr int FG = 8;  // frames per step (a multiple of 4 dividing 64; 16 measured slower)
    const BinDev bd = bins[w.bin];
    const int hi_end = WIDE && nwkb ? nwkb[w.bin] : bd.nkb;  // k-blocks whose high bytes go out
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int qi = lane & 3;        // frame offset this lane stores after the transpose
    const int q4 = (lane >> 2) * 4; // first pixel of this lane's stored dword
    // byte selectors for v_perm: lo = (y0[qi], y1[qi], 0, 0), hi = (0, 0, y2[qi], y3[qi])
    const uint32_t sel_lo = (uint32_t)qi | ((uint32_t)(4 + qi) << 8) | 0x0c0c0000u;
    const uint32_t sel_hi = 0x00000c0cu | ((uint32_t)qi << 16) | ((uint32_t)(4 + qi) << 24);
    const int64_t plane = (int64_t)bd.nkb * nrows * KB;  // bytes between digit planes

    // Wave w takes the 4 consecutive k-blocks kb0 + 4 w .. + 3 of each 16 and walks them
    // FG frames at a time together, so that at each frame the workgroup reads up to 16
    // consecutive k-blocks' pixels (up to 2 KiB of contiguous frame row for uint16).
    // Slots past the item's last k-block repeat that k-block: every step then issues
    // all 4 FG = 32 loads before its first wait (one basic block, so the waits count
    // down through them), and a repeat stores the same bytes to the same addresses.
    // Whole-line stores (an exchange across the wave halves so that each store writes
    // two 128-byte slab lines) measured slower here than these 64-byte runs.
    const int kend = w.kb0 + w.nkb;
    int8_t* dst = slab + bd.slab_off;
    for (int kbase = w.kb0 + 4 * wave; kbase < kend; kbase += 16) {
        const int nk = min(4, kend - kbase);
        int kbj[4];
        const T* src[4];
        uint32_t keep[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            kbj[j] = kbase + min(j, nk - 1);
            const int p = pix_idx[bd.idx_off + (int64_t)kbj[j] * KB + lane];
            src[j] = frames + (p >= 0 ? p : 0);
            keep[j] = p >= 0 ? 0xFFFFu : 0u;
        }
        for (int t = t0; t < tend; t += FG) {
            uint32_t v[4][FG];
#pragma unroll
            for (int i = 0; i < FG; ++i) {
                const int64_t ro = (int64_t)(t + i < nframes ? t + i : nframes - 1) * hw;
#pragma unroll
                for (int j = 0; j < 4; ++j) v[j][i] = (uint32_t)src[j][ro];
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                // dwords of frames t + 4 g + qi, low and high digits;
                // the transposes run with every lane active (DPP reads the quad's other
                // lanes, including those past the last frame)
                uint32_t d[FG / 4][2];
#pragma unroll
                for (int g = 0; g < FG / 4; ++g) {
                    uint32_t xl = 0, xh = 0;
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const uint32_t x = v[j][4 * g + i] & keep[j];
                        xl |= (x & 0xFFu) << (8 * i);
                        if (WIDE) xh |= (x >> 8) << (8 * i);
                    }
                    d[g][0] = quad_transpose(xl, sel_lo, sel_hi) ^ (BIAS ? 0x80808080u : 0u);
                    d[g][1] = WIDE ? quad_transpose(xh, sel_lo, sel_hi) ^ 0x80808080u : 0u;
                }
#pragma unroll
                for (int g = 0; g < FG / 4; ++g) {
                    const int tq = t + 4 * g + qi;
                    const int64_t o = slab_offset(kbj[j], tq, q4, nrows);
                    if (tq < nframes) {
                        if (WIDE) {
                            if (kbj[j] < hi_end) *reinterpret_cast<uint32_t*>(dst + o) = d[g][1];
                            *reinterpret_cast<uint32_t*>(dst + plane + o) = d[g][0];
                        } else {
                            *reinterpret_cast<uint32_t*>(dst + o) = d[g][0];
                        }
                    }
                }
            }
        }
    }
}

template <typename T, bool BIAS>
__global__ __launch_bounds__(256) void gather_i8_kernel(
    const T* __restrict__ frames, int64_t hw, int nframes, int nrows,
    const int32_t* __restrict__ pix_idx, const BinDev* __restrict__ bins,
    const GatherWork* __restrict__ work, int8_t* __restrict__ slab,
    const int* __restrict__ nwkb) {
    const int t0 = blockIdx.y * GATHER_TF;
    gather_general<T, BIAS>(frames, hw, nframes, nrows, pix_idx, bins, work[blockIdx.x], t0,
                            min(nframes, t0 + GATHER_TF), slab, nwkb);
}

// Band gather (uint8 / int8 frames).  A workgroup takes one band of detector rows and up
// to BAND_ITEM_KMAX full k-blocks whose pixels lie in those rows, for BAND_FRAMES frames.
// Each wave loads its k-blocks' pixel quads (16-bit offsets into the band) and slab
// addresses (in 4-byte units) once, into registers.  Then, BAND_STAGE frames at a time,
// the workgroup copies the band's rows of each frame -- one contiguous byte range, from
// its 16-byte aligned start -- into LDS with 16-byte global->LDS DMA, and lane
// (h, w, i) = (lane / 32, lane / 8 % 4, lane % 8) assembles list entries 16 (2 p + h) +
// 4 w .. +3 of frame t0 + i in pass p = 0, 1 and stores them as one dword.  In the slab
// a quarter (16 entries) of frames t .. t+7 (t % 8 = 0) is one 128-byte line, so each
// store writes two whole lines.  A slab address splits into a k-block part and a frame
// part (slab_offset(kb, t, kk) = slab_offset(kb, 0, kk) + (t/32) 2048 + (t%32) 16).  The
// work list gives each k-block to the band of its first row, with one row of overlap
// staged below the band, so k-blocks spanning two rows are covered; the others, and
// every bin's partial last k-block, go to gather_i8_kernel.  Two workgroups share a CU:
// one copies while the other assembles.
constexpr int BAND_FRAMES = 128;    // frames per workgroup (a multiple of BAND_STAGE)
constexpr int BAND_GROUP = 4;       // k-blocks whose LDS reads are issued together
static_assert(BAND_FRAMES % BAND_STAGE == 0 && BAND_STAGE == 8, "band stage");

template <bool BIAS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void gather_band_kernel(
    const uint8_t* __restrict__ frames, int64_t hw, int width, int nframes, int nrows, int srow,
    const int32_t* __restrict__ pix_idx, const BinDev* __restrict__ bins,
    const BandWork* __restrict__ bands, int n_bands, const int2* __restrict__ band_kb,
    const GatherWork* __restrict__ fall, int8_t* __restrict__ slab) {
    extern __shared__ __attribute__((aligned(16))) uint8_t stage[];  // BAND_STAGE x srow
    __shared__ int shift[BAND_STAGE];
    if ((int)blockIdx.x >= n_bands) {  // a k-block run the bands do not cover
        const int t0 = blockIdx.y * BAND_FRAMES;
        gather_general<uint8_t, BIAS>(frames, hw, nframes, nrows, pix_idx, bins,
                                      fall[blockIdx.x - n_bands], t0,
                                      min(nframes, t0 + BAND_FRAMES), slab);
        return;
    }
    const BandWork bw = bands[blockIdx.x];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int i = lane & 7;          // frame t0 + i
    const int w = (lane >> 3) & 3;   // dword w of a quarter
    const int h = lane >> 5;         // quarter 2 p + h in pass p
    const int64_t band0 = (int64_t)bw.y0 * width;
    const int64_t total = (int64_t)nframes * hw;
    const int64_t span = (int64_t)bw.rows * width;
    const int nk = (bw.k1 - bw.k0 - wave + 3) / 4;  // this wave's k-blocks: k0 + wave + 4j

    uint2 pq0[BAND_KMAX], pq1[BAND_KMAX];  // 4 x 16-bit offsets into the staged rows, pass 0/1
    uint32_t sbase[BAND_KMAX];             // slab address / 4
    auto pack = [&](int4 p) {
        return make_uint2((uint32_t)(p.x - band0) | (uint32_t)(p.y - band0) << 16,
                          (uint32_t)(p.z - band0) | (uint32_t)(p.w - band0) << 16);
    };
    if (nk > 0) {
        // slots past nk repeat the wave's last k-block: a group then needs no per-slot
        // branch, and the repeats store the same bytes to the same addresses
#pragma unroll
        for (int j = 0; j < BAND_KMAX; ++j) {
            if (j < ((nk + BAND_GROUP - 1) / BAND_GROUP) * BAND_GROUP) {
                const int2 bk = band_kb[bw.k0 + wave + 4 * min(j, nk - 1)];
                const BinDev bd = bins[bk.x];
                const int32_t* pl = pix_idx + bd.idx_off + (int64_t)bk.y * KB + 16 * h + 4 * w;
                pq0[j] = pack(*reinterpret_cast<const int4*>(pl));
                pq1[j] = pack(*reinterpret_cast<const int4*>(pl + 32));
                sbase[j] = (uint32_t)((bd.slab_off + slab_offset(bk.y, 0, 16 * h + 4 * w, nrows)) >> 2);
            }
        }
    }

    const int tb = blockIdx.y * BAND_FRAMES;
    const int te = min(nframes, tb + BAND_FRAMES);
    for (int t0 = tb; t0 < te; t0 += BAND_STAGE) {
        __syncthreads();  // the previous frames' stage has been read
        // the chunks of a 64-lane group land contiguously (lane l at base + 16 l), as the
        // stage is laid out; a chunk running past the end of the frame buffer (the last
        // frame's last chunk at most) is left to the byte loads below
        int64_t tail_a = -1;
        int tail_f = 0, tail_c = 0;
        for (int f = 0; f < BAND_STAGE && t0 + f < nframes; ++f) {
            const int64_t a0 = (int64_t)(t0 + f) * hw + band0;
            const int64_t A = a0 & ~(int64_t)15;
            if (threadIdx.x == 0) shift[f] = (int)(a0 - A);
            const int nchunk = (int)((a0 + span - A + 15) >> 4);
            uint8_t* dst = stage + f * srow;
            for (int g = wave; g * 64 < nchunk; g += 4) {
                const int c = g * 64 + lane;
                const int64_t a = A + 16 * (int64_t)c;
                if (c < nchunk && a + 16 <= total)
                    __builtin_amdgcn_global_load_lds((const void*)(frames + a),
                                                     (lds_void*)(dst + 16 * g * 64), 16, 0, 0);
                else if (c < nchunk) {
                    tail_a = a;
                    tail_f = f;
                    tail_c = c;
                }
            }
        }
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this thread's copies landed
        if (tail_a >= 0) {  // bytes inside the buffer only
            uint8_t* d = stage + tail_f * srow + 16 * tail_c;
            for (int e = 0; e < 16; ++e) d[e] = tail_a + e < total ? frames[tail_a + e] : 0;
        }
        __syncthreads();
        const int t = t0 + i;
        if (nk > 0 && t < nframes) {
            const uint8_t* st = stage + i * srow + shift[i];
            int8_t* out = slab + (int64_t)(t >> 5) * 2048 + (t & 31) * 16;
            auto bytes = [&](uint2 p) -> uint32_t {
                uint32_t v = (uint32_t)st[p.x & 0xFFFFu] | (uint32_t)st[p.x >> 16] << 8 |
                             (uint32_t)st[p.y & 0xFFFFu] << 16 | (uint32_t)st[p.y >> 16] << 24;
                return BIAS ? v ^ 0x80808080u : v;
            };
            // Groups of BAND_GROUP k-blocks: the empty asm marks a group's packed
            // registers as rewritten in every stage, so their unpacked and widened forms
            // are not hoisted out of the frame loop (that would triple the registers),
            // while the group's LDS reads can all be in flight together.
            static_assert(BAND_KMAX % BAND_GROUP == 0 && BAND_GROUP == 4, "k-block groups");
#pragma unroll
            for (int g = 0; g < BAND_KMAX; g += BAND_GROUP) {
                if (g < nk) {
                    asm volatile(""
                                 : "+v"(pq0[g].x), "+v"(pq0[g].y), "+v"(pq1[g].x), "+v"(pq1[g].y),
                                   "+v"(sbase[g]), "+v"(pq0[g + 1].x), "+v"(pq0[g + 1].y),
                                   "+v"(pq1[g + 1].x), "+v"(pq1[g + 1].y), "+v"(sbase[g + 1]));
                    asm volatile(""
                                 : "+v"(pq0[g + 2].x), "+v"(pq0[g + 2].y), "+v"(pq1[g + 2].x),
                                   "+v"(pq1[g + 2].y), "+v"(sbase[g + 2]), "+v"(pq0[g + 3].x),
                                   "+v"(pq0[g + 3].y), "+v"(pq1[g + 3].x), "+v"(pq1[g + 3].y),
                                   "+v"(sbase[g + 3]));
#pragma unroll
                    for (int j = g; j < g + BAND_GROUP; ++j) {
                        int8_t* o = out + ((int64_t)sbase[j] << 2);
                        *reinterpret_cast<uint32_t*>(o) = bytes(pq0[j]);
                        *reinterpret_cast<uint32_t*>(o + 1024) = bytes(pq1[j]);
                    }
                }
            }
        }
    }
}

void launch_gather_band(const uint8_t* frames, int dtype, int64_t hw, int width, int nframes,
                        int nrows, int srow, const int32_t* pix_idx, const BinDev* bins,
                        const BandWork* bands, int n_bands, const int2* band_kb,
                        const GatherWork* fall, int n_fall, int8_t* slab, hipStream_t s) {
    if (n_bands + n_fall <= 0 || nframes <= 0) return;
    const dim3 grid((unsigned)(n_bands + n_fall),
                    (unsigned)((nframes + BAND_FRAMES - 1) / BAND_FRAMES));
    const size_t lds = (size_t)BAND_STAGE * srow;
    if (dtype == DXG2_U8)
        gather_band_kernel<true><<<grid, dim3(256), lds, s>>>(frames, hw, width, nframes, nrows,
                                                              srow, pix_idx, bins, bands, n_bands,
                                                              band_kb, fall, slab);
    else
        gather_band_kernel<false><<<grid, dim3(256), lds, s>>>(frames, hw, width, nframes, nrows,
                                                               srow, pix_idx, bins, bands, n_bands,
                                                               band_kb, fall, slab);
}

void launch_gather_i8(const uint8_t* frames, int dtype, int64_t hw, int nframes, int nrows,
                      const int32_t* pix_idx, const BinDev* bins, const GatherWork* work,
                      int n_work, int8_t* slab, hipStream_t s, const int* nwkb) {
    if (n_work <= 0 || nframes <= 0) return;
    dim3 grid((unsigned)n_work, (unsigned)((nframes + GATHER_TF - 1) / GATHER_TF));
    if (dtype == DXG2_U16)
        gather_i8_kernel<uint16_t, true><<<grid, dim3(256), 0, s>>>(
            reinterpret_cast<const uint16_t*>(frames), hw, nframes, nrows, pix_idx, bins, work,
            slab, nwkb);
    else if (dtype == DXG2_U8)
        gather_i8_kernel<uint8_t, true><<<grid, dim3(256), 0, s>>>(frames, hw, nframes, nrows,
                                                                  pix_idx, bins, work, slab,
                                                                  nullptr);
    else
        gather_i8_kernel<uint8_t, false><<<grid, dim3(256), 0, s>>>(frames, hw, nframes, nrows,
                                                                   pix_idx, bins, work, slab,
                                                                   nullptr);
}

}  // namespace dxg2
