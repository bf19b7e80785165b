// qmask gather: frames (N, H, W) -> per-bin i8 slabs slab[kb][t][64] (uint16 frames:
// two digit planes).
//
// The reference gathers one bin at a time on the host (correlator/dense.py:48-49,
// dense.py:317-319).  Here one launch covers every bin.  A wave owns four k-blocks (64
// consecutive entries each of a bin's raster-ordered pixel list; the k-blocks of all
// bins are taken in detector order, so a workgroup's sixteen read neighbouring stretches
// of the same rows) and a frame tile:
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

// One wave's step over four slots: slot j is k-block kbj[j] of the bin whose slab starts
// at dst[j] (low plane `plane[j]` bytes further; high bytes written when hi[j]), its 64
// pixels read through src[j] (keep[j] = 0 for padded list entries).  Frames [t0, tend)
// FG at a time: each step issues all 4 FG loads before its first wait (one basic block,
// so the waits count down through them).  A slot that repeats another stores the same
// bytes to the same addresses.  Whole-line stores (an exchange across the wave halves so
// that each store writes two 128-byte slab lines) measured slower here than these
// 64-byte runs.
// FLAG (uint16): the high bytes are not stored; instead the lane ORs its pixel's counts
// over the frames and sets the pixel's byte of flagw (pixel 4j+k is byte k of word j)
// when any count exceeds 255 (pix[j] = the slot's pixel index, < 0 for padding).
template <typename T, bool BIAS, bool FLAG = false>
__device__ __forceinline__ void gather_wave4(const T* const (&src)[4], const uint32_t (&keep)[4],
                                             int8_t* const (&dst)[4], const int (&kbj)[4],
                                             const int64_t (&plane)[4], const bool (&hi)[4],
                                             int64_t hw, int nframes, int nrows, int t0,
                                             int tend, const int (&pix)[4] = {-1, -1, -1, -1},
                                             uint32_t* __restrict__ flagw = nullptr) {
    constexpr bool WIDE = sizeof(T) == 2;
    uint32_t seen[4] = {0u, 0u, 0u, 0u};
    constexpr int FG = 8;  // frames per step (a multiple of 4 dividing 64; 16 measured slower)
    const int lane = threadIdx.x & 63;
    const int qi = lane & 3;        // frame offset this lane stores after the transpose
    const int q4 = (lane >> 2) * 4; // first pixel of this lane's stored dword
    // byte selectors for v_perm: lo = (y0[qi], y1[qi], 0, 0), hi = (0, 0, y2[qi], y3[qi])
    const uint32_t sel_lo = (uint32_t)qi | ((uint32_t)(4 + qi) << 8) | 0x0c0c0000u;
    const uint32_t sel_hi = 0x00000c0cu | ((uint32_t)qi << 16) | ((uint32_t)(4 + qi) << 24);
    for (int t = t0; t < tend; t += FG) {
        uint32_t v[4][FG];
#pragma unroll
        for (int i = 0; i < FG; ++i) {
            const int64_t ro = (int64_t)(t + i < nframes ? t + i : nframes - 1) * hw;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                v[j][i] = (uint32_t)src[j][ro];
                if (FLAG) seen[j] |= v[j][i];
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            // dwords of frames t + 4 g + qi, low and high digits; the transposes run with
            // every lane active (DPP reads the quad's other lanes, including those past
            // the last frame)
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
                        if (!FLAG && hi[j]) *reinterpret_cast<uint32_t*>(dst[j] + o) = d[g][1];
                        *reinterpret_cast<uint32_t*>(dst[j] + plane[j] + o) = d[g][0];
                    } else {
                        *reinterpret_cast<uint32_t*>(dst[j] + o) = d[g][0];
                    }
                }
            }
        }
    }
    if (FLAG) {
        // slots repeating a k-block flag the same pixels again: harmless
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (pix[j] >= 0 && (seen[j] & 0xFF00u))
                atomicOr(flagw + (pix[j] >> 2), 1u << (8 * (pix[j] & 3)));
    }
}

// Work layout: blockIdx.x -> GatherWork (bin, first k-block, k-block count <= 16),
// blockIdx.y -> frame tile.  Wave w handles k-blocks kb0 + 4 w .. kb0 + 4 w + 3.
// BIAS (uint8 and uint16 input): bytes are stored as v ^ 0x80 = v - 128 (i8 MFMA
// operands); padded list entries (index -1) read as 0 and become -128 the same way.
// T = uint16_t writes the high bytes to k-block kb and the low bytes to k-block
// kb + nkb of the bin (two digit planes).
template <typename T, bool BIAS>
__device__ __forceinline__ void gather_general(const T* __restrict__ frames, int64_t hw,
                                               int nframes, int nrows,
                                               const int32_t* __restrict__ pix_idx,
                                               const BinDev* __restrict__ bins, const GatherWork w,
                                               int t0, int tend, int8_t* __restrict__ slab) {
    const BinDev bd = bins[w.bin];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int64_t plane = (int64_t)bd.nkb * nrows * KB;  // bytes between digit planes
    // Wave w takes the 4 consecutive k-blocks kb0 + 4 w .. + 3 of each 16, so that at each
    // frame the workgroup reads up to 16 consecutive k-blocks' pixels.  Slots past the
    // item's last k-block repeat that k-block.
    const int kend = w.kb0 + w.nkb;
    for (int kbase = w.kb0 + 4 * wave; kbase < kend; kbase += 16) {
        const int nk = min(4, kend - kbase);
        int kbj[4];
        const T* src[4];
        uint32_t keep[4];
        int8_t* dst[4];
        int64_t pl[4];
        bool hi[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            kbj[j] = kbase + min(j, nk - 1);
            const int p = pix_idx[bd.idx_off + (int64_t)kbj[j] * KB + lane];
            src[j] = frames + (p >= 0 ? p : 0);
            keep[j] = p >= 0 ? 0xFFFFu : 0u;
            dst[j] = slab + bd.slab_off;
            pl[j] = plane;
            hi[j] = true;
        }
        gather_wave4<T, BIAS>(src, keep, dst, kbj, pl, hi, hw, nframes, nrows, t0, tend);
    }
}

// The general gather over k-blocks of any bins: blockIdx.x -> 16 slots kbl[16 x ..],
// k-blocks (bin, kb) in detector order (a wave's four slots; a slot past the item's
// last k-block repeats the wave's last one, a wave with nothing to do has kb < 0 in
// its first slot), blockIdx.y -> frame tile.  Mixing bins lets the k-blocks that read
// the same stretch of detector rows share one workgroup and its cache lines.
// FLAG (uint16 with wide pixels handled apart): only the low plane is written, and the
// pixels with a count above 255 are flagged in flagw (zeroed beforehand).
template <typename T, bool BIAS, bool FLAG>
__global__ __launch_bounds__(256) void gather_slots_kernel(
    const T* __restrict__ frames, int64_t hw, int nframes, int nrows,
    const int32_t* __restrict__ pix_idx, const BinDev* __restrict__ bins,
    const KbWork* __restrict__ kbl, int tf, int8_t* __restrict__ slab,
    uint32_t* __restrict__ flagw) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const KbWork* sl = kbl + (int64_t)blockIdx.x * 16 + 4 * wave;
    if (sl[0].kb < 0) return;
    const int t0 = blockIdx.y * tf;
    const int tend = min(nframes, t0 + tf);
    int kbj[4], pix[4];
    const T* src[4];
    uint32_t keep[4];
    int8_t* dst[4];
    int64_t pl[4];
    bool hi[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const KbWork e = sl[j];
        const BinDev bd = bins[e.bin];
        kbj[j] = e.kb;
        const int p = pix_idx[bd.idx_off + (int64_t)e.kb * KB + lane];
        pix[j] = p;
        src[j] = frames + (p >= 0 ? p : 0);
        keep[j] = p >= 0 ? 0xFFFFu : 0u;
        dst[j] = slab + bd.slab_off;
        pl[j] = (int64_t)bd.nkb * nrows * KB;
        hi[j] = true;
    }
    gather_wave4<T, BIAS, FLAG>(src, keep, dst, kbj, pl, hi, hw, nframes, nrows, t0, tend, pix,
                                flagw);
}

// uint16 wide pixels (after gather_slots_kernel<FLAG> and launch_wide_partition).  For a
// bin with clip = nwkb[b] wide k-blocks:
//  * 2 clip <= nkb: k-blocks kb < clip of the bin's wide list (entries past its
//    n_wide = count[b] read as padding) go to the high plane's k-block kb (high bytes)
//    and k-block clip + kb (low bytes);
//  * otherwise (clip = nkb): every k-block of the list in its own order, high bytes to
//    k-block kb (and the low bytes again to the low plane, unchanged).
// Work units (4 k-blocks of a bin from kb0 = 4 level, one frame tile; each wave a quarter
// of the tile's frames) up to the level of the largest clip (nwkb[n_bins]) are taken by
// a fixed grid in turn: the wide k-blocks are a small share of the slab, so a grid over
// every k-block would be mostly empty workgroups, and scattered hot pixels make these
// loads latency-bound, so every wave of a unit has frames to load.
__global__ __launch_bounds__(256) void wide_gather_kernel(
    const uint16_t* __restrict__ frames, int64_t hw, int nframes, int nrows,
    const int32_t* __restrict__ pix, const int32_t* __restrict__ pix_w,
    const BinDev* __restrict__ bins, int n_bins, const int* __restrict__ nwkb,
    int8_t* __restrict__ slab) {
    constexpr int WF = GATHER_TF / 4;  // frames per wave of a unit
    const int n_tiles = (nframes + GATHER_TF - 1) / GATHER_TF;
    const int levels = (nwkb[n_bins] + 3) / 4;
    const int64_t units = (int64_t)levels * n_bins * n_tiles;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    for (int64_t u = blockIdx.x; u < units; u += gridDim.x) {
        const int tile = (int)(u % n_tiles);
        const int64_t rest = u / n_tiles;
        const int b = (int)(rest % n_bins);
        const int kbase = 4 * (int)(rest / n_bins);
        const BinDev bd = bins[b];
        const int clip = nwkb[b];
        if (kbase >= clip) continue;  // the same for the whole workgroup; no barrier here
        const bool compact = 2 * clip <= bd.nkb;
        const int32_t* list = compact ? pix_w : pix;
        const int n_list = compact ? nwkb[n_bins + 1 + b] : bd.npix;
        const int t0 = tile * GATHER_TF + wave * WF;
        const int tend = min(nframes, t0 + WF);
        if (t0 >= tend) continue;
        const int nk = min(4, clip - kbase);
        int kbj[4];
        const uint16_t* src[4];
        uint32_t keep[4];
        int8_t* dst[4];
        int64_t pl[4];
        bool hi[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            kbj[j] = kbase + min(j, nk - 1);
            const int e = kbj[j] * KB + lane;
            const int p = e < n_list ? list[bd.idx_off + e] : -1;
            src[j] = frames + (p >= 0 ? p : 0);
            keep[j] = p >= 0 ? 0xFFFFu : 0u;
            dst[j] = slab + bd.slab_off;
            pl[j] = (int64_t)(compact ? clip : bd.nkb) * nrows * KB;
            hi[j] = true;
        }
        gather_wave4<uint16_t, true>(src, keep, dst, kbj, pl, hi, hw, nframes, nrows, t0, tend);
    }
}

void launch_wide_gather(const uint16_t* frames, int64_t hw, int nframes, int nrows,
                        const int32_t* pix, const int32_t* pix_w, const BinDev* bins, int n_bins,
                        const int* nwkb, int8_t* slab, hipStream_t s) {
    if (n_bins <= 0 || nframes <= 0) return;
    wide_gather_kernel<<<dim3(2048), dim3(256), 0, s>>>(frames, hw, nframes, nrows, pix, pix_w,
                                                        bins, n_bins, nwkb, slab);
}

void launch_gather_slots(const uint8_t* frames, int dtype, int64_t hw, int nframes, int nrows,
                         const int32_t* pix_idx, const BinDev* bins, const KbWork* kbl,
                         int n_items, int tf, int8_t* slab, hipStream_t s, uint32_t* flagw) {
    if (n_items <= 0 || nframes <= 0 || tf <= 0) return;
    dim3 grid((unsigned)n_items, (unsigned)((nframes + tf - 1) / tf));
    const uint16_t* f16 = reinterpret_cast<const uint16_t*>(frames);
    if (dtype == DXG2_U16 && flagw)
        gather_slots_kernel<uint16_t, true, true><<<grid, dim3(256), 0, s>>>(
            f16, hw, nframes, nrows, pix_idx, bins, kbl, tf, slab, flagw);
    else if (dtype == DXG2_U16)
        gather_slots_kernel<uint16_t, true, false><<<grid, dim3(256), 0, s>>>(
            f16, hw, nframes, nrows, pix_idx, bins, kbl, tf, slab, nullptr);
    else if (dtype == DXG2_U8)
        gather_slots_kernel<uint8_t, true, false><<<grid, dim3(256), 0, s>>>(
            frames, hw, nframes, nrows, pix_idx, bins, kbl, tf, slab, nullptr);
    else
        gather_slots_kernel<uint8_t, false, false><<<grid, dim3(256), 0, s>>>(
            frames, hw, nframes, nrows, pix_idx, bins, kbl, tf, slab, nullptr);
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
// every bin's partial last k-block, go to the general gather.  Two workgroups share a CU:
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

}  // namespace dxg2
