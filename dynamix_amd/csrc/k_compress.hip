// Dense frames -> per-pixel event lists, on the GPU: the FramesCompressor layout of the
// event correlator's input (correlator/event.py:207-321).
//
// For every pixel p in raster order, its samples with value > 0 (signed for int8, as
// `frames > 0` in event.py:276,293) are listed in frame order: data[offsets[p] ..
// offsets[p+1]) and the matching frame indices times[...].
//
// Whole stack (compress_all_stack, event.py:267-284): a count pass (one thread per
// word of pixels and chunk of frames; frames read as 16-byte or 4-byte words when the
// frame size and alignment allow, so a load instruction moves 1 KiB or 256 bytes), a
// scan over the chunks and one over the pixels into offsets, and a fill pass that
// re-reads the frames and writes each pixel's events from its offset on.
//
// Streaming (process_frame / get_compacted_events, event.py:287-321): each pushed frame
// appends its events to per-pixel slots of max_nnz entries; compaction scans the
// per-pixel counters and copies the slots into the same 1-D layout.
#include "dxg2_internal.h"

namespace dxg2 {

template <typename T>
__device__ __forceinline__ bool positive(T v) {
    return v > T(0);
}

// W: the word one thread loads per frame (uint4, uint32_t or T itself); it holds
// sizeof(W) / sizeof(T) consecutive pixels.
template <typename T, typename W>
struct Lanes {
    static constexpr int N = sizeof(W) / sizeof(T);
    __device__ __forceinline__ static T get(const W& w, int k) {
        T out;
        __builtin_memcpy(&out, reinterpret_cast<const char*>(&w) + k * sizeof(T), sizeof(T));
        return out;
    }
};

// The frame axis is cut into C chunks of F frames (C <= 64) so that every thread reads a
// short run of frames: per-chunk counts cnt[c][p] -> per-pixel bases within the pixel's
// list (cnt becomes base in place, total into offsets[p]) -> exclusive scan over pixels
// -> fill, each (pixel word, chunk) thread writing from offsets[p] + base[c][p].
int compress_chunks(int nframes) {
    int c = (nframes + 63) / 64;
    return c < 1 ? 1 : (c > 64 ? 64 : c);
}

template <typename T, typename W>
__global__ __launch_bounds__(256) void compress_count_kernel(const T* __restrict__ frames,
                                                             int nframes, int64_t hw, int F,
                                                             uint32_t* __restrict__ cnt) {
    constexpr int N = Lanes<T, W>::N;
    const int64_t nv = hw / N;  // hw is a multiple of N
    const int c = blockIdx.y;
    const int t0 = c * F, t1 = min(nframes, t0 + F);
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x) {
        uint32_t n[N];
#pragma unroll
        for (int k = 0; k < N; ++k) n[k] = 0;
        const W* src = reinterpret_cast<const W*>(frames) + v;
        for (int t = t0; t < t1; ++t) {
            const W w = src[(int64_t)t * nv];
#pragma unroll
            for (int k = 0; k < N; ++k) n[k] += positive(Lanes<T, W>::get(w, k)) ? 1u : 0u;
        }
#pragma unroll
        for (int k = 0; k < N; ++k) cnt[(int64_t)c * hw + v * N + k] = n[k];
    }
}

// cnt[c][p] -> exclusive base over c (in place); offsets[p] = the pixel's total.
__global__ void compress_chunk_scan_kernel(uint32_t* __restrict__ cnt, int64_t hw, int C,
                                           uint32_t* __restrict__ total) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < hw;
         p += (int64_t)gridDim.x * blockDim.x) {
        uint32_t run = 0;
        for (int c = 0; c < C; ++c) {
            const uint32_t n = cnt[(int64_t)c * hw + p];
            cnt[(int64_t)c * hw + p] = run;
            run += n;
        }
        total[p] = run;
    }
}

template <typename T, typename W>
__global__ __launch_bounds__(256) void compress_fill_kernel(const T* __restrict__ frames,
                                                            int nframes, int64_t hw, int F,
                                                            const uint32_t* __restrict__ base,
                                                            const uint32_t* __restrict__ offsets,
                                                            int32_t* __restrict__ times,
                                                            T* __restrict__ data) {
    constexpr int N = Lanes<T, W>::N;
    const int64_t nv = hw / N;
    const int c = blockIdx.y;
    const int t0 = c * F, t1 = min(nframes, t0 + F);
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x) {
        uint32_t cur[N];
#pragma unroll
        for (int k = 0; k < N; ++k)
            cur[k] = offsets[v * N + k] + base[(int64_t)c * hw + v * N + k];
        const W* src = reinterpret_cast<const W*>(frames) + v;
        for (int t = t0; t < t1; ++t) {
            const W w = src[(int64_t)t * nv];
#pragma unroll
            for (int k = 0; k < N; ++k) {
                const T x = Lanes<T, W>::get(w, k);
                if (positive(x)) {
                    times[cur[k]] = t;
                    data[cur[k]] = x;
                    ++cur[k];
                }
            }
        }
    }
}

static unsigned compress_grid(int64_t n) {
    int64_t g = (n + 255) / 256;
    if (g < 1) g = 1;
    if (g > 65536) g = 65536;
    return (unsigned)g;
}

// Launch with the widest word that tiles every frame exactly and is aligned at the
// frames pointer (frame t starts at t * hw elements).
template <typename T>
static void compress_dispatch(bool fill, const T* frames, int nframes, int64_t hw,
                              uint32_t* cnt, const uint32_t* offsets, int32_t* times, T* data,
                              hipStream_t s) {
    const int C = compress_chunks(nframes);
    const int F = (nframes + C - 1) / C;
    const int64_t bytes = hw * (int64_t)sizeof(T);
    const uintptr_t addr = reinterpret_cast<uintptr_t>(frames);
#define DXG2_COMPRESS_LAUNCH(W)                                                                  \
    do {                                                                                         \
        const dim3 grid(compress_grid(hw / (int64_t)(sizeof(W) / sizeof(T))), (unsigned)C);      \
        if (fill)                                                                                \
            compress_fill_kernel<T, W><<<grid, 256, 0, s>>>(frames, nframes, hw, F, cnt, offsets,\
                                                            times, data);                        \
        else                                                                                     \
            compress_count_kernel<T, W><<<grid, 256, 0, s>>>(frames, nframes, hw, F, cnt);       \
    } while (0)
    if (bytes % 16 == 0 && addr % 16 == 0)
        DXG2_COMPRESS_LAUNCH(uint4);
    else if (bytes % 4 == 0 && addr % 4 == 0)
        DXG2_COMPRESS_LAUNCH(uint32_t);
    else
        DXG2_COMPRESS_LAUNCH(T);
#undef DXG2_COMPRESS_LAUNCH
}

size_t compress_scratch_bytes(int nframes, int64_t hw) {
    return ((size_t)compress_chunks(nframes) * hw + (hw + 1023) / 1024 + 2) * sizeof(uint32_t);
}

void launch_compress_count(int dtype, const void* frames, int nframes, int64_t hw,
                           uint32_t* scratch, uint32_t* offsets, hipStream_t s) {
    if (hw <= 0) return;
    const int C = compress_chunks(nframes);
    uint32_t* cnt = scratch;
    uint32_t* tot = scratch + (size_t)C * hw;
    if (dtype == DXG2_U16)
        compress_dispatch<uint16_t>(false, (const uint16_t*)frames, nframes, hw, cnt, nullptr,
                                    nullptr, nullptr, s);
    else if (dtype == DXG2_I8)
        compress_dispatch<int8_t>(false, (const int8_t*)frames, nframes, hw, cnt, nullptr, nullptr,
                                  nullptr, s);
    else
        compress_dispatch<uint8_t>(false, (const uint8_t*)frames, nframes, hw, cnt, nullptr,
                                   nullptr, nullptr, s);
    compress_chunk_scan_kernel<<<compress_grid(hw), 256, 0, s>>>(cnt, hw, C, offsets);
    launch_excl_scan(offsets, hw, tot, offsets, s);  // offsets[hw] = total
}

void launch_compress_fill(int dtype, const void* frames, int nframes, int64_t hw,
                          const uint32_t* scratch, const uint32_t* offsets, int32_t* times,
                          void* data, hipStream_t s) {
    if (hw <= 0) return;
    uint32_t* base = const_cast<uint32_t*>(scratch);
    if (dtype == DXG2_U16)
        compress_dispatch<uint16_t>(true, (const uint16_t*)frames, nframes, hw, base, offsets,
                                    times, (uint16_t*)data, s);
    else if (dtype == DXG2_I8)
        compress_dispatch<int8_t>(true, (const int8_t*)frames, nframes, hw, base, offsets, times,
                                  (int8_t*)data, s);
    else
        compress_dispatch<uint8_t>(true, (const uint8_t*)frames, nframes, hw, base, offsets, times,
                                   (uint8_t*)data, s);
}

// Streaming: append frame t's events to the per-pixel slots (pixel p owns entries
// [p * max_nnz, (p + 1) * max_nnz)).  A pixel whose slots are full sets *err and keeps
// its count at max_nnz + 1 (so compaction can refuse).
template <typename T>
__global__ void compress_push_kernel(const T* __restrict__ frame, int64_t hw, int t, int max_nnz,
                                     uint32_t* __restrict__ counter, T* __restrict__ events,
                                     int32_t* __restrict__ times, int* __restrict__ err) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < hw;
         p += (int64_t)gridDim.x * blockDim.x) {
        const T x = frame[p];
        if (!positive(x)) continue;
        const uint32_t c = counter[p];
        if (c < (uint32_t)max_nnz) {
            events[p * max_nnz + c] = x;
            times[p * max_nnz + c] = t;
            counter[p] = c + 1;
        } else {
            counter[p] = (uint32_t)max_nnz + 1;
            *err = 1;
        }
    }
}

template <typename T>
__global__ void compress_compact_kernel(const uint32_t* __restrict__ counter,
                                        const T* __restrict__ events,
                                        const int32_t* __restrict__ times, int64_t hw,
                                        int max_nnz, const uint32_t* __restrict__ offsets,
                                        T* __restrict__ out_data, int32_t* __restrict__ out_times) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < hw;
         p += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t n = min(counter[p], (uint32_t)max_nnz);
        const uint32_t o = offsets[p];
        for (uint32_t k = 0; k < n; ++k) {
            out_data[o + k] = events[p * max_nnz + k];
            out_times[o + k] = times[p * max_nnz + k];
        }
    }
}

void launch_compress_push(int dtype, const void* frame, int64_t hw, int t, int max_nnz,
                          uint32_t* counter, void* events, int32_t* times, int* err,
                          hipStream_t s) {
    if (hw <= 0) return;
    const unsigned g = compress_grid(hw);
    if (dtype == DXG2_U16)
        compress_push_kernel<uint16_t><<<g, 256, 0, s>>>((const uint16_t*)frame, hw, t, max_nnz,
                                                         counter, (uint16_t*)events, times, err);
    else if (dtype == DXG2_I8)
        compress_push_kernel<int8_t><<<g, 256, 0, s>>>((const int8_t*)frame, hw, t, max_nnz,
                                                       counter, (int8_t*)events, times, err);
    else
        compress_push_kernel<uint8_t><<<g, 256, 0, s>>>((const uint8_t*)frame, hw, t, max_nnz,
                                                        counter, (uint8_t*)events, times, err);
}

void launch_compress_compact(int dtype, const uint32_t* counter, const void* events,
                             const int32_t* times, int64_t hw, int max_nnz, uint32_t* scratch,
                             uint32_t* offsets, void* out_data, int32_t* out_times,
                             hipStream_t s) {
    if (hw <= 0) return;
    launch_copy_u32(counter, offsets, hw, s);
    launch_clamp_u32(offsets, hw, (uint32_t)max_nnz, s);
    launch_excl_scan(offsets, hw, scratch, offsets, s);
    const unsigned g = compress_grid(hw);
    if (dtype == DXG2_U16)
        compress_compact_kernel<uint16_t><<<g, 256, 0, s>>>(counter, (const uint16_t*)events, times,
                                                            hw, max_nnz, offsets,
                                                            (uint16_t*)out_data, out_times);
    else if (dtype == DXG2_I8)
        compress_compact_kernel<int8_t><<<g, 256, 0, s>>>(counter, (const int8_t*)events, times, hw,
                                                          max_nnz, offsets, (int8_t*)out_data,
                                                          out_times);
    else
        compress_compact_kernel<uint8_t><<<g, 256, 0, s>>>(counter, (const uint8_t*)events, times,
                                                           hw, max_nnz, offsets,
                                                           (uint8_t*)out_data, out_times);
}

__global__ void clamp_u32_kernel(uint32_t* __restrict__ p, int64_t n, uint32_t hi) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = min(p[i], hi);
}

void launch_clamp_u32(uint32_t* p, int64_t n, uint32_t hi, hipStream_t s) {
    if (n > 0) clamp_u32_kernel<<<compress_grid(n), 256, 0, s>>>(p, n, hi);
}

}  // namespace dxg2
