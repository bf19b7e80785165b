// Synthetic XPCS speckle frames on the device (test and benchmark inputs).
//
// Per pixel p: complex Gaussian field E_t = rho E_{t-1} + sqrt(1 - rho^2) w_t with
// w ~ CN(0, 1), rho = exp(-1/tau_c), E_0 ~ CN(0, 1), so <|E|^2> = 1; counts
// I(t, p) = Poisson(mean_map[p] |E_t|^2), clipped to the output type.  One thread per
// pixel walks the frames in order, so writes of a frame are coalesced across threads.
//
// Random numbers: Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers:
// as easy as 1, 2, 3", SC 2011), counter = (pixel, frame + 1, 0, 0), key = seed; each
// round's two 32x32 -> 64-bit products are single v_mad_u64_u32 instructions.
// Normals by the Box-Muller transform; Poisson by sequential inversion for lambda < 32
// and by a rounded normal approximation above (only the hot-pixel population).  The
// transcendentals are the hardware's (v_log_f32, v_exp_f32, v_sqrt_f32 and
// v_sin_f32 / v_cos_f32, whose argument is in turns): a test-data generator needs no
// more than their few-ulp accuracy, and they cut its instruction count.
#include "dxg2_internal.h"

namespace dxg2 {

struct U4 {
    uint32_t x, y, z, w;
};

__device__ __forceinline__ U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)M0 * c.x, p1 = (uint64_t)M1 * c.z;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += W0;
        k1 += W1;
    }
    return c;
}

__device__ __forceinline__ float u01(uint32_t x) {  // (0, 1)
    return ((float)(x >> 8) + 0.5f) * (1.0f / 16777216.0f);
}

// sqrt(-ln u) for u in (0, 1): sqrt(-2 ln u) * sqrt(1/2), the Box-Muller radius of a
// complex normal of unit variance
__device__ __forceinline__ float bm_radius(float u) {
    return __builtin_amdgcn_sqrtf(-0.69314718f * __builtin_amdgcn_logf(u));
}

// Per-pixel speckle state: field E and the count draw of frame t.
struct SpeckleWalk {
    float er, ei, mu, rho, inov;
    uint32_t k0, k1;
    int64_t p;

    __device__ __forceinline__ void init(int64_t pix, float mean, float rho_, uint32_t k0_,
                                         uint32_t k1_) {
        p = pix;
        mu = mean;
        rho = rho_;
        inov = sqrtf(fmaxf(0.f, 1.f - rho_ * rho_));
        k0 = k0_;
        k1 = k1_;
        const U4 r = philox4x32_10(U4{(uint32_t)p, 0u, (uint32_t)(p >> 32), 0u}, k0, k1);
        const float rad = bm_radius(u01(r.x));
        const float turn = u01(r.y);  // angle 2 pi turn
        er = rad * __builtin_amdgcn_cosf(turn);
        ei = rad * __builtin_amdgcn_sinf(turn);
    }

    // advance to frame t (t = 0 keeps E_0) and draw its count
    __device__ __forceinline__ uint32_t draw(int t, uint32_t maxval) {
        const U4 r = philox4x32_10(U4{(uint32_t)p, (uint32_t)(t + 1), (uint32_t)(p >> 32), 0u},
                                   k0, k1);
        if (t > 0) {
            const float rad = bm_radius(u01(r.x));
            const float turn = u01(r.y);
            er = rho * er + inov * rad * __builtin_amdgcn_cosf(turn);
            ei = rho * ei + inov * rad * __builtin_amdgcn_sinf(turn);
        }
        const float lam = mu * (er * er + ei * ei);
        uint32_t k = 0;
        if (lam < 32.f) {
            const float u = u01(r.z);
            float pk = __builtin_amdgcn_exp2f(-1.44269504f * lam);  // e^-lam
            float cdf = pk;
            while (u > cdf && k < 200u) {
                ++k;
                pk *= lam / (float)k;
                cdf += pk;
            }
        } else {
            const float z = 1.41421356f * bm_radius(u01(r.z)) * __builtin_amdgcn_cosf(u01(r.w));
            const float v = floorf(lam + sqrtf(lam) * z + 0.5f);
            k = v > 0.f ? (uint32_t)v : 0u;
        }
        return k > maxval ? maxval : k;
    }
};

template <typename T>
__global__ void synth_speckle_kernel(T* __restrict__ out, int nframes, int64_t hw,
                                     const float* __restrict__ mean_map, float rho,
                                     uint32_t k0, uint32_t k1, uint32_t maxval) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= hw) return;
    SpeckleWalk w;
    w.init(p, mean_map[p], rho, k0, k1);
    for (int t = 0; t < nframes; ++t) out[(int64_t)t * hw + p] = (T)w.draw(t, maxval);
}

// The same walk as synth_speckle_kernel, emitting only the non-zero samples in the
// pixel-major event layout: pass 0 counts events per pixel, pass 1 writes them at
// offsets[p] (frames ascending), so the events equal the compaction of the dense stack.
template <typename T>
__global__ void synth_events_kernel(int pass, int nframes, int64_t hw,
                                    const float* __restrict__ mean_map, float rho, uint32_t k0,
                                    uint32_t k1, uint32_t maxval, uint32_t* __restrict__ count,
                                    const uint32_t* __restrict__ offsets,
                                    int32_t* __restrict__ times, T* __restrict__ data) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= hw) return;
    SpeckleWalk w;
    w.init(p, mean_map[p], rho, k0, k1);
    uint32_t c = 0;
    uint32_t pos = pass ? offsets[p] : 0u;
    for (int t = 0; t < nframes; ++t) {
        const uint32_t k = w.draw(t, maxval);
        if (k) {
            if (pass) {
                times[pos] = t;
                data[pos] = (T)k;
                ++pos;
            }
            ++c;
        }
    }
    if (!pass) count[p] = c;
}

void launch_synth_events(int pass, int dtype, int nframes, int64_t hw, const float* mean_map,
                         float rho, uint64_t seed, uint32_t* count, const uint32_t* offsets,
                         int32_t* times, void* data, hipStream_t s) {
    const unsigned blocks = (unsigned)((hw + 255) / 256);
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    if (dtype == DXG2_U16)
        synth_events_kernel<uint16_t><<<blocks, 256, 0, s>>>(pass, nframes, hw, mean_map, rho, k0,
                                                             k1, 65535u, count, offsets, times,
                                                             (uint16_t*)data);
    else
        synth_events_kernel<uint8_t><<<blocks, 256, 0, s>>>(pass, nframes, hw, mean_map, rho, k0,
                                                            k1, 255u, count, offsets, times,
                                                            (uint8_t*)data);
}

// Pixel-major events -> frame-sorted COO: per-frame counts, a scan, and a scatter by
// one thread per pixel through per-frame atomic cursors.
__global__ void frame_count_kernel(const int32_t* __restrict__ times, int64_t nnz, int nframes,
                                   uint32_t* __restrict__ fcount) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nnz;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int32_t t = times[e];
        if (t >= 0 && t < nframes) atomicAdd(&fcount[t], 1u);
    }
}

template <typename T>
__global__ void to_coo_kernel(const int32_t* __restrict__ times, const T* __restrict__ data,
                              const uint32_t* __restrict__ offsets, int64_t npix, int nframes,
                              uint32_t* __restrict__ fcursor, uint32_t* __restrict__ pix,
                              uint32_t* __restrict__ frame, T* __restrict__ cnt) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= npix) return;
    for (uint32_t e = offsets[p]; e < offsets[p + 1]; ++e) {
        const int32_t t = times[e];
        if (t < 0 || t >= nframes) continue;
        const uint32_t pos = atomicAdd(&fcursor[t], 1u);
        pix[pos] = (uint32_t)p;
        frame[pos] = (uint32_t)t;
        cnt[pos] = data[e];
    }
}

void launch_frame_count(const int32_t* times, int64_t nnz, int nframes, uint32_t* fcount,
                        hipStream_t s) {
    if (nnz <= 0) return;
    int64_t g = (nnz + 255) / 256;
    g = g > 65536 ? 65536 : g;
    frame_count_kernel<<<(unsigned)g, 256, 0, s>>>(times, nnz, nframes, fcount);
}

void launch_to_coo(int dtype, const int32_t* times, const void* data, const uint32_t* offsets,
                   int64_t npix, int nframes, uint32_t* fcursor, uint32_t* pix, uint32_t* frame,
                   void* cnt, hipStream_t s) {
    const unsigned pb = (unsigned)((npix + 255) / 256);
    if (dtype == DXG2_U16)
        to_coo_kernel<uint16_t><<<pb, 256, 0, s>>>(times, (const uint16_t*)data, offsets, npix,
                                                   nframes, fcursor, pix, frame, (uint16_t*)cnt);
    else
        to_coo_kernel<uint8_t><<<pb, 256, 0, s>>>(times, (const uint8_t*)data, offsets, npix,
                                                  nframes, fcursor, pix, frame, (uint8_t*)cnt);
}

void launch_synth(void* frames, int dtype, int nframes, int64_t hw, const float* mean_map,
                  float rho, uint64_t seed, hipStream_t s) {
    const unsigned blocks = (unsigned)((hw + 255) / 256);
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    if (dtype == DXG2_U16)
        synth_speckle_kernel<uint16_t><<<blocks, 256, 0, s>>>(
            (uint16_t*)frames, nframes, hw, mean_map, rho, k0, k1, 65535u);
    else if (dtype == DXG2_I8)
        synth_speckle_kernel<int8_t><<<blocks, 256, 0, s>>>(
            (int8_t*)frames, nframes, hw, mean_map, rho, k0, k1, 127u);
    else
        synth_speckle_kernel<uint8_t><<<blocks, 256, 0, s>>>(
            (uint8_t*)frames, nframes, hw, mean_map, rho, k0, k1, 255u);
}

}  // namespace dxg2
