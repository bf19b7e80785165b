// Synthetic XPCS speckle frames on the device (test and benchmark inputs).
//
// Per pixel p: complex Gaussian field E_t = rho E_{t-1} + sqrt(1 - rho^2) w_t with
// w ~ CN(0, 1), rho = exp(-1/tau_c), E_0 ~ CN(0, 1), so <|E|^2> = 1; counts
// I(t, p) = Poisson(mean_map[p] |E_t|^2), clipped to the output type.  One thread per
// pixel walks the frames in order, so writes of a frame are coalesced across threads.
//
// Random numbers: Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers:
// as easy as 1, 2, 3", SC 2011), counter = (pixel, frame + 1, 0, 0), key = seed.
// Normals by the Box-Muller transform; Poisson by sequential inversion for lambda < 32
// and by a rounded normal approximation above (only the hot-pixel population).
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
        const uint32_t hi0 = __umulhi(M0, c.x), lo0 = M0 * c.x;
        const uint32_t hi1 = __umulhi(M1, c.z), lo1 = M1 * c.z;
        c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += W0;
        k1 += W1;
    }
    return c;
}

__device__ __forceinline__ float u01(uint32_t x) {  // (0, 1)
    return ((float)(x >> 8) + 0.5f) * (1.0f / 16777216.0f);
}

template <typename T>
__global__ void synth_speckle_kernel(T* __restrict__ out, int nframes, int64_t hw,
                                     const float* __restrict__ mean_map, float rho,
                                     uint32_t k0, uint32_t k1, uint32_t maxval) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= hw) return;
    const float mu = mean_map[p];
    const float inov = sqrtf(fmaxf(0.f, 1.f - rho * rho));
    const float two_pi = 6.283185307179586f;
    float er, ei;
    {
        const U4 r = philox4x32_10(U4{(uint32_t)p, 0u, (uint32_t)(p >> 32), 0u}, k0, k1);
        const float rad = sqrtf(-logf(u01(r.x)));  // sqrt(-2 ln u) * sqrt(1/2)
        float sn, cs;
        sincosf(two_pi * u01(r.y), &sn, &cs);
        er = rad * cs;
        ei = rad * sn;
    }
    for (int t = 0; t < nframes; ++t) {
        const U4 r = philox4x32_10(U4{(uint32_t)p, (uint32_t)(t + 1), (uint32_t)(p >> 32), 0u},
                                   k0, k1);
        if (t > 0) {
            const float rad = sqrtf(-logf(u01(r.x)));
            float sn, cs;
            sincosf(two_pi * u01(r.y), &sn, &cs);
            er = rho * er + inov * rad * cs;
            ei = rho * ei + inov * rad * sn;
        }
        const float lam = mu * (er * er + ei * ei);
        uint32_t k = 0;
        if (lam < 32.f) {
            const float u = u01(r.z);
            float pk = expf(-lam);
            float cdf = pk;
            while (u > cdf && k < 200u) {
                ++k;
                pk *= lam / (float)k;
                cdf += pk;
            }
        } else {
            const float rad = sqrtf(-2.f * logf(u01(r.z)));
            const float z = rad * cosf(two_pi * u01(r.w));
            const float v = floorf(lam + sqrtf(lam) * z + 0.5f);
            k = v > 0.f ? (uint32_t)v : 0u;
        }
        if (k > maxval) k = maxval;
        out[(int64_t)t * hw + p] = (T)k;
    }
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
