// Plain FP32 comparator: one bin per launch sequence, the structure of
// CublasMatMulCorrelator._correlate_matmul_cublas (correlator/cuda.py:101-116):
//   A = frames[:, bin pixels] as float32                       (cuda.py:102)
//   G = A A^T, full N x N square, FP32 accumulate              (cuda.py:107)
//   m = mean(A, axis=1)                                        (cuda.py:108)
//   num(tau) = sum diag(G, tau), den(tau) = sum diag(m m^T, tau)  (cuda.py:94-98,111-112)
//   g2 = num / den / n_b                                        (cuda.py:113-114)
// Written as its own HIP kernels (FP32 FMA SGEMM, no vendor BLAS).  The diagonal sum
// of m m^T is evaluated directly as sum_t m_t m_{t+tau} (same FP32 products).
#include "dxg2_internal.h"

namespace dxg2 {

// A[t][k] (row stride lda, zero padded to lda) = frames[t][pix[k]].
__global__ void plain_gather_kernel(const uint8_t* __restrict__ frames, int dtype, int64_t hw,
                                    int nframes, const int32_t* __restrict__ pix, int npix,
                                    int lda, float* __restrict__ a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= lda) return;
    const int p = k < npix ? pix[k] : -1;
    for (int t = blockIdx.y; t < nframes; t += gridDim.y) {
        float v = 0.f;
        if (p >= 0) {
            const uint8_t* fr = frames + (int64_t)t * hw * (dtype == DXG2_U16 ? 2 : 1);
            if (dtype == DXG2_U8)
                v = (float)fr[p];
            else if (dtype == DXG2_I8)
                v = (float)(int8_t)fr[p];
            else
                v = (float)reinterpret_cast<const uint16_t*>(fr)[p];
        }
        a[(int64_t)t * lda + k] = v;
    }
}

// m[t] = mean over the row (FP32 accumulate), one wave per frame.
__global__ void plain_mean_kernel(const float* __restrict__ a, int nframes, int npix, int lda,
                                  float* __restrict__ m) {
    const int t = blockIdx.x;
    if (t >= nframes) return;
    float s = 0.f;
    for (int k = threadIdx.x; k < npix; k += 64) s += a[(int64_t)t * lda + k];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (threadIdx.x == 0) m[t] = s / (float)npix;
}

// C (n x n) = A A^T, A (n x lda) row-major, lda % 8 == 0.  128 x 128 tiles, 256
// threads, 8 x 8 outputs per thread, BK = 8, LDS stored k-major.
__global__ __launch_bounds__(256) void plain_sgemm_nt_kernel(const float* __restrict__ a,
                                                             int n, int lda,
                                                             float* __restrict__ c) {
    constexpr int BM = 128, BK = 8;
    __shared__ float sa[BK][BM + 4];
    __shared__ float sb[BK][BM + 4];
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const int row0 = blockIdx.y * BM, col0 = blockIdx.x * BM;
    // loader: thread -> (row = tid >> 1, k4 = (tid & 1) * 4)
    const int lr = tid >> 1, lk = (tid & 1) * 4;
    float accum[8][8];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) accum[i][j] = 0.f;
    const int ra = row0 + lr < n ? row0 + lr : n - 1;
    const int rb = col0 + lr < n ? col0 + lr : n - 1;
    const bool va = row0 + lr < n, vb = col0 + lr < n;
    for (int k0 = 0; k0 < lda; k0 += BK) {
        float4 xa = *reinterpret_cast<const float4*>(a + (int64_t)ra * lda + k0 + lk);
        float4 xb = *reinterpret_cast<const float4*>(a + (int64_t)rb * lda + k0 + lk);
        if (!va) xa = make_float4(0.f, 0.f, 0.f, 0.f);
        if (!vb) xb = make_float4(0.f, 0.f, 0.f, 0.f);
        __syncthreads();
        sa[lk + 0][lr] = xa.x; sa[lk + 1][lr] = xa.y; sa[lk + 2][lr] = xa.z; sa[lk + 3][lr] = xa.w;
        sb[lk + 0][lr] = xb.x; sb[lk + 1][lr] = xb.y; sb[lk + 2][lr] = xb.z; sb[lk + 3][lr] = xb.w;
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; ++kk) {
            float fa[8], fb[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) fa[i] = sa[kk][ty + 16 * i];
#pragma unroll
            for (int j = 0; j < 8; ++j) fb[j] = sb[kk][tx + 16 * j];
#pragma unroll
            for (int i = 0; i < 8; ++i)
#pragma unroll
                for (int j = 0; j < 8; ++j) accum[i][j] = fmaf(fa[i], fb[j], accum[i][j]);
        }
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int r = row0 + ty + 16 * i;
        if (r >= n) continue;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int cc = col0 + tx + 16 * j;
            if (cc < n) c[(int64_t)r * n + cc] = accum[i][j];
        }
    }
}

// One thread per lag: FP32 diagonal sums of G and of m m^T, then the ratio.
__global__ void plain_diag_kernel(const float* __restrict__ g, const float* __restrict__ m,
                                  int n, int npix, float* __restrict__ out) {
    const int tau = blockIdx.x * blockDim.x + threadIdx.x;
    if (tau >= n) return;
    float sn = 0.f, sd = 0.f;
    for (int t = 0; t + tau < n; ++t) {
        sn += g[(int64_t)t * n + t + tau];
        sd += m[t] * m[t + tau];
    }
    out[tau] = sn / sd / (float)npix;
}

void launch_plain_bin(const uint8_t* frames, int dtype, int64_t hw, int nframes,
                      const int32_t* pix, int npix, float* a, float* gram, float* means,
                      float* g2_row, hipStream_t s) {
    const int lda = (npix + 7) / 8 * 8;
    const int gy = nframes < 65535 ? nframes : 65535;
    plain_gather_kernel<<<dim3((unsigned)((lda + 255) / 256), (unsigned)gy), dim3(256), 0, s>>>(
        frames, dtype, hw, nframes, pix, npix, lda, a);
    plain_mean_kernel<<<dim3((unsigned)nframes), dim3(64), 0, s>>>(a, nframes, npix, lda, means);
    const unsigned nt = (unsigned)((nframes + 127) / 128);
    plain_sgemm_nt_kernel<<<dim3(nt, nt), dim3(256), 0, s>>>(a, nframes, lda, gram);
    plain_diag_kernel<<<dim3((unsigned)((nframes + 255) / 256)), dim3(256), 0, s>>>(
        gram, means, nframes, npix, g2_row);
}

}  // namespace dxg2
