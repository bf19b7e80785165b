// FFT correlator: numerator from power spectra along the time axis.
//
// Reference: FFTWCorrelator._correlate_fft (correlator/dense.py:347-369) and
// CUFFTCorrelator (cuda.py:185-219): per bin, a float32 FFT of every pixel series
// zero-padded to Nf >= 2N (dense.py:289), times the FFT of the reversed series, one
// inverse FFT per pixel, a sum over pixels; the denominator is np.correlate of the
// integer frame sums (dense.py:365-366).
//
// Here, per bin:
//   Num(tau) = IFFT_Nf( sum_p |X_p(f)|^2 )[tau],   X_p = FFT_Nf(x_p zero-padded),
// (the product with the reversed series is the same autocorrelation; the inverse is
// linear, so one inverse per bin replaces one per pixel).  Nf = 2L with L >= N from a
// set of mixed-radix lengths (2^a 5^b), so a linear autocorrelation needs no more.
//  * gather_pm_kernel<E>: frames -> pixel-major series xs[k][t] of the frame element
//    type E (uint8, int8 or uint16; k in the bin's raster list order, t < nrows, zero
//    past the last frame), 64 pixels x 256 frames per workgroup through an LDS transpose;
//    the same workgroup adds its 64 pixels per frame from the LDS tile and sends each
//    frame's total to the exact S_b(t) with one 64-bit atomic.
//  * fft_spectra_kernel<E, L, T, R...>: one workgroup per (bin, k-block of 64 series);
//    each series is packed as z[n] = x[2n] + i x[2n+1] (L complex points) and
//    transformed by an FP32 mixed-radix Stockham FFT (Govindaraju et al., "High
//    performance discrete Fourier transforms on graphics processors", SC 2008) with
//    register DFTs of radix 4, 5, 8 or 16, then unpacked to X[k], k = 0..L, and |X[k]|^2
//    is accumulated in FP64 per thread.  The first pass reads its inputs straight from
//    HBM into registers (the next series' are loaded while the current one is
//    transformed); the twiddles of the middle passes come from small LDS tables and those
//    of the last pass and the unpacking factors are held in registers, so the series
//    loop issues no other global load.  The last pass works in place (each butterfly
//    writes the points it read).
//  * fft_inverse_kernel: Num(tau) = (P_0 + (-1)^tau P_L + 2 sum_{k=1}^{L-1} P_k
//    cos(2 pi k tau / Nf)) / Nf in FP64, the cosines by four interleaved FP64 rotation
//    recurrences.
// Twiddle tables are computed on the host in FP64 and rounded to FP32.
#include "dxg2_internal.h"

namespace dxg2 {

struct c32 {
    float x, y;
};
__device__ __forceinline__ c32 cmul(c32 a, c32 b) {
    return c32{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__device__ __forceinline__ c32 cadd(c32 a, c32 b) { return c32{a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ c32 csub(c32 a, c32 b) { return c32{a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ c32 negi(c32 a) { return c32{a.y, -a.x}; }   // -i a
__device__ __forceinline__ c32 cscale(c32 a, float s) { return c32{a.x * s, a.y * s}; }
__device__ __forceinline__ c32 c_of(float2 t) { return c32{t.x, t.y}; }

// ------------------------------------------------------------- register DFTs
// Dft<R>::run(a): in place, natural order, a[k] = sum_n a[n] exp(-2 pi i n k / R).
template <int R>
struct Dft;

template <>
struct Dft<2> {
    __device__ __forceinline__ static void run(c32* a) {
        const c32 s = cadd(a[0], a[1]), d = csub(a[0], a[1]);
        a[0] = s;
        a[1] = d;
    }
};

template <>
struct Dft<4> {
    __device__ __forceinline__ static void run(c32* a) {
        const c32 s02 = cadd(a[0], a[2]), d02 = csub(a[0], a[2]);
        const c32 s13 = cadd(a[1], a[3]), d13 = negi(csub(a[1], a[3]));
        a[0] = cadd(s02, s13);
        a[1] = cadd(d02, d13);
        a[2] = csub(s02, s13);
        a[3] = csub(d02, d13);
    }
};

template <>
struct Dft<5> {
    __device__ __forceinline__ static void run(c32* a) {
        // cos(2 pi/5), cos(4 pi/5), sin(2 pi/5), sin(4 pi/5)
        constexpr float C1 = 0.30901699437494745f, C2 = -0.80901699437494734f;
        constexpr float S1 = 0.95105651629515353f, S2 = 0.58778525229247325f;
        const c32 t1 = cadd(a[1], a[4]), t2 = cadd(a[2], a[3]);
        const c32 t3 = csub(a[1], a[4]), t4 = csub(a[2], a[3]);
        const c32 b1 = c32{a[0].x + C1 * t1.x + C2 * t2.x, a[0].y + C1 * t1.y + C2 * t2.y};
        const c32 b2 = c32{a[0].x + C2 * t1.x + C1 * t2.x, a[0].y + C2 * t1.y + C1 * t2.y};
        const c32 d1 = c32{S1 * t3.x + S2 * t4.x, S1 * t3.y + S2 * t4.y};
        const c32 d2 = c32{S2 * t3.x - S1 * t4.x, S2 * t3.y - S1 * t4.y};
        a[0] = cadd(a[0], cadd(t1, t2));
        a[1] = cadd(b1, negi(d1));
        a[4] = csub(b1, negi(d1));
        a[2] = cadd(b2, negi(d2));
        a[3] = csub(b2, negi(d2));
    }
};

// exp(-2 pi i m / 8) for m < 4
__device__ __forceinline__ c32 w8(int m) {
    constexpr float H = 0.70710678118654752f;
    switch (m) {
        case 1: return c32{H, -H};
        case 2: return c32{0.f, -1.f};
        case 3: return c32{-H, -H};
        default: return c32{1.f, 0.f};
    }
}

template <>
struct Dft<8> {
    // n = 2 m + p: E_p = DFT4 over m of a[2m + p]; a[k] = E0[k] + w8^k E1[k],
    // a[k + 4] = E0[k] - w8^k E1[k]
    __device__ __forceinline__ static void run(c32* a) {
        c32 e0[4] = {a[0], a[2], a[4], a[6]};
        c32 e1[4] = {a[1], a[3], a[5], a[7]};
        Dft<4>::run(e0);
        Dft<4>::run(e1);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const c32 t = k ? cmul(w8(k), e1[k]) : e1[k];
            a[k] = cadd(e0[k], t);
            a[k + 4] = csub(e0[k], t);
        }
    }
};

// exp(-2 pi i m / 16) for m = p k in {0, 1, 2, 3, 4, 6, 9}
__device__ __forceinline__ c32 w16(int m) {
    constexpr float C1 = 0.92387953251128674f, S1 = 0.38268343236508978f;
    constexpr float H = 0.70710678118654752f;
    switch (m) {
        case 1: return c32{C1, -S1};
        case 2: return c32{H, -H};
        case 3: return c32{S1, -C1};
        case 4: return c32{0.f, -1.f};
        case 6: return c32{-H, -H};
        case 9: return c32{-C1, S1};
        default: return c32{1.f, 0.f};
    }
}

template <>
struct Dft<16> {
    // n = 4 m + p: E_p = DFT4 over m of a[4m + p]; a[k + 4q] = DFT4 over p of
    // (w16^(p k) E_p[k]), output q
    __device__ __forceinline__ static void run(c32* a) {
        c32 e[4][4];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
#pragma unroll
            for (int m = 0; m < 4; ++m) e[p][m] = a[4 * m + p];
            Dft<4>::run(e[p]);
        }
        twist(e, a);
    }
    // a[8..15] zero (the zero-padded half of a series): the first DFT4s see
    // (a[p], a[4 + p], 0, 0), and Dft<4> on them gives these sums and differences (its
    // terms with the zeros are exact), so the result is the same as run()'s.
    __device__ __forceinline__ static void run_lo8(c32* a) {
        c32 e[4][4];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            const c32 x0 = a[p], x1 = a[4 + p], y1 = negi(x1);
            e[p][0] = cadd(x0, x1);
            e[p][1] = cadd(x0, y1);
            e[p][2] = csub(x0, x1);
            e[p][3] = csub(x0, y1);
        }
        twist(e, a);
    }
    // second stage: a[k + 4q] = DFT4 over p of (w16^(p k) E_p[k]), output q
    __device__ __forceinline__ static void twist(c32 (&e)[4][4], c32* a) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            c32 v[4];
#pragma unroll
            for (int p = 0; p < 4; ++p) v[p] = (p && k) ? cmul(w16(p * k), e[p][k]) : e[p][k];
            Dft<4>::run(v);
#pragma unroll
            for (int q = 0; q < 4; ++q) a[k + 4 * q] = v[q];
        }
    }
};

// LDS index of point i: the low 4 bits are xor-ed with the next 4, so the stride-16
// accesses of a pass hit 16 different 8-byte slots and the array needs no padding.
__device__ __forceinline__ int sw(int i) { return i ^ ((i >> 4) & 15); }
// sw(base + r S) with r known at compile time (r may be negative): for S a multiple of
// 256, r S leaves bits 0..7 of base + r S as they are in base, so the swizzle is
// sw(base) + r S; for S an odd multiple of 128 it leaves bits 0..6 and flips bit 7 when
// r is odd, which flips bit 3 of the xor term: (sw(base) ^ 8) + r S.  Either way one
// address per butterfly and constant offsets.
template <int S>
__device__ __forceinline__ int sw_step(int base, int r) {
    if constexpr (S % 256 == 0) return sw(base) + r * S;
    else if constexpr (S % 128 == 0) return (sw(base) ^ ((r & 1) << 3)) + r * S;
    else return sw(base + r * S);
}

// ----------------------------------------------------------- series layout
// Raw load of complex point i of a series: elements 2i and 2i + 1 (one 16-bit word for
// 8-bit frames, one 32-bit word for 16-bit frames), zero past the row.
template <typename E>
struct Elem;
template <>
struct Elem<uint8_t> {
    static constexpr int SIZE = 1;
    __device__ __forceinline__ static uint32_t load(const uint8_t* row, int i, int nrows) {
        return 2 * i < nrows ? (uint32_t) reinterpret_cast<const uint16_t*>(row)[i] : 0u;
    }
    __device__ __forceinline__ static c32 value(uint32_t r) {
        return c32{(float)(r & 0xFFu), (float)(r >> 8)};
    }
};
template <>
struct Elem<int8_t> {
    static constexpr int SIZE = 1;
    __device__ __forceinline__ static uint32_t load(const uint8_t* row, int i, int nrows) {
        return 2 * i < nrows ? (uint32_t) reinterpret_cast<const uint16_t*>(row)[i] : 0u;
    }
    __device__ __forceinline__ static c32 value(uint32_t r) {
        return c32{(float)(int8_t)(r & 0xFFu), (float)(int8_t)((r >> 8) & 0xFFu)};
    }
};
template <>
struct Elem<uint16_t> {
    static constexpr int SIZE = 2;
    __device__ __forceinline__ static uint32_t load(const uint8_t* row, int i, int nrows) {
        return 2 * i < nrows ? reinterpret_cast<const uint32_t*>(row)[i] : 0u;
    }
    __device__ __forceinline__ static c32 value(uint32_t r) {
        return c32{(float)(r & 0xFFFFu), (float)(r >> 16)};
    }
};

// --------------------------------------------------------------- pixel-major gather
// grid (k-blocks in gather order, frame tiles of 256).  Lane = pixel of the k-block: each
// load instruction reads one frame of the 64 pixels (the k-block's stretch of detector
// row).  A lane packs 4 / sizeof(E) consecutive frames into one 32-bit word and writes
// it to its pixel's row of an LDS tile whose row stride is an odd number of words, so
// the 64 lanes' column writes fall in distinct banks; then each thread reads 16
// contiguous bytes of one row and stores them to xs, so consecutive threads write
// consecutive 16-byte pieces of a pixel's series.  The frame sums S_b(t) (zeroed first)
// come from the same tile: a thread adds one word (4 frames of 8-bit values in 16-bit
// halves, or 2 of 16-bit values in 32-bit registers) down 16 or 32 pixel rows, the row
// groups meet in LDS, and each frame's total over the 64 pixels (below 2^31) goes to
// sums with one 64-bit atomic (the row groups' partials reuse the tile's LDS once every
// thread has read it).  SGN (int8 frames): bytes are added as v ^ 0x80 = v + 128
// and 64 x 128 is taken off the total (padded pixels and frames hold 0, so they add 128
// each and are taken off with the rest).
// quad_kb (8-bit frames, 4 | hw, frames 4-aligned; else null): per bin, the leading
// k-blocks whose list entries are 16 aligned quads of detector pixels 4a .. 4a + 3.  In
// those, lane (g, kk) = (lane % 16, lane / 16) reads quad g of frame 4k + f (f < 4) of word
// k = 4 (wave + 4 i) + kk as one aligned 4-byte load (a quarter of the load instructions),
// and a 4 x 4 byte transpose in registers turns the four frames' quads into the four
// pixels' words of the tile.
// A workgroup takes KPB k-blocks that are neighbours in detector order (work item
// blockIdx.x; kb -1: none) over 256 / KPB frames, so that the workgroups in flight read
// a larger share of each frame row, and still stores whole 128-byte lines of each series.
template <typename E, bool SGN, int KPB>
__global__ __launch_bounds__(256) void gather_pm_kernel(const uint8_t* __restrict__ frames,
                                                        int64_t hw, int nframes, int nrows,
                                                        const int32_t* __restrict__ pix_idx,
                                                        const int32_t* __restrict__ quad_kb,
                                                        const BinDev* __restrict__ bins,
                                                        const KbWork* __restrict__ work,
                                                        uint8_t* __restrict__ xs,
                                                        int64_t* __restrict__ sums) {
    constexpr int FT = 256 / KPB;      // frames per workgroup
    constexpr int ES = sizeof(E);
    constexpr int FPW = 4 / ES;        // frames per word
    constexpr int WPR = FT / FPW;      // words per pixel row of the tile
    constexpr int STRIDE = WPR + 1;    // odd
    constexpr int ROWS = KPB * KB;     // pixel rows of the tile
    constexpr int RG = 256 / WPR;      // row groups of the sums (threads = WPR x RG)
    constexpr int RPG = ROWS / RG;     // rows per group: within one k-block
    static_assert(KB % RPG == 0 && FT * KPB == 256 && WPR % 16 == 0, "gather shape");
    __shared__ __attribute__((aligned(16))) uint32_t tile[ROWS * STRIDE];
    static_assert(RG * FT <= ROWS * STRIDE, "partials fit the tile");
    const KbWork* item = work + (int64_t)blockIdx.x * KPB;
    if (item[0].kb < 0) return;  // padding of the gather order (the first slot is filled first)
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int t0 = blockIdx.y * FT;
    const int ntq = max(0, min(FT, nframes - t0));  // frames present
    constexpr uint32_t MASK = ES == 1 ? 0xFFu : 0xFFFFu;
#pragma unroll
    for (int j = 0; j < KPB; ++j) {
        const KbWork w = item[j];
        if (w.kb < 0) continue;  // the same for the whole workgroup
        const BinDev bd = bins[w.bin];
        uint32_t* rows = tile + j * KB * STRIDE;
        if (ES == 1 && quad_kb && w.kb < quad_kb[w.bin]) {
            const int g = lane & 15, kk = lane >> 4;
            const int q0 = pix_idx[bd.idx_off + (int64_t)w.kb * KB + 4 * g];  // 4 | q0
#pragma unroll
            for (int i = 0; i < WPR / 16; ++i) {
                const int k = 4 * (wave + 4 * i) + kk;
                uint32_t v[4];
#pragma unroll
                for (int f4 = 0; f4 < 4; ++f4) {
                    const int f = 4 * k + f4;
                    v[f4] = f < ntq ? *reinterpret_cast<const uint32_t*>(
                                          frames + (int64_t)(t0 + f) * hw + q0)
                                    : 0u;
                }
                // v[f] byte b = pixel 4g + b at frame 4k + f  ->  word k of pixel 4g + b
                const uint32_t l01 = __builtin_amdgcn_perm(v[1], v[0], 0x05010400u);
                const uint32_t h01 = __builtin_amdgcn_perm(v[1], v[0], 0x07030602u);
                const uint32_t l23 = __builtin_amdgcn_perm(v[3], v[2], 0x05010400u);
                const uint32_t h23 = __builtin_amdgcn_perm(v[3], v[2], 0x07030602u);
                uint32_t* row = rows + (4 * g) * STRIDE + k;
                row[0] = __builtin_amdgcn_perm(l23, l01, 0x05040100u);
                row[STRIDE] = __builtin_amdgcn_perm(l23, l01, 0x07060302u);
                row[2 * STRIDE] = __builtin_amdgcn_perm(h23, h01, 0x05040100u);
                row[3 * STRIDE] = __builtin_amdgcn_perm(h23, h01, 0x07060302u);
            }
        } else {
            const int p = pix_idx[bd.idx_off + (int64_t)w.kb * KB + lane];
            const E* fr = reinterpret_cast<const E*>(frames) + (p >= 0 ? p : 0);
            const int nt = p >= 0 ? ntq : 0;
#pragma unroll
            for (int i = 0; i < WPR / 4; ++i) {
                const int k = wave + 4 * i;  // word k: frames t0 + FPW k ..
                uint32_t word = 0u;
#pragma unroll
                for (int q = 0; q < FPW; ++q) {
                    const int f = FPW * k + q;
                    const uint32_t v = f < nt ? (uint32_t)fr[(int64_t)(t0 + f) * hw] : 0u;
                    word |= (v & MASK) << (8 * ES * q);
                }
                rows[lane * STRIDE + k] = word;
            }
        }
    }
    __syncthreads();
    constexpr int CHUNKS = WPR / 4;  // 16-byte pieces per pixel row
    for (int c = threadIdx.x; c < ROWS * CHUNKS; c += 256) {
        const int r = c / CHUNKS;
        const KbWork w = item[r / KB];
        if (w.kb < 0) continue;
        const int o = (c % CHUNKS) * 4;
        const uint32_t* src = tile + r * STRIDE + o;
        uint8_t* dst = xs + bins[w.bin].slab_off +
                       (((int64_t)w.kb * KB + r % KB) * nrows + t0) * ES + 4 * o;
        *reinterpret_cast<uint4*>(dst) = make_uint4(src[0], src[1], src[2], src[3]);
    }
    const int k = threadIdx.x % WPR, g = threadIdx.x / WPR;  // word k, rows RPG g ..
    const uint32_t* col = tile + RPG * g * STRIDE + k;
    uint32_t s0 = 0u, s1 = 0u;
    if constexpr (ES == 1) {
        // s0: frames 4k, 4k + 2 and s1: 4k + 1, 4k + 3 in 16-bit halves (16 rows x 255)
#pragma unroll
        for (int j = 0; j < RPG; ++j) {
            const uint32_t x = col[j * STRIDE] ^ (SGN ? 0x80808080u : 0u);
            s0 += x & 0x00FF00FFu;
            s1 += (x >> 8) & 0x00FF00FFu;
        }
    } else {
        // s0: frame 2k, s1: 2k + 1 (32 rows x 65535)
#pragma unroll 8
        for (int j = 0; j < RPG; ++j) {
            const uint32_t x = col[j * STRIDE];
            s0 += x & 0xFFFFu;
            s1 += x >> 16;
        }
    }
    static_assert(ES != 1 || RPG <= 256, "16-bit halves");
    __syncthreads();  // every thread has read the tile: its words now hold the partials
    int32_t* part = reinterpret_cast<int32_t*>(tile);  // [RG][FT]
    if constexpr (ES == 1)
        *reinterpret_cast<int4*>(part + FT * g + 4 * k) =
            make_int4((int)(s0 & 0xFFFFu), (int)(s1 & 0xFFFFu), (int)(s0 >> 16),
                      (int)(s1 >> 16));
    else
        *reinterpret_cast<int2*>(part + FT * g + 2 * k) = make_int2((int)s0, (int)s1);
    __syncthreads();
    // thread (j, f): frame t0 + f of k-block j, from the RG / KPB row groups of that k-block
    const int jb = threadIdx.x / FT, f = threadIdx.x % FT;
    const KbWork w = item[jb];
    const int t = t0 + f;
    if (w.kb >= 0 && t < nframes) {
        constexpr int GPK = RG / KPB;  // row groups per k-block
        int64_t tot = SGN ? -(int64_t)KB * 128 : 0;
#pragma unroll
        for (int q = 0; q < GPK; ++q) tot += part[FT * (GPK * jb + q) + f];
        if (tot)
            atomicAdd(reinterpret_cast<unsigned long long*>(sums + (int64_t)w.bin * nframes + t),
                      (unsigned long long)tot);
    }
}

// ------------------------------------------------------------------- spectra
// A plan: L complex points, T threads, radices R1..R4 (R4 = 1 when there are three
// passes).  The last pass's twiddles are held in registers, the middle passes' in LDS
// tables (pass p's table holds exp(-2 pi i m / P_p), m < P_p = R1 ... R_p).
template <int L_, int T_, int R1_, int R2_, int R3_, int R4_>
struct FftPlan {
    static constexpr int L = L_, T = T_, R1 = R1_, R2 = R2_, R3 = R3_, R4 = R4_;
    static constexpr int NP = R4 > 1 ? 4 : 3;
    static constexpr int RL = NP == 4 ? R4 : R3;        // radix of the last pass
    static constexpr int NSL = L / RL;                   // its stride
    static constexpr int P2 = R1 * R2;                   // table of pass 2
    static constexpr int P3 = NP == 4 ? R1 * R2 * R3 : 1;  // table of pass 3 (4 passes)
    static_assert(R1 * R2 * R3 * R4 == L, "radices must multiply to L");
};

template <typename Plan, int R, int NS, int PT>
__device__ __forceinline__ void mid_pass(c32* z, const c32* tab, int tid) {
    constexpr int L = Plan::L, T = Plan::T;
    constexpr int NB = L / R;
    constexpr int M = (NB + T - 1) / T;
    c32 v[M][R];
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const int b = tid + m * T;
        if (NB % T == 0 || b < NB) {
            const int k = b % NS;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                c32 a = z[sw_step<NB>(b, r)];
                if (r) a = cmul(a, tab[r * k]);      // exp(-2 pi i r k / (R NS)), PT = R NS
                v[m][r] = a;
            }
            Dft<R>::run(v[m]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const int b = tid + m * T;
        if (NB % T == 0 || b < NB) {
            const int k = b % NS;
            const int o = (b - k) * R + k;
#pragma unroll
            for (int q = 0; q < R; ++q) z[sw_step<NS>(o, q)] = v[m][q];
        }
    }
    __syncthreads();
}

template <typename E, typename Plan>
__global__ __launch_bounds__(Plan::T) void fft_spectra_kernel(
    const uint8_t* __restrict__ xs, const BinDev* __restrict__ bins,
    const KbWork* __restrict__ work, int nrows, const float2* __restrict__ tw,
    const float2* __restrict__ uw, double* __restrict__ power) {
    constexpr int L = Plan::L, T = Plan::T;
    constexpr int R1 = Plan::R1, R2 = Plan::R2, R3 = Plan::R3, RL = Plan::RL;
    constexpr int NB1 = L / R1, M1 = (NB1 + T - 1) / T;       // first pass
    constexpr int NBL = L / RL, ML = (NBL + T - 1) / T;       // last pass
    constexpr int HP = L / 2;                                 // k = HP pairs with itself
    constexpr int MH = (HP + T - 1) / T;                      // pairs (k, L - k), k < HP
    __shared__ c32 z[L];
    __shared__ c32 tab2[Plan::P2];
    __shared__ c32 tab3[Plan::P3];
    const KbWork w = work[blockIdx.x];
    const BinDev bd = bins[w.bin];
    const int tid = threadIdx.x;

    for (int m = tid; m < Plan::P2; m += T) tab2[m] = c_of(tw[m * (L / Plan::P2)]);
    if (Plan::NP == 4)
        for (int m = tid; m < Plan::P3; m += T) tab3[m] = c_of(tw[m * (L / Plan::P3)]);
    double acc_lo[MH], acc_hi[MH], acc_mid = 0.0;  // |X|^2 at k, L - k, HP
#pragma unroll
    for (int m = 0; m < MH; ++m) acc_lo[m] = acc_hi[m] = 0.0;

    const int kend = min(KB, bd.npix - w.kb * KB);
    const uint8_t* rows = xs + bd.slab_off + (int64_t)w.kb * KB * nrows * Elem<E>::SIZE;
    const int64_t row_bytes = (int64_t)nrows * Elem<E>::SIZE;
    // Points from L / 2 on are frames from L on, past every row (the launcher checks
    // nrows <= L): the zero padding.  With R1 = 16 they are the first pass's inputs
    // r >= 8, which are neither loaded nor added.
    constexpr int RLD = R1 == 16 ? 8 : R1;   // inputs loaded per first-pass butterfly
    static_assert(R1 != 16 || 8 * NB1 == L / 2, "zero half");
    uint32_t raw[M1][RLD];
#pragma unroll
    for (int m = 0; m < M1; ++m)
#pragma unroll
        for (int r = 0; r < RLD; ++r) {
            const int b = tid + m * T;
            raw[m][r] = (NB1 % T == 0 || b < NB1) ? Elem<E>::load(rows, b + r * NB1, nrows) : 0u;
        }
    __syncthreads();   // tables

    for (int j = 0; j < kend; ++j) {
        // Every address, twiddle and factor below depends on the thread only; an opaque
        // copy of the thread index per series keeps the compiler from holding them all
        // in registers across the series loop (which spills): they are recomputed and
        // the last pass's twiddles and the unpacking factors re-read from L1/L2.  By
        // then the next series' loads issued after the first pass have landed.
        int t = tid;
        asm volatile("" : "+v"(t));
        // first pass (stride 1, no twiddles) from the registers loaded during the
        // previous series: z[b + r NB1] -> outputs q at b R1 + q
#pragma unroll
        for (int m = 0; m < M1; ++m) {
            const int b = t + m * T;
            if (NB1 % T == 0 || b < NB1) {
                c32 v[R1];
#pragma unroll
                for (int r = 0; r < RLD; ++r) v[r] = Elem<E>::value(raw[m][r]);
                if constexpr (R1 == 16)
                    Dft<16>::run_lo8(v);
                else
                    Dft<R1>::run(v);
#pragma unroll
                for (int q = 0; q < R1; ++q) z[sw(b * R1 + q)] = v[q];
            }
        }
        // the next series' inputs, in flight during this series' passes
        if (j + 1 < kend) {
            const uint8_t* nxt = rows + (int64_t)(j + 1) * row_bytes;
#pragma unroll
            for (int m = 0; m < M1; ++m)
#pragma unroll
                for (int r = 0; r < RLD; ++r) {
                    const int b = t + m * T;
                    if (NB1 % T == 0 || b < NB1)
                        raw[m][r] = Elem<E>::load(nxt, b + r * NB1, nrows);
                }
        }
        __syncthreads();
        mid_pass<Plan, R2, R1, Plan::P2>(z, tab2, t);
        if (Plan::NP == 4) mid_pass<Plan, R3, R1 * R2, Plan::P3>(z, tab3, t);
        // last pass, in place: butterfly b reads and writes z[b + r NSL]; its twiddles
        // are exp(-2 pi i r b / L) (k = b < NSL)
#pragma unroll
        for (int m = 0; m < ML; ++m) {
            const int b = t + m * T;
            if (NBL % T == 0 || b < NBL) {
                c32 v[RL];
#pragma unroll
                for (int r = 0; r < RL; ++r) {
                    const c32 a = z[sw_step<NBL>(b, r)];
                    v[r] = r ? cmul(a, c_of(tw[r * b])) : a;
                }
                Dft<RL>::run(v);
#pragma unroll
                for (int q = 0; q < RL; ++q) z[sw_step<NBL>(b, q)] = v[q];
            }
        }
        __syncthreads();
        // unpack X[k] = S + D with S = (Z[k] + conj Z[L-k]) / 2 and
        // D = -i W^k (Z[k] - conj Z[L-k]) / 2, W = exp(-pi i / L), Z[L] = Z[0].  As
        // W^(L-k) = -conj W^k, the same S and D give conj X[L-k] = S - D, so each pair
        // (k, L - k), k < L/2, is read and unpacked once (k = L/2 pairs with itself), and
        // |X|^2 accumulates at both ends.  With t < T only m = 0 meets k = 0, and only
        // the last m can pass L/2: the other points need no test.  The halves of S and D
        // are left out here, and the accumulated |2X|^2 is scaled by 1/4 (exact) when it
        // is flushed.
        auto unpack = [&](c32 a, c32 c, int k, double& lo, double* hi) {
            const c32 s2 = c32{a.x + c.x, a.y - c.y};
            const c32 d = c32{a.x - c.x, a.y + c.y};
            const c32 wd = cmul(c_of(uw[k]), d);
            const double xr = s2.x + wd.y, xi = s2.y - wd.x;  // 2 X[k], rounded in FP32
            lo = fma(xi, xi, fma(xr, xr, lo));
            if (hi) {
                const double yr = s2.x - wd.y, yi = s2.y + wd.x;  // 2 conj X[L-k]
                *hi = fma(yi, yi, fma(yr, yr, *hi));
            }
        };
#pragma unroll
        for (int m = 0; m < MH; ++m) {
            const int k = t + m * T;
            if (HP % T == 0 || m < MH - 1 || k < HP) {
                const c32 a = z[sw_step<T>(t, m)];
                const c32 c = z[m == 0 ? sw(k == 0 ? 0 : L - k) : sw_step<T>(L - t, -m)];
                unpack(a, c, k, acc_lo[m], &acc_hi[m]);
            }
        }
        if (t == 0) {
            const c32 a = z[sw(HP)];
            unpack(a, a, HP, acc_mid, nullptr);
        }
        __syncthreads();
    }
    double* prow = power + (int64_t)w.bin * (L + 1);
#pragma unroll
    for (int m = 0; m < MH; ++m) {
        const int k = tid + m * T;
        if (k < HP) {
            if (acc_lo[m] != 0.0) atomicAdd(prow + k, 0.25 * acc_lo[m]);
            if (acc_hi[m] != 0.0) atomicAdd(prow + L - k, 0.25 * acc_hi[m]);
        }
    }
    if (tid == 0 && acc_mid != 0.0) atomicAdd(prow + HP, 0.25 * acc_mid);
}

// ------------------------------------------------- long series: four-step through HBM
// For N > 16384 the half-length is L = L1 L2 (L1 a plan length up to 8192, L2 in
// {2, 4, 5, 8, 16}) and the transform is split (Bailey's four-step FFT):
//   n = n1 + L1 n2, k = k2 + L2 k1,
//   Y[k2][n1] = exp(-2 pi i n1 k2 / L) DFT_L2 over n2 of z[n1 + L1 n2]   (four_step_cols)
//   Z[k2 + L2 k1] = DFT_L1 over n1 of Y[k2][n1]                          (four_step_rows)
// through a scratch Y of L complex values per series, for a chunk of k-blocks at a time.
// Z[k] pairs with Z[L - k] = Z[(L2 - k2) + L2 (L1 - 1 - k1)] for k2 > 0 (and with
// Z[L2 ((L1 - k1) mod L1)] for k2 = 0), so a workgroup of four_step_rows transforms the
// two rows k2 and L2 - k2 of a series and unpacks them together.

// grid (series of the chunk, ceil(L1 / 256)); thread = n1.  Series s of the chunk is row
// j = s % 64 of k-block work[kb0 + s / 64].
template <typename E, int L2>
__global__ __launch_bounds__(256) void four_step_cols_kernel(
    const uint8_t* __restrict__ xs, const BinDev* __restrict__ bins,
    const KbWork* __restrict__ work, int kb0, int nrows, int L1,
    const float2* __restrict__ tw, float2* __restrict__ y) {
    const int s = blockIdx.x;
    const KbWork w = work[kb0 + s / KB];
    const BinDev bd = bins[w.bin];
    const int j = s % KB;
    if (j >= bd.npix - w.kb * KB) return;     // padding rows of a bin's last k-block
    const int n1 = blockIdx.y * 256 + threadIdx.x;
    if (n1 >= L1) return;
    const uint8_t* row = xs + bd.slab_off +
                         ((int64_t)w.kb * KB + j) * nrows * Elem<E>::SIZE;
    c32 v[L2];
#pragma unroll
    for (int n2 = 0; n2 < L2; ++n2) v[n2] = Elem<E>::value(Elem<E>::load(row, n1 + L1 * n2, nrows));
    Dft<L2>::run(v);
    float2* out = y + (int64_t)s * L1 * L2;
#pragma unroll
    for (int k2 = 0; k2 < L2; ++k2) {
        const c32 r = k2 ? cmul(v[k2], c_of(tw[n1 * k2])) : v[k2];
        out[(int64_t)k2 * L1 + n1] = make_float2(r.x, r.y);
    }
}

template <int L2>
struct RowPairs {   // (row A, row B) with B = L2 - A, or B = A for a row that pairs itself
    static constexpr int N = L2 / 2 + 1;
    __device__ __forceinline__ static void get(int i, int& a, int& b) {
        a = i;
        b = (i == 0 || 2 * i == L2) ? i : L2 - i;
    }
};

// First pass of one row from registers (complex values loaded from Y).
template <typename Plan>
__device__ __forceinline__ void first_pass_c(c32* z, const c32 (&raw)[(Plan::L / Plan::R1 + Plan::T - 1) / Plan::T][Plan::R1], int t) {
    constexpr int L = Plan::L, T = Plan::T, R1 = Plan::R1;
    constexpr int NB1 = L / R1, M1 = (NB1 + T - 1) / T;
#pragma unroll
    for (int m = 0; m < M1; ++m) {
        const int b = t + m * T;
        if (NB1 % T == 0 || b < NB1) {
            c32 v[R1];
#pragma unroll
            for (int r = 0; r < R1; ++r) v[r] = raw[m][r];
            Dft<R1>::run(v);
#pragma unroll
            for (int q = 0; q < R1; ++q) z[sw(b * R1 + q)] = v[q];
        }
    }
}

template <typename Plan>
__device__ __forceinline__ void load_row(c32 (&raw)[(Plan::L / Plan::R1 + Plan::T - 1) / Plan::T][Plan::R1],
                                         const float2* __restrict__ src, int t) {
    constexpr int L = Plan::L, T = Plan::T, R1 = Plan::R1;
    constexpr int NB1 = L / R1, M1 = (NB1 + T - 1) / T;
#pragma unroll
    for (int m = 0; m < M1; ++m)
#pragma unroll
        for (int r = 0; r < R1; ++r) {
            const int b = t + m * T;
            if (NB1 % T == 0 || b < NB1) raw[m][r] = c_of(src[b + r * NB1]);
        }
}

// Last pass of one row in place; twiddles exp(-2 pi i r b / L1) = tw[r b L2].
template <typename Plan>
__device__ __forceinline__ void last_pass_row(c32* z, const float2* __restrict__ tw, int l2,
                                              int t) {
    constexpr int L = Plan::L, T = Plan::T, RL = Plan::RL;
    constexpr int NBL = L / RL, ML = (NBL + T - 1) / T;
#pragma unroll
    for (int m = 0; m < ML; ++m) {
        const int b = t + m * T;
        if (NBL % T == 0 || b < NBL) {
            c32 v[RL];
#pragma unroll
            for (int r = 0; r < RL; ++r) {
                const c32 a = z[sw(b + r * NBL)];
                v[r] = r ? cmul(a, c_of(tw[r * b * l2])) : a;
            }
            Dft<RL>::run(v);
#pragma unroll
            for (int q = 0; q < RL; ++q) z[sw(b + q * NBL)] = v[q];
        }
    }
}

// grid (k-blocks of the chunk, row pairs); Plan is the row plan (L = L1 <= 8192).  The
// power of bin b is accumulated at k = k2 + L2 k1 (and k = L by the row-0 workgroup).
template <typename Plan, int L2>
__global__ __launch_bounds__(Plan::T) void four_step_rows_kernel(
    const float2* __restrict__ y, const BinDev* __restrict__ bins,
    const KbWork* __restrict__ work, int kb0, const float2* __restrict__ tw,
    const float2* __restrict__ uw, double* __restrict__ power) {
    constexpr int L1 = Plan::L, T = Plan::T, R1 = Plan::R1, R2 = Plan::R2, R3 = Plan::R3;
    constexpr int L = L1 * L2;
    constexpr int M1 = (L1 / R1 + T - 1) / T;
    constexpr int MU = (L1 + T - 1) / T;
    __shared__ c32 za[L1];
    __shared__ c32 zb[L1];
    __shared__ c32 tab2[Plan::P2];
    __shared__ c32 tab3[Plan::P3];
    const int kbl = blockIdx.x;
    const KbWork w = work[kb0 + kbl];
    const BinDev bd = bins[w.bin];
    int ra, rb;
    RowPairs<L2>::get(blockIdx.y, ra, rb);
    const bool self = ra == rb;
    const int tid = threadIdx.x;
    for (int m = tid; m < Plan::P2; m += T) tab2[m] = c_of(tw[m * (L1 / Plan::P2) * L2]);
    if (Plan::NP == 4)
        for (int m = tid; m < Plan::P3; m += T) tab3[m] = c_of(tw[m * (L1 / Plan::P3) * L2]);
    double acc_a[MU], acc_b[MU];
    double acc_last = 0.0;   // k = L (row 0 only)
#pragma unroll
    for (int m = 0; m < MU; ++m) acc_a[m] = acc_b[m] = 0.0;
    const int kend = min(KB, bd.npix - w.kb * KB);
    const float2* ybase = y + (int64_t)kbl * KB * L;
    c32 raw_a[M1][R1], raw_b[M1][R1];
    load_row<Plan>(raw_a, ybase + (int64_t)ra * L1, tid);
    if (!self) load_row<Plan>(raw_b, ybase + (int64_t)rb * L1, tid);
    __syncthreads();
    for (int j = 0; j < kend; ++j) {
        int t = tid;
        asm volatile("" : "+v"(t));
        first_pass_c<Plan>(za, raw_a, t);
        if (!self) first_pass_c<Plan>(zb, raw_b, t);
        if (j + 1 < kend) {
            const float2* nxt = ybase + (int64_t)(j + 1) * L;
            load_row<Plan>(raw_a, nxt + (int64_t)ra * L1, t);
            if (!self) load_row<Plan>(raw_b, nxt + (int64_t)rb * L1, t);
        }
        __syncthreads();
        mid_pass<Plan, R2, R1, Plan::P2>(za, tab2, t);
        if (!self) mid_pass<Plan, R2, R1, Plan::P2>(zb, tab2, t);
        if (Plan::NP == 4) {
            mid_pass<Plan, R3, R1 * R2, Plan::P3>(za, tab3, t);
            if (!self) mid_pass<Plan, R3, R1 * R2, Plan::P3>(zb, tab3, t);
        }
        last_pass_row<Plan>(za, tw, L2, t);
        if (!self) last_pass_row<Plan>(zb, tw, L2, t);
        __syncthreads();
        // unpack: X[k] = (Z[k] + conj Z[L-k]) / 2 - i W^k (Z[k] - conj Z[L-k]) / 2
        const c32* zo = self ? za : zb;   // the row holding the partners of row A
#pragma unroll
        for (int m = 0; m < MU; ++m) {
            const int k1 = t + m * T;
            if (k1 < L1) {
                // row A point k1 and its partner
                {
                    const int ka = ra + L2 * k1;
                    const int pk = ra == 0 ? (L1 - k1) % L1 : L1 - 1 - k1;
                    const c32 a = za[sw(k1)];
                    const c32 c = zo[sw(pk)];
                    const c32 s2 = c32{a.x + c.x, a.y - c.y};
                    const c32 d = c32{a.x - c.x, a.y + c.y};
                    const c32 wd = cmul(c_of(uw[ka]), d);
                    const float xr = 0.5f * (s2.x + wd.y), xi = 0.5f * (s2.y - wd.x);
                    acc_a[m] += (double)xr * xr + (double)xi * xi;
                }
                if (!self) {   // row B point k1, partner in row A
                    const int kb_ = rb + L2 * k1;
                    const c32 a = zb[sw(k1)];
                    const c32 c = za[sw(L1 - 1 - k1)];
                    const c32 s2 = c32{a.x + c.x, a.y - c.y};
                    const c32 d = c32{a.x - c.x, a.y + c.y};
                    const c32 wd = cmul(c_of(uw[kb_]), d);
                    const float xr = 0.5f * (s2.x + wd.y), xi = 0.5f * (s2.y - wd.x);
                    acc_b[m] += (double)xr * xr + (double)xi * xi;
                }
            }
        }
        if (ra == 0 && t == 0) {   // k = L: Z[L] = Z[0]
            const c32 a = za[sw(0)];
            const c32 s2 = c32{a.x + a.x, 0.f};
            const c32 d = c32{0.f, a.y + a.y};
            const c32 wd = cmul(c_of(uw[L]), d);
            const float xr = 0.5f * (s2.x + wd.y), xi = 0.5f * (s2.y - wd.x);
            acc_last += (double)xr * xr + (double)xi * xi;
        }
        __syncthreads();
    }
    double* prow = power + (int64_t)w.bin * (L + 1);
#pragma unroll
    for (int m = 0; m < MU; ++m) {
        const int k1 = tid + m * T;
        if (k1 < L1) {
            if (acc_a[m] != 0.0) atomicAdd(prow + ra + L2 * k1, acc_a[m]);
            if (!self && acc_b[m] != 0.0) atomicAdd(prow + rb + L2 * k1, acc_b[m]);
        }
    }
    if (ra == 0 && tid == 0 && acc_last != 0.0) atomicAdd(prow + L, acc_last);
}

// Num(tau) from the power spectrum, FP64; grid (ceil(N / 256), n_bins).  Every
// thread reads the same ps[k] in step, so the spectrum is read as uniform loads.
// One block per (256 lags, group of INV_G bins): the group's power rows are staged in LDS
// in chunks of INV_CH values (every thread reads the same P_k, a broadcast), and each
// cosine serves the INV_G bins.  The cosines run as four independent rotation
// recurrences (k = k0 + 4 m + r, each turning by 4 theta, theta = 2 pi tau / Nf), so four
// dependent chains overlap; each chain is re-anchored with sincospi at every chunk start
// (512 rotations at most), and the four partial sums are joined in a fixed order.
constexpr int INV_CH = 2048;
constexpr int INV_G = 4;
__global__ __launch_bounds__(256) void fft_inverse_kernel(const double* __restrict__ power,
                                                          int L, int n_bins, int nframes,
                                                          double* __restrict__ num) {
    __shared__ double pk[INV_G][INV_CH];
    const int b0 = blockIdx.y * INV_G;
    const int ng = min(INV_G, n_bins - b0);
    const int tau0 = blockIdx.x * 256 + threadIdx.x;
    const int64_t tau = tau0 < nframes ? tau0 : 0;
    const double nf = 2.0 * L;
    const int64_t period = 2 * (int64_t)L;  // cos(2 pi k tau / Nf) depends on k tau mod Nf
    double s4, c4;
    sincospi(2.0 * (double)((4 * tau) % period) / nf, &s4, &c4);  // rotation by 4 theta
    double acc[INV_G][4];
#pragma unroll
    for (int g = 0; g < INV_G; ++g)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[g][r] = 0.0;
    for (int k0 = 1; k0 < L; k0 += INV_CH) {
        const int k1 = min(L, k0 + INV_CH);
        const int n = k1 - k0;
        __syncthreads();                     // the previous chunk has been read
        for (int i = threadIdx.x; i < INV_G * INV_CH; i += 256) {
            const int g = i / INV_CH, k = i % INV_CH;
            pk[g][k] = (g < ng && k < n) ? power[(int64_t)(b0 + g) * (L + 1) + k0 + k] : 0.0;
        }
        __syncthreads();
        double c[4], s[4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
            sincospi(2.0 * (double)(((int64_t)(k0 + r) * tau) % period) / nf, &s[r], &c[r]);
        // the staged rows are zero past n, so whole steps of four cover the chunk
        for (int i = 0; i < n; i += 4) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
#pragma unroll
                for (int g = 0; g < INV_G; ++g) acc[g][r] += pk[g][i + r] * c[r];
                const double cn = c[r] * c4 - s[r] * s4;
                s[r] = s[r] * c4 + c[r] * s4;
                c[r] = cn;
            }
        }
    }
    if (tau0 >= nframes) return;
#pragma unroll
    for (int g = 0; g < INV_G; ++g) {
        if (g >= ng) break;
        const double* ps = power + (int64_t)(b0 + g) * (L + 1);
        const double last = (tau & 1) ? -ps[L] : ps[L];
        num[(int64_t)(b0 + g) * nframes + tau] =
            (ps[0] + last + 2.0 * ((acc[g][0] + acc[g][1]) + (acc[g][2] + acc[g][3]))) / nf;
    }
}

// ------------------------------------------------------------------ launchers
template <typename E, bool SGN>
static void gather_pm_typed(const uint8_t* frames, int64_t hw, int nframes, int nrows,
                            const int32_t* pix_idx, const int32_t* quad_kb, const BinDev* bins,
                            const KbWork* work, int n_work, uint8_t* xs, int64_t* sums,
                            hipStream_t s) {
    constexpr int FT = 256 / PM_KPB;  // nrows is a multiple of 256
    gather_pm_kernel<E, SGN, PM_KPB>
        <<<dim3((unsigned)n_work, (unsigned)(nrows / FT)), 256, 0, s>>>(
            frames, hw, nframes, nrows, pix_idx, quad_kb, bins, work, xs, sums);
}

void launch_gather_pm(const uint8_t* frames, int dtype, int64_t hw, int nframes, int nrows,
                      const int32_t* pix_idx, const int32_t* quad_kb, const BinDev* bins,
                      int n_bins, const KbWork* work, int n_work, uint8_t* xs, int64_t* sums,
                      hipStream_t s) {
    if (n_bins <= 0) return;
    launch_fill_i64(sums, (int64_t)n_bins * nframes, s);
    if (n_work <= 0) return;
    // aligned quad loads need 4 | hw and a 4-aligned frame pointer
    if (hw % 4 != 0 || reinterpret_cast<uintptr_t>(frames) % 4 != 0) quad_kb = nullptr;
    if (dtype == DXG2_U16)
        gather_pm_typed<uint16_t, false>(frames, hw, nframes, nrows, pix_idx, nullptr, bins,
                                         work, n_work, xs, sums, s);
    else if (dtype == DXG2_I8)
        gather_pm_typed<uint8_t, true>(frames, hw, nframes, nrows, pix_idx, quad_kb, bins, work,
                                       n_work, xs, sums, s);
    else
        gather_pm_typed<uint8_t, false>(frames, hw, nframes, nrows, pix_idx, quad_kb, bins,
                                        work, n_work, xs, sums, s);
}

// Lengths with a plan: L (complex points), threads, radices.
using Plan1024 = FftPlan<1024, 64, 16, 16, 4, 1>;
using Plan1280 = FftPlan<1280, 128, 16, 16, 5, 1>;
using Plan2048 = FftPlan<2048, 128, 16, 16, 8, 1>;
using Plan2560 = FftPlan<2560, 256, 16, 4, 8, 5>;
using Plan4096 = FftPlan<4096, 256, 16, 16, 16, 1>;
using Plan5120 = FftPlan<5120, 320, 16, 16, 5, 4>;
using Plan8192 = FftPlan<8192, 512, 16, 16, 8, 4>;
using Plan10240 = FftPlan<10240, 640, 16, 16, 5, 8>;
using Plan16384 = FftPlan<16384, 1024, 16, 16, 4, 16>;

static const int kFftLengths[] = {1024, 1280, 2048, 2560, 4096, 5120, 8192, 10240, 16384};

int fft_lds_length(int n) {
    for (int L : kFftLengths)
        if (L >= n) return L;
    return 0;
}

template <typename E, typename Plan>
static void spectra_typed(const uint8_t* xs, const BinDev* bins, const KbWork* work, int n_work,
                          int nrows, const float2* tw, const float2* uw, double* power,
                          hipStream_t s) {
    fft_spectra_kernel<E, Plan><<<dim3((unsigned)n_work), Plan::T, 0, s>>>(xs, bins, work, nrows,
                                                                          tw, uw, power);
}

template <typename Plan>
static void spectra_dtype(int dtype, const uint8_t* xs, const BinDev* bins, const KbWork* work,
                          int n_work, int nrows, const float2* tw, const float2* uw,
                          double* power, hipStream_t s) {
    if (dtype == DXG2_U16)
        spectra_typed<uint16_t, Plan>(xs, bins, work, n_work, nrows, tw, uw, power, s);
    else if (dtype == DXG2_I8)
        spectra_typed<int8_t, Plan>(xs, bins, work, n_work, nrows, tw, uw, power, s);
    else
        spectra_typed<uint8_t, Plan>(xs, bins, work, n_work, nrows, tw, uw, power, s);
}

int launch_fft_power(int L, int dtype, const uint8_t* xs, const BinDev* bins, const KbWork* work,
                     int n_work, int nrows, const float2* tw, const float2* uw, double* power,
                     hipStream_t s) {
    if (n_work <= 0) return 0;
    if (nrows > L) return -1;  // the spectra kernels take frames from L on as zero
#define DXG2_FFT_CASE(LL)                                                                  \
    case LL:                                                                               \
        spectra_dtype<Plan##LL>(dtype, xs, bins, work, n_work, nrows, tw, uw, power, s);   \
        return 0;
    switch (L) {
        DXG2_FFT_CASE(1024)
        DXG2_FFT_CASE(1280)
        DXG2_FFT_CASE(2048)
        DXG2_FFT_CASE(2560)
        DXG2_FFT_CASE(4096)
        DXG2_FFT_CASE(5120)
        DXG2_FFT_CASE(8192)
        DXG2_FFT_CASE(10240)
        DXG2_FFT_CASE(16384)
        default:
            return -1;
    }
#undef DXG2_FFT_CASE
}

// Long series: L = L1 L2 with L1 a row plan (<= 8192) and L2 in {2, 4, 5, 8, 16}; the
// smallest such L >= n above 16384, the larger L1 on ties.  Returns 0 past 131072.
static const int kRowLengths[] = {8192, 5120, 4096, 2560, 2048, 1280, 1024};
static const int kColLengths[] = {2, 4, 5, 8, 16};

int fft_four_step_length(int n, int* L1, int* L2) {
    int best = 0;
    for (int l1 : kRowLengths)
        for (int l2 : kColLengths) {
            const int L = l1 * l2;
            if (L >= n && L > 16384 && (best == 0 || L < best)) {
                best = L;
                *L1 = l1;
                *L2 = l2;
            }
        }
    return best;
}

template <typename E, int L2>
static void cols_typed(const uint8_t* xs, const BinDev* bins, const KbWork* work, int kb0,
                       int n_kb, int nrows, int L1, const float2* tw, float2* y, hipStream_t s) {
    four_step_cols_kernel<E, L2><<<dim3((unsigned)(n_kb * KB), (unsigned)((L1 + 255) / 256)),
                                   256, 0, s>>>(xs, bins, work, kb0, nrows, L1, tw, y);
}

template <int L2>
static void cols_dtype(int dtype, const uint8_t* xs, const BinDev* bins, const KbWork* work,
                       int kb0, int n_kb, int nrows, int L1, const float2* tw, float2* y,
                       hipStream_t s) {
    if (dtype == DXG2_U16)
        cols_typed<uint16_t, L2>(xs, bins, work, kb0, n_kb, nrows, L1, tw, y, s);
    else if (dtype == DXG2_I8)
        cols_typed<int8_t, L2>(xs, bins, work, kb0, n_kb, nrows, L1, tw, y, s);
    else
        cols_typed<uint8_t, L2>(xs, bins, work, kb0, n_kb, nrows, L1, tw, y, s);
}

template <typename Plan, int L2>
static void rows_launch(const float2* y, const BinDev* bins, const KbWork* work, int kb0,
                        int n_kb, const float2* tw, const float2* uw, double* power,
                        hipStream_t s) {
    four_step_rows_kernel<Plan, L2><<<dim3((unsigned)n_kb, (unsigned)RowPairs<L2>::N), Plan::T,
                                      0, s>>>(y, bins, work, kb0, tw, uw, power);
}

template <int L2>
static int rows_plan(int L1, const float2* y, const BinDev* bins, const KbWork* work, int kb0,
                     int n_kb, const float2* tw, const float2* uw, double* power,
                     hipStream_t s) {
    switch (L1) {
        case 1024: rows_launch<Plan1024, L2>(y, bins, work, kb0, n_kb, tw, uw, power, s); return 0;
        case 1280: rows_launch<Plan1280, L2>(y, bins, work, kb0, n_kb, tw, uw, power, s); return 0;
        case 2048: rows_launch<Plan2048, L2>(y, bins, work, kb0, n_kb, tw, uw, power, s); return 0;
        case 2560: rows_launch<Plan2560, L2>(y, bins, work, kb0, n_kb, tw, uw, power, s); return 0;
        case 4096: rows_launch<Plan4096, L2>(y, bins, work, kb0, n_kb, tw, uw, power, s); return 0;
        case 5120: rows_launch<Plan5120, L2>(y, bins, work, kb0, n_kb, tw, uw, power, s); return 0;
        case 8192: rows_launch<Plan8192, L2>(y, bins, work, kb0, n_kb, tw, uw, power, s); return 0;
        default: return -1;
    }
}

// One chunk of k-blocks [kb0, kb0 + n_kb) through the scratch y (n_kb * 64 * L complex).
int launch_fft_four_step(int L1, int L2, int dtype, const uint8_t* xs, const BinDev* bins,
                         const KbWork* work, int kb0, int n_kb, int nrows, const float2* tw,
                         const float2* uw, float2* y, double* power, hipStream_t s) {
    if (n_kb <= 0) return 0;
    switch (L2) {
        case 2: cols_dtype<2>(dtype, xs, bins, work, kb0, n_kb, nrows, L1, tw, y, s);
                return rows_plan<2>(L1, y, bins, work, kb0, n_kb, tw, uw, power, s);
        case 4: cols_dtype<4>(dtype, xs, bins, work, kb0, n_kb, nrows, L1, tw, y, s);
                return rows_plan<4>(L1, y, bins, work, kb0, n_kb, tw, uw, power, s);
        case 5: cols_dtype<5>(dtype, xs, bins, work, kb0, n_kb, nrows, L1, tw, y, s);
                return rows_plan<5>(L1, y, bins, work, kb0, n_kb, tw, uw, power, s);
        case 8: cols_dtype<8>(dtype, xs, bins, work, kb0, n_kb, nrows, L1, tw, y, s);
                return rows_plan<8>(L1, y, bins, work, kb0, n_kb, tw, uw, power, s);
        case 16: cols_dtype<16>(dtype, xs, bins, work, kb0, n_kb, nrows, L1, tw, y, s);
                 return rows_plan<16>(L1, y, bins, work, kb0, n_kb, tw, uw, power, s);
        default: return -1;
    }
}

void launch_fft_inverse(const double* power, int L, int n_bins, int nframes, double* num,
                        hipStream_t s) {
    if (n_bins <= 0) return;
    fft_inverse_kernel<<<dim3((unsigned)((nframes + 255) / 256),
                              (unsigned)((n_bins + INV_G - 1) / INV_G)),
                         256, 0, s>>>(power, L, n_bins, nframes, num);
}

}  // namespace dxg2
