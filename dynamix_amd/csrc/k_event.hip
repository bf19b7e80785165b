// Event (sparse) correlator kernels.
//
// Reference: EventCorrelator (correlator/event.py:162-204) with event_correlator
// (resources/opencl/evtcorrelator.cl:3-46): one work-item per pixel walks every ordered
// pair j <= i of that pixel's frame-sorted events and does a global atomic
// res[bin][t_i - t_j] += d_i d_j, plus sums[bin][t_i] += d_i.
//
// Here:
//  * work items are (bin, range of the bin's pixel list) x (lag window);
//    a wave takes one pixel at a time (the next pixel's events load into registers
//    meanwhile), stages its events in LDS, and each lane takes
//    events i = lane + 128 r and i + 64 together; a binary search finds the first
//    partner j > i whose lag t_j - t_i reaches the window, and the pairs up to the
//    window's end go to an LDS lag histogram (EvHist below), flushed once per work
//    item with one global 64-bit atomic per non-zero lag;
//  * S_b(t) and the self pairs (lag 0, d_i^2) are added in the first window;
//  * the input is validated on the device (frames in [0, N), strictly increasing per
//    pixel, as compress_all_stack produces, event.py:267-284); a violation sets a
//    flag the host turns into an error instead of reading out of bounds.
// Frame-sorted COO input (pixel, frame, count) is first turned into this pixel-major
// layout on the device: the (frame, count) pairs, packed into one word, are ordered by
// pixel with this file's stable LSD radix sort (8-bit digits over the bits a pixel
// index needs), so each pixel's events keep the input's frame order; the offsets are
// the lower bounds of every pixel in the sorted keys, and the packed words are split
// into times and counts.
#include "dxg2_internal.h"

namespace dxg2 {

constexpr int ECAP = 256;       // staging entries of one wave in LDS
constexpr int EMAX = ECAP - 4;  // events of a staged pixel: its 4 sentinels always fit
constexpr int EPF = ECAP / 64;  // event rounds of the next pixel prefetched in registers

// LDS lag histogram of one work item, shared by the NW waves of its workgroup: 8 KiB of
// counters per wave, so that with NW = 4 four workgroups (16 waves, enough to cover the
// dependent LDS reads of the searches) share a CU with their staging, and with NW = 8
// two workgroups of twice the lag window do.
//  * 8-bit counts: 2048 NW lags of 32-bit counters.  A pixel adds at most sum_k d_k^2 to
//    any one lag (Cauchy-Schwarz: at a fixed lag each event pairs with at most one
//    partner), so each wave adds to the histogram only while its pixels' sum of d^2
//    stays below budget(NW) = 2^32 / NW (uint8) or 2^31 / NW (int8), and sends the pairs
//    of a pixel beyond it straight to the global 64-bit numerators.  The NW waves then
//    bound every counter below 2^32 (uint8) or its magnitude below 2^31 (int8, two's
//    complement), so it is exact.
//  * uint16 counts: 1024 NW lags of 64-bit counters, exact for any data.
template <typename T> struct EvHist {
    using type = uint32_t;
    static constexpr int lags_per_wave = 2048;
    // lags of one window: the last 64 counters stay outside every window, so the staged
    // partner loop can send each rejected partner's product to its own lane's counter there
    static constexpr int window(int nw) { return lags_per_wave * nw - 64; }
    static constexpr long long budget(int nw) {
        return (std::is_signed<T>::value ? (1ll << 31) : (1ll << 32)) / nw;
    }
    static __device__ __forceinline__ unsigned long long widen(uint32_t v) {
        return std::is_signed<T>::value ? (unsigned long long)(long long)(int32_t)v
                                        : (unsigned long long)v;
    }
};
template <> struct EvHist<uint16_t> {
    using type = unsigned long long;
    static constexpr int lags_per_wave = 1024;
    static constexpr int window(int nw) { return lags_per_wave * nw; }
    static constexpr long long budget(int) { return 0; }  // no bound needed
    static __device__ __forceinline__ unsigned long long widen(unsigned long long v) { return v; }
};

template <typename T>
__device__ __forceinline__ long long as_count(T v) { return (long long)v; }

__device__ __forceinline__ long long wave_sum64(long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Per-pixel bookkeeping of one wave: pixel index and event range, valid or not.
struct EvPixel {
    int p;
    uint32_t beg, end;
    int n;  // events to correlate: 0 for an empty, unbinned or invalid range
};

// The pairs of one work item (a run of one bin's pixels) in lag window blockIdx.y, with
// the workgroup's LDS: the lag histogram and each wave's staged events (uint8, int8 and
// uint16 counts fit 32 bits), times and counts in one array st[2][NW][ECAP] so that the
// two reads of a step share one address.
template <typename T, int NW>
__device__ __forceinline__ void event_pairs_item(
    const int32_t* __restrict__ times, const T* __restrict__ data,
    const uint32_t* __restrict__ offsets, const int32_t* __restrict__ pix_idx,
    const BinDev* __restrict__ bins, const EventWork w, int nframes, int64_t nnz,
    int64_t* __restrict__ num, int64_t* __restrict__ sums, int* __restrict__ err,
    typename EvHist<T>::type* __restrict__ hist, int (*__restrict__ st)[NW][ECAP]) {
    int (*const st_t)[ECAP] = st[0];
    int (*const st_d)[ECAP] = st[1];
    using H = typename EvHist<T>::type;
    constexpr int EW = EvHist<T>::lags_per_wave * NW;  // counters
    constexpr int WS = EvHist<T>::window(NW);          // lags of a window
    constexpr int NT = 64 * NW;
    constexpr bool BOUNDED = sizeof(H) == 4;
    const BinDev bd = bins[w.bin];
    const int w0 = blockIdx.y * WS;
    if (w0 >= nframes) return;
    const int w1 = min(nframes, w0 + WS);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int k = threadIdx.x; k < EW; k += NT) hist[k] = 0;
    __syncthreads();
    int64_t* srow = sums + (int64_t)w.bin * nframes;
    unsigned long long* nwin = reinterpret_cast<unsigned long long*>(num + (int64_t)w.bin * nframes + w0);
    long long used = 0;  // this wave's sum of d^2 added to the histogram (BOUNDED)

    // The wave takes pixels w.k0 + wave + NW r, r < nr.  Their list entries and offsets are
    // loaded 64 at a time (lane r % 64 holds pixel r), and while the pairs of pixel r are
    // counted, the first ECAP events of pixel r + 1 are already loading into registers,
    // so the chain list entry -> offsets -> events is not waited for once per pixel.
    const int npx = w.k1 - w.k0 - wave;
    const int nr = npx > 0 ? (npx + NW - 1) / NW : 0;
    int mp = -1;
    uint32_t mb = 0u, me = 0u;
    auto load_meta = [&](int r0) {
        const int r = r0 + lane;
        mp = -1;
        mb = me = 0u;
        if (r < nr) {
            mp = pix_idx[bd.idx_off + w.k0 + wave + NW * r];
            if (mp >= 0) {
                mb = offsets[mp];
                me = offsets[mp + 1];
            }
        }
    };
    auto meta = [&](int r) {
        EvPixel x;
        x.p = __builtin_amdgcn_readlane(mp, r & 63);
        x.beg = (uint32_t)__builtin_amdgcn_readlane((int)mb, r & 63);
        x.end = (uint32_t)__builtin_amdgcn_readlane((int)me, r & 63);
        const bool bad = x.p >= 0 && (x.end < x.beg || x.end > (uint64_t)nnz);
        x.n = (x.p < 0 || bad || x.end <= x.beg) ? 0 : (int)(x.end - x.beg);
        if (bad && lane == 0) atomicOr(err, 2);
        return x;
    };
    int ct[EPF];
    int cd[EPF];
    auto prefetch = [&](const EvPixel& x) {
#pragma unroll
        for (int q = 0; q < EPF; ++q) {
            const int k = lane + 64 * q;
            if (k < x.n && x.n <= EMAX) {
                ct[q] = times[x.beg + k];
                cd[q] = (int)data[x.beg + k];
            }
        }
    };

    EvPixel cur = {-1, 0u, 0u, 0};
    if (nr > 0) {
        load_meta(0);
        cur = meta(0);
        prefetch(cur);
    }
    for (int r = 0; r < nr; ++r) {
        // wave-uniform (the compiler keeps them, and the loops they bound, scalar)
        const int n = __builtin_amdgcn_readfirstlane(cur.n);
        const uint32_t beg = (uint32_t)__builtin_amdgcn_readfirstlane((int)cur.beg);
        const bool staged = n <= EMAX;
        long long sq = 0;  // this lane's share of the pixel's sum of d^2
        if (n > 0 && staged) {
#pragma unroll
            for (int q = 0; q < EPF; ++q) {
                const int k = lane + 64 * q;
                if (k < n) {
                    st_t[wave][k] = ct[q];
                    st_d[wave][k] = cd[q];
                    sq += (long long)cd[q] * cd[q];
                } else if (k < n + 4) {
                    st_t[wave][k] = 0x7fffffff;  // sentinel time: no partner lies there
                    st_d[wave][k] = 0;
                }
            }
        }
        // the next pixel's range, and its events into the registers just staged
        if (r + 1 < nr) {
            if (((r + 1) & 63) == 0) load_meta(r + 1);
            cur = meta(r + 1);
            prefetch(cur);
        }
        if (n == 0) continue;
        __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): staged values visible to the wave
        __builtin_amdgcn_wave_barrier();
        // The pixel's pairs, with its events read from LDS (staged) or from HBM; the two
        // are separate instantiations so that each reads through its own address space.
        auto pixel = [&](auto staged_tag) {
            constexpr bool S = decltype(staged_tag)::value;
            auto time_of = [&](int k) -> int {
                if constexpr (S) return st_t[wave][k];
                else return times[beg + k];
            };
            // 8-bit counts multiply in 32 bits (|d_i d_j| <= 255^2), 16-bit in 64
            using C = typename std::conditional<sizeof(T) == 1, int, long long>::type;
            auto count_of = [&](int k) -> C {
                if constexpr (S) return (C)st_d[wave][k];
                else return (C)as_count<T>(data[beg + k]);
            };
            // Validation, in the first lag window (the one that also writes the frame
            // sums): frames in [0, N) and strictly increasing.  An invalid pixel is
            // skipped there and flagged, and the call reports the error; the other
            // windows do not repeat the check, and their lag test below keeps every
            // write inside the window whatever the input.
            if (w0 == 0) {
                int bad = 0;
                for (int k = lane; k < n; k += 64) {
                    const int t = time_of(k);
                    if (t < 0 || t >= nframes) bad = 1;
                    if (k + 1 < n && time_of(k + 1) <= t) bad = 1;
                }
                if (bad) atomicOr(err, 1);
                if (__any(bad)) return;
            }
            long long sq_lane = sq;
            if constexpr (!S)
                for (int k = lane; k < n; k += 64) {
                    const long long d = count_of(k);
                    sq_lane += d * d;
                }
            const long long sqw = wave_sum64(sq_lane);  // sum of d^2: the lag-0 term
            bool to_lds = true;
            if constexpr (BOUNDED) {
                to_lds = used + sqw < EvHist<T>::budget(NW);
                if (to_lds) used += sqw;
            }
            if (w0 == 0 && lane == 0 && sqw != 0) {  // lag 0
                if (to_lds)
                    atomicAdd(&hist[0], (H)sqw);
                else
                    atomicAdd(nwin, (unsigned long long)sqw);
            }
            const unsigned span = (unsigned)(w1 - w0);
            // the rest in one of two instantiations: LDS counters, or (past the budget)
            // the global 64-bit numerators, so the pair loop carries no branch on it
            auto events = [&](auto lds_tag) {
            constexpr bool LDS = decltype(lds_tag)::value;
            auto add = [&](int lag, C prod) {  // lag in [0, w1 - w0)
                if constexpr (LDS)
                    atomicAdd(&hist[lag], (H)prod);
                else
                    atomicAdd(nwin + lag, (unsigned long long)(long long)prod);
            };
            // pairs (i, j), j > i, with lag t_j - t_i in [max(w0, 1), w1): partners from
            // the first j with t_j >= t_i + max(w0, 1) while t_j < t_i + w1, four per
            // step so that their reads and atomics overlap.  The lag test is one unsigned
            // compare, which also rejects a negative lag.
            auto partners = [&](int ti, C di, int j) {
                const int base = ti + w0;
                if constexpr (S && LDS && sizeof(T) == 1) {
                    // Staged 8-bit counts into the LDS histogram: the four partners of a
                    // step come from one aligned 16-byte read of each staged array, from
                    // j & ~3 (j <= n <= EMAX, so the reads stay inside the wave's n + 4
                    // staged entries, the sentinels included, and a step that reaches
                    // them ends the loop: no test of j against n).  With lo = max(w0, 1), a partner has t - (t_i + lo) in
                    // [0, span1); everything else lies outside as an unsigned value:
                    // entries before j lie below t_i + lo (times strictly increase; in the
                    // first window entry i itself sits at lag 0 < lo), and the entries
                    // from n on hold the staging's sentinel time INT_MAX.  The counter
                    // index min(t - (t_i + lo), span1 + lane) sends every non-partner to
                    // a counter past the window (the 64 that no window covers), so every
                    // slot adds without a test or a branch.  Products of two 8-bit counts
                    // fit 24 bits.
                    const int lo = w0 > 0 ? w0 : 1;
                    const int base1 = ti + lo;
                    const unsigned span1 = span - (unsigned)(lo - w0);
                    const unsigned park = span1 + (unsigned)lane;
                    H* const h1 = hist + (lo - w0);
                    const int4* t4p = reinterpret_cast<const int4*>(st_t[wave]);
                    const int4* d4p = reinterpret_cast<const int4*>(st_d[wave]);
                    for (int j4 = j & ~3;; j4 += 4) {
                        const int4 t4 = t4p[j4 >> 2];
                        const int4 d4 = d4p[j4 >> 2];
                        const int tt[4] = {t4.x, t4.y, t4.z, t4.w};
                        const int dd[4] = {d4.x, d4.y, d4.z, d4.w};
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const unsigned lag = min((unsigned)(tt[u] - base1), park);
                            const int prod = std::is_signed<T>::value
                                                 ? __mul24((int)di, dd[u])
                                                 : (int)__umul24((unsigned)di, (unsigned)dd[u]);
                            atomicAdd(&h1[lag], (H)prod);
                        }
                        if (tt[3] - base1 >= (int)span1) break;  // INT_MAX past n
                    }
                    return;
                }
                for (; j < n; j += 4) {
                    int tj[4];
                    C dj[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const bool in = j + u < n;
                        const int jj = in ? j + u : n - 1;
                        const int t = time_of(jj);
                        const C d = count_of(jj);
                        tj[u] = in ? t : base + (int)span;
                        dj[u] = in ? d : (C)0;
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const unsigned lag = (unsigned)(tj[u] - base);
                        const C prod = di * dj[u];
                        if (lag < span && prod != 0) add((int)lag, prod);
                    }
                    if (tj[3] - base >= (int)span) break;  // times increase: the rest are later
                }
            };
            // wave-uniform steps of the first-partner search: powers of two from the
            // largest not above n
            const int top = 1 << (31 - __builtin_clz(n));
            // Each lane takes events i and i + 64 together, so that their two searches
            // for the first partner overlap.
            for (int i0 = lane; i0 < n; i0 += 128) {
                const int i1 = i0 + 64;
                const bool has1 = i1 < n;
                const int ti0 = time_of(i0);
                const int ti1 = has1 ? time_of(i1) : 0;
                const C d0 = count_of(i0);
                const C d1 = has1 ? count_of(i1) : (C)0;
                if (w0 == 0) {
                    if (d0 != 0)
                        atomicAdd(reinterpret_cast<unsigned long long*>(srow + ti0),
                                  (unsigned long long)(long long)d0);
                    if (d1 != 0)
                        atomicAdd(reinterpret_cast<unsigned long long*>(srow + ti1),
                                  (unsigned long long)(long long)d1);
                }
                // first partner: i + 1 in the first window (times strictly increase),
                // else one past the last p >= i with t_p < t_i + w0, by a descent over
                // power-of-two steps (a monotone test; reads clamped inside the list)
                int p0 = i0, p1 = has1 ? i1 : n - 1;
                if (w0 > 0) {
                    const int g0 = ti0 + w0, g1 = ti1 + w0;
                    for (int step = top; step > 0; step >>= 1) {
                        const int q0 = p0 + step, q1 = p1 + step;
                        const int t0 = time_of(min(q0, n - 1));
                        const int t1 = time_of(min(q1, n - 1));
                        if (q0 < n && t0 < g0) p0 = q0;
                        if (q1 < n && t1 < g1) p1 = q1;
                    }
                }
                partners(ti0, d0, p0 + 1);
                if (has1) partners(ti1, d1, p1 + 1);
            }
            };
            if (to_lds)
                events(std::true_type{});
            else
                events(std::false_type{});
        };
        if (staged)
            pixel(std::true_type{});
        else
            pixel(std::false_type{});
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    for (int k = threadIdx.x; k < w1 - w0; k += NT) {
        const H v = hist[k];
        if (v != 0) atomicAdd(nwin + k, EvHist<T>::widen(v));
    }
}

template <typename T, int NW>
__global__ __launch_bounds__(64 * NW) void event_pairs_kernel(
    const int32_t* __restrict__ times, const T* __restrict__ data,
    const uint32_t* __restrict__ offsets, const int32_t* __restrict__ pix_idx,
    const BinDev* __restrict__ bins, const EventWork* __restrict__ work, int nframes,
    int64_t nnz, int64_t* __restrict__ num, int64_t* __restrict__ sums, int* __restrict__ err) {
    __shared__ typename EvHist<T>::type hist[EvHist<T>::lags_per_wave * NW];
    __shared__ __attribute__((aligned(16))) int st[2][NW][ECAP];
    event_pairs_item<T, NW>(times, data, offsets, pix_idx, bins, work[blockIdx.x], nframes, nnz,
                            num, sums, err, hist, st);
}

// A pointer read from a job table, marked as global memory (a generic pointer would
// turn every access into a flat one).
template <typename P>
__device__ __forceinline__ P* global_ptr(P* p) {
    using G = __attribute__((address_space(1))) P*;
    return (P*)reinterpret_cast<G>(reinterpret_cast<uintptr_t>(p));
}

// Several jobs in one launch: work item w belongs to job w.pad (padding items: bin < 0);
// lag windows past a job's frame count return at once.
template <typename T, int NW>
__global__ __launch_bounds__(64 * NW) void event_pairs_jobs_kernel(
    const EventJob* __restrict__ jobs, const EventWork* __restrict__ work) {
    __shared__ typename EvHist<T>::type hist[EvHist<T>::lags_per_wave * NW];
    __shared__ __attribute__((aligned(16))) int st[2][NW][ECAP];
    const EventWork w = work[blockIdx.x];
    if (w.bin < 0) return;
    const EventJob j = jobs[w.pad];
    event_pairs_item<T, NW>(global_ptr(j.times), global_ptr(reinterpret_cast<const T*>(j.data)),
                            global_ptr(j.offsets), global_ptr(j.pix), global_ptr(j.bins), w,
                            j.nframes, j.nnz, global_ptr(j.num), global_ptr(j.sums),
                            global_ptr(j.err), hist, st);
}

// Waves per workgroup of the pair kernels: 16 (one workgroup per CU, 32,768 counters and
// windows of 32,704 lags for 8-bit counts: the fewest windows, so the least per-window
// work per pixel), or DXG2_EVENT_WAVES = 8 (two per CU, 16,384 counters) or 4 (four per
// CU, 8,192 counters).
// Measured on the events workload (1030x1065 x 100,000 frames): pairs 8.4 ms at 16
// waves, 9.8 ms at 8, 13.2 ms at 4 (largest work items each).
int event_waves() {
    static const int nw = [] {
        const char* v = getenv("DXG2_EVENT_WAVES");
        const int x = v ? atoi(v) : 16;
        return (x == 4 || x == 8) ? x : 16;
    }();
    return nw;
}

int event_lag_window(int dtype) {
    return dtype == DXG2_U16 ? EvHist<uint16_t>::window(event_waves())
                             : EvHist<uint8_t>::window(event_waves());
}

template <typename T, int NW>
void pairs_jobs_launch(const EventJob* jobs, const EventWork* work, int n_work, int max_frames,
                       hipStream_t s) {
    constexpr int WS = EvHist<T>::window(NW);
    event_pairs_jobs_kernel<T, NW><<<dim3((unsigned)n_work, (unsigned)((max_frames + WS - 1) / WS)),
                                     64 * NW, 0, s>>>(jobs, work);
}

template <typename T>
void pairs_jobs_typed(const EventJob* jobs, const EventWork* work, int n_work, int max_frames,
                      hipStream_t s) {
    if (event_waves() == 4)
        pairs_jobs_launch<T, 4>(jobs, work, n_work, max_frames, s);
    else if (event_waves() == 16)
        pairs_jobs_launch<T, 16>(jobs, work, n_work, max_frames, s);
    else
        pairs_jobs_launch<T, 8>(jobs, work, n_work, max_frames, s);
}

void launch_event_pairs_jobs(int dtype, const EventJob* jobs, const EventWork* work, int n_work,
                             int max_frames, hipStream_t s) {
    if (n_work <= 0 || max_frames <= 0) return;
    if (dtype == DXG2_U16)
        pairs_jobs_typed<uint16_t>(jobs, work, n_work, max_frames, s);
    else if (dtype == DXG2_I8)
        pairs_jobs_typed<int8_t>(jobs, work, n_work, max_frames, s);
    else
        pairs_jobs_typed<uint8_t>(jobs, work, n_work, max_frames, s);
}

template <typename T, int NW>
void pairs_launch(const int32_t* times, const void* data, const uint32_t* offsets,
                  const int32_t* pix_idx, const BinDev* bins, const EventWork* work, int n_work,
                  int nframes, int64_t nnz, int64_t* num, int64_t* sums, int* err, hipStream_t s) {
    constexpr int WS = EvHist<T>::window(NW);
    event_pairs_kernel<T, NW><<<dim3((unsigned)n_work, (unsigned)((nframes + WS - 1) / WS)),
                                64 * NW, 0, s>>>(times, (const T*)data, offsets, pix_idx, bins,
                                                 work, nframes, nnz, num, sums, err);
}

template <typename T>
void pairs_typed(const int32_t* times, const void* data, const uint32_t* offsets,
                 const int32_t* pix_idx, const BinDev* bins, const EventWork* work, int n_work,
                 int nframes, int64_t nnz, int64_t* num, int64_t* sums, int* err, hipStream_t s) {
    if (event_waves() == 4)
        pairs_launch<T, 4>(times, data, offsets, pix_idx, bins, work, n_work, nframes, nnz, num,
                           sums, err, s);
    else if (event_waves() == 16)
        pairs_launch<T, 16>(times, data, offsets, pix_idx, bins, work, n_work, nframes, nnz, num,
                            sums, err, s);
    else
        pairs_launch<T, 8>(times, data, offsets, pix_idx, bins, work, n_work, nframes, nnz, num,
                           sums, err, s);
}

void launch_event_pairs(int dtype, const int32_t* times, const void* data,
                        const uint32_t* offsets, const int32_t* pix_idx, const BinDev* bins,
                        const EventWork* work, int n_work, int nframes, int64_t nnz,
                        int64_t* num, int64_t* sums, int* err, hipStream_t s) {
    if (n_work <= 0 || nframes <= 0) return;
    if (dtype == DXG2_U16)
        pairs_typed<uint16_t>(times, data, offsets, pix_idx, bins, work, n_work, nframes, nnz,
                              num, sums, err, s);
    else if (dtype == DXG2_I8)
        pairs_typed<int8_t>(times, data, offsets, pix_idx, bins, work, n_work, nframes, nnz, num,
                            sums, err, s);
    else
        pairs_typed<uint8_t>(times, data, offsets, pix_idx, bins, work, n_work, nframes, nnz,
                             num, sums, err, s);
}

// LDS atomic-add throughput, the ceiling the pair kernel's histogram updates are measured
// against: workgroups of 256 threads with that kernel's LDS footprint (8192 32-bit
// counters and 8 KiB of staging, four workgroups per CU), each lane adding 1 `iters`
// times at pseudo-random counters (mode 0; a linear congruential sequence per lane) or
// at lane-consecutive ones (mode 1).
template <int MODE>
__global__ __launch_bounds__(256) void lds_atomic_rate_kernel(uint32_t* __restrict__ out,
                                                              int iters) {
    __shared__ uint32_t h[(EvHist<uint8_t>::lags_per_wave * 4)];
    __shared__ int stage_share[2 * 4 * ECAP];  // the staging arrays of four waves
    for (int i = threadIdx.x; i < (EvHist<uint8_t>::lags_per_wave * 4); i += 256) h[i] = 0;
    if (threadIdx.x < 8) stage_share[threadIdx.x] = 0;
    __syncthreads();
    uint32_t x = blockIdx.x * 2654435761u + threadIdx.x * 40503u + 1u;
    for (int i = 0; i < iters; ++i) {
        uint32_t a;
        if (MODE == 0) {
            x = x * 1664525u + 1013904223u;
            a = x >> 19;  // 0 .. 8191
        } else {
            a = (threadIdx.x + 64u * (uint32_t)i) & ((EvHist<uint8_t>::lags_per_wave * 4) - 1);
        }
        atomicAdd(&h[a], 1u);
    }
    __syncthreads();
    if (threadIdx.x == 0)
        out[blockIdx.x] = h[blockIdx.x & ((EvHist<uint8_t>::lags_per_wave * 4) - 1)] + (uint32_t)stage_share[0];
}

void launch_lds_atomic_rate(int mode, int blocks, int iters, uint32_t* out, hipStream_t s) {
    if (mode == 0)
        lds_atomic_rate_kernel<0><<<blocks, 256, 0, s>>>(out, iters);
    else
        lds_atomic_rate_kernel<1><<<blocks, 256, 0, s>>>(out, iters);
}

// ----------------------------------------------------------------- COO -> CSR
__global__ void fill_u32_kernel(uint32_t* __restrict__ p, int64_t n, uint32_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void coo_count_kernel(const uint32_t* __restrict__ pix, int64_t nnz, uint32_t npix,
                                 uint32_t* __restrict__ count, int* __restrict__ err) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nnz;
         e += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t p = pix[e];
        if (p >= npix) {
            atomicOr(err, 4);
            continue;
        }
        atomicAdd(&count[p], 1u);
    }
}

// Exclusive scan of n counts into out[0..n] (out[n] = total), three passes over
// blocks of 1024: per-block totals, a scan of the totals, then the block scans.
constexpr int SCAN_B = 1024;

__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* sh, uint32_t* total) {
    // 256 threads x 4 values handled by the caller; here one value per thread
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const uint32_t x = tid >= off ? sh[tid - off] : 0u;
        __syncthreads();
        sh[tid] += x;
        __syncthreads();
    }
    const uint32_t incl = sh[tid];
    *total = sh[255];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(256) void scan_block_totals(const uint32_t* __restrict__ in,
                                                         int64_t n, uint32_t* __restrict__ tot) {
    __shared__ uint32_t sh[256];
    const int64_t base = (int64_t)blockIdx.x * SCAN_B + threadIdx.x * 4;
    uint32_t s = 0;
    for (int k = 0; k < 4; ++k)
        if (base + k < n) s += in[base + k];
    uint32_t total;
    block_excl_scan(s, sh, &total);
    if (threadIdx.x == 0) tot[blockIdx.x] = total;
}

__global__ __launch_bounds__(256) void scan_totals(uint32_t* __restrict__ tot, int nb) {
    __shared__ uint32_t sh[256];
    uint32_t carry = 0;
    for (int b0 = 0; b0 < nb; b0 += 256) {
        const int b = b0 + threadIdx.x;
        const uint32_t v = b < nb ? tot[b] : 0u;
        uint32_t total;
        const uint32_t ex = block_excl_scan(v, sh, &total);
        if (b < nb) tot[b] = carry + ex;
        carry += total;
    }
}

__global__ __launch_bounds__(256) void scan_apply(const uint32_t* __restrict__ in, int64_t n,
                                                  const uint32_t* __restrict__ tot,
                                                  uint32_t* __restrict__ out) {
    __shared__ uint32_t sh[256];
    const int64_t base = (int64_t)blockIdx.x * SCAN_B + threadIdx.x * 4;
    uint32_t v[4], s = 0;
    for (int k = 0; k < 4; ++k) {
        v[k] = base + k < n ? in[base + k] : 0u;
        s += v[k];
    }
    uint32_t total;
    uint32_t ex = block_excl_scan(s, sh, &total) + tot[blockIdx.x];
    for (int k = 0; k < 4; ++k) {
        if (base + k < n) out[base + k] = ex;
        ex += v[k];
    }
    if (base <= n - 1 && n - 1 < base + 4) out[n] = ex;  // grand total
}

__global__ void copy_u32_kernel(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst,
                                int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// (frame, count) of event e packed as frame << CB | count; a frame past the series sets
// flag bit 0 (the packing keeps CB bits of count and the rest for the frame).
template <typename V, typename T, int CB>
__global__ void coo_pack_kernel(const uint32_t* __restrict__ pix, const uint32_t* __restrict__ frame,
                                const T* __restrict__ cnt, int64_t nnz, uint32_t npix, int nframes,
                                V* __restrict__ out, int* __restrict__ err) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nnz;
         e += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t f = frame[e];
        if (f >= (uint32_t)nframes) atomicOr(err, 1);
        if (pix[e] >= npix) atomicOr(err, 4);
        out[e] = ((V)f << CB) | (V)cnt[e];
    }
}

template <typename V, typename T, int CB>
__global__ void coo_unpack_kernel(const V* __restrict__ in, int64_t nnz,
                                  int32_t* __restrict__ times, T* __restrict__ data) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nnz;
         e += (int64_t)gridDim.x * blockDim.x) {
        const V v = in[e];
        times[e] = (int32_t)(uint32_t)(v >> CB);
        data[e] = (T)(v & (((V)1 << CB) - 1));
    }
}

// Key bits of pixel indices below npix.
static int pixel_key_bits(uint32_t npix) {
    int b = 1;
    while (b < 32 && ((uint64_t)1 << b) < npix) ++b;
    return b;
}

static unsigned grid_for(int64_t n, int64_t cap = 65536);

// First index in times[b, e) (ascending) whose frame is >= t.
__device__ __forceinline__ uint32_t lower_frame(const int32_t* __restrict__ times, uint32_t b,
                                                uint32_t e, int t) {
    while (b < e) {
        const uint32_t m = b + (e - b) / 2;
        if (times[m] < t)
            b = m + 1;
        else
            e = m;
    }
    return b;
}

// Frame window [t0, t1) of a pixel-major event list: per-pixel counts, then (after a scan
// of the counts into new offsets) the copy with frames re-based to t - t0.
__global__ void window_count_kernel(const uint32_t* __restrict__ offsets,
                                    const int32_t* __restrict__ times, int64_t npix, int t0,
                                    int t1, uint32_t* __restrict__ count) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < npix;
         p += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t b = offsets[p], e = offsets[p + 1];
        count[p] = lower_frame(times, b, e, t1) - lower_frame(times, b, e, t0);
    }
}

template <typename T>
__global__ void window_copy_kernel(const uint32_t* __restrict__ offsets,
                                   const int32_t* __restrict__ times, const T* __restrict__ data,
                                   int64_t npix, int t0, const uint32_t* __restrict__ out_off,
                                   int32_t* __restrict__ out_times, T* __restrict__ out_data) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < npix;
         p += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t lo = lower_frame(times, offsets[p], offsets[p + 1], t0);
        const uint32_t o = out_off[p], n = out_off[p + 1] - o;
        for (uint32_t i = 0; i < n; ++i) {
            out_times[o + i] = times[lo + i] - t0;
            out_data[o + i] = data[lo + i];
        }
    }
}

void launch_events_window(int dtype, const uint32_t* offsets, const int32_t* times,
                          const void* data, int64_t npix, int t0, int t1, uint32_t* scratch,
                          uint32_t* out_offsets, int32_t* out_times, void* out_data,
                          hipStream_t s) {
    if (npix <= 0) return;
    window_count_kernel<<<grid_for(npix), 256, 0, s>>>(offsets, times, npix, t0, t1, out_offsets);
    launch_excl_scan(out_offsets, npix, scratch, out_offsets, s);  // in place
    if (dtype == DXG2_U16)
        window_copy_kernel<uint16_t><<<grid_for(npix), 256, 0, s>>>(
            offsets, times, (const uint16_t*)data, npix, t0, out_offsets, out_times,
            (uint16_t*)out_data);
    else
        window_copy_kernel<uint8_t><<<grid_for(npix), 256, 0, s>>>(
            offsets, times, (const uint8_t*)data, npix, t0, out_offsets, out_times,
            (uint8_t*)out_data);
}

// Events of listed pixels -> dense columns: out[t * n_cols + k] = the count of pixel
// cols[k] at frame t (zero where it has no event).  One thread per listed pixel walks its
// events; a frame outside [0, nframes) sets bit 0 of *err and is skipped, a listed pixel
// outside [0, npix) sets bit 1.  The caller's buffer is cleared first by fill_bytes_kernel.
__global__ void fill_bytes_kernel(uint32_t* __restrict__ p, int64_t words) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < words;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

__global__ void fill_tail_kernel(uint8_t* __restrict__ p, int64_t from, int64_t to) {
    const int64_t i = from + threadIdx.x;
    if (i < to) p[i] = 0;
}

template <typename T>
__global__ void events_to_columns_kernel(const uint32_t* __restrict__ offsets,
                                         const int32_t* __restrict__ times,
                                         const T* __restrict__ data, int64_t npix,
                                         const int32_t* __restrict__ cols, int64_t n_cols,
                                         int nframes, T* __restrict__ out, int* __restrict__ err) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n_cols;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int32_t p = cols[k];
        if (p < 0 || p >= npix) {
            if (err) atomicOr(err, 2);
            continue;
        }
        const uint32_t e1 = offsets[p + 1];
        for (uint32_t i = offsets[p]; i < e1; ++i) {
            const int32_t t = times[i];
            if (t < 0 || t >= nframes) {
                if (err) atomicOr(err, 1);
                continue;
            }
            out[(int64_t)t * n_cols + k] = data[i];
        }
    }
}

void launch_events_to_columns(int dtype, const uint32_t* offsets, const int32_t* times,
                              const void* data, int64_t npix, const int32_t* cols,
                              int64_t n_cols, int nframes, void* out, int* err, hipStream_t s) {
    const int64_t bytes = n_cols * nframes * (dtype == DXG2_U16 ? 2 : 1);
    const int64_t words = bytes / 4;
    if (words > 0) fill_bytes_kernel<<<grid_for(words), 256, 0, s>>>((uint32_t*)out, words);
    if (bytes > words * 4)
        fill_tail_kernel<<<1, 64, 0, s>>>((uint8_t*)out, words * 4, bytes);
    if (n_cols <= 0) return;
    if (dtype == DXG2_U16)
        events_to_columns_kernel<uint16_t><<<grid_for(n_cols), 256, 0, s>>>(
            offsets, times, (const uint16_t*)data, npix, cols, n_cols, nframes, (uint16_t*)out,
            err);
    else
        events_to_columns_kernel<uint8_t><<<grid_for(n_cols), 256, 0, s>>>(
            offsets, times, (const uint8_t*)data, npix, cols, n_cols, nframes, (uint8_t*)out,
            err);
}

void launch_excl_scan(const uint32_t* in, int64_t n, uint32_t* tot, uint32_t* out, hipStream_t s) {
    if (n <= 0) return;
    const int nb = (int)((n + SCAN_B - 1) / SCAN_B);
    scan_block_totals<<<nb, 256, 0, s>>>(in, n, tot);
    scan_totals<<<1, 256, 0, s>>>(tot, nb);
    scan_apply<<<nb, 256, 0, s>>>(in, n, tot, out);
}

void launch_copy_u32(const uint32_t* src, uint32_t* dst, int64_t n, hipStream_t s) {
    if (n > 0) copy_u32_kernel<<<grid_for(n), 256, 0, s>>>(src, dst, n);
}

void launch_fill_u32(uint32_t* p, int64_t n, hipStream_t s) {
    if (n > 0) fill_u32_kernel<<<grid_for(n), 256, 0, s>>>(p, n, 0u);
}

static unsigned grid_for(int64_t n, int64_t cap) {
    int64_t g = (n + 255) / 256;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

bool coo_narrow_values(int dtype, int nframes) {
    return dtype != DXG2_U16 && nframes <= (1 << 24);
}

// ------------------------------------------------------- stable LSD radix sort
// Sorts (key, value) pairs by the key bits [0, 8 * passes), 8 bits per pass, stably.
// Per pass over tiles of RS_TILE = 4096 pairs (4 waves x 16 rounds x 64 lanes, element
// index = wave * 1024 + round * 64 + lane):
//   radix_hist_kernel     digit histogram of each tile -> hist[digit * n_tiles + tile];
//   exclusive scan of hist (digit-major): the output position of each tile's first
//                         element of each digit;
//   radix_scatter_kernel  each wave ranks its 16 rounds in order: the lanes holding one
//                         digit find each other with 8 ballots (one per digit bit), the
//                         rank is the wave's running count of that digit plus the lanes
//                         below with it; the waves' counts give the cross-wave base; the
//                         tile is then reordered by digit in LDS and written out in runs.
constexpr int RS_TILE = 4096;
constexpr int RS_ROUNDS = 16;   // per wave

__global__ __launch_bounds__(256) void radix_hist_kernel(const uint32_t* __restrict__ keys,
                                                         int64_t n, int shift,
                                                         uint32_t* __restrict__ hist,
                                                         int n_tiles) {
    __shared__ uint32_t cnt[256];
    cnt[threadIdx.x] = 0u;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
    for (int i = threadIdx.x; i < RS_TILE; i += 256) {
        const int64_t e = base + i;
        if (e < n) atomicAdd(&cnt[(keys[e] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * n_tiles + blockIdx.x] = cnt[threadIdx.x];
}

template <typename V>
__global__ __launch_bounds__(256) void radix_scatter_kernel(
    const uint32_t* __restrict__ keys_in, const V* __restrict__ vals_in, int64_t n, int shift,
    const uint32_t* __restrict__ pos, int n_tiles, uint32_t* __restrict__ keys_out,
    V* __restrict__ vals_out) {
    __shared__ uint32_t wave_cnt[4][256];     // per wave: count of each digit so far
    __shared__ uint32_t local_off[256];       // the tile's exclusive digit offsets
    __shared__ uint32_t tot;
    __shared__ uint32_t s_keys[RS_TILE];
    __shared__ V s_vals[RS_TILE];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    for (int w = 0; w < 4; ++w) wave_cnt[w][tid] = 0u;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RS_TILE + wave * (RS_TILE / 4);
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t key[RS_ROUNDS];
    V val[RS_ROUNDS];
    uint32_t rank[RS_ROUNDS];
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const int64_t e = base + r * 64 + lane;
        const bool valid = e < n;
        key[r] = valid ? keys_in[e] : 0u;
        val[r] = valid ? vals_in[e] : (V)0;
    }
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const int64_t e = base + r * 64 + lane;
        const bool valid = e < n;
        const uint32_t d = (key[r] >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const uint64_t m = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? m : ~m;
        }
        const uint32_t before = wave_cnt[wave][d];      // same value for every peer
        rank[r] = before + (uint32_t)__popcll(peers & below);
        if (valid && (peers & below) == 0ull)           // the lowest peer updates
            wave_cnt[wave][d] = before + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // per digit (thread = digit): the waves' counts, their prefix, the tile's total
    const uint32_t c0 = wave_cnt[0][tid], c1 = wave_cnt[1][tid], c2 = wave_cnt[2][tid],
                   c3 = wave_cnt[3][tid];
    const uint32_t total = c0 + c1 + c2 + c3;
    __syncthreads();
    wave_cnt[0][tid] = 0u;
    wave_cnt[1][tid] = c0;
    wave_cnt[2][tid] = c0 + c1;
    wave_cnt[3][tid] = c0 + c1 + c2;
    // exclusive scan of the 256 totals into local_off (Hillis-Steele in LDS)
    local_off[tid] = total;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const uint32_t x = tid >= off ? local_off[tid - off] : 0u;
        __syncthreads();
        local_off[tid] += x;
        __syncthreads();
    }
    const uint32_t incl = local_off[tid];
    __syncthreads();
    local_off[tid] = incl - total;
    if (tid == 255) tot = incl;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const int64_t e = base + r * 64 + lane;
        if (e < n) {
            const uint32_t d = (key[r] >> shift) & 255u;
            const uint32_t p = local_off[d] + wave_cnt[wave][d] + rank[r];
            s_keys[p] = key[r];
            s_vals[p] = val[r];
        }
    }
    __syncthreads();
    const uint32_t count = tot;
    for (uint32_t i = tid; i < count; i += 256) {
        const uint32_t k = s_keys[i];
        const uint32_t d = (k >> shift) & 255u;
        const int64_t o = (int64_t)pos[(int64_t)d * n_tiles + blockIdx.x] + (i - local_off[d]);
        keys_out[o] = k;
        vals_out[o] = s_vals[i];
    }
}

static int radix_passes(uint32_t npix) { return (pixel_key_bits(npix) + 7) / 8; }

static int64_t radix_tiles(int64_t n) { return (n + RS_TILE - 1) / RS_TILE; }

// Workspace of the sort beyond the caller's keys / value buffers: a second key buffer,
// the histogram and its scan totals.
size_t coo_sort_bytes(int dtype, int nframes, int64_t cap, uint32_t npix) {
    (void)dtype;
    (void)nframes;
    (void)npix;
    const int64_t nt = radix_tiles(std::max<int64_t>(cap, 1));
    const int64_t hist = 256 * nt;
    // second key buffer (64-aligned), histogram + the scan's grand total, scan totals
    return (size_t)((cap + 63) / 64 * 64) * 4 + (size_t)(hist + 1) * 4 +
           (size_t)((hist + 1023) / 1024 + 1) * 4 + 256;
}

// Sorts (pix, v0) by pixel; returns the buffers holding the sorted keys and values.
template <typename V>
static void radix_sort_by_pixel(const uint32_t* pix, V* v0, V* v1, int64_t n, uint32_t npix,
                                const CooSortBufs& b, const uint32_t** keys_sorted,
                                const V** vals_sorted, hipStream_t s) {
    const int64_t nt = radix_tiles(n);
    uint32_t* keys2 = reinterpret_cast<uint32_t*>(b.tmp);
    uint32_t* hist = keys2 + ((n + 63) / 64) * 64;
    uint32_t* tot = hist + 256 * nt + 1;   // the in-place scan writes hist[256 nt]
    const uint32_t* kin = pix;
    const V* vin = v0;
    uint32_t* kbuf[2] = {b.keys, keys2};
    V* vbuf[2] = {v1, v0};
    const int passes = radix_passes(npix);
    for (int p = 0; p < passes; ++p) {
        uint32_t* kout = kbuf[p & 1];
        V* vout = vbuf[p & 1];
        radix_hist_kernel<<<(unsigned)nt, 256, 0, s>>>(kin, n, 8 * p, hist, (int)nt);
        launch_excl_scan(hist, 256 * nt, tot, hist, s);
        radix_scatter_kernel<V><<<(unsigned)nt, 256, 0, s>>>(kin, vin, n, 8 * p, hist, (int)nt,
                                                             kout, vout);
        kin = kout;
        vin = vout;
    }
    *keys_sorted = kin;
    *vals_sorted = vin;
}

// offsets[p] = the first position of a key >= p in the sorted keys, p = 0 .. npix.
__global__ void offsets_from_keys_kernel(const uint32_t* __restrict__ keys, int64_t n,
                                         uint32_t npix, uint32_t* __restrict__ offsets) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p <= (int64_t)npix;
         p += (int64_t)gridDim.x * blockDim.x) {
        int64_t lo = 0, hi = n;
        while (lo < hi) {
            const int64_t m = lo + (hi - lo) / 2;
            if (keys[m] < (uint32_t)p)
                lo = m + 1;
            else
                hi = m;
        }
        offsets[p] = (uint32_t)lo;
    }
}

template <typename V, typename T, int CB>
static void coo_sort_split(const uint32_t* pix, const uint32_t* frame, const T* cnt, int64_t nnz,
                           uint32_t npix, int nframes, const CooSortBufs& b, uint32_t* offsets,
                           int32_t* times, T* data, int* err, hipStream_t s) {
    V* v0 = reinterpret_cast<V*>(b.vals0);
    V* v1 = reinterpret_cast<V*>(b.vals1);
    coo_pack_kernel<V, T, CB><<<grid_for(nnz), 256, 0, s>>>(pix, frame, cnt, nnz, npix, nframes,
                                                           v0, err);
    const uint32_t* keys = nullptr;
    const V* vals = nullptr;
    radix_sort_by_pixel<V>(pix, v0, v1, nnz, npix, b, &keys, &vals, s);
    offsets_from_keys_kernel<<<grid_for((int64_t)npix + 1), 256, 0, s>>>(keys, nnz, npix,
                                                                        offsets);
    coo_unpack_kernel<V, T, CB><<<grid_for(nnz), 256, 0, s>>>(vals, nnz, times, data);
}

hipError_t launch_coo_to_csr(int dtype, const uint32_t* pix, const uint32_t* frame,
                             const void* cnt, int64_t nnz, uint32_t npix, int nframes,
                             uint32_t* offsets, const CooSortBufs& sort, int32_t* times,
                             void* data, int* err, hipStream_t s) {
    if (nnz <= 0) {
        fill_u32_kernel<<<grid_for((int64_t)npix + 1), 256, 0, s>>>(offsets, (int64_t)npix + 1, 0u);
        return hipGetLastError();
    }
    if (dtype == DXG2_U16)
        coo_sort_split<uint64_t, uint16_t, 16>(pix, frame, (const uint16_t*)cnt, nnz, npix,
                                               nframes, sort, offsets, times, (uint16_t*)data,
                                               err, s);
    else if (coo_narrow_values(dtype, nframes))
        coo_sort_split<uint32_t, uint8_t, 8>(pix, frame, (const uint8_t*)cnt, nnz, npix, nframes,
                                             sort, offsets, times, (uint8_t*)data, err, s);
    else
        coo_sort_split<uint64_t, uint8_t, 16>(pix, frame, (const uint8_t*)cnt, nnz, npix,
                                              nframes, sort, offsets, times, (uint8_t*)data, err,
                                              s);
    return hipGetLastError();
}

}  // namespace dxg2
