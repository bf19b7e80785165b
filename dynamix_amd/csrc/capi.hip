// Plan construction (host) and the extern "C" entry points of libdxg2.so.
//
// A plan turns (shape, nframes, qmask, dtype, scale factors) into: per-bin raster-
// ordered pixel index lists padded to k-blocks of 64, the i8 slab layout, the gather
// work list and the Gram tile work list (one grouped launch for all bins), plus the
// device workspaces.  Bins follow BaseCorrelator._set_qmask (correlator/common.py:36-46):
// qmask NULL is one bin of every pixel, otherwise bins are 1..max(qmask); scale factors
// follow _set_scale_factor (common.py:56-72): default pixel count per bin.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "dxg2_internal.h"

#include <algorithm>
#include <memory>
#include <mutex>

using namespace dxg2;

namespace {

thread_local std::string g_last_error;

int set_error(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define DXG2_HIP(call)                                                                     \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return set_error(DXG2_EHIP, "%s failed: %s", #call, hipGetErrorString(e_));     \
    } while (0)

// Device memory of plans and batches comes from an allocator: by default hipMalloc /
// hipFree (which waits for every kernel in flight, on every stream), or the caller's
// (dxg2_set_allocator), e.g. a caching allocator whose frees never touch the device.
// Every plan and batch keeps the allocator it was created with.
struct DevAllocator {
    dxg2_alloc_fn alloc = nullptr;
    dxg2_free_fn free = nullptr;
    void* ctx = nullptr;
};

std::mutex g_alloc_mu;
DevAllocator g_alloc;

DevAllocator current_allocator() {
    std::lock_guard<std::mutex> lock(g_alloc_mu);
    return g_alloc;
}

hipError_t dev_alloc(const DevAllocator& a, void** ptr, size_t bytes, hipStream_t s) {
    bytes = std::max<size_t>(bytes, 1);
    if (!a.alloc) return hipMalloc(ptr, bytes);
    *ptr = a.alloc(bytes, (void*)s, a.ctx);
    if (!*ptr) return hipErrorOutOfMemory;
    if (reinterpret_cast<uintptr_t>(*ptr) % 256 != 0) {  // the kernels' vector accesses
        if (a.free) a.free(*ptr, (void*)s, a.ctx);
        *ptr = nullptr;
        return hipErrorInvalidValue;
    }
    return hipSuccess;
}

void dev_free(const DevAllocator& a, void* ptr, hipStream_t s) {
    if (!ptr) return;
    if (a.free)
        a.free(ptr, (void*)s, a.ctx);
    else
        (void)hipFree(ptr);
}

// Plan set-up (workspace allocation and the upload of index lists and work lists) runs
// on a library-owned non-blocking stream per device, through a pinned staging buffer:
// building a plan never touches the null stream, so it never waits for work queued on
// the caller's streams.  Set-ups are serialised by a mutex and wait only for their own
// copies.
struct SetupContext {
    std::mutex mu;
    hipStream_t stream = nullptr;
    char* staging = nullptr;  // pinned
    size_t cap = 0;
};

SetupContext& setup_context(int dev) {
    static std::mutex mu;
    static std::vector<std::unique_ptr<SetupContext>> ctx;
    std::lock_guard<std::mutex> lock(mu);
    if ((int)ctx.size() <= dev) ctx.resize(dev + 1);
    if (!ctx[dev]) ctx[dev].reset(new SetupContext());
    return *ctx[dev];
}

// Order in which the hardware hands workgroups to the 8 XCDs of an MI355X: workgroup
// b runs on XCD b % 8.  remap_for_xcds() permutes a list so that XCD x receives the
// chunks x, x+8, x+16, ... of `group` consecutive items of the original order: items
// that share frame panels then run concurrently on one XCD and share its L2.
constexpr int N_XCD = 8;

template <typename T>
std::vector<T> remap_for_xcds(const std::vector<T>& in, int group, const T& pad) {
    const size_t n = in.size();
    if (n == 0) return {};
    const size_t n_groups = (n + group - 1) / group;
    std::vector<std::vector<T>> per(N_XCD);
    for (size_t g = 0; g < n_groups; ++g) {
        auto& dst = per[g % N_XCD];
        for (size_t i = g * group; i < std::min(n, (g + 1) * (size_t)group); ++i)
            dst.push_back(in[i]);
    }
    size_t len = 0;
    for (auto& v : per) len = std::max(len, v.size());
    std::vector<T> out(len * N_XCD, pad);
    for (int x = 0; x < N_XCD; ++x)
        for (size_t r = 0; r < per[x].size(); ++r) out[r * N_XCD + x] = per[x][r];
    return out;
}

// Gather order: items sorted by the detector position (raster index) of their first
// pixel, then dealt to the XCDs in chunks of `group`, so the k-blocks of all bins that
// read the same stretch of detector rows run together on one XCD and share the frame
// lines in its L2 instead of each XCD fetching them again.  The list is padded to a
// multiple of 8 with `pad` (a no-op item): a 2-D grid then keeps item i on XCD i % 8
// for every frame tile.
template <typename T, typename Key>
std::vector<T> raster_order_for_xcds(const std::vector<T>& in, Key key, int group, const T& pad) {
    std::vector<std::pair<int64_t, size_t>> k(in.size());
    for (size_t i = 0; i < in.size(); ++i) k[i] = {key(in[i]), i};
    std::stable_sort(k.begin(), k.end(),
                     [](const std::pair<int64_t, size_t>& a, const std::pair<int64_t, size_t>& b) {
                         return a.first < b.first;
                     });
    std::vector<T> sorted(in.size());
    for (size_t i = 0; i < in.size(); ++i) sorted[i] = in[k[i].second];
    return remap_for_xcds(sorted, group, pad);
}

// The same chunks as remap_for_xcds(), as 8 unpadded per-XCD queues laid end to end:
// queue x (chunks x, x+8, ...) is out[off[x] .. off[x+1]).
template <typename T>
std::vector<T> split_for_xcds(const std::vector<T>& in, int group, int* off) {
    std::vector<std::vector<T>> per(N_XCD);
    const size_t g = std::max(group, 1);
    for (size_t c = 0; c * g < in.size(); ++c)
        for (size_t i = c * g; i < std::min(in.size(), (c + 1) * g); ++i)
            per[c % N_XCD].push_back(in[i]);
    std::vector<T> out;
    for (int x = 0; x < N_XCD; ++x) {
        off[x] = (int)out.size();
        out.insert(out.end(), per[x].begin(), per[x].end());
    }
    off[N_XCD] = (int)out.size();
    return out;
}

// Work of the launches that store whole Gram tiles (calc_std, two-time): every
// upper-triangle tile of the bins concerned, its work items (k-ranges; for uint16 the
// HH, LL and cross items) numbered by scratch slot, cut into chunks whose slots fit the
// scratch.
struct TileWork {
    struct Chunk {
        int tile0, tile1;  // [tile0, tile1) of tiles
        int m0, m1;        // of mwork
        int x0, x1;        // of xwork
    };
    std::vector<GramWork> mwork, xwork;
    std::vector<StdTile> tiles;
    std::vector<Chunk> chunks;
    size_t slots = 0;  // scratch slots of the largest chunk
    GramWork* d_mwork = nullptr;
    GramWork* d_xwork = nullptr;
    StdTile* d_tiles = nullptr;
    int* d_scratch = nullptr;
};

}  // namespace

struct dxg2_plan {
    int H = 0, W = 0, N = 0, dtype = 0, n_bins = 0;
    int nrows = 0;  // slab rows per k-block: N rounded up to SLAB_ROW_ALIGN
    int device = 0;
    int gram_tile = 256;
    int tune_band = 4;    // tile rows per raster band of the Gram work list
    int tune_group = 32;  // consecutive items per XCD chunk (0: no XCD permutation)
    int tune_kchunk = MAX_KBLOCKS_PER_ITEM;
    int gather_frames = 64;  // frames per workgroup of the general gather (a multiple of 8)
    // graph replay of the dense chain (dxg2_correlate_dense_graph)
    hipStream_t cap_stream = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::array<const void*, 5> gkey{};
    std::vector<int64_t> npix;
    std::vector<double> scale;
    std::vector<int32_t> pix_idx;
    std::vector<BinDev> bins;
    std::vector<GatherWork> gather_work;
    // general gather slots: every k-block in detector order, 16 per item (kb -1: none)
    std::vector<KbWork> gather_slots;
    // band gather (uint8/int8): bands of detector rows, their (bin, k-block) list, the
    // k-blocks left to the general gather, and the LDS bytes staged per frame
    std::vector<BandWork> gband;
    std::vector<int2> gband_kb;
    std::vector<GatherWork> gfall;
    int64_t gfall_kb = 0;  // k-blocks on the fallback
    int gband_srow = 0;
    BandWork* d_gband = nullptr;
    int2* d_gband_kb = nullptr;
    GatherWork* d_gfall = nullptr;
    std::vector<GramWork> gram_work;    // plain kernel: every item of uint8/int8 plans;
                                        // HH and LL items of uint16 plans
    std::vector<GramWork> gram_xwork;   // cross kernel: [H|L][L|H]^T items of uint16 plans
    std::vector<GramWork> gram_raster, gram_xraster;  // the same before the XCD permutation
    std::vector<float2> twiddles;                     // FFT twiddle table (upload source)
    DevAllocator alloc;                               // allocator of every device buffer
    int gram_group = 0;
    size_t slab_bytes = 0;

    // device (owned)
    int32_t* d_pix = nullptr;
    BinDev* d_bins = nullptr;
    GatherWork* d_gwork = nullptr;
    KbWork* d_gslots = nullptr;
    GramWork* d_mwork = nullptr;
    GramWork* d_xwork = nullptr;
    // persistent launch: per-XCD queues of the plain and cross items, their offsets
    // (plain [0, 9), cross [9, 18)) and heads (plain [0, 8), cross [8, 16))
    bool use_queue = true;
    std::vector<GramWork> qm_work, qx_work;
    int q_off[18] = {0};
    GramWork* d_qmwork = nullptr;
    GramWork* d_qxwork = nullptr;
    int* d_qoff = nullptr;
    int* d_heads = nullptr;
    int8_t* d_slab = nullptr;
    int64_t* d_ugram = nullptr;
    int64_t* d_sums = nullptr;
    int64_t* d_psums = nullptr;   // digit planes: S_H, S_L, then S_Lw rows (n_bins each)
    uint8_t* d_wflag = nullptr;   // uint16: per detector pixel, a count > 255 in some frame
    int32_t* d_pixw = nullptr;    // uint16: per bin, its wide pixels (count[b] of them)
    int* d_nwkb = nullptr;        // uint16: per bin, k-blocks holding the wide pixels; then
                                  // their maximum over the bins; then per bin the number
                                  // of wide pixels
    int64_t* d_prefix = nullptr;  // per bin and plane: prefix sums of S(t), N + 1 values
    int64_t* d_smax = nullptr;    // per bin: max |S(t)|
    double* d_dpre = nullptr;     // per bin and lag: FP64 Den(tau) before the finalise
    DenMfma dmf;                  // i8-MFMA denominator workspace (k_den.hip)
    float* d_pa = nullptr;
    float* d_pg = nullptr;
    float* d_pm = nullptr;
    size_t pa_elems = 0;

    // calc_std (built on the first dxg2_correlate_dense_std call) and the two-time output
    // (for one bin, rebuilt when the bin changes): per-tile work with scratch slots
    bool std_built = false;
    TileWork stdw;
    size_t std_parts = 0;         // StdPart entries: n_bins * T(T+1)/2 tiles * (2T - 1)
    StdPart* d_std_part = nullptr;
    int ttcf_bin = -1;
    TileWork ttw;

    // event path
    std::vector<EventWork> event_work;
    EventWork* d_ework = nullptr;
    int* d_err = nullptr;
    uint32_t* d_cnt = nullptr;     // per-pixel counts / scatter cursors (COO input)
    uint32_t* d_scan = nullptr;    // scan block totals
    uint32_t* d_off = nullptr;     // CSR offsets built from COO input
    int32_t* d_etimes = nullptr;   // CSR frames built from COO input
    void* d_edata = nullptr;       // CSR counts built from COO input
    int64_t ev_cap = 0;            // capacity of d_etimes / d_edata / the sort buffers in events
    CooSortBufs coo_sort{nullptr, nullptr, nullptr, nullptr, 0};  // COO -> CSR sort workspace

    // FFT path
    std::vector<KbWork> kb_work;   // every k-block of every bin
    KbWork* d_kbwork = nullptr;
    std::vector<KbWork> kb_gather; // the same in gather order (kb -1: padding)
    KbWork* d_kbgather = nullptr;
    // the FFT path's pixel order within each bin (its sums do not depend on it): aligned
    // quads of detector pixels 4a .. 4a + 3 that lie wholly in the bin first, then the
    // rest, in raster order; fft_qkb[b] leading k-blocks of bin b are all such quads
    std::vector<int32_t> pix_fft;
    std::vector<int32_t> fft_qkb;
    int32_t* d_pixfft = nullptr;
    int32_t* d_fftqkb = nullptr;
    uint8_t* d_xs = nullptr;       // pixel-major bytes, same size as the slab
    float2* d_tw = nullptr;        // exp(-2 pi i m / L), m < L; then exp(-pi i k / L), k <= L
    double* d_power = nullptr;     // (n_bins, L + 1)
    double* d_fnum = nullptr;      // (n_bins, N)
    int fft_L = 0;
    int fft_L1 = 0, fft_L2 = 0;    // long series: L = L1 L2 through the scratch below
    int fft_chunk = 0;             // k-blocks per four-step chunk
    float2* d_fy = nullptr;        // four-step scratch: fft_chunk * 64 * L complex

    // Profiling: each correlate call records 4 events on its stream (no host sync);
    // dxg2_plan_stage_times() waits for the last one and averages every recorded call.
    bool profiling = false;
    std::vector<std::array<hipEvent_t, 4>> ev_pool;
    size_t ev_used = 0;
};

namespace {

int env_int(const char* name, int dflt) {
    const char* v = getenv(name);
    return (v && *v) ? atoi(v) : dflt;
}

// K-blocks per Gram work item by default.  uint16 plans split the k-range into chunks of
// 512: with N = 5,000 a tile row's frame panel of a large bin (256 frames x 3,000
// k-blocks x 64 bytes, 49 MB) outgrows an XCD's L2, and tiles that walk the same chunk
// at the same time share it there (cfg3 Gram 49.1 -> 44.3 ms); 8-bit plans keep whole
// items (at most 2,047 k-blocks: the int32 accumulator bound).
int default_kchunk(int dtype) { return dtype == DXG2_U16 ? 512 : MAX_KBLOCKS_PER_ITEM; }

// Slots of the general gather (gather_slots_kernel): every k-block of every bin sorted by
// the detector position of its first pixel, 16 per workgroup item (four per wave; a slot
// past the item's last k-block repeats its wave's last one, a wave with none gets kb -1),
// items dealt to the XCDs in chunks of GATHER_XCD_GROUP.
void build_gather_slots(dxg2_plan* p) {
    std::vector<std::pair<int64_t, KbWork>> all;
    for (int b = 0; b < p->n_bins; ++b)
        for (int kb = 0; kb < p->bins[b].nkb; ++kb)
            all.push_back({(int64_t)p->pix_idx[p->bins[b].idx_off + (int64_t)kb * KB], KbWork{b, kb}});
    std::stable_sort(all.begin(), all.end(),
                     [](const std::pair<int64_t, KbWork>& a, const std::pair<int64_t, KbWork>& b) {
                         return a.first < b.first;
                     });
    using Item = std::array<KbWork, 16>;
    std::vector<Item> items;
    for (size_t i = 0; i < all.size(); i += 16) {
        const int c = (int)std::min<size_t>(16, all.size() - i);
        Item it;
        for (int w = 0; w < 4; ++w)
            for (int j = 0; j < 4; ++j)
                it[4 * w + j] = 4 * w < c ? all[i + std::min(4 * w + j, c - 1)].second
                                          : KbWork{0, -1};
        items.push_back(it);
    }
    Item pad;
    pad.fill(KbWork{0, -1});
    items = remap_for_xcds(items, std::max(1, env_int("DXG2_GATHER_GROUP", GATHER_XCD_GROUP)), pad);
    p->gather_slots.clear();
    for (const Item& it : items) p->gather_slots.insert(p->gather_slots.end(), it.begin(), it.end());
}

// Band gather lists (uint8/int8 frames): rows per band RB such that 4 frames x (RB + 2)
// rows fit in 64 KiB of LDS; a k-block goes to the band of its first row if its last
// row is within the band's 2 rows of overlap, else to the fallback list.
void build_band_work(dxg2_plan* p, const std::vector<std::vector<int32_t>>& lists) {
    p->gband.clear();
    p->gband_kb.clear();
    p->gfall.clear();
    p->gfall_kb = 0;
    p->gband_srow = 0;
    const int W = p->W, H = p->H;
    // bands of RB rows, staged with one row of overlap below: RB + 1 rows of a frame fit
    // BAND_ROW_BYTES (with the 16-byte alignment slack)
    const int RB = (BAND_ROW_BYTES - 32) / W - 1;
    // the band kernel keeps 16-bit offsets into the staged rows and slab addresses / 4
    // in 32 bits
    if (p->dtype == DXG2_U16 || RB < 1 || p->slab_bytes >= ((size_t)1 << 34) ||
        env_int("DXG2_GATHER_BAND", 1) == 0)
        return;
    const int nbands = (H + RB - 1) / RB;
    std::vector<std::vector<int2>> per(nbands);
    for (int b = 0; b < p->n_bins; ++b) {
        const auto& L = lists[b];
        const int n = (int)L.size();
        int fall0 = -1;
        auto flush = [&](int kb_end) {
            for (int k = fall0; k < kb_end; k += 16)
                p->gfall.push_back(GatherWork{b, k, std::min(16, kb_end - k), 0});
            p->gfall_kb += kb_end - fall0;
            fall0 = -1;
        };
        const int nkb = (n + KB - 1) / KB;
        for (int kb = 0; kb < nkb; ++kb) {
            const int ymin = L[kb * KB] / W;
            const int ymax = L[std::min(n, (kb + 1) * KB) - 1] / W;
            const int band = ymin / RB;
            // a partial last k-block (entries past the end of the list) goes to the
            // general gather: the band kernel reads every entry of its k-blocks
            if ((kb + 1) * KB <= n && ymax - band * RB <= RB) {
                if (fall0 >= 0) flush(kb);
                per[band].push_back(make_int2(b, kb));
            } else if (fall0 < 0) {
                fall0 = kb;
            }
        }
        if (fall0 >= 0) flush(nkb);
    }
    // one work item per band and run of at most BAND_ITEM_KMAX k-blocks (the band kernel
    // keeps BAND_KMAX k-blocks per wave in registers)
    const int kmax = BAND_ITEM_KMAX;
    for (int band = 0; band < nbands; ++band) {
        const int y0 = band * RB;
        for (size_t k = 0; k < per[band].size(); k += kmax) {
            const int k0 = (int)p->gband_kb.size();
            const size_t k1 = std::min(per[band].size(), k + kmax);
            p->gband_kb.insert(p->gband_kb.end(), per[band].begin() + k, per[band].begin() + k1);
            p->gband.push_back(BandWork{y0, std::min(RB + 1, H - y0), k0, (int)p->gband_kb.size()});
        }
    }
    p->gband_srow = ((RB + 1) * W + 16 + 15) / 16 * 16;
    // the band kernel stages BAND_STAGE frames of rows in LDS, two workgroups per CU
    // (gfx950: 160 KiB of LDS per CU)
    if ((size_t)BAND_STAGE * p->gband_srow + 64 > (size_t)GFX950_LDS_BYTES / 2) {
        p->gband.clear();
        p->gband_kb.clear();
        p->gfall.clear();
        p->gfall_kb = 0;
        p->gband_srow = 0;
    }
}

// Tile order of the grouped Gram launch.  Within a bin (largest bins first) the
// upper-triangle tiles are walked in bands of `band` tile rows, column by column, so
// consecutive items share row and column frame panels; the list is then dealt to the
// XCDs in chunks of `group` items (0: no XCD remap).  Both are tuning parameters
// (environment DXG2_GRAM_BAND, DXG2_GRAM_GROUP).
void build_gram_work(dxg2_plan* p) {
    const int T = p->gram_tile;
    const int nT = (p->N + T - 1) / T;
    const int band = p->tune_band;
    const int group = p->tune_group;
    const int kchunk = p->tune_kchunk;
    std::vector<int> order(p->n_bins);
    for (int b = 0; b < p->n_bins; ++b) order[b] = b;
    // Largest bins first so their long tiles start early (LPT).
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return p->bins[a].nkb > p->bins[b].nkb; });
    std::vector<GramWork> raster, xraster;
    // One pass of upper-triangle tiles over k-blocks [ka, kend) of bin b, adding into
    // output row `out`.
    auto triangle = [&](int b, int ka, int kend, int out, int cross) {
        auto& dst = cross ? xraster : raster;
        for (int k0 = ka; k0 < kend; k0 += kchunk) {
            const int k1 = std::min(kend, k0 + kchunk);
            for (int i0 = 0; i0 < nT; i0 += band)
                for (int j = i0; j < nT; ++j)
                    for (int i = i0; i < std::min(nT, i0 + band) && i <= j; ++i)
                        dst.push_back(GramWork{b, i, j, k0, k1, out, cross, 0, 0, 0});
        }
    };
    const int nb = p->n_bins;
    for (int b : order) {
        const int nkb = p->bins[b].nkb;
        if (nkb == 0) continue;
        if (slab_planes(p->dtype) == 1) {
            triangle(b, 0, nkb, b, 0);
        } else {  // digit planes: H H^T, L L^T, [H|L][L|H]^T
            triangle(b, 0, nkb, b, 0);
            triangle(b, nkb, 2 * nkb, nb + b, 0);
            // the cross items of each half apart, so that an item clipped to the wide
            // k-blocks keeps one k-range
            triangle(b, 0, nkb, 2 * nb + b, 1);
            triangle(b, nkb, 2 * nkb, 2 * nb + b, 1);
        }
    }
    GramWork pad{-1, 0, 0, 0, 0, 0, 0, 0, 0, 0};  // job 0
    p->gram_work = group > 0 ? remap_for_xcds(raster, group, pad) : raster;
    p->gram_xwork = group > 0 ? remap_for_xcds(xraster, group, pad) : xraster;
    p->qm_work = split_for_xcds(raster, group > 0 ? group : 1, p->q_off);
    p->qx_work = split_for_xcds(xraster, group > 0 ? group : 1, p->q_off + 9);
    p->gram_raster = std::move(raster);
    p->gram_xraster = std::move(xraster);
    p->gram_group = group;
}

// Tile work (see TileWork) for every bin, or for bin `only` (0-based) alone.  Tiles
// follow the band order of the plain launch; each chunk's items are dealt to the XCDs.
void build_tile_work(const dxg2_plan* p, int only, size_t cap_bytes, TileWork& tw) {
    const int T = p->gram_tile;
    const int nT = (p->N + T - 1) / T;
    const int64_t ntri = (int64_t)nT * (nT + 1) / 2;
    const int band = p->tune_band;
    const int group = p->tune_group;
    const int kchunk = p->tune_kchunk;
    const bool wide = slab_planes(p->dtype) == 2;
    const size_t tile_bytes = (size_t)T * T * 4;
    int max_items = 0;
    for (int b = 0; b < p->n_bins; ++b) {
        if (only >= 0 && b != only) continue;
        const int nkb = p->bins[b].nkb;
        const int n1 = (nkb + kchunk - 1) / kchunk;
        max_items = std::max(max_items, wide ? 2 * n1 + (2 * nkb + kchunk - 1) / kchunk : n1);
    }
    const size_t cap_slots = std::max<size_t>((size_t)std::max(max_items, 1), cap_bytes / tile_bytes);
    std::vector<GramWork> cm, cx;  // current chunk's items, in tile order
    size_t slots = 0;
    GramWork pad{-1, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    auto close_chunk = [&]() {  // once per chunk, before the next one opens and at the end
        if (tw.chunks.empty()) return;
        auto& ch = tw.chunks.back();
        ch.tile1 = (int)tw.tiles.size();
        auto m = group > 0 ? remap_for_xcds(cm, group, pad) : cm;
        auto x = group > 0 ? remap_for_xcds(cx, group, pad) : cx;
        ch.m0 = (int)tw.mwork.size();
        tw.mwork.insert(tw.mwork.end(), m.begin(), m.end());
        ch.m1 = (int)tw.mwork.size();
        ch.x0 = (int)tw.xwork.size();
        tw.xwork.insert(tw.xwork.end(), x.begin(), x.end());
        ch.x1 = (int)tw.xwork.size();
        tw.slots = std::max(tw.slots, slots);
        cm.clear();
        cx.clear();
        slots = 0;
    };
    const int nb = p->n_bins;
    for (int b = 0; b < nb; ++b) {
        if (only >= 0 && b != only) continue;
        const int nkb = p->bins[b].nkb;
        if (nkb == 0) continue;
        const int n_hh = (nkb + kchunk - 1) / kchunk;
        const int n_x = wide ? (2 * nkb + kchunk - 1) / kchunk : 0;
        const int n_items = n_hh * (wide ? 2 : 1) + n_x;
        for (int i0 = 0; i0 < nT; i0 += band)
            for (int j = i0; j < nT; ++j)
                for (int i = i0; i < std::min(nT, i0 + band) && i <= j; ++i) {
                    if (tw.chunks.empty() || slots + n_items > cap_slots) {
                        close_chunk();
                        tw.chunks.push_back({(int)tw.tiles.size(), -1, 0, 0, 0, 0});
                    }
                    const int idx = (int)(b * ntri + (int64_t)i * nT - (int64_t)i * (i - 1) / 2 + (j - i));
                    tw.tiles.push_back(StdTile{b, i, j, (int)slots, n_hh, wide ? n_hh : 0, n_x, idx});
                    auto add = [&](int k0, int k1, int out, int cross) {
                        (cross ? cx : cm).push_back(
                            GramWork{b, i, j, k0, k1, out, cross, 0, (int)slots++, 0});
                    };
                    for (int k0 = 0; k0 < nkb; k0 += kchunk) add(k0, std::min(nkb, k0 + kchunk), b, 0);
                    if (wide) {
                        for (int k0 = nkb; k0 < 2 * nkb; k0 += kchunk)
                            add(k0, std::min(2 * nkb, k0 + kchunk), nb + b, 0);
                        for (int k0 = 0; k0 < 2 * nkb; k0 += kchunk)
                            add(k0, std::min(2 * nkb, k0 + kchunk), 2 * nb + b, 1);
                    }
                }
    }
    close_chunk();
}

// Free a TileWork's device copies and scratch and clear its lists.
void drop_tile_work(const DevAllocator& a, TileWork& tw, hipStream_t s) {
    void* bufs[] = {tw.d_mwork, tw.d_xwork, tw.d_tiles, tw.d_scratch};
    for (void* b : bufs) dev_free(a, b, s);
    tw = TileWork();
}

// Upload a TileWork's lists (pageable host memory: the copies complete before return)
// and allocate its scratch slots.
int upload_tile_work(const DevAllocator& a, TileWork& tw, hipStream_t s, int T) {
    auto upload = [&](void** dst, const void* src, size_t bytes) -> int {
        if (bytes == 0) return DXG2_OK;
        if (dev_alloc(a, dst, bytes, s) != hipSuccess)
            return set_error(DXG2_ENOMEM, "cannot allocate %zu bytes of tile work", bytes);
        DXG2_HIP(hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, s));
        return DXG2_OK;
    };
    int rc = upload((void**)&tw.d_mwork, tw.mwork.data(), tw.mwork.size() * sizeof(GramWork));
    if (!rc) rc = upload((void**)&tw.d_xwork, tw.xwork.data(), tw.xwork.size() * sizeof(GramWork));
    if (!rc) rc = upload((void**)&tw.d_tiles, tw.tiles.data(), tw.tiles.size() * sizeof(StdTile));
    if (!rc && dev_alloc(a, (void**)&tw.d_scratch, std::max<size_t>(tw.slots, 1) * T * T * 4, s) !=
                   hipSuccess)
        rc = set_error(DXG2_ENOMEM, "cannot allocate %zu Gram tile slots", tw.slots);
    if (!rc) DXG2_HIP(hipStreamSynchronize(s));
    return rc;
}

size_t std_scratch_cap() {
    return (size_t)std::max(1, env_int("DXG2_STD_SCRATCH_MB", 4096)) << 20;
}

int ensure_std_buffers(dxg2_plan* p, hipStream_t s) {
    if (p->std_built) return DXG2_OK;
    build_tile_work(p, -1, std_scratch_cap(), p->stdw);
    const int T = p->gram_tile;
    const int nT = (p->N + T - 1) / T;
    p->std_parts = (size_t)std::max(p->n_bins, 1) * ((size_t)nT * (nT + 1) / 2) * (2 * T - 1);
    int rc = upload_tile_work(p->alloc, p->stdw, s, T);
    if (!rc && dev_alloc(p->alloc, (void**)&p->d_std_part, p->std_parts * sizeof(StdPart), s) != hipSuccess)
        rc = set_error(DXG2_ENOMEM, "cannot allocate calc_std diagonal segments");
    if (rc) return rc;
    p->std_built = true;
    return DXG2_OK;
}

// Tile work of the two-time output for bin b (0-based), rebuilt when b changes.
int ensure_ttcf_buffers(dxg2_plan* p, int b, hipStream_t s) {
    if (p->ttcf_bin == b) return DXG2_OK;
    drop_tile_work(p->alloc, p->ttw, s);
    p->ttcf_bin = -1;
    build_tile_work(p, b, std_scratch_cap(), p->ttw);
    int rc = upload_tile_work(p->alloc, p->ttw, s, p->gram_tile);
    if (rc) return rc;
    p->ttcf_bin = b;
    return DXG2_OK;
}

// Bytes of the i8-MFMA denominator workspace: per bin, two digit planes of the frame
// sums, each 64 rows of dmf.up bytes.
size_t den_mfma_bytes(const dxg2_plan* p) {
    return (size_t)std::max(p->n_bins, 1) * 2 * 64 * (size_t)p->dmf.up;
}

size_t std_footprint(const dxg2_plan* p) {
    const size_t T = p->gram_tile;
    const size_t nT = (p->N + T - 1) / T;
    const size_t tiles = (size_t)std::max(p->n_bins, 1) * nT * (nT + 1) / 2;
    if (p->std_built)
        return (p->stdw.mwork.size() + p->stdw.xwork.size()) * sizeof(GramWork) +
               p->stdw.tiles.size() * sizeof(StdTile) + std::max<size_t>(p->stdw.slots, 1) * T * T * 4 +
               p->std_parts * sizeof(StdPart);
    // before the first call: every tile's items bounded by the scratch cap
    const int planes = slab_planes(p->dtype);
    const size_t items = tiles * (planes == 2 ? 4 : 1);
    return items * (sizeof(GramWork) + 4) + tiles * sizeof(StdTile) +
           std::min(items * T * T * 4, std_scratch_cap() + (size_t)T * T * 4 * 8) +
           tiles * (2 * T - 1) * sizeof(StdPart);
}

// Next unused event quad of the profiling pool (created on first use).
int next_events(dxg2_plan* p, hipEvent_t** out) {
    if (p->ev_used == p->ev_pool.size()) {
        std::array<hipEvent_t, 4> q{};
        for (auto& e : q) DXG2_HIP(hipEventCreate(&e));
        p->ev_pool.push_back(q);
    }
    *out = p->ev_pool[p->ev_used++].data();
    return DXG2_OK;
}

// Pixels of one bin's list handled by one event work item.
constexpr int EVENT_ITEM_PIXELS = 1024;

// FFT length: Nf = 2L with L the smallest length of the mixed-radix set (2^a 5^b, at
// least 1024) with L >= N; any Nf >= 2N - 1 gives the linear autocorrelation.  Up to
// 16384 a series is transformed in LDS; above, L = L1 L2 goes through the four-step
// path (L <= 131072).  *L1 / *L2 are 0 for the in-LDS path.
int fft_length(int n, int* L1 = nullptr, int* L2 = nullptr) {
    int l1 = 0, l2 = 0;
    int L = fft_lds_length(std::max(n, 1));
    if (L == 0) L = fft_four_step_length(n, &l1, &l2);
    if (L1) *L1 = l1;
    if (L2) *L2 = l2;
    return L;
}

// k-blocks per four-step chunk: the scratch holds 64 series of L complex points per
// k-block, at most DXG2_FFT_SCRATCH_MB MiB (default 1024).
int fft_chunk_kblocks(int L, int total_kb) {
    const size_t cap = (size_t)std::max(1, env_int("DXG2_FFT_SCRATCH_MB", 1024)) << 20;
    const size_t per = (size_t)KB * L * sizeof(float2);
    return (int)std::max<size_t>(1, std::min<size_t>((size_t)std::max(total_kb, 1), cap / per));
}

int ensure_fft_buffers(dxg2_plan* p, hipStream_t s) {
    if (p->d_power) return DXG2_OK;
    p->kb_work.clear();
    int L1 = 0, L2 = 0;
    const int L = fft_length(p->N, &L1, &L2);
    if (L == 0)
        return set_error(DXG2_EUNSUPPORTED, "FFT path supports up to 131072 frames (got %d)",
                         p->N);
    p->fft_L = L;
    p->fft_L1 = L1;
    p->fft_L2 = L2;
    for (int b = 0; b < p->n_bins; ++b)
        for (int kb = 0; kb < p->bins[b].nkb; ++kb) p->kb_work.push_back(KbWork{b, kb});
    // quads are loaded as aligned 4-byte words, which needs 4 | H W (and a 4-aligned
    // frame pointer, checked per call)
    const bool quads = ((int64_t)p->H * p->W) % 4 == 0;
    p->pix_fft = p->pix_idx;
    p->fft_qkb.assign((size_t)std::max(p->n_bins, 1), 0);
    for (int b = 0; b < p->n_bins && quads; ++b) {
        const int32_t* list = p->pix_idx.data() + p->bins[b].idx_off;
        const int n = p->bins[b].npix;
        std::vector<int32_t> quad, rest;
        for (int i = 0; i < n;) {
            if (list[i] % 4 == 0 && i + 3 < n && list[i + 3] == list[i] + 3) {
                quad.insert(quad.end(), list + i, list + i + 4);  // sorted, distinct
                i += 4;
            } else {
                rest.push_back(list[i++]);
            }
        }
        int32_t* out = p->pix_fft.data() + p->bins[b].idx_off;
        std::copy(quad.begin(), quad.end(), out);
        std::copy(rest.begin(), rest.end(), out + quad.size());
        p->fft_qkb[b] = (int32_t)(quad.size() / KB);
    }
    // gather items: PM_KPB k-blocks consecutive in detector order (kb -1 past the end),
    // dealt to the XCDs as units
    {
        std::vector<std::pair<int64_t, KbWork>> order;
        for (const KbWork& w : p->kb_work)
            order.push_back({(int64_t)p->pix_fft[p->bins[w.bin].idx_off + (int64_t)w.kb * KB], w});
        std::stable_sort(order.begin(), order.end(),
                         [](const std::pair<int64_t, KbWork>& a,
                            const std::pair<int64_t, KbWork>& b) { return a.first < b.first; });
        using Item = std::array<KbWork, PM_KPB>;
        std::vector<Item> items;
        for (size_t i = 0; i < order.size(); i += PM_KPB) {
            Item it;
            for (int j = 0; j < PM_KPB; ++j)
                it[j] = i + j < order.size() ? order[i + j].second : KbWork{0, -1};
            items.push_back(it);
        }
        Item pad;
        pad.fill(KbWork{0, -1});
        items = remap_for_xcds(items, PM_XCD_GROUP / PM_KPB, pad);
        p->kb_gather.clear();
        for (const Item& it : items) p->kb_gather.insert(p->kb_gather.end(), it.begin(), it.end());
    }
    p->fft_chunk = L1 ? fft_chunk_kblocks(L, (int)p->kb_work.size()) : 0;
    // [0, L): exp(-2 pi i m / L), the transform's twiddles; [L, 2L]: exp(-pi i k / L),
    // k = 0..L, the factor that unpacks the half-length complex transform
    p->twiddles.resize(2 * (size_t)L + 1);
    for (int m = 0; m < L; ++m) {
        const double a = -2.0 * M_PI * (double)m / (double)L;
        p->twiddles[m] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    for (int k = 0; k <= L; ++k) {
        const double a = -M_PI * (double)k / (double)L;
        p->twiddles[L + k] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    // d_power is allocated last: it marks the workspace complete.  A failure releases
    // what this call allocated, so a later call starts again from nothing.
    uint8_t* xs = nullptr;
    float2* tw = nullptr;
    double *power = nullptr, *fnum = nullptr;
    KbWork* kbw = nullptr;
    KbWork* kbg = nullptr;
    float2* fy = nullptr;
    int32_t* pixf = nullptr;
    int32_t* qkb = nullptr;
    if ((L1 && dev_alloc(p->alloc, (void**)&fy, (size_t)p->fft_chunk * KB * L * sizeof(float2),
                         s) != hipSuccess) ||
        (!p->pix_fft.empty() &&
         dev_alloc(p->alloc, (void**)&pixf, p->pix_fft.size() * sizeof(int32_t), s) !=
             hipSuccess) ||
        dev_alloc(p->alloc, (void**)&qkb, p->fft_qkb.size() * sizeof(int32_t), s) != hipSuccess ||
        dev_alloc(p->alloc, (void**)&xs, p->slab_bytes, s) != hipSuccess ||
        dev_alloc(p->alloc, (void**)&tw, p->twiddles.size() * sizeof(float2), s) != hipSuccess ||
        dev_alloc(p->alloc, (void**)&fnum, (size_t)std::max(p->n_bins, 1) * p->N * sizeof(double), s) !=
            hipSuccess ||
        (!p->kb_work.empty() &&
         (dev_alloc(p->alloc, (void**)&kbw, p->kb_work.size() * sizeof(KbWork), s) != hipSuccess ||
          dev_alloc(p->alloc, (void**)&kbg, p->kb_gather.size() * sizeof(KbWork), s) !=
              hipSuccess)) ||
        dev_alloc(p->alloc, (void**)&power, (size_t)std::max(p->n_bins, 1) * (L + 1) * sizeof(double), s) !=
            hipSuccess) {
        for (void* b : {(void*)xs, (void*)tw, (void*)fnum, (void*)kbw, (void*)kbg, (void*)power,
                        (void*)fy, (void*)pixf, (void*)qkb})
            dev_free(p->alloc, b, s);
        return set_error(DXG2_ENOMEM, "cannot allocate FFT workspace");
    }
    p->d_fy = fy;
    p->d_pixfft = pixf;
    p->d_fftqkb = qkb;
    p->d_xs = xs;
    p->d_tw = tw;
    p->d_fnum = fnum;
    p->d_kbwork = kbw;
    p->d_kbgather = kbg;
    p->d_power = power;
    DXG2_HIP(hipMemcpyAsync(p->d_tw, p->twiddles.data(), p->twiddles.size() * sizeof(float2),
                            hipMemcpyHostToDevice, s));
    if (!p->pix_fft.empty())
        DXG2_HIP(hipMemcpyAsync(p->d_pixfft, p->pix_fft.data(),
                                p->pix_fft.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
    DXG2_HIP(hipMemcpyAsync(p->d_fftqkb, p->fft_qkb.data(), p->fft_qkb.size() * sizeof(int32_t),
                            hipMemcpyHostToDevice, s));
    if (!p->kb_work.empty()) {
        DXG2_HIP(hipMemcpyAsync(p->d_kbwork, p->kb_work.data(),
                                p->kb_work.size() * sizeof(KbWork), hipMemcpyHostToDevice, s));
        DXG2_HIP(hipMemcpyAsync(p->d_kbgather, p->kb_gather.data(),
                                p->kb_gather.size() * sizeof(KbWork), hipMemcpyHostToDevice, s));
    }
    return DXG2_OK;
}

// Larger items flush fewer lag histograms and repeat less per-item work; the size is
// the largest power of two from EVENT_ITEM_PIXELS up to 8 times that for which the
// items still give about 512 workgroups over the series' lag windows (two per CU).
// DXG2_EVENT_ITEM_PIXELS fixes it instead.
void build_event_work(dxg2_plan* p) {
    int64_t total = 0;
    for (const auto& n : p->npix) total += n;
    const int windows = (p->N + event_lag_window(p->dtype) - 1) / event_lag_window(p->dtype);
    const int64_t want_items = std::max<int64_t>(1, (512 + windows - 1) / windows);
    int item = EVENT_ITEM_PIXELS;
    while (item < 8 * EVENT_ITEM_PIXELS && total / (2 * item) >= want_items) item *= 2;
    item = std::max(64, env_int("DXG2_EVENT_ITEM_PIXELS", item));
    for (int b = 0; b < p->n_bins; ++b)
        for (int k0 = 0; k0 < p->bins[b].npix; k0 += item)
            p->event_work.push_back(
                EventWork{b, k0, std::min<int>(p->bins[b].npix, k0 + item), 0});
}

int ensure_event_buffers(dxg2_plan* p, int64_t nnz_coo, hipStream_t s) {
    if (!p->d_err) {
        // d_err is set last: it marks the pair below as complete
        int* err = nullptr;
        EventWork* ework = nullptr;
        const size_t wbytes = p->event_work.size() * sizeof(EventWork);
        if (dev_alloc(p->alloc, (void**)&err, sizeof(int), s) != hipSuccess ||
            (wbytes && dev_alloc(p->alloc, (void**)&ework, wbytes, s) != hipSuccess)) {
            dev_free(p->alloc, err, s);
            dev_free(p->alloc, ework, s);
            return set_error(DXG2_ENOMEM, "cannot allocate the event work list");
        }
        if (wbytes) {
            const hipError_t e = hipMemcpyAsync(ework, p->event_work.data(), wbytes,
                                                hipMemcpyHostToDevice, s);
            if (e != hipSuccess) {
                dev_free(p->alloc, err, s);
                dev_free(p->alloc, ework, s);
                return set_error(DXG2_EHIP, "event work upload failed: %s", hipGetErrorString(e));
            }
        }
        p->d_ework = ework;
        p->d_err = err;
    }
    if (nnz_coo >= 0) {
        const int64_t npix = (int64_t)p->H * p->W;
        if (!p->d_cnt) {
            // all three or none: a partial allocation is released so a later call retries
            uint32_t *cnt = nullptr, *off = nullptr, *scan = nullptr;
            if (dev_alloc(p->alloc, (void**)&cnt, npix * 4, s) != hipSuccess ||
                dev_alloc(p->alloc, (void**)&off, (npix + 1) * 4, s) != hipSuccess ||
                dev_alloc(p->alloc, (void**)&scan, ((npix + 1023) / 1024 + 1) * 4, s) !=
                    hipSuccess) {
                for (void* b : {(void*)cnt, (void*)off, (void*)scan}) dev_free(p->alloc, b, s);
                return set_error(DXG2_ENOMEM, "cannot allocate the COO conversion workspace");
            }
            p->d_cnt = cnt;
            p->d_off = off;
            p->d_scan = scan;
        }
        if (nnz_coo > p->ev_cap) {
            CooSortBufs& cs = p->coo_sort;
            dev_free(p->alloc, p->d_etimes, s);
            dev_free(p->alloc, p->d_edata, s);
            dev_free(p->alloc, cs.keys, s);
            dev_free(p->alloc, cs.vals0, s);
            dev_free(p->alloc, cs.vals1, s);
            dev_free(p->alloc, cs.tmp, s);
            p->d_etimes = nullptr;
            p->d_edata = nullptr;
            cs = CooSortBufs{nullptr, nullptr, nullptr, nullptr, 0};
            p->ev_cap = 0;
            const int64_t cap = std::max<int64_t>(nnz_coo, 1);
            const int64_t vb = coo_narrow_values(p->dtype, p->N) ? 4 : 8;
            const size_t tmp = coo_sort_bytes(p->dtype, p->N, cap, (uint32_t)npix);
            if (tmp == 0)
                return set_error(DXG2_EHIP, "cannot size the sort of %lld events",
                                 (long long)cap);
            if (dev_alloc(p->alloc, (void**)&p->d_etimes, cap * 4, s) != hipSuccess ||
                dev_alloc(p->alloc, &p->d_edata, cap * 2, s) != hipSuccess ||
                dev_alloc(p->alloc, (void**)&cs.keys, cap * 4, s) != hipSuccess ||
                dev_alloc(p->alloc, &cs.vals0, cap * vb, s) != hipSuccess ||
                dev_alloc(p->alloc, &cs.vals1, cap * vb, s) != hipSuccess ||
                dev_alloc(p->alloc, &cs.tmp, tmp, s) != hipSuccess)
                return set_error(DXG2_ENOMEM, "cannot allocate %lld events", (long long)cap);
            cs.tmp_bytes = tmp;
            p->ev_cap = cap;
        }
    }
    return DXG2_OK;
}

// Slab of the dense path, allocated on the first dense call and cleared on its stream:
// rows past the last frame stay zero for the life of the plan.
int ensure_dense_buffers(dxg2_plan* p, hipStream_t s) {
    if (!p->d_slab && p->slab_bytes) {
        if (dev_alloc(p->alloc, (void**)&p->d_slab, p->slab_bytes, s) != hipSuccess) {
            p->d_slab = nullptr;
            return set_error(DXG2_ENOMEM, "cannot allocate %zu bytes of slab", p->slab_bytes);
        }
        DXG2_HIP(hipMemsetAsync(p->d_slab, 0, p->slab_bytes, s));
    }
    return DXG2_OK;
}

// Clear the accumulators of a dense call and gather the frames into the slab.  Returns
// the frame-sum rows the Gram's diagonal tiles add into: the caller's sums (uint8/int8)
// or the plan's digit-plane sums S_H, S_L (uint16).
// wide_first (uint16 plans; the plain dense calls): the low plane in list order with the
// wide pixels flagged in the same pass, then the wide pixels' own k-blocks (k_wide.hip),
// for a Gram clipped to them (dense_gram and dense_finish with the same flag).
int64_t* dense_gather(dxg2_plan* p, const void* d_frames, int64_t* d_sums, hipStream_t s,
                      bool wide_first = false) {
    int64_t* sums = d_sums ? d_sums : p->d_sums;
    const bool wide = p->dtype == DXG2_U16;
    int64_t* gsums = wide ? p->d_psums : sums;
    const int64_t nbn = (int64_t)p->n_bins * p->N;
    launch_fill_i64(gsums, (wide ? 2 : 1) * nbn, s);
    launch_fill_i64(p->d_ugram, (wide ? 3 : 1) * nbn, s);
    const int64_t hw = (int64_t)p->H * p->W;
    uint32_t* flagw = wide && wide_first ? reinterpret_cast<uint32_t*>(p->d_wflag) : nullptr;
    if (flagw) launch_fill_u32(flagw, (hw + 3) / 4, s);
    if (!p->gband.empty() && reinterpret_cast<uintptr_t>(d_frames) % 16 == 0) {
        launch_gather_band((const uint8_t*)d_frames, p->dtype, hw, p->W, p->N, p->nrows,
                           p->gband_srow, p->d_pix, p->d_bins, p->d_gband, (int)p->gband.size(),
                           p->d_gband_kb, p->d_gfall, (int)p->gfall.size(), p->d_slab, s);
    } else {
        launch_gather_slots((const uint8_t*)d_frames, p->dtype, hw, p->N, p->nrows, p->d_pix,
                            p->d_bins, p->d_gslots, (int)(p->gather_slots.size() / 16),
                            p->gather_frames, p->d_slab, s, flagw);
    }
    if (flagw) {
        int64_t max_npix = 0;
        for (auto n : p->npix) max_npix = std::max(max_npix, n);
        launch_wide_partition(p->d_pix, p->d_bins, p->n_bins, (int)max_npix, p->d_wflag,
                              p->d_pixw, p->d_nwkb, s);
        launch_wide_gather((const uint16_t*)d_frames, hw, p->N, p->nrows, p->d_pix, p->d_pixw,
                           p->d_bins, p->n_bins, p->d_nwkb, p->d_slab, s);
        launch_wide_low_sums(p->d_slab, p->d_bins, p->d_nwkb, p->n_bins, p->N, p->nrows,
                             gsums + 2 * nbn, s);
    }
    return gsums;
}

// The Gram of every bin: the persistent launch over the per-XCD work queues (heads
// cleared by a fill kernel, so the sequence can be captured in a graph), or one
// workgroup per tile.
void dense_gram(dxg2_plan* p, int64_t* gsums, hipStream_t s, bool wide_first = false) {
    const int biased = p->dtype == DXG2_I8 ? 0 : 1;
    const int* nwkb = wide_first && p->dtype == DXG2_U16 ? p->d_nwkb : nullptr;
    if (p->use_queue) {
        launch_fill_i32(p->d_heads, 16, s);
        launch_gram_i8_queue(p->gram_tile, false, p->d_slab, p->d_bins, p->d_qmwork,
                             (int)p->qm_work.size(), p->d_qoff, p->d_heads, p->N, p->nrows,
                             biased, nwkb, p->d_ugram, gsums, s);
        launch_gram_i8_queue(p->gram_tile, true, p->d_slab, p->d_bins, p->d_qxwork,
                             (int)p->qx_work.size(), p->d_qoff + 9, p->d_heads + 8, p->N,
                             p->nrows, biased, nwkb, p->d_ugram, gsums, s);
    } else {
        launch_gram_i8(p->gram_tile, false, p->d_slab, p->d_bins, p->d_mwork,
                       (int)p->gram_work.size(), p->N, p->nrows, biased, nwkb, p->d_ugram, gsums,
                       s);
        launch_gram_i8(p->gram_tile, true, p->d_slab, p->d_bins, p->d_xwork,
                       (int)p->gram_xwork.size(), p->N, p->nrows, biased, nwkb, p->d_ugram, gsums,
                       s);
    }
}

// Bias corrections, exact denominators and g2 after the Gram.
void dense_finish(dxg2_plan* p, int64_t* gsums, int64_t* d_num, int64_t* d_sums, double* d_den,
                  float* d_g2, hipStream_t s, bool wide_first = false) {
    int64_t* sums = d_sums ? d_sums : p->d_sums;
    const bool wide = p->dtype == DXG2_U16;
    const int mode = wide ? (wide_first ? 3 : 2) : (p->dtype == DXG2_U8 ? 1 : 0);
    launch_finalize(gsums, sums, p->d_ugram, p->d_bins, p->n_bins, p->N, mode, p->d_prefix,
                    p->d_smax, p->d_dpre, wide ? p->d_nwkb : nullptr, d_num, d_den, d_g2,
                    p->dmf, s);
}

}  // namespace

struct dxg2_batch {
    hipEvent_t staged = nullptr;  // the last batch's descriptor copies have run
    GramJob* h_jobs = nullptr;    // pinned staging
    GramWork* h_work = nullptr;
    GramJob* d_jobs = nullptr;
    GramWork* d_work = nullptr;
    size_t jobs_cap = 0, work_cap = 0;
    EventJob* h_ejobs = nullptr;  // grouped event launches
    EventWork* h_ework = nullptr;
    EventJob* d_ejobs = nullptr;
    EventWork* d_ework = nullptr;
    size_t ejobs_cap = 0, ework_cap = 0;
    std::vector<void*> retired;   // outgrown pinned buffers, freed with the batch (a
                                  // pinned free can wait for the whole device)
    DevAllocator alloc;
};

extern "C" {

const char* dxg2_last_error(void) { return g_last_error.c_str(); }

int dxg2_set_allocator(dxg2_alloc_fn alloc, dxg2_free_fn free_fn, void* ctx) {
    if ((alloc == nullptr) != (free_fn == nullptr))
        return set_error(DXG2_EINVAL, "give both allocator functions or neither");
    std::lock_guard<std::mutex> lock(g_alloc_mu);
    g_alloc = DevAllocator{alloc, free_fn, ctx};
    return DXG2_OK;
}

const char* dxg2_version(void) { return "dxg2 0.1.0 gfx950"; }

namespace {

// Host side of plan construction (no device call): validates the description and builds
// the bins, padded pixel lists, slab layout, gather, band, Gram and event work lists.
// Returns the plan (or nullptr with the error recorded).
dxg2_plan* build_plan_host(const dxg2_plan_desc* desc, int* rc) {
    *rc = DXG2_OK;
    if (desc->height <= 0 || desc->width <= 0) {
        *rc = set_error(DXG2_EINVAL, "bad frame shape (%d, %d)", desc->height, desc->width);
        return nullptr;
    }
    if (desc->nframes <= 0) {
        *rc = set_error(DXG2_EINVAL, "nframes must be > 0");
        return nullptr;
    }
    if (desc->dtype != DXG2_U8 && desc->dtype != DXG2_I8 && desc->dtype != DXG2_U16) {
        *rc = set_error(DXG2_EINVAL, "unsupported dtype code %d", desc->dtype);
        return nullptr;
    }
    const int64_t hw = (int64_t)desc->height * desc->width;
    if (hw > (int64_t)INT32_MAX) {
        *rc = set_error(DXG2_EINVAL, "frame too large");
        return nullptr;
    }

    dxg2_plan* p = new dxg2_plan();
    p->alloc = current_allocator();
    p->H = desc->height;
    p->W = desc->width;
    p->N = desc->nframes;
    p->nrows = (desc->nframes + SLAB_ROW_ALIGN - 1) / SLAB_ROW_ALIGN * SLAB_ROW_ALIGN;
    p->dtype = desc->dtype;
    p->gram_tile = (desc->flags & DXG2_FLAG_TILE128) ? 128 : 256;
    // Tuning defaults (environment overrides: DXG2_GRAM_BAND, _GROUP, _KCHUNK, _QUEUE);
    // dxg2_plan_set_tuning replaces them.
    p->tune_band = std::max(1, env_int("DXG2_GRAM_BAND", 4));
    p->tune_group = std::max(0, env_int("DXG2_GRAM_GROUP", 32));
    p->gather_frames = std::max(8, env_int("DXG2_GATHER_FRAMES", 64) / 8 * 8);
    p->tune_kchunk = std::min(MAX_KBLOCKS_PER_ITEM,
                              std::max(1, env_int("DXG2_GRAM_KCHUNK", default_kchunk(p->dtype))));
    p->use_queue = env_int("DXG2_GRAM_QUEUE", 1) != 0 && !(desc->flags & DXG2_FLAG_NO_QUEUE);
    // i8-MFMA denominators for the bins whose frame sums fit two digits; DXG2_DEN_MFMA=0
    // leaves every bin to the VALU kernels
    p->dmf.up = env_int("DXG2_DEN_MFMA", 1) ? den_mfma_columns(p->N) : 0;
    p->dmf.umax = p->dtype == DXG2_I8 ? 127 : 255;

    // qmask -> bins (common.py:36-46)
    std::vector<std::vector<int32_t>> lists;
    if (!desc->qmask) {
        p->n_bins = 1;
        lists.resize(1);
        lists[0].resize(hw);
        for (int64_t i = 0; i < hw; ++i) lists[0][i] = (int32_t)i;
    } else {
        int32_t mx = 0;
        for (int64_t i = 0; i < hw; ++i) mx = std::max(mx, desc->qmask[i]);
        p->n_bins = mx;
        lists.resize(mx);
        for (int64_t i = 0; i < hw; ++i) {
            const int32_t q = desc->qmask[i];
            if (q > 0) lists[q - 1].push_back((int32_t)i);
        }
    }
    // scale factors (common.py:56-72)
    if (desc->scale_factors && desc->n_scale_factors != p->n_bins) {
        *rc = set_error(DXG2_EINVAL, "got %d scale factors for %d bins", desc->n_scale_factors,
                        p->n_bins);
        delete p;
        return nullptr;
    }
    int64_t idx_off = 0, slab_off = 0;
    for (int b = 0; b < p->n_bins; ++b) {
        const int64_t n = (int64_t)lists[b].size();
        const int nkb = (int)((n + KB - 1) / KB);
        BinDev bd;
        bd.slab_off = slab_off;
        bd.idx_off = idx_off;
        bd.npix = (int32_t)n;
        bd.nkb = nkb;
        bd.scale = desc->scale_factors ? desc->scale_factors[b] : (double)n;
        p->bins.push_back(bd);
        p->npix.push_back(n);
        p->scale.push_back(bd.scale);
        for (int64_t i = 0; i < (int64_t)nkb * KB; ++i)
            p->pix_idx.push_back(i < n ? lists[b][i] : -1);
        for (int kb0 = 0; kb0 < nkb; kb0 += 16)
            p->gather_work.push_back(GatherWork{b, kb0, std::min(16, nkb - kb0), 0});
        idx_off += (int64_t)nkb * KB;
        slab_off += (int64_t)slab_planes(p->dtype) * nkb * KB * p->nrows;
    }
    p->slab_bytes = (size_t)slab_off;
    p->gather_work = raster_order_for_xcds(
        p->gather_work,
        [&](const GatherWork& g) {
            return (int64_t)p->pix_idx[p->bins[g.bin].idx_off + (int64_t)g.kb0 * KB];
        },
        GATHER_XCD_GROUP, GatherWork{0, 0, 0, 0});
    build_gather_slots(p);
    build_band_work(p, lists);
    build_gram_work(p);
    build_event_work(p);
    return p;
}

}  // namespace

int dxg2_plan_dry_run(const dxg2_plan_desc* desc, int64_t* summary, int32_t n) {
    if (!desc || (n > 0 && !summary)) return set_error(DXG2_EINVAL, "null argument");
    int rc = DXG2_OK;
    dxg2_plan* p = build_plan_host(desc, &rc);
    if (!p) return rc;
    TileWork tw;
    build_tile_work(p, -1, std_scratch_cap(), tw);
    int64_t items = 0;
    for (const auto& c : tw.chunks) items += (c.m1 - c.m0) + (c.x1 - c.x0);
    const int64_t v[] = {p->n_bins,
                         (int64_t)p->pix_idx.size(),
                         (int64_t)p->slab_bytes,
                         (int64_t)p->gather_work.size(),
                         (int64_t)p->gband.size(),
                         (int64_t)p->gband_kb.size(),
                         p->gfall_kb,
                         (int64_t)(p->gram_work.size() + p->gram_xwork.size()),
                         (int64_t)(p->qm_work.size() + p->qx_work.size()),
                         (int64_t)p->event_work.size(),
                         (int64_t)tw.tiles.size(),
                         items};
    for (int32_t i = 0; i < n && i < (int32_t)(sizeof(v) / sizeof(v[0])); ++i) summary[i] = v[i];
    delete p;
    return DXG2_OK;
}

int dxg2_plan_create(const dxg2_plan_desc* desc, dxg2_plan** out) {
    if (!desc || !out) return set_error(DXG2_EINVAL, "null argument");
    *out = nullptr;
    int rc0 = DXG2_OK;
    dxg2_plan* p = build_plan_host(desc, &rc0);
    if (!p) return rc0;

    hipError_t e = hipGetDevice(&p->device);
    if (e != hipSuccess) {
        delete p;
        return set_error(DXG2_EHIP, "hipGetDevice: %s", hipGetErrorString(e));
    }
    // Workspaces come from the stream-ordered pool on the set-up stream; the uploads go
    // through the pinned staging buffer, and the plan waits for them before it is
    // returned.
    SetupContext& sc = setup_context(p->device);
    std::lock_guard<std::mutex> setup_lock(sc.mu);
    if (!sc.stream && hipStreamCreateWithFlags(&sc.stream, hipStreamNonBlocking) != hipSuccess) {
        delete p;
        return set_error(DXG2_EHIP, "cannot create the plan set-up stream");
    }
    hipStream_t s0 = sc.stream;
    const size_t staged = p->pix_idx.size() * 4 + p->bins.size() * sizeof(BinDev) +
                          (p->gather_work.size() + p->gfall.size() + 2) * sizeof(GatherWork) +
                          (p->gather_slots.size() + 1) * sizeof(KbWork) +
                          (p->gband.size() + 1) * sizeof(BandWork) +
                          (p->gband_kb.size() + 1) * sizeof(int2) +
                          (p->gram_work.size() + p->gram_xwork.size() + 2) * sizeof(GramWork) +
                          (p->qm_work.size() + p->qx_work.size() + 2) * sizeof(GramWork) +
                          sizeof(p->q_off) +
                          (slab_planes(p->dtype) == 2 ? p->pix_idx.size() * 4 : 0) +
                          16 * 256;  // each of the (at most 16) uploads starts 256-aligned
    if (sc.cap < staged) {
        if (sc.staging) {
            (void)hipStreamSynchronize(s0);
            (void)hipHostFree(sc.staging);
        }
        sc.staging = nullptr;
        sc.cap = 0;
        const size_t cap = std::max<size_t>(staged, (size_t)1 << 24);
        if (hipHostMalloc((void**)&sc.staging, cap, 0) != hipSuccess) {
            delete p;
            return set_error(DXG2_ENOMEM, "cannot allocate %zu bytes of pinned staging", cap);
        }
        sc.cap = cap;
    }
    size_t staged_at = 0;
    auto upload = [&](void** dst, const void* src, size_t bytes) -> int {
        if (bytes == 0) return DXG2_OK;
        if (dev_alloc(p->alloc, dst, bytes, s0) != hipSuccess)
            return set_error(DXG2_ENOMEM, "cannot allocate %zu bytes", bytes);
        char* stage = sc.staging + staged_at;
        staged_at += (bytes + 255) / 256 * 256;
        memcpy(stage, src, bytes);
        if (hipMemcpyAsync(*dst, stage, bytes, hipMemcpyHostToDevice, s0) != hipSuccess)
            return set_error(DXG2_EHIP, "upload of %zu bytes failed", bytes);
        return DXG2_OK;
    };
    int rc = DXG2_OK;
    if (!rc) rc = upload((void**)&p->d_pix, p->pix_idx.data(), p->pix_idx.size() * 4);
    if (!rc) rc = upload((void**)&p->d_bins, p->bins.data(), p->bins.size() * sizeof(BinDev));
    if (!rc)
        rc = upload((void**)&p->d_gwork, p->gather_work.data(),
                    p->gather_work.size() * sizeof(GatherWork));
    if (!rc)
        rc = upload((void**)&p->d_gslots, p->gather_slots.data(),
                    p->gather_slots.size() * sizeof(KbWork));
    if (!rc) rc = upload((void**)&p->d_gband, p->gband.data(), p->gband.size() * sizeof(BandWork));
    if (!rc)
        rc = upload((void**)&p->d_gband_kb, p->gband_kb.data(), p->gband_kb.size() * sizeof(int2));
    if (!rc) rc = upload((void**)&p->d_gfall, p->gfall.data(), p->gfall.size() * sizeof(GatherWork));
    if (!rc)
        rc = upload((void**)&p->d_mwork, p->gram_work.data(),
                    p->gram_work.size() * sizeof(GramWork));
    if (!rc)
        rc = upload((void**)&p->d_xwork, p->gram_xwork.data(),
                    p->gram_xwork.size() * sizeof(GramWork));
    if (!rc)
        rc = upload((void**)&p->d_qmwork, p->qm_work.data(), p->qm_work.size() * sizeof(GramWork));
    if (!rc)
        rc = upload((void**)&p->d_qxwork, p->qx_work.data(), p->qx_work.size() * sizeof(GramWork));
    if (!rc) rc = upload((void**)&p->d_qoff, p->q_off, sizeof(p->q_off));
    if (!rc && dev_alloc(p->alloc, (void**)&p->d_heads, 16 * sizeof(int), s0) != hipSuccess)
        rc = set_error(DXG2_ENOMEM, "cannot allocate work-queue heads");
    const size_t nb_n = (size_t)std::max(p->n_bins, 1) * p->N * sizeof(int64_t);
    const int planes = slab_planes(p->dtype);
    // numerator rows: one per bin, or HH, LL and cross rows for digit planes
    if (!rc && dev_alloc(p->alloc, (void**)&p->d_ugram, (planes == 2 ? 3 : 1) * nb_n, s0) != hipSuccess)
        rc = set_error(DXG2_ENOMEM, "cannot allocate numerator workspace");
    if (!rc && dev_alloc(p->alloc, (void**)&p->d_sums, nb_n, s0) != hipSuccess)
        rc = set_error(DXG2_ENOMEM, "cannot allocate sums workspace");
    if (!rc && planes == 2 && dev_alloc(p->alloc, (void**)&p->d_psums, 3 * nb_n, s0) != hipSuccess)
        rc = set_error(DXG2_ENOMEM, "cannot allocate digit-plane sums workspace");
    if (!rc && planes == 2 &&
        (dev_alloc(p->alloc, (void**)&p->d_wflag, (size_t)p->H * p->W + 16, s0) != hipSuccess ||
         dev_alloc(p->alloc, (void**)&p->d_nwkb, (size_t)(2 * p->n_bins + 1) * sizeof(int),
                   s0) != hipSuccess))
        rc = set_error(DXG2_ENOMEM, "cannot allocate wide-pixel workspace");
    if (!rc && planes == 2) rc = upload((void**)&p->d_pixw, p->pix_idx.data(), p->pix_idx.size() * 4);
    if (!rc && (dev_alloc(p->alloc, (void**)&p->d_prefix, (size_t)(planes == 2 ? 3 : 1) *
                                                     std::max(p->n_bins, 1) *
                                                     (p->N + 1) * sizeof(int64_t), s0) !=
                    hipSuccess ||
                dev_alloc(p->alloc, (void**)&p->d_smax, (size_t)std::max(p->n_bins, 1) * sizeof(int64_t),
                           s0) != hipSuccess ||
                dev_alloc(p->alloc, (void**)&p->d_dpre,
                          (size_t)std::max(p->n_bins, 1) * p->N * sizeof(double), s0) != hipSuccess))
        rc = set_error(DXG2_ENOMEM, "cannot allocate denominator workspace");
    if (!rc && p->dmf.up > 0 &&
        dev_alloc(p->alloc, (void**)&p->dmf.v, den_mfma_bytes(p), s0) != hipSuccess) {
        p->dmf.v = nullptr;
        rc = set_error(DXG2_ENOMEM, "cannot allocate denominator workspace");
    }
    if (!rc && hipStreamSynchronize(s0) != hipSuccess)
        rc = set_error(DXG2_EHIP, "plan upload failed");
    if (rc) {
        dxg2_plan_destroy(p);
        return rc;
    }
    *out = p;
    return DXG2_OK;
}

// Release the captured dense graph (its work-list pointers are about to change or go).
static void drop_graph(dxg2_plan* p) {
    if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
    p->gexec = nullptr;
    p->gkey = {};
}

// Free the calc_std and two-time work lists and workspaces; the next call rebuilds them.
static void drop_std(dxg2_plan* p, hipStream_t s) {
    drop_tile_work(p->alloc, p->stdw, s);
    drop_tile_work(p->alloc, p->ttw, s);
    dev_free(p->alloc, p->d_std_part, s);
    p->d_std_part = nullptr;
    p->std_parts = 0;
    p->std_built = false;
    p->ttcf_bin = -1;
}

static void plan_release(dxg2_plan* p, hipStream_t s) {
    drop_graph(p);
    if (p->cap_stream) (void)hipStreamDestroy(p->cap_stream);
    void* bufs[] = {p->d_pix,   p->d_bins, p->d_gwork, p->d_gslots, p->d_mwork, p->d_xwork, p->d_slab,
                    p->d_ugram, p->d_sums, p->d_psums, p->d_prefix, p->d_smax, p->d_dpre,
                    p->dmf.v, p->d_pa,
                    p->d_pg, p->d_pm,
                    p->d_ework, p->d_err,  p->d_cnt,   p->d_scan,  p->d_off,
                    p->d_etimes, p->d_edata, p->d_kbwork, p->d_kbgather, p->d_xs, p->d_tw,
                    p->d_pixfft, p->d_fftqkb,
                    p->d_power, p->d_fnum, p->d_fy, p->d_qmwork, p->d_qxwork, p->d_qoff,
                    p->d_heads, p->d_gband, p->d_gband_kb, p->d_gfall,
                    p->coo_sort.keys, p->coo_sort.vals0, p->coo_sort.vals1, p->coo_sort.tmp,
                    p->d_wflag, p->d_pixw, p->d_nwkb};
    drop_std(p, s);
    for (void* b : bufs) dev_free(p->alloc, b, s);
    for (auto& q : p->ev_pool)
        for (auto e : q)
            if (e) (void)hipEventDestroy(e);
    delete p;
}

int dxg2_plan_destroy(dxg2_plan* p) {
    if (!p) return DXG2_OK;
    (void)hipDeviceSynchronize();  // no call on the plan may still be running
    plan_release(p, nullptr);
    return DXG2_OK;
}

int dxg2_plan_destroy_async(dxg2_plan* p, void* stream) {
    if (!p) return DXG2_OK;
    plan_release(p, (hipStream_t)stream);
    return DXG2_OK;
}

int dxg2_plan_query(const dxg2_plan* p, int32_t* n_bins, int64_t* pixel_counts, double* scale) {
    if (!p) return set_error(DXG2_EINVAL, "null plan");
    if (n_bins) *n_bins = p->n_bins;
    for (int b = 0; b < p->n_bins; ++b) {
        if (pixel_counts) pixel_counts[b] = p->npix[b];
        if (scale) scale[b] = p->scale[b];
    }
    return DXG2_OK;
}

int dxg2_plan_footprint(const dxg2_plan* p, int32_t path, size_t* bytes) {
    if (!p || !bytes) return set_error(DXG2_EINVAL, "null argument");
    const size_t nb_n = (size_t)std::max(p->n_bins, 1) * p->N * sizeof(int64_t);
    const int planes = slab_planes(p->dtype);
    size_t common = p->pix_idx.size() * 4 + p->bins.size() * sizeof(BinDev) +
                    (p->dmf.up > 0 ? den_mfma_bytes(p) : 0) + (planes == 2 ? 8 : 3) * nb_n +
                    (size_t)std::max(p->n_bins, 1) * ((planes == 2 ? 3 : 1) * (p->N + 1) + 1) *
                        sizeof(int64_t) +
                    (planes == 2 ? p->pix_idx.size() * 4 + (size_t)p->H * p->W +
                                       (size_t)std::max(p->n_bins, 1) * 4
                                 : 0);
    switch (path) {
        case 0:
            *bytes = common + p->slab_bytes + p->gather_work.size() * sizeof(GatherWork) +
                     p->gather_slots.size() * sizeof(KbWork) +
                     (p->gram_work.size() + p->gram_xwork.size() + p->qm_work.size() +
                      p->qx_work.size()) * sizeof(GramWork) + sizeof(p->q_off) + 16 * sizeof(int);
            return DXG2_OK;
        case 1: {
            int64_t mx = 0;
            for (auto n : p->npix) mx = std::max(mx, n);
            *bytes = common + ((size_t)p->N * ((mx + 7) / 8 * 8) + (size_t)p->N * p->N + p->N) * 4;
            return DXG2_OK;
        }
        case 2: {
            int L1 = 0, L2 = 0;
            const int L = fft_length(p->N, &L1, &L2);
            int64_t nkb = 0;
            for (const auto& bd : p->bins) nkb += bd.nkb;
            const size_t scratch =
                L1 ? (size_t)fft_chunk_kblocks(L, (int)nkb) * KB * L * sizeof(float2) : 0;
            *bytes = common + p->slab_bytes + (2 * (size_t)L + 1) * sizeof(float2) + scratch +
                     (size_t)std::max(p->n_bins, 1) * ((L + 1) + p->N) * sizeof(double) +
                     (2 * (size_t)nkb + (size_t)N_XCD * PM_XCD_GROUP) * sizeof(KbWork) +
                     (p->pix_idx.size() + (size_t)std::max(p->n_bins, 1)) * sizeof(int32_t);
            return DXG2_OK;
        }
        case 3: {
            const int64_t npix = (int64_t)p->H * p->W;
            *bytes = common + p->event_work.size() * sizeof(EventWork) + sizeof(int) +
                     (size_t)(2 * npix + 1 + (npix + 1023) / 1024 + 1) * 4 +
                     (size_t)p->ev_cap * (6 + 4 + 2 * (coo_narrow_values(p->dtype, p->N) ? 4 : 8)) +
                     p->coo_sort.tmp_bytes;
            return DXG2_OK;
        }
        case 4:
            *bytes = std_footprint(p);
            return DXG2_OK;
        default:
            return set_error(DXG2_EINVAL, "unknown path %d", path);
    }
}

int dxg2_plan_set_profiling(dxg2_plan* p, int32_t on) {
    if (!p) return set_error(DXG2_EINVAL, "null plan");
    p->profiling = on != 0;
    p->ev_used = 0;
    return DXG2_OK;
}

int dxg2_plan_stage_times(dxg2_plan* p, float* times, int32_t n) {
    if (!p || !times) return set_error(DXG2_EINVAL, "null argument");
    float acc[5] = {0, 0, 0, 0, 0};
    if (p->ev_used) {
        DXG2_HIP(hipEventSynchronize(p->ev_pool[p->ev_used - 1][3]));
        for (size_t c = 0; c < p->ev_used; ++c) {
            auto& q = p->ev_pool[c];
            float ms;
            for (int k = 0; k < 3; ++k) {
                DXG2_HIP(hipEventElapsedTime(&ms, q[k], q[k + 1]));
                acc[k] += ms;
            }
            DXG2_HIP(hipEventElapsedTime(&ms, q[0], q[3]));
            acc[3] += ms;
        }
        for (int k = 0; k < 4; ++k) acc[k] /= (float)p->ev_used;
        acc[4] = (float)p->ev_used;
    }
    for (int i = 0; i < n && i < 5; ++i) times[i] = acc[i];
    p->ev_used = 0;
    return DXG2_OK;
}

int dxg2_correlate_dense(dxg2_plan* p, const void* d_frames, int64_t* d_num, int64_t* d_sums,
                         double* d_den, float* d_g2, void* stream) {
    if (!p) return set_error(DXG2_EINVAL, "null plan");
    if (!d_frames || !d_g2) return set_error(DXG2_EINVAL, "frames and g2 must be non-null");
    if (p->n_bins == 0) return DXG2_OK;
    hipStream_t s = (hipStream_t)stream;
    int rc = ensure_dense_buffers(p, s);
    if (rc) return rc;
    (void)hipGetLastError();
    hipEvent_t* ev = nullptr;
    if (p->profiling && (rc = next_events(p, &ev))) return rc;
    if (ev) DXG2_HIP(hipEventRecord(ev[0], s));
    int64_t* gsums = dense_gather(p, d_frames, d_sums, s, true);
    if (ev) DXG2_HIP(hipEventRecord(ev[1], s));
    dense_gram(p, gsums, s, true);
    if (ev) DXG2_HIP(hipEventRecord(ev[2], s));
    dense_finish(p, gsums, d_num, d_sums, d_den, d_g2, s, true);
    if (ev) DXG2_HIP(hipEventRecord(ev[3], s));
    DXG2_HIP(hipGetLastError());
    return DXG2_OK;
}

int dxg2_plan_get_tuning(const dxg2_plan* p, dxg2_tuning* t) {
    if (!p || !t) return set_error(DXG2_EINVAL, "null argument");
    t->tile = p->gram_tile;
    t->band = p->tune_band;
    t->group = p->tune_group;
    t->queue = p->use_queue ? 1 : 0;
    t->kchunk = p->tune_kchunk;
    return DXG2_OK;
}

int dxg2_plan_set_tuning(dxg2_plan* p, const dxg2_tuning* t) {
    if (!p || !t) return set_error(DXG2_EINVAL, "null argument");
    if ((t->tile != 128 && t->tile != 256) || t->band < 1 || t->group < 0 ||
        (t->queue != 0 && t->queue != 1) || t->kchunk < 0 || t->kchunk > MAX_KBLOCKS_PER_ITEM)
        return set_error(DXG2_EINVAL,
                         "bad tuning (tile %d, band %d, group %d, queue %d, kchunk %d)", t->tile,
                         t->band, t->group, t->queue, t->kchunk);
    DXG2_HIP(hipDeviceSynchronize());  // calls on the plan may still read its work lists
    // The new work lists are built on a copy of the plan's tuning state and uploaded into
    // new buffers; the plan switches to them only when every upload succeeded, so a
    // failure leaves it whole, with its previous tuning.
    struct TuningState {
        int tile, band, group, kchunk;
        bool queue;
        std::vector<GramWork> mwork, xwork, raster, xraster, qm, qx;
        int q_off[18];
        int gram_group;
    };
    auto save = [&]() {
        TuningState st{p->gram_tile, p->tune_band, p->tune_group, p->tune_kchunk, p->use_queue,
                       p->gram_work, p->gram_xwork, p->gram_raster, p->gram_xraster,
                       p->qm_work, p->qx_work, {}, p->gram_group};
        memcpy(st.q_off, p->q_off, sizeof(p->q_off));
        return st;
    };
    auto restore = [&](TuningState& st) {
        p->gram_tile = st.tile;
        p->tune_band = st.band;
        p->tune_group = st.group;
        p->tune_kchunk = st.kchunk;
        p->use_queue = st.queue;
        p->gram_work.swap(st.mwork);
        p->gram_xwork.swap(st.xwork);
        p->gram_raster.swap(st.raster);
        p->gram_xraster.swap(st.xraster);
        p->qm_work.swap(st.qm);
        p->qx_work.swap(st.qx);
        memcpy(p->q_off, st.q_off, sizeof(p->q_off));
        p->gram_group = st.gram_group;
    };
    TuningState old = save();
    p->gram_tile = t->tile;
    p->tune_band = t->band;
    p->tune_group = t->group;
    p->tune_kchunk = t->kchunk > 0 ? t->kchunk : default_kchunk(p->dtype);
    p->use_queue = t->queue != 0;
    p->gram_work.clear();
    p->gram_xwork.clear();
    build_gram_work(p);
    const std::vector<GramWork>* src[4] = {&p->gram_work, &p->gram_xwork, &p->qm_work,
                                           &p->qx_work};
    GramWork** dst[4] = {&p->d_mwork, &p->d_xwork, &p->d_qmwork, &p->d_qxwork};
    GramWork* fresh[4] = {nullptr, nullptr, nullptr, nullptr};
    int rc = DXG2_OK;
    for (int i = 0; i < 4 && !rc; ++i) {
        if (src[i]->empty()) continue;
        const size_t bytes = src[i]->size() * sizeof(GramWork);
        if (dev_alloc(p->alloc, (void**)&fresh[i], bytes, nullptr) != hipSuccess) {
            fresh[i] = nullptr;
            rc = set_error(DXG2_ENOMEM, "cannot allocate %zu bytes of work list; the plan "
                           "keeps its previous tuning", bytes);
        } else if (hipMemcpy(fresh[i], src[i]->data(), bytes, hipMemcpyHostToDevice) !=
                   hipSuccess) {
            rc = set_error(DXG2_EHIP, "work list upload failed; the plan keeps its previous "
                           "tuning");
        }
    }
    // the queue offsets live in a fixed 18-entry buffer: written only on success below
    if (rc) {
        for (GramWork* f : fresh) dev_free(p->alloc, f, nullptr);
        restore(old);
        return rc;
    }
    const hipError_t e = hipMemcpy(p->d_qoff, p->q_off, sizeof(p->q_off), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        for (GramWork* f : fresh) dev_free(p->alloc, f, nullptr);
        restore(old);
        (void)hipMemcpy(p->d_qoff, p->q_off, sizeof(p->q_off), hipMemcpyHostToDevice);
        return set_error(DXG2_EHIP, "queue offsets upload failed: %s", hipGetErrorString(e));
    }
    drop_graph(p);
    drop_std(p, nullptr);
    for (int i = 0; i < 4; ++i) {
        dev_free(p->alloc, *dst[i], nullptr);
        *dst[i] = fresh[i];
    }
    return DXG2_OK;
}

int dxg2_correlate_dense_graph(dxg2_plan* p, const void* d_frames, int64_t* d_num,
                               int64_t* d_sums, double* d_den, float* d_g2, void* stream) {
    if (!p) return set_error(DXG2_EINVAL, "null plan");
    if (!d_frames || !d_g2) return set_error(DXG2_EINVAL, "frames and g2 must be non-null");
    if (p->n_bins == 0) return DXG2_OK;
    hipStream_t s = (hipStream_t)stream;
    int rc = ensure_dense_buffers(p, s);  // the one-time slab clear runs eagerly, on s
    if (rc) return rc;
    const std::array<const void*, 5> key{d_frames, d_num, d_sums, d_den, d_g2};
    if (!p->gexec || key != p->gkey) {
        drop_graph(p);
        if (!p->cap_stream)
            DXG2_HIP(hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking));
        gram_queue_prepare(p->gram_tile);  // device queries outside the capture
        (void)hipGetLastError();
        DXG2_HIP(hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeThreadLocal));
        int64_t* gsums = dense_gather(p, d_frames, d_sums, p->cap_stream, true);
        dense_gram(p, gsums, p->cap_stream, true);
        dense_finish(p, gsums, d_num, d_sums, d_den, d_g2, p->cap_stream, true);
        hipGraph_t g = nullptr;
        const hipError_t e1 = hipStreamEndCapture(p->cap_stream, &g);
        const hipError_t e0 = hipGetLastError();
        if (e0 != hipSuccess || e1 != hipSuccess || !g) {
            if (g) (void)hipGraphDestroy(g);
            return set_error(DXG2_EHIP, "capture of the dense chain failed: %s",
                             hipGetErrorString(e1 != hipSuccess ? e1 : e0));
        }
        const hipError_t e2 = hipGraphInstantiate(&p->gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e2 != hipSuccess) {
            p->gexec = nullptr;
            return set_error(DXG2_EHIP, "graph instantiation failed: %s", hipGetErrorString(e2));
        }
        p->gkey = key;
    }
    DXG2_HIP(hipGraphLaunch(p->gexec, s));
    return DXG2_OK;
}

int dxg2_correlate_dense_ttcf(dxg2_plan* p, const void* d_frames, int32_t q, float* d_trc,
                              void* stream) {
    if (!p) return set_error(DXG2_EINVAL, "null plan");
    if (!d_frames || !d_trc) return set_error(DXG2_EINVAL, "frames and trc must be non-null");
    if (q < 1 || q > p->n_bins)
        return set_error(DXG2_EINVAL, "bin %d is not in 1..%d", q, p->n_bins);
    hipStream_t s = (hipStream_t)stream;
    const int b = q - 1;
    const int64_t N = p->N;
    (void)hipGetLastError();
    if (p->bins[b].npix == 0) {  // 0 / 0 everywhere
        launch_fill_f32(d_trc, N * N, __builtin_nanf(""), s);
        DXG2_HIP(hipGetLastError());
        return DXG2_OK;
    }
    int rc = ensure_dense_buffers(p, s);
    if (!rc) rc = ensure_ttcf_buffers(p, b, s);
    if (rc) return rc;
    int64_t* gsums = dense_gather(p, d_frames, nullptr, s);
    launch_slab_sums(p->d_slab, p->dtype, p->d_bins, p->d_gwork, (int)p->gather_work.size(),
                     p->n_bins, p->N, p->nrows, gsums, s);
    const int biased = p->dtype == DXG2_I8 ? 0 : 1;
    const int mode = p->dtype == DXG2_U16 ? 2 : (p->dtype == DXG2_U8 ? 1 : 0);
    const TileWork& tw = p->ttw;
    for (const auto& ch : tw.chunks) {
        launch_gram_i8_store(p->gram_tile, false, p->d_slab, p->d_bins, tw.d_mwork + ch.m0,
                             ch.m1 - ch.m0, p->N, p->nrows, biased, p->d_ugram, tw.d_scratch, s);
        launch_gram_i8_store(p->gram_tile, true, p->d_slab, p->d_bins, tw.d_xwork + ch.x0,
                             ch.x1 - ch.x0, p->N, p->nrows, biased, p->d_ugram, tw.d_scratch, s);
        launch_ttcf_tiles(p->gram_tile, tw.d_tiles + ch.tile0, ch.tile1 - ch.tile0, p->d_bins,
                          tw.d_scratch, gsums, mode, p->n_bins, p->N, d_trc, s);
    }
    launch_ttcf_diag(d_trc, p->N, gsums, mode, p->n_bins, b, s);
    DXG2_HIP(hipGetLastError());
    return DXG2_OK;
}

int dxg2_correlate_dense_std(dxg2_plan* p, const void* d_frames, int64_t* d_num,
                             int64_t* d_sums, double* d_den, float* d_g2, float* d_dev,
                             void* stream) {
    if (!p) return set_error(DXG2_EINVAL, "null plan");
    if (!d_frames || !d_g2 || !d_dev)
        return set_error(DXG2_EINVAL, "frames, g2 and dev must be non-null");
    if (p->n_bins == 0) return DXG2_OK;
    hipStream_t s = (hipStream_t)stream;
    int rc = ensure_dense_buffers(p, s);
    if (!rc) rc = ensure_std_buffers(p, s);
    if (rc) return rc;
    (void)hipGetLastError();
    hipEvent_t* ev = nullptr;
    if (p->profiling && (rc = next_events(p, &ev))) return rc;
    if (ev) DXG2_HIP(hipEventRecord(ev[0], s));
    int64_t* gsums = dense_gather(p, d_frames, d_sums, s);
    launch_slab_sums(p->d_slab, p->dtype, p->d_bins, p->d_gwork, (int)p->gather_work.size(),
                     p->n_bins, p->N, p->nrows, gsums, s);
    if (ev) DXG2_HIP(hipEventRecord(ev[1], s));
    const int biased = p->dtype == DXG2_I8 ? 0 : 1;
    const int mode = p->dtype == DXG2_U16 ? 2 : (p->dtype == DXG2_U8 ? 1 : 0);
    const TileWork& tw = p->stdw;
    for (const auto& ch : tw.chunks) {
        launch_gram_i8_store(p->gram_tile, false, p->d_slab, p->d_bins, tw.d_mwork + ch.m0,
                             ch.m1 - ch.m0, p->N, p->nrows, biased, p->d_ugram, tw.d_scratch, s);
        launch_gram_i8_store(p->gram_tile, true, p->d_slab, p->d_bins, tw.d_xwork + ch.x0,
                             ch.x1 - ch.x0, p->N, p->nrows, biased, p->d_ugram, tw.d_scratch, s);
        launch_std_tiles(p->gram_tile, tw.d_tiles + ch.tile0, ch.tile1 - ch.tile0, p->d_bins,
                         tw.d_scratch, gsums, mode, p->n_bins, p->N, p->d_std_part, s);
    }
    if (ev) DXG2_HIP(hipEventRecord(ev[2], s));
    dense_finish(p, gsums, d_num, d_sums, d_den, d_g2, s);
    launch_dev_finalize(p->d_std_part, p->d_bins, p->n_bins, p->N, p->gram_tile, d_dev, s);
    if (ev) DXG2_HIP(hipEventRecord(ev[3], s));
    DXG2_HIP(hipGetLastError());
    return DXG2_OK;
}

int dxg2_batch_create(dxg2_batch** out) {
    if (!out) return set_error(DXG2_EINVAL, "null argument");
    dxg2_batch* b = new dxg2_batch();
    b->alloc = current_allocator();
    if (hipEventCreateWithFlags(&b->staged, hipEventDisableTiming) != hipSuccess) {
        delete b;
        return set_error(DXG2_EHIP, "hipEventCreate failed");
    }
    *out = b;
    return DXG2_OK;
}

int dxg2_batch_destroy(dxg2_batch* b) {
    if (!b) return DXG2_OK;
    (void)hipEventSynchronize(b->staged);
    (void)hipEventDestroy(b->staged);
    (void)hipDeviceSynchronize();
    dev_free(b->alloc, b->d_jobs, nullptr);
    dev_free(b->alloc, b->d_work, nullptr);
    dev_free(b->alloc, b->d_ejobs, nullptr);
    dev_free(b->alloc, b->d_ework, nullptr);
    b->retired.push_back(b->h_jobs);
    b->retired.push_back(b->h_work);
    b->retired.push_back(b->h_ejobs);
    b->retired.push_back(b->h_ework);
    for (void* h : b->retired)
        if (h) (void)hipHostFree(h);
    delete b;
    return DXG2_OK;
}

int dxg2_batch_correlate_dense(dxg2_batch* b, int32_t n, dxg2_plan* const* plans,
                               const void* const* d_frames, int64_t* const* d_num,
                               int64_t* const* d_sums, double* const* d_den,
                               float* const* d_g2, void* stream) {
    if (!b || n < 0 || (n > 0 && (!plans || !d_frames || !d_g2)))
        return set_error(DXG2_EINVAL, "null argument");
    if (n == 0) return DXG2_OK;
    const int tile = plans[0] ? plans[0]->gram_tile : 0;
    for (int i = 0; i < n; ++i) {
        if (!plans[i] || !d_frames[i] || !d_g2[i])
            return set_error(DXG2_EINVAL, "job %d: null plan, frames or g2", i);
        if (plans[i]->gram_tile != tile)
            return set_error(DXG2_EINVAL, "job %d: Gram tile %d differs from %d", i,
                             plans[i]->gram_tile, tile);
        for (int k = 0; k < i; ++k)
            if (plans[k] == plans[i])
                return set_error(DXG2_EINVAL, "plan of job %d appears twice in the batch", i);
        int rc = ensure_dense_buffers(plans[i], (hipStream_t)stream);
        if (rc) return rc;
    }
    hipStream_t s = (hipStream_t)stream;
    (void)hipGetLastError();
    // Work list: every job's items, largest jobs first, one XCD permutation over the lot.
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    auto ops = [&](int i) {
        double w = 0;
        for (const auto& it : plans[i]->gram_raster) w += it.kb1 - it.kb0;
        for (const auto& it : plans[i]->gram_xraster) w += it.kb1 - it.kb0;
        return w;
    };
    std::vector<double> wk(n);
    for (int i = 0; i < n; ++i) wk[i] = ops(i);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return wk[x] > wk[y]; });
    std::vector<GramWork> plain, cross;
    std::vector<GramJob> jobs(n);
    for (int i : order) {
        dxg2_plan* p = plans[i];
        for (GramWork w : p->gram_raster) {
            w.job = i;
            plain.push_back(w);
        }
        for (GramWork w : p->gram_xraster) {
            w.job = i;
            cross.push_back(w);
        }
    }
    const int group = plans[0]->gram_group;
    GramWork pad{-1, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (group > 0) {
        plain = remap_for_xcds(plain, group, pad);
        cross = remap_for_xcds(cross, group, pad);
    }
    // Per-job gathers first: they also clear the outputs the Gram accumulates into.
    for (int i = 0; i < n; ++i) {
        dxg2_plan* p = plans[i];
        if (p->n_bins == 0) continue;
        int64_t* gsums = dense_gather(p, d_frames[i], d_sums ? d_sums[i] : nullptr, s);
        jobs[i] = GramJob{p->d_slab, p->d_bins, p->d_ugram, gsums, p->N, p->nrows,
                          p->dtype == DXG2_I8 ? 0 : 1, 0};
    }
    // Stage descriptors and work lists through pinned memory (reused once the previous
    // batch's copies have run).
    const size_t nw = plain.size() + cross.size();
    DXG2_HIP(hipEventSynchronize(b->staged));
    if (b->jobs_cap < (size_t)n) {  // grow geometrically; the outgrown buffers are kept
        const size_t cap = std::max<size_t>(std::max<size_t>(2 * b->jobs_cap, n), 16);
        b->retired.push_back(b->h_jobs);
        dev_free(b->alloc, b->d_jobs, s);
        b->h_jobs = nullptr;
        b->d_jobs = nullptr;
        b->jobs_cap = 0;
        if (hipHostMalloc((void**)&b->h_jobs, cap * sizeof(GramJob), 0) != hipSuccess ||
            dev_alloc(b->alloc, (void**)&b->d_jobs, cap * sizeof(GramJob), s) != hipSuccess)
            return set_error(DXG2_ENOMEM, "cannot allocate batch job table");
        b->jobs_cap = cap;
    }
    if (b->work_cap < nw) {
        const size_t cap = std::max<size_t>(std::max<size_t>(2 * b->work_cap, nw), 1 << 16);
        b->retired.push_back(b->h_work);
        dev_free(b->alloc, b->d_work, s);
        b->h_work = nullptr;
        b->d_work = nullptr;
        b->work_cap = 0;
        if (hipHostMalloc((void**)&b->h_work, cap * sizeof(GramWork), 0) != hipSuccess ||
            dev_alloc(b->alloc, (void**)&b->d_work, cap * sizeof(GramWork), s) != hipSuccess)
            return set_error(DXG2_ENOMEM, "cannot allocate batch work list");
        b->work_cap = cap;
    }
    std::copy(jobs.begin(), jobs.end(), b->h_jobs);
    std::copy(plain.begin(), plain.end(), b->h_work);
    std::copy(cross.begin(), cross.end(), b->h_work + plain.size());
    DXG2_HIP(hipMemcpyAsync(b->d_jobs, b->h_jobs, n * sizeof(GramJob), hipMemcpyHostToDevice, s));
    if (nw)
        DXG2_HIP(hipMemcpyAsync(b->d_work, b->h_work, nw * sizeof(GramWork),
                                hipMemcpyHostToDevice, s));
    DXG2_HIP(hipEventRecord(b->staged, s));
    launch_gram_i8_jobs(tile, false, b->d_jobs, b->d_work, (int)plain.size(), s);
    launch_gram_i8_jobs(tile, true, b->d_jobs, b->d_work + plain.size(), (int)cross.size(), s);
    for (int i = 0; i < n; ++i) {
        dxg2_plan* p = plans[i];
        if (p->n_bins == 0) continue;
        dense_finish(p, jobs[i].sums, d_num ? d_num[i] : nullptr, d_sums ? d_sums[i] : nullptr,
                     d_den ? d_den[i] : nullptr, d_g2[i], s);
    }
    DXG2_HIP(hipGetLastError());
    return DXG2_OK;
}

}  // extern "C"

namespace {

// Pinned staging + device copy of a batch table, grown geometrically; outgrown pinned
// buffers are kept until the batch is destroyed (a pinned free can wait for the device).
template <typename T>
int grow_table(dxg2_batch* b, T** h, T** d, size_t* cap, size_t need, size_t min_cap,
               hipStream_t s) {
    if (*cap >= need) return DXG2_OK;
    const size_t c = std::max<size_t>(std::max<size_t>(2 * *cap, need), min_cap);
    b->retired.push_back(*h);
    dev_free(b->alloc, *d, s);
    *h = nullptr;
    *d = nullptr;
    *cap = 0;
    if (hipHostMalloc((void**)h, c * sizeof(T), 0) != hipSuccess ||
        dev_alloc(b->alloc, (void**)d, c * sizeof(T), s) != hipSuccess)
        return set_error(DXG2_ENOMEM, "cannot allocate batch table");
    *cap = c;
    return DXG2_OK;
}

}  // namespace

extern "C" {

int dxg2_batch_correlate_events(dxg2_batch* b, int32_t n, dxg2_plan* const* plans,
                                const int32_t* const* d_times, const void* const* d_data,
                                const uint32_t* const* d_offsets, const int64_t* nnz,
                                int64_t* const* d_num, int64_t* const* d_sums,
                                double* const* d_den, float* const* d_g2, void* stream) {
    if (!b || n < 0 || (n > 0 && (!plans || !d_times || !d_data || !d_offsets || !nnz || !d_g2)))
        return set_error(DXG2_EINVAL, "null argument");
    if (n == 0) return DXG2_OK;
    hipStream_t s = (hipStream_t)stream;
    const int dtype = plans[0] ? plans[0]->dtype : 0;
    for (int i = 0; i < n; ++i) {
        dxg2_plan* p = plans[i];
        if (!p || !d_offsets[i] || !d_g2[i] || (nnz[i] > 0 && (!d_times[i] || !d_data[i])))
            return set_error(DXG2_EINVAL, "job %d: null plan, times, data, offsets or g2", i);
        if (nnz[i] < 0 || nnz[i] > (int64_t)UINT32_MAX)
            return set_error(DXG2_EINVAL, "job %d: event count %lld out of range", i,
                             (long long)nnz[i]);
        if (p->dtype != dtype)
            return set_error(DXG2_EINVAL, "job %d: count type differs from job 0's", i);
        for (int k = 0; k < i; ++k)
            if (plans[k] == p)
                return set_error(DXG2_EINVAL, "plan of job %d appears twice in the batch", i);
        int rc = ensure_event_buffers(p, -1, s);
        if (rc) return rc;
    }
    (void)hipGetLastError();
    // one work list over every job's items, each tagged with its job; the launch covers
    // the lag windows of the longest series
    std::vector<EventJob> jobs(n);
    std::vector<EventWork> work;
    int max_frames = 0;
    for (int i = 0; i < n; ++i) {
        dxg2_plan* p = plans[i];
        const int64_t nbn = (int64_t)p->n_bins * p->N;
        int64_t* sums = d_sums && d_sums[i] ? d_sums[i] : p->d_sums;
        launch_fill_i64(sums, nbn, s);
        launch_fill_i64(p->d_ugram, nbn, s);
        launch_fill_i32(p->d_err, 1, s);
        jobs[i] = EventJob{d_times[i], d_data[i], d_offsets[i], p->d_pix, p->d_bins, nnz[i],
                           p->d_ugram, sums, p->d_err, p->N, 0};
        if (p->n_bins == 0) continue;
        max_frames = std::max(max_frames, p->N);
        for (EventWork w : p->event_work) {
            w.pad = i;
            work.push_back(w);
        }
    }
    DXG2_HIP(hipEventSynchronize(b->staged));
    int rc = grow_table(b, &b->h_ejobs, &b->d_ejobs, &b->ejobs_cap, (size_t)n, 16, s);
    if (!rc) rc = grow_table(b, &b->h_ework, &b->d_ework, &b->ework_cap, work.size(), 4096, s);
    if (rc) return rc;
    std::copy(jobs.begin(), jobs.end(), b->h_ejobs);
    std::copy(work.begin(), work.end(), b->h_ework);
    DXG2_HIP(hipMemcpyAsync(b->d_ejobs, b->h_ejobs, n * sizeof(EventJob), hipMemcpyHostToDevice, s));
    if (!work.empty())
        DXG2_HIP(hipMemcpyAsync(b->d_ework, b->h_ework, work.size() * sizeof(EventWork),
                                hipMemcpyHostToDevice, s));
    DXG2_HIP(hipEventRecord(b->staged, s));
    launch_event_pairs_jobs(dtype, b->d_ejobs, b->d_ework, (int)work.size(), max_frames, s);
    for (int i = 0; i < n; ++i) {
        dxg2_plan* p = plans[i];
        if (p->n_bins == 0) continue;
        launch_finalize(jobs[i].sums, jobs[i].sums, p->d_ugram, p->d_bins, p->n_bins, p->N, 0,
                        p->d_prefix, p->d_smax, p->d_dpre, nullptr, d_num ? d_num[i] : nullptr,
                        d_den ? d_den[i] : nullptr, d_g2[i], p->dmf, s);
    }
    DXG2_HIP(hipGetLastError());
    return DXG2_OK;
}

int dxg2_correlate_dense_plain(dxg2_plan* p, const void* d_frames, float* d_g2, void* stream) {
    if (!p) return set_error(DXG2_EINVAL, "null plan");
    if (!d_frames || !d_g2) return set_error(DXG2_EINVAL, "frames and g2 must be non-null");
    hipStream_t s = (hipStream_t)stream;
    int64_t mx = 0;
    for (auto n : p->npix) mx = std::max(mx, n);
    const size_t lda = (size_t)((mx + 7) / 8 * 8);
    if (!p->d_pa && lda > 0) {
        p->pa_elems = (size_t)p->N * lda;
        if (dev_alloc(p->alloc, (void**)&p->d_pa, p->pa_elems * 4, s) != hipSuccess ||
            dev_alloc(p->alloc, (void**)&p->d_pg, (size_t)p->N * p->N * 4, s) != hipSuccess ||
            dev_alloc(p->alloc, (void**)&p->d_pm, (size_t)p->N * 4, s) != hipSuccess)
            return set_error(DXG2_ENOMEM, "cannot allocate plain-path workspace");
    }
    (void)hipGetLastError();
    hipEvent_t* ev = nullptr;
    int rc;
    if (p->profiling && (rc = next_events(p, &ev))) return rc;
    if (ev) DXG2_HIP(hipEventRecord(ev[0], s));
    for (int b = 0; b < p->n_bins; ++b) {
        float* row = d_g2 + (int64_t)b * p->N;
        if (p->npix[b] == 0) {
            std::vector<float> nan(p->N, std::nanf(""));
            DXG2_HIP(hipMemcpyAsync(row, nan.data(), p->N * 4, hipMemcpyHostToDevice, s));
            DXG2_HIP(hipStreamSynchronize(s));
            continue;
        }
        launch_plain_bin((const uint8_t*)d_frames, p->dtype, (int64_t)p->H * p->W, p->N,
                         p->d_pix + p->bins[b].idx_off, (int)p->npix[b], p->d_pa, p->d_pg,
                         p->d_pm, row, s);
    }
    if (ev) {
        DXG2_HIP(hipEventRecord(ev[1], s));
        DXG2_HIP(hipEventRecord(ev[2], s));
        DXG2_HIP(hipEventRecord(ev[3], s));
    }
    DXG2_HIP(hipGetLastError());
    return DXG2_OK;
}

// Shared tail of both event entry points: pairs + sums from pixel-major CSR, then the
// exact denominator and g2 (no bias: counts are used as they are).
// With profiling on, `ev` holds the call's events: [0] and [1] are already recorded
// around the input conversion (none for CSR input), [2] follows the pairs, [3] the rest.
static int events_from_csr(dxg2_plan* p, const int32_t* times, const void* data,
                           const uint32_t* offsets, int64_t nnz, int64_t* d_num, int64_t* d_sums,
                           double* d_den, float* d_g2, hipStream_t s, hipEvent_t* ev) {
    int64_t* sums = d_sums ? d_sums : p->d_sums;
    launch_event_pairs(p->dtype, times, data, offsets, p->d_pix, p->d_bins, p->d_ework,
                       (int)p->event_work.size(), p->N, nnz, p->d_ugram, sums, p->d_err, s);
    if (ev) DXG2_HIP(hipEventRecord(ev[2], s));
    launch_finalize(sums, sums, p->d_ugram, p->d_bins, p->n_bins, p->N, 0, p->d_prefix,
                    p->d_smax, p->d_dpre, nullptr, d_num, d_den, d_g2, p->dmf, s);
    if (ev) DXG2_HIP(hipEventRecord(ev[3], s));
    DXG2_HIP(hipGetLastError());
    return DXG2_OK;
}

int dxg2_correlate_events_csr(dxg2_plan* p, const int32_t* d_times, const void* d_data,
                              const uint32_t* d_offsets, int64_t nnz, int64_t* d_num,
                              int64_t* d_sums, double* d_den, float* d_g2, void* stream) {
    if (!p) return set_error(DXG2_EINVAL, "null plan");
    if (!d_offsets || !d_g2 || (nnz > 0 && (!d_times || !d_data)))
        return set_error(DXG2_EINVAL, "times, data, offsets and g2 must be non-null");
    if (nnz < 0 || nnz > (int64_t)UINT32_MAX)
        return set_error(DXG2_EINVAL, "event count %lld out of range", (long long)nnz);
    if (p->n_bins == 0) return DXG2_OK;
    int rc = ensure_event_buffers(p, -1, (hipStream_t)stream);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    (void)hipGetLastError();
    const int64_t nbn = (int64_t)p->n_bins * p->N;
    launch_fill_i64(d_sums ? d_sums : p->d_sums, nbn, s);
    launch_fill_i64(p->d_ugram, nbn, s);
    launch_fill_i32(p->d_err, 1, s);
    hipEvent_t* ev = nullptr;
    if (p->profiling && (rc = next_events(p, &ev))) return rc;
    if (ev) {
        DXG2_HIP(hipEventRecord(ev[0], s));
        DXG2_HIP(hipEventRecord(ev[1], s));
    }
    return events_from_csr(p, d_times, d_data, d_offsets, nnz, d_num, d_sums, d_den, d_g2, s,
                           ev);
}

int dxg2_correlate_events_coo(dxg2_plan* p, const uint32_t* d_pixel, const uint32_t* d_frame,
                              const void* d_count, int64_t nnz, int64_t* d_num, int64_t* d_sums,
                              double* d_den, float* d_g2, void* stream) {
    if (!p) return set_error(DXG2_EINVAL, "null plan");
    if (!d_g2 || (nnz > 0 && (!d_pixel || !d_frame || !d_count)))
        return set_error(DXG2_EINVAL, "pixel, frame, count and g2 must be non-null");
    if (nnz < 0 || nnz > (int64_t)UINT32_MAX)
        return set_error(DXG2_EINVAL, "event count %lld out of range", (long long)nnz);
    if (p->dtype == DXG2_I8)
        return set_error(DXG2_EUNSUPPORTED, "COO event counts are unsigned (uint8 / uint16)");
    if (p->n_bins == 0) return DXG2_OK;
    int rc = ensure_event_buffers(p, nnz, (hipStream_t)stream);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    (void)hipGetLastError();
    const int64_t nbn = (int64_t)p->n_bins * p->N;
    launch_fill_i64(d_sums ? d_sums : p->d_sums, nbn, s);
    launch_fill_i64(p->d_ugram, nbn, s);
    launch_fill_i32(p->d_err, 1, s);
    hipEvent_t* ev = nullptr;
    if (p->profiling && (rc = next_events(p, &ev))) return rc;
    if (ev) DXG2_HIP(hipEventRecord(ev[0], s));
    if (launch_coo_to_csr(p->dtype, d_pixel, d_frame, d_count, nnz,
                          (uint32_t)((int64_t)p->H * p->W), p->N, p->d_off, p->coo_sort,
                          p->d_etimes, p->d_edata, p->d_err, s) != hipSuccess)
        return set_error(DXG2_EHIP, "sort of %lld COO events failed", (long long)nnz);
    if (ev) DXG2_HIP(hipEventRecord(ev[1], s));
    return events_from_csr(p, p->d_etimes, p->d_edata, p->d_off, nnz, d_num, d_sums, d_den,
                           d_g2, s, ev);
}

int dxg2_correlate_fft(dxg2_plan* p, const void* d_frames, double* d_num, int64_t* d_sums,
                       double* d_den, float* d_g2, void* stream) {
    if (!p) return set_error(DXG2_EINVAL, "null plan");
    if (!d_frames || !d_g2) return set_error(DXG2_EINVAL, "frames and g2 must be non-null");
    if (p->n_bins == 0) return DXG2_OK;
    int rc = ensure_fft_buffers(p, (hipStream_t)stream);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    int64_t* sums = d_sums ? d_sums : p->d_sums;
    (void)hipGetLastError();
    hipEvent_t* ev = nullptr;
    if (p->profiling && (rc = next_events(p, &ev))) return rc;
    if (ev) DXG2_HIP(hipEventRecord(ev[0], s));
    launch_fill_i64(reinterpret_cast<int64_t*>(p->d_power), (int64_t)p->n_bins * (p->fft_L + 1), s);
    const bool quads = reinterpret_cast<uintptr_t>(d_frames) % 4 == 0;
    launch_gather_pm((const uint8_t*)d_frames, p->dtype, (int64_t)p->H * p->W, p->N, p->nrows,
                     p->d_pixfft, quads ? p->d_fftqkb : nullptr, p->d_bins, p->n_bins,
                     p->d_kbgather, (int)(p->kb_gather.size() / PM_KPB), p->d_xs, sums, s);
    if (ev) DXG2_HIP(hipEventRecord(ev[1], s));
    if (p->fft_L1) {
        const int total = (int)p->kb_work.size();
        for (int kb0 = 0; kb0 < total; kb0 += p->fft_chunk)
            if (launch_fft_four_step(p->fft_L1, p->fft_L2, p->dtype, p->d_xs, p->d_bins,
                                     p->d_kbwork, kb0, std::min(p->fft_chunk, total - kb0),
                                     p->nrows, p->d_tw, p->d_tw + p->fft_L, p->d_fy,
                                     p->d_power, s))
                return set_error(DXG2_EUNSUPPORTED, "no FFT kernel for length %d x %d",
                                 p->fft_L1, p->fft_L2);
    } else if (launch_fft_power(p->fft_L, p->dtype, p->d_xs, p->d_bins, p->d_kbwork,
                                (int)p->kb_work.size(), p->nrows, p->d_tw, p->d_tw + p->fft_L,
                                p->d_power, s)) {
        return set_error(DXG2_EUNSUPPORTED, "no FFT kernel for length %d", p->fft_L);
    }
    launch_fft_inverse(p->d_power, p->fft_L, p->n_bins, p->N, p->d_fnum, s);
    if (ev) DXG2_HIP(hipEventRecord(ev[2], s));
    launch_finalize_f64(sums, p->d_fnum, p->d_bins, p->n_bins, p->N, p->d_prefix, p->d_smax,
                        p->d_dpre, d_num, d_den, d_g2, p->dmf, s);
    if (ev) DXG2_HIP(hipEventRecord(ev[3], s));
    DXG2_HIP(hipGetLastError());
    return DXG2_OK;
}

int dxg2_plan_event_status(dxg2_plan* p, int32_t* flags, void* stream) {
    if (!p || !flags) return set_error(DXG2_EINVAL, "null argument");
    *flags = 0;
    if (!p->d_err) return DXG2_OK;
    int v = 0;
    DXG2_HIP(hipMemcpyAsync(&v, p->d_err, sizeof(int), hipMemcpyDeviceToHost, (hipStream_t)stream));
    DXG2_HIP(hipStreamSynchronize((hipStream_t)stream));
    *flags = v;
    return DXG2_OK;
}

int dxg2_synth_speckle(void* d_frames, int32_t dtype, int32_t nframes, int32_t height,
                       int32_t width, const float* d_mean_map, double tau_c, uint64_t seed,
                       void* stream) {
    if (!d_frames || !d_mean_map) return set_error(DXG2_EINVAL, "null argument");
    if (nframes <= 0 || height <= 0 || width <= 0) return set_error(DXG2_EINVAL, "bad sizes");
    if (dtype != DXG2_U8 && dtype != DXG2_I8 && dtype != DXG2_U16)
        return set_error(DXG2_EINVAL, "unsupported dtype %d", dtype);
    const float rho = tau_c > 0 ? (float)std::exp(-1.0 / tau_c) : 0.f;
    (void)hipGetLastError();
    launch_synth(d_frames, dtype, nframes, (int64_t)height * width, d_mean_map, rho, seed,
                 (hipStream_t)stream);
    DXG2_HIP(hipGetLastError());
    return DXG2_OK;
}

}  // extern "C"

extern "C" {

int dxg2_synth_events(const float* d_mean_map, int32_t height, int32_t width, int32_t nframes,
                      double tau_c, uint64_t seed, int32_t dtype, uint32_t* d_offsets,
                      int32_t* d_times, void* d_data, int64_t capacity, int64_t* nnz,
                      void* stream) {
    if (!d_mean_map || !d_offsets || !nnz) return set_error(DXG2_EINVAL, "null argument");
    if (height <= 0 || width <= 0 || nframes <= 0) return set_error(DXG2_EINVAL, "bad sizes");
    if (dtype != DXG2_U8 && dtype != DXG2_U16)
        return set_error(DXG2_EINVAL, "event counts are uint8 or uint16");
    const int64_t hw = (int64_t)height * width;
    hipStream_t s = (hipStream_t)stream;
    const float rho = tau_c > 0 ? (float)std::exp(-1.0 / tau_c) : 0.f;
    uint32_t *count = nullptr, *tot = nullptr;
    DXG2_HIP(dev_alloc(DevAllocator{}, (void**)&count, hw * 4, s));
    DXG2_HIP(dev_alloc(DevAllocator{}, (void**)&tot, ((hw + 1023) / 1024 + 1) * 4, s));
    (void)hipGetLastError();
    launch_synth_events(0, dtype, nframes, hw, d_mean_map, rho, seed, count, nullptr, nullptr,
                        nullptr, s);
    launch_excl_scan(count, hw, tot, d_offsets, s);
    uint32_t total = 0;
    hipError_t e = hipMemcpyAsync(&total, d_offsets + hw, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    dev_free(DevAllocator{}, count, s);
    dev_free(DevAllocator{}, tot, s);
    if (e != hipSuccess) return set_error(DXG2_EHIP, "event count failed: %s", hipGetErrorString(e));
    *nnz = total;
    if (d_times) {
        if (capacity < (int64_t)total)
            return set_error(DXG2_ECAPACITY, "capacity %lld < %u events", (long long)capacity,
                             total);
        launch_synth_events(1, dtype, nframes, hw, d_mean_map, rho, seed, nullptr, d_offsets,
                            d_times, d_data, s);
    }
    DXG2_HIP(hipGetLastError());
    return DXG2_OK;
}

int dxg2_events_window(const uint32_t* d_offsets, const int32_t* d_times, const void* d_data,
                       int64_t nnz, int64_t npix, int32_t dtype, int32_t t0, int32_t t1,
                       uint32_t* d_scratch, uint32_t* d_out_offsets, int32_t* d_out_times,
                       void* d_out_data, int64_t capacity, void* stream) {
    if (!d_offsets || !d_times || !d_data || !d_scratch || !d_out_offsets || !d_out_times ||
        !d_out_data)
        return set_error(DXG2_EINVAL, "null argument");
    if (npix <= 0 || nnz < 0 || t0 < 0 || t1 < t0)
        return set_error(DXG2_EINVAL, "bad window [%d, %d) or sizes", t0, t1);
    if (dtype != DXG2_U8 && dtype != DXG2_U16)
        return set_error(DXG2_EINVAL, "event counts are uint8 or uint16");
    if (capacity < nnz)
        return set_error(DXG2_ECAPACITY, "capacity %lld < %lld input events (the window is "
                         "at most the input)", (long long)capacity, (long long)nnz);
    (void)hipGetLastError();
    launch_events_window(dtype, d_offsets, d_times, d_data, npix, t0, t1, d_scratch,
                         d_out_offsets, d_out_times, d_out_data, (hipStream_t)stream);
    DXG2_HIP(hipGetLastError());
    return DXG2_OK;
}

int dxg2_events_to_columns(const uint32_t* d_offsets, const int32_t* d_times,
                           const void* d_data, int64_t npix, int32_t nframes, int32_t dtype,
                           const int32_t* d_cols, int64_t n_cols, void* d_out, int32_t* d_err,
                           void* stream) {
    if (!d_offsets || !d_times || !d_data || !d_out || (n_cols > 0 && !d_cols))
        return set_error(DXG2_EINVAL, "null argument");
    if (npix <= 0 || nframes <= 0 || n_cols < 0)
        return set_error(DXG2_EINVAL, "bad sizes (npix %lld, nframes %d, columns %lld)",
                         (long long)npix, nframes, (long long)n_cols);
    if (dtype != DXG2_U8 && dtype != DXG2_I8 && dtype != DXG2_U16)
        return set_error(DXG2_EINVAL, "unsupported dtype code %d", dtype);
    (void)hipGetLastError();
    launch_events_to_columns(dtype, d_offsets, d_times, d_data, npix, d_cols, n_cols, nframes,
                             d_out, d_err, (hipStream_t)stream);
    DXG2_HIP(hipGetLastError());
    return DXG2_OK;
}

// Allocation-free variant of dxg2_events_window: the scan's block totals live in a
// file-scope device array instead of a caller workspace.  One 32-bit total per 1024
// pixels, so the array bounds the pixel count at DXG2_STATIC_SCRATCH_WORDS * 1024.
// Each GPU holds its own copy at its own address; host code reaches it only through the
// address hipGetSymbolAddress returns for the current device, fetched once per device.
// The scan overwrites every total it reads, so nothing needs clearing between calls;
// calls that overlap on one GPU (different streams, or threads) race on the array.
static __device__ __align__(256) uint32_t scratch[DXG2_STATIC_SCRATCH_WORDS];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[kMaxDevices] = {};

int static_scratch(uint32_t** out) {
    int dev = 0;
    DXG2_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return set_error(DXG2_EUNSUPPORTED, "device %d", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        DXG2_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = static_cast<uint32_t*>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return DXG2_OK;
}
}  // namespace

int dxg2_events_window_noalloc(const uint32_t* d_offsets, const int32_t* d_times,
                               const void* d_data, int64_t nnz, int64_t npix, int32_t dtype,
                               int32_t t0, int32_t t1, uint32_t* d_out_offsets,
                               int32_t* d_out_times, void* d_out_data, int64_t capacity,
                               void* stream) {
    if (!d_offsets || !d_times || !d_data || !d_out_offsets || !d_out_times || !d_out_data)
        return set_error(DXG2_EINVAL, "null argument");
    if (npix <= 0 || nnz < 0 || t0 < 0 || t1 < t0)
        return set_error(DXG2_EINVAL, "bad window [%d, %d) or sizes", t0, t1);
    if (npix > (int64_t)DXG2_STATIC_SCRATCH_WORDS * 1024)
        return set_error(DXG2_EINVAL, "%lld pixels exceed the static scratch limit of %lld "
                         "pixels; use dxg2_events_window with a workspace", (long long)npix,
                         (long long)DXG2_STATIC_SCRATCH_WORDS * 1024);
    if (dtype != DXG2_U8 && dtype != DXG2_U16)
        return set_error(DXG2_EINVAL, "event counts are uint8 or uint16");
    if (capacity < nnz)
        return set_error(DXG2_ECAPACITY, "capacity %lld < %lld input events (the window is "
                         "at most the input)", (long long)capacity, (long long)nnz);
    uint32_t* ws = nullptr;
    const int rc = static_scratch(&ws);
    if (rc != DXG2_OK) return rc;
    (void)hipGetLastError();
    launch_events_window(dtype, d_offsets, d_times, d_data, npix, t0, t1, ws, d_out_offsets,
                         d_out_times, d_out_data, (hipStream_t)stream);
    DXG2_HIP(hipGetLastError());
    return DXG2_OK;
}

int dxg2_compress_scratch_bytes(int32_t nframes, int32_t height, int32_t width, size_t* bytes) {
    if (!bytes || nframes <= 0 || height <= 0 || width <= 0)
        return set_error(DXG2_EINVAL, "bad compressor geometry");
    *bytes = compress_scratch_bytes(nframes, (int64_t)height * width);
    return DXG2_OK;
}

static int compress_args(const void* frames, int32_t dtype, int32_t nframes, int32_t height,
                         int32_t width) {
    if (!frames) return set_error(DXG2_EINVAL, "null frames");
    if (nframes <= 0 || height <= 0 || width <= 0)
        return set_error(DXG2_EINVAL, "bad frame stack (%d, %d, %d)", nframes, height, width);
    if (dtype != DXG2_U8 && dtype != DXG2_I8 && dtype != DXG2_U16)
        return set_error(DXG2_EINVAL, "unsupported dtype code %d", dtype);
    return DXG2_OK;
}

int dxg2_compress_count(const void* d_frames, int32_t dtype, int32_t nframes, int32_t height,
                        int32_t width, uint32_t* d_scratch, uint32_t* d_offsets, void* stream) {
    int rc = compress_args(d_frames, dtype, nframes, height, width);
    if (rc) return rc;
    if (!d_scratch || !d_offsets) return set_error(DXG2_EINVAL, "null scratch or offsets");
    (void)hipGetLastError();
    launch_compress_count(dtype, d_frames, nframes, (int64_t)height * width, d_scratch,
                          d_offsets, (hipStream_t)stream);
    DXG2_HIP(hipGetLastError());
    return DXG2_OK;
}

int dxg2_compress_fill(const void* d_frames, int32_t dtype, int32_t nframes, int32_t height,
                       int32_t width, const uint32_t* d_scratch, const uint32_t* d_offsets,
                       int32_t* d_times, void* d_data, void* stream) {
    int rc = compress_args(d_frames, dtype, nframes, height, width);
    if (rc) return rc;
    if (!d_scratch || !d_offsets || !d_times || !d_data)
        return set_error(DXG2_EINVAL, "null scratch, offsets or outputs");
    (void)hipGetLastError();
    launch_compress_fill(dtype, d_frames, nframes, (int64_t)height * width, d_scratch, d_offsets,
                         d_times, d_data, (hipStream_t)stream);
    DXG2_HIP(hipGetLastError());
    return DXG2_OK;
}

int dxg2_compress_push(const void* d_frame, int32_t dtype, int32_t height, int32_t width,
                       int32_t t, int32_t max_nnz, uint32_t* d_counter, void* d_events,
                       int32_t* d_times, int32_t* d_err, void* stream) {
    int rc = compress_args(d_frame, dtype, 1, height, width);
    if (rc) return rc;
    if (!d_counter || !d_events || !d_times || !d_err || max_nnz <= 0 || t < 0)
        return set_error(DXG2_EINVAL, "bad compressor state");
    (void)hipGetLastError();
    launch_compress_push(dtype, d_frame, (int64_t)height * width, t, max_nnz, d_counter,
                         d_events, d_times, d_err, (hipStream_t)stream);
    DXG2_HIP(hipGetLastError());
    return DXG2_OK;
}

int dxg2_compress_compact(const uint32_t* d_counter, const void* d_events,
                          const int32_t* d_times, int32_t dtype, int32_t height, int32_t width,
                          int32_t max_nnz, uint32_t* d_scratch, uint32_t* d_offsets,
                          void* d_out_data, int32_t* d_out_times, void* stream) {
    if (!d_counter || !d_events || !d_times || !d_scratch || !d_offsets || !d_out_data ||
        !d_out_times)
        return set_error(DXG2_EINVAL, "null argument");
    if (height <= 0 || width <= 0 || max_nnz <= 0)
        return set_error(DXG2_EINVAL, "bad compressor geometry");
    if (dtype != DXG2_U8 && dtype != DXG2_I8 && dtype != DXG2_U16)
        return set_error(DXG2_EINVAL, "unsupported dtype code %d", dtype);
    (void)hipGetLastError();
    launch_compress_compact(dtype, d_counter, d_events, d_times, (int64_t)height * width,
                            max_nnz, d_scratch, d_offsets, d_out_data, d_out_times,
                            (hipStream_t)stream);
    DXG2_HIP(hipGetLastError());
    return DXG2_OK;
}

int dxg2_measure_lds_atomics(int32_t mode, int32_t iters, double* adds_per_s, void* stream) {
    if (!adds_per_s || (mode != 0 && mode != 1) || iters <= 0)
        return set_error(DXG2_EINVAL, "bad argument");
    hipStream_t s = (hipStream_t)stream;
    int dev = 0, cus = 0;
    DXG2_HIP(hipGetDevice(&dev));
    DXG2_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int blocks = std::max(cus, 1) * 4 * 8;  // eight rounds of four workgroups per CU
    uint32_t* out = nullptr;
    DXG2_HIP(hipMalloc((void**)&out, (size_t)blocks * sizeof(uint32_t)));
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipError_t e = hipEventCreate(&ev[0]);
    if (e == hipSuccess) e = hipEventCreate(&ev[1]);
    float ms = 0.f;
    if (e == hipSuccess) {
        launch_lds_atomic_rate(mode, blocks, iters, out, s);  // warm-up
        e = hipEventRecord(ev[0], s);
    }
    if (e == hipSuccess) {
        launch_lds_atomic_rate(mode, blocks, iters, out, s);
        e = hipEventRecord(ev[1], s);
    }
    if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev[0], ev[1]);
    for (hipEvent_t x : ev)
        if (x) (void)hipEventDestroy(x);
    (void)hipFree(out);
    if (e != hipSuccess)
        return set_error(DXG2_EHIP, "LDS atomic measurement: %s", hipGetErrorString(e));
    *adds_per_s = (double)blocks * 256.0 * iters / (std::max(ms, 1e-6f) * 1e-3);
    return DXG2_OK;
}

int dxg2_events_to_coo(const int32_t* d_times, const void* d_data, const uint32_t* d_offsets,
                       int64_t nnz, int32_t height, int32_t width, int32_t nframes, int32_t dtype,
                       uint32_t* d_pixel, uint32_t* d_frame, void* d_count, void* stream) {
    if (!d_offsets || (nnz > 0 && (!d_times || !d_data || !d_pixel || !d_frame || !d_count)))
        return set_error(DXG2_EINVAL, "null argument");
    if (nframes <= 0) return set_error(DXG2_EINVAL, "bad frame count");
    hipStream_t s = (hipStream_t)stream;
    const int64_t hw = (int64_t)height * width;
    uint32_t *fcount = nullptr, *fcursor = nullptr, *tot = nullptr;
    DXG2_HIP(dev_alloc(DevAllocator{}, (void**)&fcount, (size_t)nframes * 4, s));
    DXG2_HIP(dev_alloc(DevAllocator{}, (void**)&fcursor, (size_t)(nframes + 1) * 4, s));
    DXG2_HIP(dev_alloc(DevAllocator{}, (void**)&tot, (size_t)((nframes + 1023) / 1024 + 1) * 4, s));
    (void)hipGetLastError();
    launch_fill_u32(fcount, nframes, s);
    launch_frame_count(d_times, nnz, nframes, fcount, s);
    launch_excl_scan(fcount, nframes, tot, fcursor, s);
    launch_to_coo(dtype, d_times, d_data, d_offsets, hw, nframes, fcursor, d_pixel, d_frame,
                  d_count, s);
    hipError_t e = hipStreamSynchronize(s);
    dev_free(DevAllocator{}, fcount, s);
    dev_free(DevAllocator{}, fcursor, s);
    dev_free(DevAllocator{}, tot, s);
    if (e != hipSuccess) return set_error(DXG2_EHIP, "to_coo failed: %s", hipGetErrorString(e));
    DXG2_HIP(hipGetLastError());
    return DXG2_OK;
}

}  // extern "C"
