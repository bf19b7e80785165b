// Internal declarations shared by the dxg2 translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/dxg2.h"

namespace dxg2 {

// Slab geometry: every bin's gathered pixels are stored as i8, grouped in k-blocks of
// KB = 64 pixels; a k-block holds all nrows frames (nframes rounded up to
// SLAB_ROW_ALIGN; rows t >= nframes are zero).  Inside a k-block the bytes are laid out
// for the MFMA operand loads: frames form groups of 32, and each (group g, half ks) is
// one 1 KiB run in which lane l = 32 h + r holds the 16 bytes 16 (2 ks + h) .. +16 of
// frame 32 g + r (slab_offset()).  A k-block is nrows * 64 contiguous bytes.
constexpr int KB = 64;
constexpr int SLAB_ROW_ALIGN = 256;
// Largest number of k-blocks one Gram work item accumulates in int32 before it
// flushes: |a*b| <= 2^14 for i8 operands, 2^14 * 2047 * 64 < 2^31.
constexpr int MAX_KBLOCKS_PER_ITEM = 2047;

// Byte offset of (k-block kb, frame t, byte kk < 64 of the k-block) in a bin's slab.
__host__ __device__ inline int64_t slab_offset(int kb, int t, int kk, int nrows) {
    const int q = kk >> 4;
    return (((int64_t)kb * (nrows >> 5) + (t >> 5)) * 2 + (q >> 1)) * 1024 +
           ((q & 1) * 32 + (t & 31)) * 16 + (kk & 15);
}

struct BinDev {
    int64_t slab_off;   // byte offset of the bin's slab
    int64_t idx_off;    // offset of the bin's pixel index list (padded to kpad)
    int32_t npix;       // n_b
    int32_t nkb;        // k-blocks (kpad / KB)
    double scale;       // g2 scale factor
};

struct GatherWork {      // one workgroup of the gather kernel (x dimension)
    int32_t bin;
    int32_t kb0;         // first k-block
    int32_t nkb;         // k-blocks handled (<= 16)
    int32_t pad;
};

// Band gather: a workgroup of 4 waves stages BAND_STAGE frames of a band's rows in LDS
// (BAND_STAGE x srow bytes, srow <= BAND_ROW_BYTES so that two workgroups share a CU);
// each wave keeps BAND_KMAX k-blocks in registers, a work item at most BAND_ITEM_KMAX.
constexpr int BAND_STAGE = 8;
// Gather work in detector order, dealt to the XCDs in chunks of this many items:
// dense gather items (up to 16 k-blocks each) and pixel-major gather k-blocks.
constexpr int GATHER_XCD_GROUP = 16;
constexpr int PM_XCD_GROUP = 128;
constexpr int PM_KPB = 2;  // k-blocks per workgroup of the pixel-major gather
constexpr int BAND_ROW_BYTES = 9728;
constexpr int BAND_KMAX = 36;
constexpr int BAND_ITEM_KMAX = 4 * BAND_KMAX;
constexpr int GFX950_LDS_BYTES = 160 * 1024;  // LDS per CU

struct BandWork {        // one band of detector rows of the band gather (x dimension)
    int32_t y0;          // first staged row
    int32_t rows;        // rows staged (the band plus its overlap, clipped at H)
    int32_t k0;          // k-blocks [k0, k1) of the (bin, k-block) list
    int32_t k1;
};

struct GramWork {        // one output tile (i, j >= i) over a k-block range
    int32_t bin;
    int32_t ti;          // row tile
    int32_t tj;          // column tile
    int32_t kb0;
    int32_t kb1;         // exclusive
    int32_t out;         // row of the numerator (and frame-sum) outputs it adds into
    int32_t cross;       // 1: B operand is the other digit plane ([H|L] x [L|H]^T)
    int32_t job;         // grouped launches over several plans: index into GramJob[]
    int32_t slot;        // calc_std launches: tile scratch slot of this item
    int32_t pad;
};

struct GramJob {         // one plan's buffers in a grouped Gram launch
    const int8_t* slab;
    const BinDev* bins;
    int64_t* ugram;
    int64_t* sums;
    int32_t nframes;
    int32_t nrows;
    int32_t biased;
    int32_t pad;
};

struct StdTile {         // calc_std: one upper-triangle tile of one bin
    int32_t bin;
    int32_t ti;
    int32_t tj;
    int32_t slot0;       // first scratch slot of its work items (HH, then LL, then cross)
    int32_t n_hh;        // items of each kind (uint8/int8: all items are n_hh)
    int32_t n_ll;
    int32_t n_x;
    int32_t idx;         // bin * T(T+1)/2 + row-major upper-triangle index of (ti, tj)
};

struct StdPart {         // one diagonal segment of one tile: count, mean, sum of squared
    double n;            // deviations of its ratios
    double mean;
    double m2;
};

// Digit planes of a bin's slab: uint16 frames are split into a high-byte plane H
// (k-blocks 0..nkb-1) and a low-byte plane L (k-blocks nkb..2nkb-1), each stored
// biased by -128 like uint8.  Numerators combine as 65536 HH + 256 (HL + LH) + LL.
inline int slab_planes(int dtype) { return dtype == DXG2_U16 ? 2 : 1; }

struct EventWork {       // one workgroup of the event pair kernel (x dimension)
    int32_t bin;
    int32_t k0;          // range [k0, k1) of the bin's pixel list
    int32_t k1;
    int32_t pad;         // job of a grouped launch (0 otherwise)
};

struct EventJob {        // one job of a grouped event launch
    const int32_t* times;
    const void* data;
    const uint32_t* offsets;
    const int32_t* pix;
    const BinDev* bins;
    int64_t nnz;
    int64_t* num;        // the job's int64 numerator accumulator (n_bins, nframes)
    int64_t* sums;
    int* err;
    int32_t nframes;
    int32_t pad;
};

struct KbWork {          // one k-block (64 pixels) of one bin
    int32_t bin;
    int32_t kb;
};

// ------------------------------------------------------------------ launchers
void launch_lds_atomic_rate(int mode, int blocks, int iters, uint32_t* out, hipStream_t s);
void launch_fill_i64(int64_t* p, int64_t n, hipStream_t s);
void launch_gather_slots(const uint8_t* frames, int dtype, int64_t hw, int nframes, int nrows,
                         const int32_t* pix_idx, const BinDev* bins, const KbWork* kbl,
                         int n_items, int tf, int8_t* slab, hipStream_t s, uint32_t* flagw);
void launch_wide_gather(const uint16_t* frames, int64_t hw, int nframes, int nrows,
                        const int32_t* pix, const int32_t* pix_w, const BinDev* bins, int n_bins,
                        const int* nwkb, int8_t* slab, hipStream_t s);
// Band gather of uint8/int8 frames: stage `rows` detector rows of 4 frames in LDS
// (srow bytes per frame), write every k-block of the band from there.
void launch_gather_band(const uint8_t* frames, int dtype, int64_t hw, int width, int nframes,
                        int nrows, int srow, const int32_t* pix_idx, const BinDev* bins,
                        const BandWork* bands, int n_bands, const int2* band_kb,
                        const GatherWork* fall, int n_fall, int8_t* slab, hipStream_t s);
// cross: the work items are cross-plane ([H|L] [L|H]^T) items of uint16 frames.
// nwkb (uint16 slabs with the wide pixels apart, else null): per bin, the k-blocks the
// HH and cross items are clipped to (item_range in k_gram.hip).
void launch_gram_i8(int tile, bool cross, const int8_t* slab, const BinDev* bins,
                    const GramWork* work, int n_work, int nframes, int nrows, int biased,
                    const int* nwkb, int64_t* ugram, int64_t* sums, hipStream_t s);
// Persistent launch over per-XCD work queues: queue x is items [qoff[x], qoff[x+1]) of
// `work`; heads (8 ints) must be zero.
void launch_gram_i8_queue(int tile, bool cross, const int8_t* slab, const BinDev* bins,
                          const GramWork* work, int n_work, const int* qoff, int* heads,
                          int nframes, int nrows, int biased, const int* nwkb, int64_t* ugram,
                          int64_t* sums, hipStream_t s);
// uint16 wide pixels apart (k_wide.hip): from the flags of the pixels with a count > 255
// in some frame, each bin's wide pixels listed in pix_w, their count and clip = nwkb per
// bin (nwkb[n_bins] the largest clip, nwkb[n_bins + 1 + b] the counts).
void launch_wide_partition(const int32_t* pix, const BinDev* bins, int n_bins, int max_npix,
                           const uint8_t* flag, int32_t* pix_w, int* nwkb, hipStream_t s);
// Low-byte frame sums over the nwkb k-blocks each bin's cross term reads them from.
void launch_wide_low_sums(const int8_t* slab, const BinDev* bins, const int* nwkb, int n_bins,
                          int nframes, int nrows, int64_t* lsums, hipStream_t s);
// Device queries of the persistent launch's grid size, done once per device (outside
// any stream capture).
void gram_queue_prepare(int tile);
// Grouped launch: work items of several plans, each naming its job descriptor.
void launch_gram_i8_jobs(int tile, bool cross, const GramJob* jobs, const GramWork* work,
                         int n_work, hipStream_t s);
// Workspace of the i8-MFMA denominator (k_den.hip): v holds n_bins x 2 x 64 x up bytes
// (two digit planes per bin); umax is the largest max |S| taken as one biased digit (255
// for unsigned frames, 127 for signed: none); 0 switches the path off.
struct DenMfma {
    int8_t* v = nullptr;
    int64_t up = 0;
    int umax = 0;
};
// Columns of V per row for N frames (0: more than 32,768 rows of 64 frames, not taken).
int64_t den_mfma_columns(int nframes);
// Largest max |S| of a bin the i8-MFMA path takes (two digits).
int den_mfma_max();
void launch_den_mfma(const int64_t* sums, int n_bins, int nframes, const int64_t* smax,
                     const int64_t* prefix, int planes, const DenMfma& dm, double* den,
                     hipStream_t s);
// mode 0: numerators as accumulated; 1: uint8 bias correction; 2: uint16 digit planes
// (ugram holds HH, LL, cross rows; plane_sums the per-plane S_H, S_L rows, and `sums`
// receives S = 256 S_H + S_L).
void launch_finalize(const int64_t* plane_sums, int64_t* sums, const int64_t* ugram,
                     const BinDev* bins, int n_bins, int nframes, int mode, int64_t* prefix,
                     int64_t* smax, double* dpre, const int* nwkb, int64_t* num, double* den,
                     float* g2, const DenMfma& dm, hipStream_t s);
void launch_plain_bin(const uint8_t* frames, int dtype, int64_t hw, int nframes,
                      const int32_t* pix, int npix, float* a, float* gram, float* means,
                      float* g2_row, hipStream_t s);
// Lags per workgroup of the event pair kernel (its lag window) for a count type.
int event_lag_window(int dtype);
void launch_event_pairs_jobs(int dtype, const EventJob* jobs, const EventWork* work, int n_work,
                             int max_frames, hipStream_t s);
void launch_event_pairs(int dtype, const int32_t* times, const void* data,
                        const uint32_t* offsets, const int32_t* pix_idx, const BinDev* bins,
                        const EventWork* work, int n_work, int nframes, int64_t nnz,
                        int64_t* num, int64_t* sums, int* err, hipStream_t s);
// Workspace of the COO -> CSR sort: sorted keys, packed (frame, count) words before and
// after the sort (4 bytes each when coo_narrow_values, else 8), and the sort's scratch.
struct CooSortBufs {
    uint32_t* keys;
    void* vals0;
    void* vals1;
    void* tmp;
    size_t tmp_bytes;
};
bool coo_narrow_values(int dtype, int nframes);
size_t coo_sort_bytes(int dtype, int nframes, int64_t cap, uint32_t npix);
hipError_t launch_coo_to_csr(int dtype, const uint32_t* pix, const uint32_t* frame,
                             const void* cnt, int64_t nnz, uint32_t npix, int nframes,
                             uint32_t* offsets, const CooSortBufs& sort, int32_t* times,
                             void* data, int* err, hipStream_t s);
void launch_fill_i32(int* p, int64_t n, hipStream_t s);
void launch_finalize_f64(const int64_t* sums, const double* fnum, const BinDev* bins, int n_bins,
                         int nframes, int64_t* prefix, int64_t* smax, double* dpre, double* num,
                         double* den, float* g2, const DenMfma& dm, hipStream_t s);
// Pixel-major copy xs of every bin's series (in the order of pix_idx) and the exact frame
// sums (sums[b][t] for t < nframes, zeroed here first).  work: n_work items of PM_KPB
// k-blocks (kb -1: none).  quad_kb (8-bit frames; may be null): per bin, the leading
// k-blocks whose entries are 16 aligned pixel quads.
void launch_gather_pm(const uint8_t* frames, int dtype, int64_t hw, int nframes, int nrows,
                      const int32_t* pix_idx, const int32_t* quad_kb, const BinDev* bins,
                      int n_bins, const KbWork* work, int n_work, uint8_t* xs, int64_t* sums,
                      hipStream_t s);
int fft_lds_length(int n);
int fft_four_step_length(int n, int* L1, int* L2);
int launch_fft_four_step(int L1, int L2, int dtype, const uint8_t* xs, const BinDev* bins,
                         const KbWork* work, int kb0, int n_kb, int nrows, const float2* tw,
                         const float2* uw, float2* y, double* power, hipStream_t s);
int launch_fft_power(int L, int dtype, const uint8_t* xs, const BinDev* bins, const KbWork* work,
                     int n_work, int nrows, const float2* tw, const float2* uw, double* power,
                     hipStream_t s);
void launch_fft_inverse(const double* power, int L, int n_bins, int nframes, double* num,
                        hipStream_t s);
void launch_excl_scan(const uint32_t* in, int64_t n, uint32_t* tot, uint32_t* out, hipStream_t s);
void launch_events_window(int dtype, const uint32_t* offsets, const int32_t* times,
                          const void* data, int64_t npix, int t0, int t1, uint32_t* scratch,
                          uint32_t* out_offsets, int32_t* out_times, void* out_data,
                          hipStream_t s);
void launch_events_to_columns(int dtype, const uint32_t* offsets, const int32_t* times,
                              const void* data, int64_t npix, const int32_t* cols,
                              int64_t n_cols, int nframes, void* out, int* err, hipStream_t s);
void launch_copy_u32(const uint32_t* src, uint32_t* dst, int64_t n, hipStream_t s);
void launch_fill_u32(uint32_t* p, int64_t n, hipStream_t s);
void launch_synth_events(int pass, int dtype, int nframes, int64_t hw, const float* mean_map,
                         float rho, uint64_t seed, uint32_t* count, const uint32_t* offsets,
                         int32_t* times, void* data, hipStream_t s);
void launch_frame_count(const int32_t* times, int64_t nnz, int nframes, uint32_t* fcount,
                        hipStream_t s);
void launch_to_coo(int dtype, const int32_t* times, const void* data, const uint32_t* offsets,
                   int64_t npix, int nframes, uint32_t* fcursor, uint32_t* pix, uint32_t* frame,
                   void* cnt, hipStream_t s);
void launch_synth(void* frames, int dtype, int nframes, int64_t hw, const float* mean_map,
                  float rho, uint64_t seed, hipStream_t s);

// calc_std launches: the Gram that also stores every item's int32 tile in its slot,
// frame sums from the slab, per-tile diagonal segments of the ratios, and dev.
void launch_gram_i8_store(int tile, bool cross, const int8_t* slab, const BinDev* bins,
                          const GramWork* work, int n_work, int nframes, int nrows, int biased,
                          int64_t* ugram, int* tiles, hipStream_t s);
void launch_slab_sums(const int8_t* slab, int dtype, const BinDev* bins, const GatherWork* work,
                      int n_work, int n_bins, int nframes, int nrows, int64_t* sums,
                      hipStream_t s);
void launch_std_tiles(int tile, const StdTile* tiles, int n_tiles, const BinDev* bins,
                      const int* scratch, const int64_t* sums, int mode, int n_bins, int nframes,
                      StdPart* part, hipStream_t s);
void launch_dev_finalize(const StdPart* part, const BinDev* bins, int n_bins, int nframes,
                         int tile, float* dev, hipStream_t s);

// Two-time correlation matrix of one bin from stored Gram tiles, then its diagonal.
void launch_ttcf_tiles(int tile, const StdTile* tiles, int n_tiles, const BinDev* bins,
                       const int* scratch, const int64_t* sums, int mode, int n_bins,
                       int nframes, float* trc, hipStream_t s);
void launch_ttcf_diag(float* trc, int nframes, const int64_t* sums, int mode, int n_bins, int b,
                      hipStream_t s);
void launch_fill_f32(float* p, int64_t n, float v, hipStream_t s);
// Frames -> per-pixel event lists (FramesCompressor layout).  scratch holds
// compress_scratch_bytes(); the count pass leaves offsets (hw + 1 entries) and the
// per-chunk bases in it for the fill pass.
size_t compress_scratch_bytes(int nframes, int64_t hw);
void launch_compress_count(int dtype, const void* frames, int nframes, int64_t hw,
                           uint32_t* scratch, uint32_t* offsets, hipStream_t s);
void launch_compress_fill(int dtype, const void* frames, int nframes, int64_t hw,
                          const uint32_t* scratch, const uint32_t* offsets, int32_t* times,
                          void* data, hipStream_t s);
void launch_compress_push(int dtype, const void* frame, int64_t hw, int t, int max_nnz,
                          uint32_t* counter, void* events, int32_t* times, int* err,
                          hipStream_t s);
void launch_compress_compact(int dtype, const uint32_t* counter, const void* events,
                             const int32_t* times, int64_t hw, int max_nnz, uint32_t* scratch,
                             uint32_t* offsets, void* out_data, int32_t* out_times,
                             hipStream_t s);
void launch_clamp_u32(uint32_t* p, int64_t n, uint32_t hi, hipStream_t s);

int gram_tile_size(int variant);

}  // namespace dxg2
