/*
 * dxg2 -- MI355X (gfx950) XPCS g2 correlators: C ABI.
 *
 * One shared library, libdxg2.so, built from dynamix_amd/csrc with hipcc.  Every entry
 * point takes plain pointers and sizes; device pointers are borrowed (never freed) and
 * every call is ordered on the HIP stream passed as `stream` (a hipStream_t, NULL for the
 * default stream).  Functions return 0 on success or a DXG2_E* status; dxg2_last_error()
 * then holds a thread-local message.  A plan is used by one host thread at a time.
 *
 * Output layout of every correlate call (device memory, row-major):
 *   num  int64  (n_bins, nframes)   Num_b(tau), exact
 *   sums int64  (n_bins, nframes)   S_b(t) per-frame bin sums, exact
 *   den  double (n_bins, nframes)   Den_b(tau) = sum_t S_b(t) S_b(t+tau) (exact below 2^53)
 *   g2   float  (n_bins, nframes)   scale_b * Num_b(tau) / Den_b(tau), tau = 0 in column 0
 * Any of num/sums/den may be NULL.
 *
 * Reference interfaces replaced (silx-kit/dynamix 0.2.0, paths under dynamix/):
 *   see the comment above each declaration.
 */
#ifndef DXG2_H
#define DXG2_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define DXG2_OK 0
#define DXG2_EINVAL 1      /* bad argument: maps to AssertionError/ValueError in Python */
#define DXG2_ENOMEM 2      /* device allocation failed */
#define DXG2_EHIP 3        /* HIP runtime error */
#define DXG2_EUNSUPPORTED 4
#define DXG2_ECAPACITY 5   /* more events than the plan's capacity (event.py:110-115) */

/* element types of frames / event counts */
#define DXG2_U8 1
#define DXG2_I8 2
#define DXG2_U16 3

/* plan flags */
#define DXG2_FLAG_NONE 0u
#define DXG2_FLAG_TILE128 1u   /* Gram tiles of 128 frames (default 256) */
#define DXG2_FLAG_NO_QUEUE 2u  /* one workgroup per Gram tile instead of the persistent
                                  launch: leaves CUs free between tiles for kernels of
                                  other streams (serving) */

typedef struct dxg2_plan dxg2_plan;

/* Device memory of plans and batches.  alloc returns at least `bytes` of device memory
 * for use on `stream` (a hipStream_t), aligned to 256 bytes, or NULL (a pointer that is
 * not so aligned is freed and counts as a failed allocation); free releases a pointer
 * alloc returned.
 * The library frees a plan's buffers only when no call on the plan can still be running
 * (dxg2_plan_destroy waits for the device; with dxg2_plan_destroy_async the caller
 * guarantees it).  dxg2_set_allocator(NULL, NULL, NULL) restores the built-in hipMalloc /
 * hipFree (hipFree waits for every kernel in flight).  Plans and batches keep the
 * allocator that was current when they were created.  A caching allocator whose frees
 * never touch the device lets a serving loop create and release plans while other jobs
 * run.  No reference counterpart. */
typedef void* (*dxg2_alloc_fn)(size_t bytes, void* stream, void* ctx);
typedef void (*dxg2_free_fn)(void* ptr, void* stream, void* ctx);
int dxg2_set_allocator(dxg2_alloc_fn alloc, dxg2_free_fn free_fn, void* ctx);

typedef struct dxg2_plan_desc {
    int32_t height;               /* frame rows (H) */
    int32_t width;                /* frame columns (W) */
    int32_t nframes;              /* N */
    int32_t dtype;                /* DXG2_U8 / DXG2_I8 / DXG2_U16 */
    const int32_t* qmask;         /* host, H*W; bins 1..max, <= 0 excluded; NULL = one bin */
    const double* scale_factors;  /* host, n_bins values; NULL = pixel count of each bin */
    int32_t n_scale_factors;      /* length of scale_factors (must equal n_bins if given) */
    uint32_t flags;
} dxg2_plan_desc;

/* Plan: qmask -> bins, per-bin pixel lists, slab layout, tile work lists, workspaces.
 * Replaces BaseCorrelator._set_parameters/_set_qmask/_set_scale_factor
 * (correlator/common.py:22-72) and the constructors MatMulCorrelator.__init__
 * (correlator/dense.py:120-126), DenseCorrelator.__init__ (dense.py:149-213),
 * FFTCorrelator.__init__ (dense.py:277-300), EventCorrelator.__init__ (event.py:15-101). */
int dxg2_plan_create(const dxg2_plan_desc* desc, dxg2_plan** out);
/* Runs the host half of dxg2_plan_create (validation, bins, pixel lists, slab layout,
 * every work list) with no device call, and writes up to n of: n_bins, padded pixel
 * count, slab bytes, gather items, band items, band k-blocks, k-blocks on the band
 * gather's fallback, Gram items, Gram items of the queue schedule, event items, calc_std
 * tiles, calc_std chunk items.  Returns the error dxg2_plan_create would return for the description.
 * For checking the plan builder off the GPU (sanitizers, fuzzing).  No reference
 * counterpart. */
int dxg2_plan_dry_run(const dxg2_plan_desc* desc, int64_t* summary, int32_t n);
/* Waits for the device, then releases the plan. */
int dxg2_plan_destroy(dxg2_plan* plan);
/* Releases the plan without waiting for the device: its buffers go to the plan's
 * allocator's free with `stream`.  The caller guarantees that no call on the plan is
 * still running.  With the built-in allocator this waits like dxg2_plan_destroy (hipFree);
 * with a caching allocator it returns at once.  For serving loops.  No reference
 * counterpart. */
int dxg2_plan_destroy_async(dxg2_plan* plan, void* stream);

/* Host-only queries (no device work).  pixel_counts / scale may be NULL.
 * Mirrors the n_bins / bins / scale_factors attributes (common.py:36-72). */
int dxg2_plan_query(const dxg2_plan* plan, int32_t* n_bins, int64_t* pixel_counts,
                    double* scale);

/* Device bytes the plan holds for a path (0 dense, 1 plain, 2 fft, 3 events, 4 the
 * extra workspaces of dense calc_std), for admission control by memory footprint.
 * No reference counterpart. */
int dxg2_plan_footprint(const dxg2_plan* plan, int32_t path, size_t* bytes);

/* Dense batched path: qmask gather into per-bin i8 slabs (+ S_b(t)), one grouped
 * i8-MFMA Gram launch over every bin with the diagonal (tau) reduction fused, exact
 * denominator and FP64 normalisation.  d_frames: (nframes, H, W) of the plan dtype.
 * uint16 frames are byte-split into two i8 digit planes and the numerator is combined
 * as 65536 HH + 256 (HL + LH) + LL from three Gram passes in the same launch; num is
 * exact while it stays below 2^63 (g2 is formed from the exact 128-bit value).
 * Replaces MatMulCorrelator.correlate (dense.py:128-142, numerator dense.py:32-70) and
 * DenseCorrelator.correlate (dense.py:223-272 with densecorrelator.cl:6-176). */
int dxg2_correlate_dense(dxg2_plan* plan, const void* d_frames, int64_t* d_num,
                         int64_t* d_sums, double* d_den, float* d_g2, void* stream);

/* Dense path with calc_std: everything dxg2_correlate_dense computes, plus d_dev
 * (n_bins, N) float32, dev(tau) = population std over t of the ratios
 * r_t = n_b G(t, t+tau) / (S(t) S(t+tau)) divided by sqrt(N - tau), NaN where a frame sum
 * on the diagonal is zero or the bin is empty.  Each Gram work item also writes its
 * int32 tile to a scratch buffer (at most DXG2_STD_SCRATCH_MB MiB, default 4096; the
 * launch is split into chunks of tiles that fit), and a per-tile pass forms r in FP64
 * and keeps per-diagonal (count, mean, M2) segments, combined in a fixed order: dev is
 * reproducible bit for bit.  d_num, d_sums, d_den, d_g2 are as for dxg2_correlate_dense.
 * Replaces MatMulCorrelator.correlate(frames, calc_std=True) (dense.py:128-142; the
 * std at dense.py:58-59,65-66). */
int dxg2_correlate_dense_std(dxg2_plan* plan, const void* d_frames, int64_t* d_num,
                             int64_t* d_sums, double* d_den, float* d_g2, float* d_dev,
                             void* stream);

/* Two-time correlation matrix of bin q (1-based), d_trc: (N, N) float32 row-major,
 * trc(t, u) = n_b G(t, u) / (S(t) S(u)) for t != u, with n_b every pixel of the bin; the
 * diagonal 0 .. N-2 holds the mean of the positive entries of the first off-diagonal and
 * entry (N-1, N-1) is n_b (G(N-1, N-1) - S(N-1)) / S(N-1)^2, the photon pairs within the
 * last frame (NaN wherever a frame sum is zero).  Computed in FP64 from the exact integer
 * Gram tiles of that bin (the tiles are stored in a scratch of at most
 * DXG2_STD_SCRATCH_MB MiB).  Replaces the ttcf_par output `trc` of ecorrts_q
 * (correlator/event_y.py:169-230 with WXPCS/fecorrt.f:1-26). */
int dxg2_correlate_dense_ttcf(dxg2_plan* plan, const void* d_frames, int32_t q, float* d_trc,
                              void* stream);

/* The dense chain of dxg2_correlate_dense (gather, persistent Gram, finalise) captured
 * once into a HIP graph and replayed with one hipGraphLaunch on `stream`.  The graph is
 * kept in the plan for the given buffer pointers: a call with other pointers captures
 * again; dxg2_plan_set_tuning drops it.  Results are identical to dxg2_correlate_dense.
 * The capture contains kernels only (every clear is a fill kernel).
 * Replaces the per-call launch sequence of DenseCorrelator.correlate (dense.py:223-247);
 * the graph itself has no reference counterpart. */
int dxg2_correlate_dense_graph(dxg2_plan* plan, const void* d_frames, int64_t* d_num,
                               int64_t* d_sums, double* d_den, float* d_g2, void* stream);

/* Gram launch tuning of a plan: tile size (128 or 256 frames), raster band (tile rows),
 * XCD chunk (consecutive items per XCD, 0 = none), launch form (1 = persistent
 * workgroups over per-XCD work queues, 0 = one workgroup per tile) and k-chunk (k-blocks
 * of 64 pixels per work item, 1..2047; 0 on set = the dtype's default: 512 for uint16,
 * 2047 otherwise).  set_tuning waits for the device, rebuilds the work lists and drops
 * the plan's graph and calc_std workspaces.  Used by the autotuner
 * (dynamix_amd.autotune).  No reference counterpart: the reference takes block_size for
 * its OpenCL kernels (dense.py:153,163). */
typedef struct dxg2_tuning {
    int32_t tile;
    int32_t band;
    int32_t group;
    int32_t queue;
    int32_t kchunk;
} dxg2_tuning;
int dxg2_plan_get_tuning(const dxg2_plan* plan, dxg2_tuning* out);
int dxg2_plan_set_tuning(dxg2_plan* plan, const dxg2_tuning* tuning);

/* Grouped dense launch over several plans (continuous batching of dense jobs): each
 * plan's gather, then ONE Gram launch over the work items of every plan (plus one for the
 * cross-plane items of uint16 plans), then each plan's finalise, all on `stream`.  Plans
 * must share the Gram tile size and appear once each.  Array arguments hold one entry per
 * job (d_num, d_sums, d_den may be NULL arrays).  Results equal those of
 * dxg2_correlate_dense on each plan alone.  The batch object owns the device job table
 * and work list (grown as needed) and is used by one host thread at a time.
 * No reference counterpart: the serving loop's grouped submission. */
typedef struct dxg2_batch dxg2_batch;
int dxg2_batch_create(dxg2_batch** out);
int dxg2_batch_destroy(dxg2_batch* batch);
int dxg2_batch_correlate_dense(dxg2_batch* batch, int32_t n, dxg2_plan* const* plans,
                               const void* const* d_frames, int64_t* const* d_num,
                               int64_t* const* d_sums, double* const* d_den,
                               float* const* d_g2, void* stream);
/* Grouped event launch over several plans (continuous batching of event jobs): ONE pair
 * launch over the work items of every job (each item tagged with its job, lag windows
 * up to the longest series), then each plan's denominator and finalise, on `stream`.
 * Per job: pixel-major CSR input as for dxg2_correlate_events_csr (d_times, d_data,
 * d_offsets, nnz[i] events on the host).  Plans must share the count type and appear
 * once each; d_num, d_sums, d_den may be NULL arrays.  Results, and each plan's
 * dxg2_plan_event_status, equal those of dxg2_correlate_events_csr on each job alone.
 * No reference counterpart: the serving loop's grouped submission. */
int dxg2_batch_correlate_events(dxg2_batch* batch, int32_t n, dxg2_plan* const* plans,
                                const int32_t* const* d_times, const void* const* d_data,
                                const uint32_t* const* d_offsets, const int64_t* nnz,
                                int64_t* const* d_num, int64_t* const* d_sums,
                                double* const* d_den, float* const* d_g2, void* stream);

/* Plain comparator: one bin per launch, FP32 gather, FP32 full-square Gram, FP32
 * diagonal sums and FP32 mean outer-product denominator.
 * Replaces CublasMatMulCorrelator.correlate (cuda.py:101-125). */
int dxg2_correlate_dense_plain(dxg2_plan* plan, const void* d_frames, float* d_g2,
                               void* stream);

/* FFT path for uint8 frames: per bin, the power spectra of every pixel's time series
 * (FP32 Stockham FFT of length Nf = 2L, L = smallest power of two >= max(N, 1024),
 * L <= 16384) summed in FP64, one FP64 inverse per bin, exact integer denominators.
 * d_num receives the FP64 numerators (n_bins, N) (may be NULL); other outputs as for
 * the dense path.  Numerators carry FP32 FFT rounding: see the documented tolerance.
 * Replaces FFTWCorrelator.correlate / _correlate_fft (correlator/dense.py:314-369) and
 * CUFFTCorrelator (correlator/cuda.py:129-219). */
int dxg2_correlate_fft(dxg2_plan* plan, const void* d_frames, double* d_num, int64_t* d_sums,
                       double* d_den, float* d_g2, void* stream);

/* Event (sparse) path, pixel-major CSR exactly as the reference's EventCorrelator takes
 * it: d_times int32[nnz] frame indices, strictly increasing within a pixel; d_data
 * [nnz] counts of the plan dtype; d_offsets uint32[H*W+1], pixel p owns events
 * [offsets[p], offsets[p+1]).  Numerators are exact (64-bit), sums/denominators as for
 * the dense path, g2 uses the plan's scale factors (EventCorrelator honours
 * scale_factor, evtcorrelator.cl:74).  Input errors are recorded on the device and
 * reported by dxg2_plan_event_status.  Numerators are recomputed from zero on every
 * call (the reference keeps accumulating res_int across calls, event.py:98,170-171).
 * Replaces EventCorrelator.correlate (correlator/event.py:162-204) with
 * event_correlator/normalize_correlation (resources/opencl/evtcorrelator.cl:3-76). */
int dxg2_correlate_events_csr(dxg2_plan* plan, const int32_t* d_times, const void* d_data,
                              const uint32_t* d_offsets, int64_t nnz, int64_t* d_num,
                              int64_t* d_sums, double* d_den, float* d_g2, void* stream);

/* Event path from frame-sorted COO: d_pixel uint32 raster pixel index, d_frame uint32,
 * d_count uint8 (or uint16 for a DXG2_U16 plan), nnz events.  Converted to the CSR
 * layout above on the device (per-pixel counts and scan; a stable radix sort by pixel,
 * which keeps each pixel's events in input order), then as above.  Events of one pixel
 * must therefore come in increasing frame order; a list that breaks this is reported by
 * dxg2_plan_event_status (flag bit 0).
 * The layout FramesCompressor.compress_all_stack produces (correlator/event.py:267-284)
 * is the CSR form; this COO form has no reference counterpart. */
int dxg2_correlate_events_coo(dxg2_plan* plan, const uint32_t* d_pixel, const uint32_t* d_frame,
                              const void* d_count, int64_t nnz, int64_t* d_num, int64_t* d_sums,
                              double* d_den, float* d_g2, void* stream);

/* Waits for `stream` and returns the input-error flags of the last event call: bit 0
 * frame out of [0, nframes) or not strictly increasing within a pixel, bit 1 offsets
 * decreasing or past nnz, bit 2 COO pixel index past H*W.  0 = input valid.  The
 * reference raises AssertionError/ValueError on bad shapes (event.py:104-115). */
int dxg2_plan_event_status(dxg2_plan* plan, int32_t* flags, void* stream);

/* Stage timing with HIP events recorded on the call's stream (no host sync per call).
 * dxg2_plan_stage_times waits for the last recorded call and returns the averages over
 * every call since profiling was switched on or last read, in ms: times[0] gather,
 * [1] gram, [2] denominator+finalise, [3] total, [4] number of calls averaged; then
 * it resets the record.  Event calls: [0] COO -> CSR conversion (0 for CSR input),
 * [1] pairs and frame sums, [2] denominator+finalise.  Counterpart of the reference's OpenCL event profiling
 * (profile=True, profile_add in correlator/dense.py:244,261,272). */
int dxg2_plan_set_profiling(dxg2_plan* plan, int32_t on);
int dxg2_plan_stage_times(dxg2_plan* plan, float* times, int32_t n);

/* LDS atomic-add throughput of this GPU (32-bit adds per second, summed over the
 * chip), measured with workgroups of the event pair kernel's LDS footprint: mode 0
 * random counters, mode 1 lane-consecutive counters, `iters` adds per lane.  The
 * ceiling the event path's histogram updates are compared with.  Waits for the
 * device.  No reference counterpart. */
int dxg2_measure_lds_atomics(int32_t mode, int32_t iters, double* adds_per_s, void* stream);

/* Synthetic speckle frames on the device: per-pixel AR(1) complex Gaussian field with
 * correlation time tau_c frames, counts ~ Poisson(mean_map[p] * |E|^2) clipped to the
 * dtype, counter-based Philox-4x32-10 random numbers keyed by seed.  Test-data
 * generator; no reference counterpart. d_mean_map: float, H*W (device). */
int dxg2_synth_speckle(void* d_frames, int32_t dtype, int32_t nframes, int32_t height,
                       int32_t width, const float* d_mean_map, double tau_c,
                       uint64_t seed, void* stream);

/* Synthetic events: the same per-pixel walk as dxg2_synth_speckle, emitting only the
 * non-zero samples in the pixel-major layout (so they equal the compaction of the dense
 * stack with the same seed).  Always writes d_offsets (H*W+1) and *nnz (waits for the
 * stream); writes d_times/d_data too when d_times is non-NULL and capacity >= *nnz.
 * Test-data generator; no reference counterpart. */
int dxg2_synth_events(const float* d_mean_map, int32_t height, int32_t width, int32_t nframes,
                      double tau_c, uint64_t seed, int32_t dtype, uint32_t* d_offsets,
                      int32_t* d_times, void* d_data, int64_t capacity, int64_t* nnz,
                      void* stream);

/* Frame window of a pixel-major event list: the events with frames in [t0, t1), frames
 * re-based to t - t0, into d_out_* (offsets H*W+1 = npix+1 entries; capacity >= nnz, the
 * window never holds more events than the input).  d_scratch: (npix + 1023) / 1024 + 1
 * uint32 of device workspace.  Needs the input's frames ascending within each pixel.
 * Stream-ordered; no host synchronisation.  No reference counterpart (feeds jobs that
 * cover part of a recorded series). */
int dxg2_events_window(const uint32_t* d_offsets, const int32_t* d_times, const void* d_data,
                       int64_t nnz, int64_t npix, int32_t dtype, int32_t t0, int32_t t1,
                       uint32_t* d_scratch, uint32_t* d_out_offsets, int32_t* d_out_times,
                       void* d_out_data, int64_t capacity, void* stream);

/* Events of a list of pixels as dense columns: d_out (nframes, n_cols) of the count dtype,
 * d_out[t * n_cols + k] = the count of pixel d_cols[k] at frame t, 0 where it has no event
 * (the buffer is cleared first).  Input: the pixel-major layout of
 * dxg2_correlate_events_csr (npix = H*W pixels).  *d_err (int32, may be NULL, zeroed by
 * the caller) gets bit 0 for an event frame outside [0, nframes) (skipped) and bit 1 for
 * a listed pixel outside [0, npix).  Feeds the two-time output of event input: the
 * reference builds its two-time matrix from event lists (ecorrts_q / nbecorrts_q,
 * correlator/event_y.py:104-230); here a bin's events become a dense (nframes, n_b)
 * stack for the exact Gram path (dxg2_correlate_dense_ttcf). */
int dxg2_events_to_columns(const uint32_t* d_offsets, const int32_t* d_times,
                           const void* d_data, int64_t npix, int32_t nframes, int32_t dtype,
                           const int32_t* d_cols, int64_t n_cols, void* d_out, int32_t* d_err,
                           void* stream);

/* dxg2_events_window without the workspace argument, for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops that must not call
 * hipMalloc).  No reference counterpart.  The scan's workspace is a static device array
 * of DXG2_STATIC_SCRATCH_WORDS uint32 (256 KiB, held by each GPU the library runs on
 * until the process ends), so npix is at most DXG2_STATIC_SCRATCH_WORDS * 1024
 * (67,108,864 pixels); larger inputs return DXG2_EINVAL.  The call allocates nothing and
 * launches kernels only.  Every call on one GPU shares that array: calls that overlap on
 * one GPU (concurrent streams or host threads) race and must be serialised by the caller,
 * e.g. by issuing them on one stream.  Results equal dxg2_events_window's. */
#define DXG2_STATIC_SCRATCH_WORDS 65536
int dxg2_events_window_noalloc(const uint32_t* d_offsets, const int32_t* d_times,
                               const void* d_data, int64_t nnz, int64_t npix, int32_t dtype,
                               int32_t t0, int32_t t1, uint32_t* d_out_offsets,
                               int32_t* d_out_times, void* d_out_data, int64_t capacity,
                               void* stream);

/* Dense frames -> per-pixel event lists on the GPU (FramesCompressor, event.py:207-321):
 * for every pixel p in raster order, its samples with value > 0 (signed for int8) in
 * frame order, data[offsets[p] .. offsets[p+1]) and times[...] (int32).  This is the
 * input layout of dxg2_correlate_events_csr.
 *   count: offsets (H*W + 1 uint32, offsets[H*W] = number of events); the caller reads
 *          the total to size times and data (the frame dtype);
 *   fill:  writes times and data; it needs the same scratch, still holding what count
 *          left in it (dxg2_compress_scratch_bytes: per-chunk counts of the frame axis).
 * Replaces FramesCompressor.compress_all_stack (event.py:267-284). */
int dxg2_compress_scratch_bytes(int32_t nframes, int32_t height, int32_t width,
                                size_t* bytes);
int dxg2_compress_count(const void* d_frames, int32_t dtype, int32_t nframes, int32_t height,
                        int32_t width, uint32_t* d_scratch, uint32_t* d_offsets, void* stream);
int dxg2_compress_fill(const void* d_frames, int32_t dtype, int32_t nframes, int32_t height,
                       int32_t width, const uint32_t* d_scratch, const uint32_t* d_offsets,
                       int32_t* d_times, void* d_data, void* stream);
/* Streaming compressor: frame t appends its events to per-pixel slots (pixel p owns
 * entries p*max_nnz .. (p+1)*max_nnz of events/times; counter: H*W uint32, zeroed by the
 * caller).  A pixel with more than max_nnz events sets *d_err (int32, zeroed by the
 * caller).  compact: the 1-D layout above from the slots; scratch holds
 * (H*W + 1023)/1024 + 2 uint32.
 * Replace FramesCompressor.process_frame / get_compacted_events (event.py:287-321). */
int dxg2_compress_push(const void* d_frame, int32_t dtype, int32_t height, int32_t width,
                       int32_t t, int32_t max_nnz, uint32_t* d_counter, void* d_events,
                       int32_t* d_times, int32_t* d_err, void* stream);
int dxg2_compress_compact(const uint32_t* d_counter, const void* d_events,
                          const int32_t* d_times, int32_t dtype, int32_t height, int32_t width,
                          int32_t max_nnz, uint32_t* d_scratch, uint32_t* d_offsets,
                          void* d_out_data, int32_t* d_out_times, void* stream);

/* Pixel-major events -> frame-sorted COO (pixel, frame, count), pixel order within a
 * frame unspecified.  Waits for the stream.  Test-data utility; no reference counterpart. */
int dxg2_events_to_coo(const int32_t* d_times, const void* d_data, const uint32_t* d_offsets,
                       int64_t nnz, int32_t height, int32_t width, int32_t nframes, int32_t dtype,
                       uint32_t* d_pixel, uint32_t* d_frame, void* d_count, void* stream);

const char* dxg2_last_error(void);
const char* dxg2_version(void);

#ifdef __cplusplus
}
#endif

#endif /* DXG2_H */
