/*
 * fdas.h -- C ABI of the MI355X-native FDAS hot path.
 *
 * This is the drop-in boundary for the device path of UOA-PARC-SKA/FDAS: each
 * entry point below replaces one method of the reference's host class
 * `FDAS` (fdas/src/host/FDAS.h:36-187, implemented in fdas/src/host/FDAS.cpp),
 * cited per function.  Plain pointers and sizes only; no OpenCL, HIP or torch
 * types cross the boundary.  Complex values are interleaved float pairs
 * (re, im), the layout of cl_float2 / std::complex<float> / numpy complex64.
 *
 * Conventions (kept from the reference, FDAS.cpp:34-52, FDAS.h:115-181):
 *  - every call returns a status; on failure the message is written to the
 *    log callback as "[ERROR] ..." and kept for fdas_last_error();
 *  - the library owns every device object, the caller owns host arrays;
 *  - one host thread per context (not thread-safe), like the reference class.
 */
#ifndef FDAS_AMD_FDAS_H
#define FDAS_AMD_FDAS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FDAS_ABI_VERSION 1

typedef struct fdas_ctx fdas_ctx;

typedef enum fdas_status {
  FDAS_OK = 0,
  FDAS_EINVAL = 1,        /* bad argument or size */
  FDAS_ENOMEM = 2,        /* device or host allocation failed, or HBM budget exceeded */
  FDAS_EDEVICE = 3,       /* a HIP call failed */
  FDAS_EUNSUPPORTED = 4,  /* valid request outside what this library or the format supports */
  FDAS_EBUSY = 5,         /* resource in use (e.g. serving queue full) */
  FDAS_ESTATE = 6,        /* call out of order (e.g. convolution before templates) */
  FDAS_EOVERFLOW = 7,     /* detections exceeded the spill capacity; lists are incomplete */
  FDAS_ECANCELED = 8      /* a served request was withdrawn before it reached the device */
} fdas_status;

/* FDAS::FOPPart (FDAS.h:38): t in [-c, 0], [0, c] or [-c, c]. */
typedef enum fdas_part { FDAS_NEGATIVE = 0, FDAS_POSITIVE = 1, FDAS_ALL = 2 } fdas_part;

/* FDAS::BufferSet (FDAS.h:39). */
typedef enum fdas_set { FDAS_SET_A = 0, FDAS_SET_B = 1 } fdas_set;

/* Which element count fdas_get_size returns (FDAS.h:78-88). */
typedef enum fdas_size_kind {
  FDAS_SIZE_INPUT = 0,       /* N complex bins            (get_input_sz)          */
  FDAS_SIZE_TILES = 1,       /* n_tiles * S complex       (get_tiles_sz)          */
  FDAS_SIZE_TEMPLATES = 2,   /* T * S complex             (get_templates_sz)      */
  FDAS_SIZE_THRESHOLDS = 3,  /* H floats                  (get_thresholds_sz)     */
  FDAS_SIZE_FOP = 4,         /* T * N floats              (get_fop_sz)            */
  FDAS_SIZE_CANDIDATES = 5   /* H * cands_per_harmonic    (get_candidate_list_sz) */
} fdas_size_kind;

/* Config flags. */
#define FDAS_FLAG_REFERENCE_FFT_ORDER 0x1u /* templates/tiles exchanged in the reference's
                                              4-parallel FFT order (tmpl-gen.py:145-153) */
#define FDAS_FLAG_SYNC_PIPELINE 0x2u       /* hold each launch's harmonic summing until the
                                              next launch or end_pipeline (FDAS.cpp:461-464,549-550) */

/* Run-time replacement of the reference's compile-time GenInfo (gen_info.h:26-74)
 * plus the arguments of initialise_accelerator (FDAS.h:45-48). */
typedef struct fdas_config {
  uint32_t n_bins;             /* N, input_sz of initialise_accelerator; 1 .. 2^22 */
  uint32_t n_templates;        /* T, odd, <= 255 (85 in the reference) */
  uint32_t max_taps;           /* longest template, sets the overlap (421) */
  uint32_t tile;               /* FFT size S: 1024, 2048 or 4096 (2048) */
  uint32_t n_harmonics;        /* H planes, 1 .. 16 (8) */
  uint32_t cands_per_harmonic; /* candidates kept per plane, <= 8192 (512 -> 4096-entry lists) */
  uint32_t spill_per_harmonic; /* detections buffered per plane before the cap
                                  (0: max(262144, T * N / 512)) */
  int32_t device;              /* HIP device ordinal */
  uint32_t flags;              /* FDAS_FLAG_* */
} fdas_config;

/* A candidate in wide form (any T, H, N).  template_num = t in [-c, c]. */
typedef struct fdas_candidate {
  uint32_t harmonic;
  int32_t template_num;
  uint32_t bin;
  float power;
} fdas_candidate;

typedef void (*fdas_log_fn)(int level, const char* message, void* user);
#define FDAS_LOG_INFO 0
#define FDAS_LOG_WARN 1
#define FDAS_LOG_ERROR 2

/* Reference defaults: 85 templates, 421 taps, 2048-point tiles, 8 planes,
 * 512 candidates per plane (4096-entry lists), gen_info.h:28-72. */
void fdas_default_config(fdas_config* cfg);

/* FDAS::FDAS(std::ostream&) + initialise_accelerator (FDAS.h:43-48, FDAS.cpp:54-261):
 * selects the device, fixes N and allocates both buffer sets A and B. */
fdas_status fdas_create(const fdas_config* cfg, fdas_log_fn log, void* user, fdas_ctx** out);
void fdas_destroy(fdas_ctx* ctx);

/* FDAS::upload_templates (FDAS.h:50, FDAS.cpp:263-269): T x S complex template
 * spectra, natural order (or 4-parallel FFT order with FDAS_FLAG_REFERENCE_FFT_ORDER).
 * Blocking. */
fdas_status fdas_upload_templates(fdas_ctx* ctx, const float* spectra_c64, int set);

/* Time-domain taps (T x n_taps complex, zero padded) instead of spectra; the
 * library transforms them to the configured tile size.  No reference counterpart:
 * lets the tile size differ from 2048 (tmpl-gen.py:139-143 does this offline). */
fdas_status fdas_upload_template_taps(fdas_ctx* ctx, const float* taps_c64, uint32_t n_taps, int set);

/* FDAS::upload_thresholds (FDAS.h:52, FDAS.cpp:271-280): H floats, used when a
 * harmonic-summing call passes NULL thresholds. */
fdas_status fdas_upload_thresholds(fdas_ctx* ctx, const float* thresholds, int set);

/* FDAS::perform_input_tiling (FDAS.h:54, FDAS.cpp:282-328): uploads N complex
 * bins and materialises the transformed tiles (readable with fdas_retrieve_tiles).
 * Blocking.  The fused convolution re-tiles on the fly and does not need this
 * buffer; fdas_launch skips it. */
fdas_status fdas_perform_input_tiling(fdas_ctx* ctx, const float* spectrum_c64, int set);

/* As fdas_perform_input_tiling, with the spectrum already in device memory
 * (device-to-device copy, no tiles materialised).  No reference counterpart. */
fdas_status fdas_set_input_device(fdas_ctx* ctx, const void* d_spectrum_c64, int set);

/* FDAS::perform_ft_convolution (FDAS.h:56, FDAS.cpp:330-405): FOP rows of the part.  Blocking. */
fdas_status fdas_perform_ft_convolution(fdas_ctx* ctx, fdas_part part, int set);

/* FDAS::perform_harmonic_summing (FDAS.h:58, FDAS.cpp:407-546): harmonic
 * summing, thresholding and candidate selection over the part.  `thresholds`
 * (H floats) may be NULL to use the uploaded ones.  Blocking. */
fdas_status fdas_perform_harmonic_summing(fdas_ctx* ctx, const float* thresholds, fdas_part part, int set);

/* FDAS::launch (FDAS.h:60, FDAS.cpp:548-583): asynchronous upload -> FT
 * convolution -> harmonic summing -> candidate read-back.  The host arrays
 * must stay valid until fdas_wait(set); loc/pwr receive
 * fdas_get_size(FDAS_SIZE_CANDIDATES) entries in the reference layout. */
fdas_status fdas_launch(fdas_ctx* ctx, const float* spectrum_c64, const float* thresholds, uint32_t* det_loc,
                        float* det_pwr, fdas_part part, int set);

/* As fdas_launch, with the spectrum already resident in the set's input buffer
 * (from fdas_perform_input_tiling or fdas_set_input_device): FT convolution ->
 * harmonic summing -> candidate read-back, no host-to-device copy.  No
 * reference counterpart; it is the device-resident throughput path. */
fdas_status fdas_launch_resident(fdas_ctx* ctx, const float* thresholds, uint32_t* det_loc, float* det_pwr,
                                 fdas_part part, int set);

/* FDAS::wait (FDAS.h:62, FDAS.cpp:585-588). */
fdas_status fdas_wait(fdas_ctx* ctx, int set);

/* FDAS::end_pipeline (FDAS.h:64, FDAS.cpp:590-594): releases the held
 * harmonic summing of the other set (FDAS_FLAG_SYNC_PIPELINE). */
fdas_status fdas_end_pipeline(fdas_ctx* ctx, int set);

/* FDAS::retrieve_tiles / retrieve_FOP / inject_FOP / retrieve_candidates (FDAS.h:66-72,
 * FDAS.cpp:596-632).  Candidate lists use the reference's 32-bit layout: plane-major,
 * cands_per_harmonic entries per plane, sorted by (template, bin), padded with
 * encode(1, c+1, 0) / -1.0f.  Returns FDAS_EUNSUPPORTED when T > 127 or H > 8
 * (the 7-bit / 3-bit fields of gen_info.h:62-63 cannot hold them). */
fdas_status fdas_retrieve_tiles(fdas_ctx* ctx, float* tiles_c64, int set);
fdas_status fdas_retrieve_fop(fdas_ctx* ctx, float* fop, int set);
fdas_status fdas_inject_fop(fdas_ctx* ctx, const float* fop, int set);
fdas_status fdas_retrieve_candidates(fdas_ctx* ctx, uint32_t* det_loc, float* det_pwr, int set);

/* Wide candidate lists (any T, H).  Writes up to `capacity` entries, plane-major,
 * each plane sorted by (template, bin); *n_written = entries written;
 * n_detected[h] (H entries, may be NULL) = detections above threshold before the cap.
 * No reference counterpart (the reference format cannot encode T > 127 or H > 8). */
fdas_status fdas_retrieve_candidates_wide(fdas_ctx* ctx, fdas_candidate* out, uint32_t capacity,
                                          uint32_t* n_written, uint32_t* n_detected, int set);

/* Harmonic plane HP_k (k = 1..H) of the current FOP, T x N floats, rows outside
 * the part left untouched.  Debug / parity hook; the hot path never materialises
 * planes (the reference oracle writes them as hps_ref.npy, hsum.py:123). */
fdas_status fdas_retrieve_harmonic_plane(fdas_ctx* ctx, uint32_t k, float* plane, fdas_part part, int set);

/* FDAS::print_stats / print_events (FDAS.h:74-76, FDAS.cpp:634-686): per-phase
 * device durations to the log ("[INFO] ..." and "@@@," CSV lines). */
void fdas_print_stats(fdas_ctx* ctx, int set, int reset);
void fdas_print_events(fdas_ctx* ctx, int set);

/* hsum.py:125-131: the reference's "mock-up" thresholds, thr_k = quantile(HP_k,
 * 1 - n / M) over the M values of plane k in the part (linear interpolation,
 * FP64, rounded to float32), giving about n candidates per plane, for k = 1..H,
 * from the FOP in `set` (after fdas_perform_ft_convolution or fdas_inject_fop).
 * Overwrites the set's candidate lists. */
fdas_status fdas_quantile_thresholds(fdas_ctx* ctx, uint32_t n_per_plane, fdas_part part, int set, float* thresholds);

/* Device durations in ms of the last run on `set` (HIP events on the set's
 * stream): [0] input, [1] FT-convolution kernel, [2] harmonic-summing kernel,
 * [3] candidate selection, [4] candidate read-back, [5] total. */
fdas_status fdas_get_stage_times(fdas_ctx* ctx, int set, float* ms6);

/* Device enumeration for the platform / device selectors of
 * initialise_accelerator (FDAS.h:45-47, FDAS.cpp:65-115).  fdas_device_info
 * writes the device name into `name` (NUL-terminated, at most `len` bytes) and
 * returns FDAS_EINVAL for an ordinal outside [0, fdas_device_count()). */
int fdas_device_count(void);
fdas_status fdas_device_info(int device, char* name, size_t len, char* arch, size_t arch_len);
const char* fdas_platform_version(void);

/* FDAS::get_*_sz (FDAS.h:78-88). */
uint64_t fdas_get_size(const fdas_ctx* ctx, fdas_size_kind which);

const char* fdas_last_error(const fdas_ctx* ctx);
const char* fdas_status_string(fdas_status s);

/* Template bank generation (tmpl-gen.py:45-78, 154-159): Fourier response of a
 * tone drifting by drifts[i] bins, zero padded to max_taps, written as
 * n x max_taps complex floats.  Fresnel integrals are the library's own
 * (power series + continued fraction).  Host-only; no device needed. */
fdas_status fdas_make_drift_bank(const double* drifts, uint32_t n, uint32_t max_taps, float* out_c64);

/* Drifts of the reference bank generator (tmpl-gen.py:52-56, 127-133). */
fdas_status fdas_reference_drifts(uint32_t n_templates, double max_accel, double unaccel_freq, double t_obs,
                                  double speed_of_light, double* out);

/* ---------------------------------------------------------------------------
 * Allocation-free harmonic summing.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).
 *
 * fdas_hsum_noalloc runs the kernels of fdas_perform_harmonic_summing over a
 * FOP in the caller's device memory (T x N floats, row = t + T/2) and writes
 * the caller's device arrays: d_key / d_power [H][cands_per_harmonic]
 * ((row << 22) | bin and HP_k, sorted by key per plane; entries past a plane's
 * count are left as they were) and d_counts [2H] (entries kept per plane, then
 * detections per plane before the cap).  Asynchronous on `stream` (a
 * hipStream_t, NULL for the default stream); the caller synchronises.  It
 * calls no allocating routine: job descriptors, spill lists, counters,
 * histograms and compaction lists live in one device array of
 * FDAS_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends.  Sizes
 * that need more (8 bytes per spill entry per plane, plus the cap lists) are
 * rejected with FDAS_EINVAL and a message giving the limit; spill_per_harmonic
 * 0 means max(262144, T * N / 512), as in fdas_config.  Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU (other streams, other
 * threads) race; calls in one stream are ordered.  Errors go to `log`
 * ("[ERROR] ...", may be NULL). */
#define FDAS_NOALLOC_SCRATCH_BYTES (32u << 20)
fdas_status fdas_hsum_noalloc(const float* d_fop, uint32_t n_bins, uint32_t n_templates, fdas_part part,
                              uint32_t n_harmonics, const float* thresholds, uint32_t cands_per_harmonic,
                              uint32_t spill_per_harmonic, uint32_t* d_key, float* d_power, uint32_t* d_counts,
                              void* stream, fdas_log_fn log, void* user);

/* ---------------------------------------------------------------------------
 * Front end: time series -> spectrum (upstream of the reference's device path)
 * ------------------------------------------------------------------------- */

/* ft_conv.py:39-67 read_tim_file: a SIGPROC time series.  With read_header and a
 * header ("HEADER_START" after a 4-byte length, within the first 512 bytes, up to
 * "HEADER_END"), data_type must be 2 and nbits 32, and *t_samp receives tsamp
 * (else 0).  float32 samples follow.  samples == NULL only counts them.  Host-only;
 * the message of a failure is in fdas_io_last_error(). */
fdas_status fdas_read_tim(const char* path, int read_header, float* samples, uint64_t capacity,
                          uint64_t* n_samples, double* t_samp);
const char* fdas_io_last_error(void);

/* ft_conv.py:123-146 (the CXFT): X[f] = DFT(s)[f] / Ns for f < N from Ns = 2N
 * real samples, FP32 on the device (the reference's scipy.fft.fft of float32
 * samples is FP32 too).  N is a power of two, 2 .. 2^22.  No reference
 * counterpart as a device call: the reference computes it offline in Python. */
typedef struct fdas_frontend fdas_frontend;
#define FDAS_FE_DEVICE_INPUT 0x1u  /* series is device memory */
#define FDAS_FE_DEVICE_OUTPUT 0x2u /* spectrum is device memory */
fdas_status fdas_frontend_create(uint32_t n_bins, int device, fdas_frontend** out);
void fdas_frontend_destroy(fdas_frontend* fe);
/* Blocking.  series: 2N floats; spectrum_c64: N complex64 (e.g. an input of
 * fdas_set_input_device or an engine request with FDAS_REQ_DEVICE_SPECTRUM). */
fdas_status fdas_frontend_run(fdas_frontend* fe, const float* series, uint32_t flags, void* spectrum_c64);
const char* fdas_frontend_last_error(const fdas_frontend* fe);

/* ---------------------------------------------------------------------------
 * Batched and served search (no counterpart in the reference, which runs one
 * trial per launch on one buffer set; these replace its A/B double buffering
 * for callers that search many dedispersion trials, FDAS.cpp:282-615).
 *
 * An engine holds template banks (any T <= 255, any tap count), the FFT
 * twiddles of every tile size and a device-memory arena bounded by an HBM
 * budget.  A trial is one spectrum with its own N, bank, harmonic count, FOP
 * part and thresholds; its candidate lists are the same as those of the
 * single-trial path (fdas_launch + fdas_retrieve_candidates_wide) run with the
 * tile size the engine chose for that shape.
 * ------------------------------------------------------------------------- */
typedef struct fdas_engine fdas_engine;
typedef struct fdas_plan fdas_plan;

#define FDAS_ENGINE_MAX_HARMONICS 16

typedef struct fdas_engine_config {
  uint32_t tile;               /* FFT size for every shape; 0: autotuned per shape from 1024/2048/4096 */
  uint32_t rows_per_item;      /* templates per FT-convolution workgroup; 0: autotuned (8/16/32/64) */
  uint32_t hsum_rows;          /* rows per harmonic-summing workgroup; 0: autotuned (11/16/22) */
  uint32_t cands_per_harmonic; /* candidates kept per plane, <= 8192 (default 4096) */
  uint32_t spill_per_harmonic; /* detections buffered per plane before the cap
                                  (0, the default: max(262144, T * N / 512) per trial) */
  uint32_t group_trials;       /* most trials launched together (default 8) */
  uint32_t groups_in_flight;   /* groups whose state is resident at once, 2..8 (default 3) */
  int32_t device;
  uint64_t hbm_budget_bytes;   /* device memory for per-trial state (default 200 GB) */
  uint32_t ftc_workgroups_per_cu; /* FT-convolution workgroups resident per CU (a persistent grid), leaving
                                     room for the harmonic summing that runs beside it; 0: one per item */
  uint32_t plan_schedule;      /* how plans order group g's FT convolution against group g-1's harmonic
                                  summing: FDAS_SCHEDULE_LAGGED (default), _FUSED or _STREAMS */
} fdas_engine_config;

/* Plan schedules (fdas_engine_config.plan_schedule).
 *  LAGGED:  one stream; launch g runs group g's FT convolution, then group g-1's summing, back to back.
 *  FUSED:   as LAGGED, with both in one fused kernel whose workgroups interleave the two kinds of work,
 *           where the shapes allow (1024-point tiles, 8 harmonics, or 16 with N % 8 == 0).
 *  STREAMS: convolution on one stream, summing on another; group g's convolution runs beside group
 *           g-1's summing, sharing the CUs. */
#define FDAS_SCHEDULE_LAGGED 0u
#define FDAS_SCHEDULE_FUSED 1u
#define FDAS_SCHEDULE_STREAMS 2u

/* Request flags. */
#define FDAS_REQ_DEVICE_SPECTRUM 0x1u /* `spectrum` is device memory (else host memory) */

typedef struct fdas_request {
  const void* spectrum;        /* n_bins complex64 bins (re, im interleaved); must stay valid until the
                                  trial completes */
  uint32_t n_bins;             /* 1 .. 2^22 */
  uint32_t bank;               /* id from fdas_engine_add_bank */
  uint32_t n_harmonics;        /* 1 .. 16 */
  uint32_t part;               /* fdas_part */
  uint32_t flags;              /* FDAS_REQ_* */
  uint32_t pad_;
  float thresholds[FDAS_ENGINE_MAX_HARMONICS]; /* absolute threshold of plane k = 1..n_harmonics */
  uint64_t tag;                /* copied to the result */
} fdas_request;

typedef struct fdas_result {
  uint32_t* key;               /* caller-owned, n_harmonics * cands_per_harmonic entries: plane-major,
                                  (row << 22) | bin, row = t + T/2, sorted ascending per plane */
  float* power;                /* caller-owned, same layout: HP_k at that (row, bin) */
  uint64_t tag;
  int32_t status;              /* fdas_status of this trial (FDAS_EOVERFLOW: spill capacity exceeded) */
  uint32_t n_harmonics;
  uint32_t n_detected[FDAS_ENGINE_MAX_HARMONICS]; /* detections per plane before the cap */
  uint32_t n_kept[FDAS_ENGINE_MAX_HARMONICS];     /* entries written per plane */
  uint32_t tile;               /* FFT size the trial ran with */
  uint32_t pad_;
  double latency_ms;           /* served trials: submission to completion */
} fdas_result;

typedef struct fdas_engine_stats {
  uint64_t submitted, completed, groups_launched;
  uint64_t arena_reserved_bytes, arena_in_use_bytes, arena_peak_bytes;
  uint32_t queued, in_flight;  /* served trials waiting for admission / on the device */
  uint32_t max_group;          /* largest group launched */
  uint32_t pad_;
} fdas_engine_stats;

typedef struct fdas_plan_info {
  uint32_t n_trials, n_groups, n_launches, graph;
  uint64_t region_bytes;       /* device bytes of the plan's per-trial state */
  float last_ms;               /* device time of the last execution (start to end events) */
  float ftc_ms, hsum_ms;       /* summed FT-convolution / harmonic-summing kernel spans of the last eager
                                  execution, outside fused kernels (0 under a graph) */
  uint32_t lagged;             /* 1: the plan runs a lagged schedule (LAGGED or FUSED) */
  uint32_t n_fused;            /* fused launches in the last eager execution */
  float fused_ms;              /* their summed spans (each: one group's convolution + the previous group's
                                  summing) */
} fdas_plan_info;

#define FDAS_PLAN_GRAPH 0x1u   /* capture the plan's launches once as a HIP graph and replay it */

void fdas_engine_default_config(fdas_engine_config* cfg);
fdas_status fdas_engine_create(const fdas_engine_config* cfg, fdas_log_fn log, void* user, fdas_engine** out);
/* Completes every submitted trial, destroys the plans still alive (their handles become invalid:
 * do not pass them to fdas_plan_destroy afterwards), then releases all device state. */
void fdas_engine_destroy(fdas_engine* eng);
const char* fdas_engine_last_error(const fdas_engine* eng);

/* Registers a bank of n_templates (odd) x n_taps complex64 taps (the reference's
 * fdas_templates_*.npy layout, tmpl-gen.py:154-159); their spectra are formed
 * for every usable tile size and stay resident. */
fdas_status fdas_engine_add_bank(fdas_engine* eng, const float* taps_c64, uint32_t n_templates, uint32_t n_taps,
                                 uint32_t* bank_id);

/* Fixes the tile size and work shapes of (n_bins, bank, part) ahead of traffic,
 * timing the candidates not fixed by the configuration or the tuning cache on a
 * seeded noise spectrum: tile x rows per item for the FT convolution, then rows
 * per summing workgroup for summing + selection; *tile receives the tile size
 * (may be NULL). */
fdas_status fdas_engine_prepare(fdas_engine* eng, uint32_t n_bins, uint32_t bank, uint32_t part, uint32_t* tile);

/* The work shape the engine fixed for (n_bins, bank, part) (after
 * fdas_engine_prepare, a plan or a submission of that shape), with the
 * transform counts of one trial: forward transforms (one per work item) and
 * inverse transforms (one per tile and template row). */
typedef struct fdas_shape_info {
  uint32_t tile, rows_per_item, hsum_rows, n_row_groups;
  uint64_t forward_transforms, inverse_transforms;
  uint64_t stored_outputs;     /* FOP values written: rows x n_bins */
} fdas_shape_info;
fdas_status fdas_engine_shape_info(fdas_engine* eng, uint32_t n_bins, uint32_t bank, uint32_t part,
                                   fdas_shape_info* info);

/* Tuning cache: shapes listed in `path` (written by earlier engines on the same
 * device) are used without timing, and every shape this engine times is
 * appended to it.  Key: device, N, bank contents (digest of the taps), part and
 * cands_per_harmonic; value: tile, rows per item, rows per summing workgroup.
 * The tile decides the FOP rounding, so engines sharing a cache return
 * identical lists for identical requests.  A shape first seen while requests
 * are being served is not timed: it takes the cache's choice, else a cost
 * model's (fewest transform points per output). */
fdas_status fdas_engine_tune_cache(fdas_engine* eng, const char* path);

/* Batch: a fixed list of trials, split into groups of at most group_trials; the
 * FT convolution of one group runs on one stream while harmonic summing and
 * selection of the previous group run on another. */
fdas_status fdas_plan_create(fdas_engine* eng, const fdas_request* reqs, uint32_t n, uint32_t flags,
                             fdas_plan** out);
/* Replaces trial i's spectrum pointer and thresholds (same shape) before the next launch. */
fdas_status fdas_plan_set_request(fdas_plan* plan, uint32_t i, const void* spectrum, const float* thresholds);
fdas_status fdas_plan_launch(fdas_plan* plan);
/* Waits for the launch and fills results[0..n) (NULL: only waits). */
fdas_status fdas_plan_wait(fdas_plan* plan, fdas_result* results);
fdas_status fdas_plan_get_info(const fdas_plan* plan, fdas_plan_info* info);
/* Waits for the plan's launch and releases its memory; only before fdas_engine_destroy of its engine. */
void fdas_plan_destroy(fdas_plan* plan);

/* Serving: trials are queued, admitted while the HBM budget and the groups in
 * flight allow, and launched in groups formed from whatever is waiting, so
 * requests that arrive while others run join the next group.  `result` (and
 * its key/power arrays and the request's spectrum) must stay valid until the
 * result is returned by fdas_engine_poll.  Calls are thread-safe. */
fdas_status fdas_engine_submit(fdas_engine* eng, const fdas_request* req, fdas_result* result);
/* Returns up to `max` completed results (their addresses as passed to submit),
 * waiting up to timeout_ms for the first (0: no wait; < 0: until one completes
 * or nothing is pending). */
fdas_status fdas_engine_poll(fdas_engine* eng, fdas_result** done, uint32_t max, uint32_t* n_done,
                             int32_t timeout_ms);
fdas_status fdas_engine_get_stats(fdas_engine* eng, fdas_engine_stats* stats);
/* Withdraws the submitted requests with this tag that are still queued (not yet
 * admitted to the device): each completes at once with status FDAS_ECANCELED
 * and is returned by fdas_engine_poll like any other result.  Requests already
 * on the device run to completion.  *n_canceled (may be NULL) receives the
 * number withdrawn. */
fdas_status fdas_engine_cancel(fdas_engine* eng, uint64_t tag, uint32_t* n_canceled);

#ifdef __cplusplus
}
#endif

#endif /* FDAS_AMD_FDAS_H */
