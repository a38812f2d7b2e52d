// Host runtime behind the C ABI in include/fdas.h.
//
// Mirrors the reference host class FDAS (fdas/src/host/FDAS.cpp): a context
// owns the device, two buffer sets A/B with their own stream, the resident
// template spectra, the FOP and the candidate buffers.  Where the reference
// sequences ~40 OpenCL kernels over on-chip channels, one set here runs three
// HIP kernels in order on its stream: fused FT convolution, harmonic summing
// with compaction, and per-plane selection.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "fdas_geometry.h"
#include "fdas_host.h"
#include "fdas_jobs.h"
#include "fdas_kernels.h"
#include "fdas_plan.h"

using namespace fdas;

namespace {

constexpr uint32_t kMaxFtcJobs = 64;  // row groups per FT-convolution launch

// Events per launch: START, INPUT (spectrum in place), FTC_K (FT-convolution kernel start), FTC (its end),
// HSUM_K0 / HSUM_K1 (harmonic-summing kernel start / end), HSUM (selection end), XFER (lists read back).
enum Ev { EV_START = 0, EV_INPUT, EV_FTC_K, EV_FTC, EV_HSUM_K0, EV_HSUM_K1, EV_HSUM, EV_XFER, EV_COUNT };

struct DevSet {
  float2* input = nullptr;
  float2* hperm = nullptr;
  float* fop = nullptr;
  float2* tiles = nullptr;
  float* plane = nullptr;
  uint32_t* spill_key = nullptr;
  float* spill_pwr = nullptr;
  uint32_t* counts = nullptr;      // [H] detections, then [H] compact counts, then [H][kSelectBins] histograms
  uint32_t* compact_key = nullptr;
  float* compact_pwr = nullptr;
  uint32_t* out_key = nullptr;
  float* out_pwr = nullptr;
  uint32_t* out_n = nullptr;
  FtcJob* d_ftc = nullptr;
  HsumJob* d_hsum = nullptr;
  SelectJob* d_sel = nullptr;

  FtcJob* h_ftc = nullptr;
  HsumJob* h_hsum = nullptr;
  SelectJob* h_sel = nullptr;
  uint32_t* h_out_key = nullptr;
  float* h_out_pwr = nullptr;
  uint32_t* h_out_n = nullptr;
  uint32_t* h_counts = nullptr;

  hipStream_t stream = nullptr;
  hipEvent_t ev[EV_COUNT] = {};
  hipEvent_t staged_ftc = nullptr;   // h_ftc may be rewritten once this completes
  hipEvent_t staged_hsum = nullptr;  // h_hsum / h_sel may be rewritten once this completes
  hipEvent_t hsum_done = nullptr;    // the set's last harmonic-summing kernel has finished
  bool hsum_issued = false;          // hsum_done has been recorded at least once
  bool ftc_pending = false;
  bool hsum_pending = false;
  bool ev_recorded[EV_COUNT] = {};

  std::vector<uint32_t> row_taps;  // non-zero extent of each template row (sets per-group overlaps)
  bool templates = false;
  bool has_input = false;
  bool fop_valid = false;
  bool cands_valid = false;
  bool tiles_valid = false;
  float thr[kMaxHarmonics] = {};
  bool thr_set = false;

  bool launched = false;
  uint32_t* user_loc = nullptr;
  float* user_pwr = nullptr;
  bool gated = false;
  fdas_part gated_part = FDAS_ALL;
  float gated_thr[kMaxHarmonics] = {};
};

}  // namespace

struct fdas_ctx {
  fdas_config cfg{};
  Geometry geo{};
  uint32_t c = 0;
  uint32_t spill = 0;
  fdas_log_fn log = nullptr;
  void* user = nullptr;
  std::string last_error;
  float2* tw1 = nullptr;
  float2* tw2 = nullptr;
  DevSet set[2];
  hipEvent_t epoch = nullptr;
  bool sync_pipeline = false;
  // print_stats state; the reference keeps it file-static (FDAS.cpp:640-641)
  float pipeline_start = 0.f;
  float pipeline_accum = 0.f;
  unsigned pipeline_launches = 0;
};

namespace {

void emit(fdas_ctx* ctx, int level, const std::string& msg) {
  if (ctx && ctx->log) ctx->log(level, msg.c_str(), ctx->user);
}

fdas_status fail(fdas_ctx* ctx, fdas_status st, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (ctx) {
    ctx->last_error = std::string("[ERROR] ") + buf;
    emit(ctx, FDAS_LOG_ERROR, ctx->last_error);
  }
  return st;
}

void info(fdas_ctx* ctx, const char* fmt, ...) {
  if (!ctx || !ctx->log) return;
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  emit(ctx, FDAS_LOG_INFO, std::string("[INFO] ") + buf);
}

#define HIPCK(ctx, expr)                                                                                       \
  do {                                                                                                         \
    hipError_t e_ = (expr);                                                                                    \
    if (e_ != hipSuccess)                                                                                      \
      return fail(ctx, FDAS_EDEVICE, "HIP call failed with status %d (%s):\n          %s [%s:%d]", int(e_), \
                  hipGetErrorString(e_), #expr, __FILE__, __LINE__);                                           \
  } while (0)

#define CHECK_SET(ctx, set)                                                      \
  do {                                                                           \
    if (!(ctx)) return FDAS_EINVAL;                                              \
    if ((set) != FDAS_SET_A && (set) != FDAS_SET_B)                              \
      return fail(ctx, FDAS_EINVAL, "invalid buffer set %d", int(set));          \
  } while (0)

bool valid_part(fdas_part p) { return p == FDAS_NEGATIVE || p == FDAS_POSITIVE || p == FDAS_ALL; }

void part_rows(const fdas_ctx* ctx, fdas_part p, uint32_t* begin, uint32_t* count) {
  switch (p) {
    case FDAS_NEGATIVE: *begin = 0; *count = ctx->c + 1; break;
    case FDAS_POSITIVE: *begin = ctx->c; *count = ctx->c + 1; break;
    default: *begin = 0; *count = ctx->cfg.n_templates; break;
  }
}

template <typename T>
fdas_status dev_alloc(fdas_ctx* ctx, T** p, size_t count) {
  *p = nullptr;
  if (count == 0) return FDAS_OK;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return fail(ctx, FDAS_ENOMEM, "device allocation of %zu bytes failed (%s)", count * sizeof(T),
                hipGetErrorString(e));
  }
  return FDAS_OK;
}

template <typename T>
fdas_status host_alloc(fdas_ctx* ctx, T** p, size_t count) {
  *p = nullptr;
  hipError_t e = hipHostMalloc(reinterpret_cast<void**>(p), std::max<size_t>(count, 1) * sizeof(T), 0);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return fail(ctx, FDAS_ENOMEM, "pinned host allocation of %zu bytes failed (%s)", count * sizeof(T),
                hipGetErrorString(e));
  }
  return FDAS_OK;
}

// Wait until a host job staging buffer may be rewritten (its last copy to the device is done).
fdas_status wait_staging(fdas_ctx* ctx, hipEvent_t ev, bool& pending) {
  if (pending) {
    HIPCK(ctx, hipEventSynchronize(ev));
    pending = false;
  }
  return FDAS_OK;
}

fdas_status record(fdas_ctx* ctx, DevSet& s, Ev e) {
  HIPCK(ctx, hipEventRecord(s.ev[e], s.stream));
  s.ev_recorded[e] = true;
  return FDAS_OK;
}

std::vector<FtcJob> plan_ftc_jobs(const fdas_ctx* ctx, const DevSet& s, uint32_t rb, uint32_t nr) {
  std::vector<FtcJob> jobs = plan_ftc_rows(ctx->cfg.tile, ctx->geo.n_bins, ctx->cfg.max_taps, s.row_taps, rb, nr,
                                           kMaxFtcJobs);
  uint32_t item = 0;
  for (FtcJob& j : jobs) {
    j.spectrum = s.input;
    j.fop = s.fop;
    j.hperm = s.hperm;
    j.item_begin = item;
    item += j.n_tiles * j.n_chunks;
  }
  return jobs;
}

fdas_status enqueue_ftc(fdas_ctx* ctx, DevSet& s, fdas_part part) {
  uint32_t rb, nr;
  part_rows(ctx, part, &rb, &nr);
  fdas_status st = wait_staging(ctx, s.staged_ftc, s.ftc_pending);
  if (st != FDAS_OK) return st;
  const std::vector<FtcJob> jobs = plan_ftc_jobs(ctx, s, rb, nr);
  std::copy(jobs.begin(), jobs.end(), s.h_ftc);
  const uint32_t n_items = jobs.back().item_begin + jobs.back().n_tiles * jobs.back().n_chunks;
  HIPCK(ctx, hipMemcpyAsync(s.d_ftc, s.h_ftc, sizeof(FtcJob) * jobs.size(), hipMemcpyHostToDevice, s.stream));
  HIPCK(ctx, hipEventRecord(s.staged_ftc, s.stream));
  s.ftc_pending = true;
  // The two sets' kernels take turns: this convolution starts once the other
  // set's last summing kernel has finished (both kinds of kernel are VALU-bound,
  // and side by side they only slow each other down), while that set's
  // selection and read-back, and this set's input copy, still overlap.
  const DevSet& other = &s == &ctx->set[0] ? ctx->set[1] : ctx->set[0];
  if (other.hsum_issued) HIPCK(ctx, hipStreamWaitEvent(s.stream, other.hsum_done, 0));
  if ((st = record(ctx, s, EV_FTC_K)) != FDAS_OK) return st;
  HIPCK(ctx, launch_ftconv(int(ctx->cfg.tile), s.d_ftc, uint32_t(jobs.size()), n_items, ctx->tw1, ctx->tw2,
                           n_items, s.stream));
  return FDAS_OK;
}

fdas_status enqueue_hsum(fdas_ctx* ctx, DevSet& s, fdas_part part, const float* thr) {
  uint32_t rb, nr;
  part_rows(ctx, part, &rb, &nr);
  const uint32_t H = ctx->cfg.n_harmonics;
  fdas_status st = wait_staging(ctx, s.staged_hsum, s.hsum_pending);
  if (st != FDAS_OK) return st;
  HsumJob& j = *s.h_hsum;
  j = HsumJob{};
  j.fop = s.fop;
  j.n_bins = ctx->geo.n_bins;
  j.n_templates = ctx->cfg.n_templates;
  j.row_begin = rb;
  j.n_rows = nr;
  j.n_harmonics = H;
  j.spill_cap = ctx->spill;
  j.block_begin = 0;
  hsum_row_chunks(nr, &j.rows_per_block, &j.n_row_chunks);
  j.spill_key = s.spill_key;
  j.spill_pwr = s.spill_pwr;
  j.counts = s.counts;
  for (uint32_t k = 0; k < H; ++k) j.thr[k] = thr[k];
  const uint32_t cap = ctx->cfg.cands_per_harmonic, ccap = cap + kSelectPool;
  for (uint32_t p = 0; p < H; ++p) {
    SelectJob& q = s.h_sel[p];
    q.spill_key = s.spill_key + size_t(p) * ctx->spill;
    q.spill_pwr = s.spill_pwr + size_t(p) * ctx->spill;
    q.count = s.counts + p;
    q.spill_cap = ctx->spill;
    q.cap = cap;
    q.thr = thr[p];
    q.compact_cap = ccap;
    q.hist = s.counts + 2 * H + size_t(p) * kSelectBins;
    q.compact_key = s.compact_key + size_t(p) * ccap;
    q.compact_pwr = s.compact_pwr + size_t(p) * ccap;
    q.compact_n = s.counts + H + p;
    q.out_key = s.out_key + size_t(p) * cap;
    q.out_pwr = s.out_pwr + size_t(p) * cap;
    q.out_n = s.out_n + p;
  }
  HIPCK(ctx, hipMemcpyAsync(s.d_hsum, s.h_hsum, sizeof(HsumJob), hipMemcpyHostToDevice, s.stream));
  HIPCK(ctx, hipMemcpyAsync(s.d_sel, s.h_sel, sizeof(SelectJob) * H, hipMemcpyHostToDevice, s.stream));
  HIPCK(ctx, hipEventRecord(s.staged_hsum, s.stream));
  s.hsum_pending = true;
  HIPCK(ctx, launch_zero_u32(s.counts, H * (2 + kSelectBins), s.stream));
  const uint32_t blocks = (ctx->geo.n_bins + kHsumBinsPerBlock - 1) / kHsumBinsPerBlock * j.n_row_chunks;
  if ((st = record(ctx, s, EV_HSUM_K0)) != FDAS_OK) return st;
  HIPCK(ctx, launch_hsum(s.d_hsum, 1, blocks, H, hsum_vec_ok(ctx->geo.n_bins), s.stream));
  if ((st = record(ctx, s, EV_HSUM_K1)) != FDAS_OK) return st;
  HIPCK(ctx, hipEventRecord(s.hsum_done, s.stream));
  s.hsum_issued = true;
  HIPCK(ctx, launch_select(s.d_sel, H, s.stream));
  return FDAS_OK;
}

fdas_status enqueue_readback(fdas_ctx* ctx, DevSet& s) {
  const uint32_t H = ctx->cfg.n_harmonics, cap = ctx->cfg.cands_per_harmonic;
  HIPCK(ctx, hipMemcpyAsync(s.h_out_key, s.out_key, sizeof(uint32_t) * H * cap, hipMemcpyDeviceToHost, s.stream));
  HIPCK(ctx, hipMemcpyAsync(s.h_out_pwr, s.out_pwr, sizeof(float) * H * cap, hipMemcpyDeviceToHost, s.stream));
  HIPCK(ctx, hipMemcpyAsync(s.h_out_n, s.out_n, sizeof(uint32_t) * H, hipMemcpyDeviceToHost, s.stream));
  HIPCK(ctx, hipMemcpyAsync(s.h_counts, s.counts, sizeof(uint32_t) * H, hipMemcpyDeviceToHost, s.stream));
  return FDAS_OK;
}

// Converts the staged lists to the reference's 32-bit layout (gen_info.h:62-73).
void to_reference_lists(const fdas_ctx* ctx, const DevSet& s, uint32_t* loc, float* pwr) {
  const uint32_t H = ctx->cfg.n_harmonics, cap = ctx->cfg.cands_per_harmonic;
  const int c = int(ctx->c);
  const uint32_t inv = invalid_location(c);
  for (uint32_t p = 0; p < H; ++p) {
    const uint32_t n = std::min(s.h_out_n[p], cap);
    for (uint32_t i = 0; i < cap; ++i) {
      const size_t o = size_t(p) * cap + i;
      if (i < n) {
        const uint32_t key = s.h_out_key[o];
        const int row = int(key >> kKeyBinBits);
        const uint32_t bin = key & ((1u << kKeyBinBits) - 1);
        loc[o] = encode_location(p + 1, row - c, bin, c);
        pwr[o] = s.h_out_pwr[o];
      } else {
        loc[o] = inv;
        pwr[o] = -1.0f;
      }
    }
  }
}

bool reference_format_ok(const fdas_ctx* ctx) { return ctx->cfg.n_templates <= 127 && ctx->cfg.n_harmonics <= 8; }

fdas_status overflow_status(fdas_ctx* ctx, const DevSet& s) {
  for (uint32_t p = 0; p < ctx->cfg.n_harmonics; ++p)
    if (s.h_counts[p] > ctx->spill)
      return fail(ctx, FDAS_EOVERFLOW,
                  "plane %u: %u detections exceed the spill capacity %u; lists are incomplete (raise the "
                  "threshold or spill_per_harmonic)",
                  p + 1, s.h_counts[p], ctx->spill);
  return FDAS_OK;
}

fdas_status flush_gated(fdas_ctx* ctx, int set) {
  DevSet& s = ctx->set[set];
  if (!s.gated) return FDAS_OK;
  s.gated = false;
  fdas_status st = enqueue_hsum(ctx, s, s.gated_part, s.gated_thr);
  if (st != FDAS_OK) return st;
  if ((st = record(ctx, s, EV_HSUM)) != FDAS_OK) return st;
  if ((st = enqueue_readback(ctx, s)) != FDAS_OK) return st;
  return record(ctx, s, EV_XFER);
}

void free_set(DevSet& s) {
  (void)hipFree(s.input);
  (void)hipFree(s.hperm);
  (void)hipFree(s.fop);
  (void)hipFree(s.tiles);
  (void)hipFree(s.plane);
  (void)hipFree(s.spill_key);
  (void)hipFree(s.spill_pwr);
  (void)hipFree(s.counts);
  (void)hipFree(s.compact_key);
  (void)hipFree(s.compact_pwr);
  (void)hipFree(s.out_key);
  (void)hipFree(s.out_pwr);
  (void)hipFree(s.out_n);
  (void)hipFree(s.d_ftc);
  (void)hipFree(s.d_hsum);
  (void)hipFree(s.d_sel);
  (void)hipHostFree(s.h_ftc);
  (void)hipHostFree(s.h_hsum);
  (void)hipHostFree(s.h_sel);
  (void)hipHostFree(s.h_out_key);
  (void)hipHostFree(s.h_out_pwr);
  (void)hipHostFree(s.h_out_n);
  (void)hipHostFree(s.h_counts);
  for (auto& e : s.ev)
    if (e) (void)hipEventDestroy(e);
  if (s.staged_ftc) (void)hipEventDestroy(s.staged_ftc);
  if (s.staged_hsum) (void)hipEventDestroy(s.staged_hsum);
  if (s.hsum_done) (void)hipEventDestroy(s.hsum_done);
  if (s.stream) (void)hipStreamDestroy(s.stream);
  s = DevSet{};
}

fdas_status validate_config(fdas_ctx* ctx, const fdas_config& c) {
  if (c.n_bins == 0 || c.n_bins > kMaxBins)
    return fail(ctx, FDAS_EINVAL, "n_bins %u outside [1, %u]", c.n_bins, kMaxBins);
  if (c.n_templates == 0 || (c.n_templates & 1u) == 0 || c.n_templates > kMaxTemplates)
    return fail(ctx, FDAS_EINVAL, "n_templates %u must be odd and <= %u", c.n_templates, kMaxTemplates);
  if (c.tile != 1024 && c.tile != 2048 && c.tile != 4096)
    return fail(ctx, FDAS_EINVAL, "tile %u must be 1024, 2048 or 4096", c.tile);
  if (c.max_taps == 0 || overlap_for(c.max_taps) >= c.tile)
    return fail(ctx, FDAS_EINVAL, "max_taps %u leaves no payload in a %u-point tile", c.max_taps, c.tile);
  if (c.n_harmonics == 0 || c.n_harmonics > kMaxHarmonics)
    return fail(ctx, FDAS_EINVAL, "n_harmonics %u outside [1, %u]", c.n_harmonics, kMaxHarmonics);
  if (c.cands_per_harmonic == 0 || c.cands_per_harmonic > kSelectMaxCap)
    return fail(ctx, FDAS_EINVAL, "cands_per_harmonic %u outside [1, %u]", c.cands_per_harmonic, kSelectMaxCap);
  return FDAS_OK;
}

}  // namespace

// ---------------------------------------------------------------------------

extern "C" void fdas_default_config(fdas_config* cfg) {
  if (!cfg) return;
  *cfg = fdas_config{};
  cfg->n_bins = 0;
  cfg->n_templates = 85;
  cfg->max_taps = 421;
  cfg->tile = 2048;
  cfg->n_harmonics = 8;
  cfg->cands_per_harmonic = 512;
  cfg->spill_per_harmonic = 0;
  cfg->device = 0;
  cfg->flags = 0;
}

extern "C" const char* fdas_status_string(fdas_status s) {
  switch (s) {
    case FDAS_OK: return "ok";
    case FDAS_EINVAL: return "invalid argument";
    case FDAS_ENOMEM: return "out of memory";
    case FDAS_EDEVICE: return "device error";
    case FDAS_EUNSUPPORTED: return "unsupported";
    case FDAS_EBUSY: return "busy";
    case FDAS_ESTATE: return "call out of order";
    case FDAS_EOVERFLOW: return "candidate spill overflow";
    case FDAS_ECANCELED: return "request canceled";
  }
  return "unknown status";
}

extern "C" const char* fdas_last_error(const fdas_ctx* ctx) { return ctx ? ctx->last_error.c_str() : ""; }

extern "C" fdas_status fdas_create(const fdas_config* cfg, fdas_log_fn log, void* user, fdas_ctx** out) {
  if (!cfg || !out) return FDAS_EINVAL;
  *out = nullptr;
  fdas_ctx* ctx = new (std::nothrow) fdas_ctx();
  if (!ctx) return FDAS_ENOMEM;
  ctx->log = log;
  ctx->user = user;
  ctx->cfg = *cfg;
  fdas_status st = validate_config(ctx, ctx->cfg);
  if (st != FDAS_OK) {
    delete ctx;
    return st;
  }
  ctx->geo = make_geometry(cfg->n_bins, cfg->tile, cfg->max_taps);
  ctx->c = cfg->n_templates / 2;
  ctx->spill = cfg->spill_per_harmonic ? cfg->spill_per_harmonic : default_spill(cfg->n_templates, cfg->n_bins);
  ctx->sync_pipeline = (cfg->flags & FDAS_FLAG_SYNC_PIPELINE) != 0;

  auto bail = [&](fdas_status s) {
    for (auto& set : ctx->set) free_set(set);
    (void)hipFree(ctx->tw1);
    (void)hipFree(ctx->tw2);
    if (ctx->epoch) (void)hipEventDestroy(ctx->epoch);
    delete ctx;
    return s;
  };

  int n_dev = 0;
  hipError_t e = hipGetDeviceCount(&n_dev);
  if (e != hipSuccess || n_dev == 0) {
    fail(ctx, FDAS_EDEVICE, "no HIP device found (%s)", hipGetErrorString(e));
    return bail(FDAS_EDEVICE);
  }
  if (cfg->device < 0 || cfg->device >= n_dev) {
    fail(ctx, FDAS_EINVAL, "device %d outside [0, %d)", cfg->device, n_dev);
    return bail(FDAS_EINVAL);
  }
  if ((e = hipSetDevice(cfg->device)) != hipSuccess) {
    fail(ctx, FDAS_EDEVICE, "hipSetDevice(%d) failed (%s)", cfg->device, hipGetErrorString(e));
    return bail(FDAS_EDEVICE);
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, cfg->device) == hipSuccess)
    info(ctx, "Device[%d]: %s (%s, %d CUs)", cfg->device, prop.name, prop.gcnArchName, prop.multiProcessorCount);

  // Twiddle tables in FP64, rounded once.
  const uint32_t S = cfg->tile;
  std::vector<float2> tw1, tw2;
  make_twiddles(S, tw1, tw2);
  if ((st = dev_alloc(ctx, &ctx->tw1, tw1.size())) != FDAS_OK) return bail(st);
  if ((st = dev_alloc(ctx, &ctx->tw2, tw2.size())) != FDAS_OK) return bail(st);
  if (hipMemcpy(ctx->tw1, tw1.data(), tw1.size() * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(ctx->tw2, tw2.data(), tw2.size() * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess) {
    fail(ctx, FDAS_EDEVICE, "twiddle upload failed");
    return bail(FDAS_EDEVICE);
  }
  if (hipEventCreate(&ctx->epoch) != hipSuccess) return bail(fail(ctx, FDAS_EDEVICE, "event creation failed"));

  const size_t N = cfg->n_bins, T = cfg->n_templates, H = cfg->n_harmonics, cap = cfg->cands_per_harmonic;
  for (auto& s : ctx->set) {
    if ((st = dev_alloc(ctx, &s.input, N)) != FDAS_OK) return bail(st);
    if ((st = dev_alloc(ctx, &s.hperm, T * S)) != FDAS_OK) return bail(st);
    if ((st = dev_alloc(ctx, &s.fop, T * N)) != FDAS_OK) return bail(st);
    if ((st = dev_alloc(ctx, &s.spill_key, H * ctx->spill)) != FDAS_OK) return bail(st);
    if ((st = dev_alloc(ctx, &s.spill_pwr, H * ctx->spill)) != FDAS_OK) return bail(st);
    if ((st = dev_alloc(ctx, &s.counts, H * (2 + kSelectBins))) != FDAS_OK) return bail(st);
    if ((st = dev_alloc(ctx, &s.compact_key, H * (cap + kSelectPool))) != FDAS_OK) return bail(st);
    if ((st = dev_alloc(ctx, &s.compact_pwr, H * (cap + kSelectPool))) != FDAS_OK) return bail(st);
    if ((st = dev_alloc(ctx, &s.out_key, H * cap)) != FDAS_OK) return bail(st);
    if ((st = dev_alloc(ctx, &s.out_pwr, H * cap)) != FDAS_OK) return bail(st);
    if ((st = dev_alloc(ctx, &s.out_n, H)) != FDAS_OK) return bail(st);
    if ((st = dev_alloc(ctx, &s.d_ftc, kMaxFtcJobs)) != FDAS_OK) return bail(st);
    if ((st = dev_alloc(ctx, &s.d_hsum, 1)) != FDAS_OK) return bail(st);
    if ((st = dev_alloc(ctx, &s.d_sel, H)) != FDAS_OK) return bail(st);
    if ((st = host_alloc(ctx, &s.h_ftc, kMaxFtcJobs)) != FDAS_OK) return bail(st);
    if ((st = host_alloc(ctx, &s.h_hsum, 1)) != FDAS_OK) return bail(st);
    if ((st = host_alloc(ctx, &s.h_sel, H)) != FDAS_OK) return bail(st);
    if ((st = host_alloc(ctx, &s.h_out_key, H * cap)) != FDAS_OK) return bail(st);
    if ((st = host_alloc(ctx, &s.h_out_pwr, H * cap)) != FDAS_OK) return bail(st);
    if ((st = host_alloc(ctx, &s.h_out_n, H)) != FDAS_OK) return bail(st);
    if ((st = host_alloc(ctx, &s.h_counts, H)) != FDAS_OK) return bail(st);
    std::memset(s.h_out_n, 0, sizeof(uint32_t) * H);
    std::memset(s.h_counts, 0, sizeof(uint32_t) * H);
    if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess)
      return bail(fail(ctx, FDAS_EDEVICE, "stream creation failed"));
    for (auto& ev : s.ev)
      if (hipEventCreate(&ev) != hipSuccess) return bail(fail(ctx, FDAS_EDEVICE, "event creation failed"));
    if (hipEventCreateWithFlags(&s.staged_ftc, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&s.staged_hsum, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&s.hsum_done, hipEventDisableTiming) != hipSuccess)
      return bail(fail(ctx, FDAS_EDEVICE, "event creation failed"));
    // a FOP that was never written reads as zeros
    if (hipMemset(s.fop, 0, T * N * sizeof(float)) != hipSuccess ||
        hipMemset(s.out_n, 0, H * sizeof(uint32_t)) != hipSuccess ||
        hipMemset(s.counts, 0, H * (2 + kSelectBins) * sizeof(uint32_t)) != hipSuccess)
      return bail(fail(ctx, FDAS_EDEVICE, "buffer initialisation failed"));
  }
  if (hipDeviceSynchronize() != hipSuccess) return bail(fail(ctx, FDAS_EDEVICE, "device synchronisation failed"));
  (void)hipEventRecord(ctx->epoch, ctx->set[0].stream);
  info(ctx, "Geometry: N=%u, T=%u, S=%u, overlap=%u, payload=%u, tiles=%u, H=%u, cap=%u, spill=%u", cfg->n_bins,
       cfg->n_templates, S, ctx->geo.overlap, ctx->geo.payload, ctx->geo.n_tiles, cfg->n_harmonics,
       cfg->cands_per_harmonic, ctx->spill);
  *out = ctx;
  return FDAS_OK;
}

extern "C" void fdas_destroy(fdas_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->cfg.device);
  for (auto& s : ctx->set)
    if (s.stream) (void)hipStreamSynchronize(s.stream);
  for (auto& s : ctx->set) free_set(s);
  (void)hipFree(ctx->tw1);
  (void)hipFree(ctx->tw2);
  if (ctx->epoch) (void)hipEventDestroy(ctx->epoch);
  delete ctx;
}

extern "C" uint64_t fdas_get_size(const fdas_ctx* ctx, fdas_size_kind which) {
  if (!ctx) return 0;
  switch (which) {
    case FDAS_SIZE_INPUT: return ctx->geo.n_bins;
    case FDAS_SIZE_TILES: return uint64_t(ctx->geo.n_tiles) * ctx->cfg.tile;
    case FDAS_SIZE_TEMPLATES: return uint64_t(ctx->cfg.n_templates) * ctx->cfg.tile;
    case FDAS_SIZE_THRESHOLDS: return ctx->cfg.n_harmonics;
    case FDAS_SIZE_FOP: return uint64_t(ctx->cfg.n_templates) * ctx->geo.n_bins;
    case FDAS_SIZE_CANDIDATES: return uint64_t(ctx->cfg.n_harmonics) * ctx->cfg.cands_per_harmonic;
  }
  return 0;
}

static fdas_status upload_spectra(fdas_ctx* ctx, int set, const std::vector<std::complex<double>>& spec) {
  // Stored as [t][slot][tid] (kernel register order after the forward tile transform), scaled by 1/S.
  const std::vector<float2> perm = kernel_template_layout(spec, ctx->cfg.n_templates, ctx->cfg.tile);
  DevSet& s = ctx->set[set];
  HIPCK(ctx, hipSetDevice(ctx->cfg.device));
  HIPCK(ctx, hipStreamSynchronize(s.stream));
  HIPCK(ctx, hipMemcpy(s.hperm, perm.data(), perm.size() * sizeof(float2), hipMemcpyHostToDevice));
  s.templates = true;
  info(ctx, "Uploaded template coefficients");
  return FDAS_OK;
}

// Natural bin of entry i of a row stored in 4-parallel FFT order:
// stored[4m + j] = natural[bitrev2(j) * S/4 + m] (tmpl-gen.py:145-153, ft_conv.py:185-194).
static inline uint32_t p4_natural_index(uint32_t i, uint32_t S) {
  const uint32_t j = i & 3u, m = i >> 2;
  const uint32_t br = ((j & 1u) << 1) | (j >> 1);
  return br * (S / 4) + m;
}

extern "C" fdas_status fdas_upload_templates(fdas_ctx* ctx, const float* spectra_c64, int set) {
  CHECK_SET(ctx, set);
  if (!spectra_c64) return fail(ctx, FDAS_EINVAL, "null template pointer");
  // spectra do not reveal the tap extent exactly: every row uses the configured max_taps
  ctx->set[set].row_taps.assign(ctx->cfg.n_templates, ctx->cfg.max_taps);
  const uint32_t S = ctx->cfg.tile, T = ctx->cfg.n_templates;
  const bool p4 = (ctx->cfg.flags & FDAS_FLAG_REFERENCE_FFT_ORDER) != 0;
  std::vector<std::complex<double>> spec(size_t(T) * S);
  for (uint32_t t = 0; t < T; ++t)
    for (uint32_t i = 0; i < S; ++i) {
      const float* v = spectra_c64 + 2 * (size_t(t) * S + i);
      const uint32_t dst = p4 ? p4_natural_index(i, S) : i;
      spec[size_t(t) * S + dst] = std::complex<double>(v[0], v[1]);
    }
  return upload_spectra(ctx, set, spec);
}

extern "C" fdas_status fdas_upload_template_taps(fdas_ctx* ctx, const float* taps_c64, uint32_t n_taps, int set) {
  CHECK_SET(ctx, set);
  if (!taps_c64 || n_taps == 0) return fail(ctx, FDAS_EINVAL, "null or empty template taps");
  const uint32_t S = ctx->cfg.tile, T = ctx->cfg.n_templates;
  // Taps beyond max_taps must be zero: the tile overlap covers max_taps - 1 samples only.
  for (uint32_t t = 0; t < T; ++t)
    for (uint32_t m = ctx->cfg.max_taps; m < n_taps; ++m) {
      const float* v = taps_c64 + 2 * (size_t(t) * n_taps + m);
      if (v[0] != 0.f || v[1] != 0.f)
        return fail(ctx, FDAS_EINVAL, "template %u has a non-zero tap %u beyond max_taps %u", t, m,
                    ctx->cfg.max_taps);
    }
  std::vector<uint32_t> extent;
  const std::vector<std::complex<double>> spec = taps_to_spectra(taps_c64, T, n_taps, ctx->cfg.max_taps, S, extent);
  fdas_status st = upload_spectra(ctx, set, spec);
  if (st == FDAS_OK) ctx->set[set].row_taps = extent;
  return st;
}

extern "C" fdas_status fdas_upload_thresholds(fdas_ctx* ctx, const float* thresholds, int set) {
  CHECK_SET(ctx, set);
  if (!thresholds) return fail(ctx, FDAS_EINVAL, "null thresholds");
  DevSet& s = ctx->set[set];
  for (uint32_t k = 0; k < ctx->cfg.n_harmonics; ++k) s.thr[k] = thresholds[k];
  s.thr_set = true;
  info(ctx, "Uploaded thresholds");
  return FDAS_OK;
}

static fdas_status upload_input(fdas_ctx* ctx, DevSet& s, const void* src, hipMemcpyKind kind) {
  fdas_status st;
  if ((st = record(ctx, s, EV_START)) != FDAS_OK) return st;
  HIPCK(ctx, hipMemcpyAsync(s.input, src, sizeof(float2) * ctx->geo.n_bins, kind, s.stream));
  s.has_input = true;
  s.fop_valid = false;
  s.cands_valid = false;
  s.tiles_valid = false;
  for (int e = EV_INPUT; e < EV_COUNT; ++e) s.ev_recorded[e] = false;
  return FDAS_OK;
}

extern "C" fdas_status fdas_perform_input_tiling(fdas_ctx* ctx, const float* spectrum_c64, int set) {
  CHECK_SET(ctx, set);
  if (!spectrum_c64) return fail(ctx, FDAS_EINVAL, "null spectrum");
  DevSet& s = ctx->set[set];
  HIPCK(ctx, hipSetDevice(ctx->cfg.device));
  fdas_status st;
  if (!s.tiles && (st = dev_alloc(ctx, &s.tiles, size_t(ctx->geo.n_tiles) * ctx->cfg.tile)) != FDAS_OK) return st;
  if ((st = upload_input(ctx, s, spectrum_c64, hipMemcpyHostToDevice)) != FDAS_OK) return st;
  HIPCK(ctx, launch_tiles(int(ctx->cfg.tile), s.input, ctx->geo.n_bins, ctx->geo.overlap, ctx->geo.payload,
                          ctx->geo.n_tiles, ctx->tw1, ctx->tw2, s.tiles, s.stream));
  if ((st = record(ctx, s, EV_INPUT)) != FDAS_OK) return st;
  HIPCK(ctx, hipEventSynchronize(s.ev[EV_INPUT]));
  s.tiles_valid = true;
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, s.ev[EV_START], s.ev[EV_INPUT]);
  info(ctx, "Preparation duration: %.3f ms", double(ms));
  return FDAS_OK;
}

extern "C" fdas_status fdas_set_input_device(fdas_ctx* ctx, const void* d_spectrum_c64, int set) {
  CHECK_SET(ctx, set);
  if (!d_spectrum_c64) return fail(ctx, FDAS_EINVAL, "null device spectrum");
  DevSet& s = ctx->set[set];
  HIPCK(ctx, hipSetDevice(ctx->cfg.device));
  fdas_status st;
  if ((st = upload_input(ctx, s, d_spectrum_c64, hipMemcpyDeviceToDevice)) != FDAS_OK) return st;
  return record(ctx, s, EV_INPUT);
}

static fdas_status check_ready_ftc(fdas_ctx* ctx, const DevSet& s, fdas_part part) {
  if (!valid_part(part)) return fail(ctx, FDAS_EINVAL, "invalid FOP part %d", int(part));
  if (!s.templates) return fail(ctx, FDAS_ESTATE, "templates not uploaded for this buffer set");
  if (!s.has_input) return fail(ctx, FDAS_ESTATE, "no input spectrum in this buffer set");
  return FDAS_OK;
}

extern "C" fdas_status fdas_perform_ft_convolution(fdas_ctx* ctx, fdas_part part, int set) {
  CHECK_SET(ctx, set);
  DevSet& s = ctx->set[set];
  fdas_status st = check_ready_ftc(ctx, s, part);
  if (st != FDAS_OK) return st;
  HIPCK(ctx, hipSetDevice(ctx->cfg.device));
  if ((st = record(ctx, s, EV_INPUT)) != FDAS_OK) return st;
  if ((st = enqueue_ftc(ctx, s, part)) != FDAS_OK) return st;
  if ((st = record(ctx, s, EV_FTC)) != FDAS_OK) return st;
  HIPCK(ctx, hipEventSynchronize(s.ev[EV_FTC]));
  s.fop_valid = true;
  s.cands_valid = false;
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, s.ev[EV_FTC_K], s.ev[EV_FTC]);
  info(ctx, "FT convolution and IFFT duration: %.3f ms", double(ms));
  return FDAS_OK;
}

static const float* pick_thresholds(fdas_ctx* ctx, const DevSet& s, const float* thr) {
  if (thr) return thr;
  if (s.thr_set) return s.thr;
  fail(ctx, FDAS_ESTATE, "no thresholds given or uploaded");
  return nullptr;
}

extern "C" fdas_status fdas_perform_harmonic_summing(fdas_ctx* ctx, const float* thresholds, fdas_part part,
                                                     int set) {
  CHECK_SET(ctx, set);
  DevSet& s = ctx->set[set];
  if (!valid_part(part)) return fail(ctx, FDAS_EINVAL, "invalid FOP part %d", int(part));
  const float* thr = pick_thresholds(ctx, s, thresholds);
  if (!thr) return FDAS_ESTATE;
  HIPCK(ctx, hipSetDevice(ctx->cfg.device));
  fdas_status st;
  if ((st = record(ctx, s, EV_FTC)) != FDAS_OK) return st;
  if ((st = enqueue_hsum(ctx, s, part, thr)) != FDAS_OK) return st;
  if ((st = record(ctx, s, EV_HSUM)) != FDAS_OK) return st;
  if ((st = enqueue_readback(ctx, s)) != FDAS_OK) return st;
  if ((st = record(ctx, s, EV_XFER)) != FDAS_OK) return st;
  HIPCK(ctx, hipEventSynchronize(s.ev[EV_XFER]));
  s.cands_valid = true;
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, s.ev[EV_HSUM_K0], s.ev[EV_HSUM]);
  info(ctx, "Harmonic summing duration: %.3f ms", double(ms));
  return overflow_status(ctx, s);
}

extern "C" fdas_status fdas_launch(fdas_ctx* ctx, const float* spectrum_c64, const float* thresholds, uint32_t* det_loc,
                                   float* det_pwr, fdas_part part, int set) {
  CHECK_SET(ctx, set);
  DevSet& s = ctx->set[set];
  if (!spectrum_c64 || !det_loc || !det_pwr) return fail(ctx, FDAS_EINVAL, "null launch argument");
  if (!valid_part(part)) return fail(ctx, FDAS_EINVAL, "invalid FOP part %d", int(part));
  if (!reference_format_ok(ctx))
    return fail(ctx, FDAS_EUNSUPPORTED, "T=%u/H=%u do not fit the 32-bit candidate layout; use the wide API",
                ctx->cfg.n_templates, ctx->cfg.n_harmonics);
  if (!s.templates) return fail(ctx, FDAS_ESTATE, "templates not uploaded for this buffer set");
  const float* thr = pick_thresholds(ctx, s, thresholds);
  if (!thr) return FDAS_ESTATE;
  HIPCK(ctx, hipSetDevice(ctx->cfg.device));
  fdas_status st;
  // Releases the other set's held harmonic summing (FDAS.cpp:549-550).
  if (ctx->sync_pipeline && (st = flush_gated(ctx, 1 - set)) != FDAS_OK) return st;
  if (s.gated && (st = flush_gated(ctx, set)) != FDAS_OK) return st;
  if ((st = upload_input(ctx, s, spectrum_c64, hipMemcpyHostToDevice)) != FDAS_OK) return st;
  if ((st = record(ctx, s, EV_INPUT)) != FDAS_OK) return st;
  if ((st = enqueue_ftc(ctx, s, part)) != FDAS_OK) return st;
  if ((st = record(ctx, s, EV_FTC)) != FDAS_OK) return st;
  s.launched = true;
  s.user_loc = det_loc;
  s.user_pwr = det_pwr;
  if (ctx->sync_pipeline) {
    s.gated = true;
    s.gated_part = part;
    for (uint32_t k = 0; k < ctx->cfg.n_harmonics; ++k) s.gated_thr[k] = thr[k];
    return FDAS_OK;
  }
  if ((st = enqueue_hsum(ctx, s, part, thr)) != FDAS_OK) return st;
  if ((st = record(ctx, s, EV_HSUM)) != FDAS_OK) return st;
  if ((st = enqueue_readback(ctx, s)) != FDAS_OK) return st;
  return record(ctx, s, EV_XFER);
}

extern "C" fdas_status fdas_launch_resident(fdas_ctx* ctx, const float* thresholds, uint32_t* det_loc,
                                            float* det_pwr, fdas_part part, int set) {
  CHECK_SET(ctx, set);
  DevSet& s = ctx->set[set];
  if (!det_loc || !det_pwr) return fail(ctx, FDAS_EINVAL, "null launch argument");
  if (!valid_part(part)) return fail(ctx, FDAS_EINVAL, "invalid FOP part %d", int(part));
  if (!reference_format_ok(ctx))
    return fail(ctx, FDAS_EUNSUPPORTED, "T=%u/H=%u do not fit the 32-bit candidate layout; use the wide API",
                ctx->cfg.n_templates, ctx->cfg.n_harmonics);
  if (!s.templates) return fail(ctx, FDAS_ESTATE, "templates not uploaded for this buffer set");
  if (!s.has_input) return fail(ctx, FDAS_ESTATE, "no input spectrum in this buffer set");
  const float* thr = pick_thresholds(ctx, s, thresholds);
  if (!thr) return FDAS_ESTATE;
  HIPCK(ctx, hipSetDevice(ctx->cfg.device));
  fdas_status st;
  if (ctx->sync_pipeline && (st = flush_gated(ctx, 1 - set)) != FDAS_OK) return st;
  if (s.gated && (st = flush_gated(ctx, set)) != FDAS_OK) return st;
  if ((st = record(ctx, s, EV_START)) != FDAS_OK) return st;
  if ((st = record(ctx, s, EV_INPUT)) != FDAS_OK) return st;
  if ((st = enqueue_ftc(ctx, s, part)) != FDAS_OK) return st;
  if ((st = record(ctx, s, EV_FTC)) != FDAS_OK) return st;
  s.launched = true;
  s.user_loc = det_loc;
  s.user_pwr = det_pwr;
  if (ctx->sync_pipeline) {
    s.gated = true;
    s.gated_part = part;
    for (uint32_t k = 0; k < ctx->cfg.n_harmonics; ++k) s.gated_thr[k] = thr[k];
    return FDAS_OK;
  }
  if ((st = enqueue_hsum(ctx, s, part, thr)) != FDAS_OK) return st;
  if ((st = record(ctx, s, EV_HSUM)) != FDAS_OK) return st;
  if ((st = enqueue_readback(ctx, s)) != FDAS_OK) return st;
  return record(ctx, s, EV_XFER);
}

extern "C" fdas_status fdas_wait(fdas_ctx* ctx, int set) {
  CHECK_SET(ctx, set);
  DevSet& s = ctx->set[set];
  if (!s.launched) return fail(ctx, FDAS_ESTATE, "wait on a buffer set without a launch");
  HIPCK(ctx, hipSetDevice(ctx->cfg.device));
  fdas_status st;
  if (s.gated && (st = flush_gated(ctx, set)) != FDAS_OK) return st;
  HIPCK(ctx, hipEventSynchronize(s.ev[EV_XFER]));
  s.launched = false;
  s.fop_valid = true;
  s.cands_valid = true;
  to_reference_lists(ctx, s, s.user_loc, s.user_pwr);
  return overflow_status(ctx, s);
}

extern "C" fdas_status fdas_end_pipeline(fdas_ctx* ctx, int set) {
  CHECK_SET(ctx, set);
  HIPCK(ctx, hipSetDevice(ctx->cfg.device));
  return flush_gated(ctx, 1 - set);
}

extern "C" fdas_status fdas_retrieve_tiles(fdas_ctx* ctx, float* tiles_c64, int set) {
  CHECK_SET(ctx, set);
  DevSet& s = ctx->set[set];
  if (!tiles_c64) return fail(ctx, FDAS_EINVAL, "null tiles pointer");
  if (!s.tiles_valid) return fail(ctx, FDAS_ESTATE, "tiles not materialised (call perform_input_tiling)");
  HIPCK(ctx, hipSetDevice(ctx->cfg.device));
  const size_t n = size_t(ctx->geo.n_tiles) * ctx->cfg.tile;
  const bool p4 = (ctx->cfg.flags & FDAS_FLAG_REFERENCE_FFT_ORDER) != 0;
  if (!p4) {
    HIPCK(ctx, hipMemcpyAsync(tiles_c64, s.tiles, n * sizeof(float2), hipMemcpyDeviceToHost, s.stream));
    HIPCK(ctx, hipStreamSynchronize(s.stream));
    return FDAS_OK;
  }
  std::vector<float2> nat(n);
  HIPCK(ctx, hipMemcpyAsync(nat.data(), s.tiles, n * sizeof(float2), hipMemcpyDeviceToHost, s.stream));
  HIPCK(ctx, hipStreamSynchronize(s.stream));
  const uint32_t S = ctx->cfg.tile;
  for (size_t t = 0; t < ctx->geo.n_tiles; ++t)
    for (uint32_t i = 0; i < S; ++i) {
      const float2 v = nat[t * S + p4_natural_index(i, S)];
      tiles_c64[2 * (t * S + i)] = v.x;
      tiles_c64[2 * (t * S + i) + 1] = v.y;
    }
  return FDAS_OK;
}

extern "C" fdas_status fdas_retrieve_fop(fdas_ctx* ctx, float* fop, int set) {
  CHECK_SET(ctx, set);
  DevSet& s = ctx->set[set];
  if (!fop) return fail(ctx, FDAS_EINVAL, "null FOP pointer");
  HIPCK(ctx, hipSetDevice(ctx->cfg.device));
  const size_t n = size_t(ctx->cfg.n_templates) * ctx->geo.n_bins;
  HIPCK(ctx, hipMemcpyAsync(fop, s.fop, n * sizeof(float), hipMemcpyDeviceToHost, s.stream));
  HIPCK(ctx, hipStreamSynchronize(s.stream));
  return FDAS_OK;
}

extern "C" fdas_status fdas_inject_fop(fdas_ctx* ctx, const float* fop, int set) {
  CHECK_SET(ctx, set);
  DevSet& s = ctx->set[set];
  if (!fop) return fail(ctx, FDAS_EINVAL, "null FOP pointer");
  HIPCK(ctx, hipSetDevice(ctx->cfg.device));
  const size_t n = size_t(ctx->cfg.n_templates) * ctx->geo.n_bins;
  fdas_status st;
  if ((st = record(ctx, s, EV_START)) != FDAS_OK) return st;
  if ((st = record(ctx, s, EV_INPUT)) != FDAS_OK) return st;
  HIPCK(ctx, hipMemcpyAsync(s.fop, fop, n * sizeof(float), hipMemcpyHostToDevice, s.stream));
  if ((st = record(ctx, s, EV_FTC)) != FDAS_OK) return st;
  s.fop_valid = true;
  s.cands_valid = false;
  return FDAS_OK;
}

extern "C" fdas_status fdas_retrieve_candidates(fdas_ctx* ctx, uint32_t* det_loc, float* det_pwr, int set) {
  CHECK_SET(ctx, set);
  DevSet& s = ctx->set[set];
  if (!det_loc || !det_pwr) return fail(ctx, FDAS_EINVAL, "null candidate pointer");
  if (!reference_format_ok(ctx))
    return fail(ctx, FDAS_EUNSUPPORTED, "T=%u/H=%u do not fit the 32-bit candidate layout; use the wide API",
                ctx->cfg.n_templates, ctx->cfg.n_harmonics);
  HIPCK(ctx, hipSetDevice(ctx->cfg.device));
  fdas_status st;
  if ((st = enqueue_readback(ctx, s)) != FDAS_OK) return st;
  HIPCK(ctx, hipStreamSynchronize(s.stream));
  to_reference_lists(ctx, s, det_loc, det_pwr);
  return FDAS_OK;
}

extern "C" fdas_status fdas_retrieve_candidates_wide(fdas_ctx* ctx, fdas_candidate* out, uint32_t capacity,
                                                     uint32_t* n_written, uint32_t* n_detected, int set) {
  CHECK_SET(ctx, set);
  DevSet& s = ctx->set[set];
  if (!out && capacity) return fail(ctx, FDAS_EINVAL, "null candidate pointer");
  HIPCK(ctx, hipSetDevice(ctx->cfg.device));
  fdas_status st;
  if ((st = enqueue_readback(ctx, s)) != FDAS_OK) return st;
  HIPCK(ctx, hipStreamSynchronize(s.stream));
  const uint32_t H = ctx->cfg.n_harmonics, cap = ctx->cfg.cands_per_harmonic;
  const int c = int(ctx->c);
  uint32_t w = 0;
  for (uint32_t p = 0; p < H; ++p) {
    if (n_detected) n_detected[p] = s.h_counts[p];
    const uint32_t n = std::min(s.h_out_n[p], cap);
    for (uint32_t i = 0; i < n && w < capacity; ++i, ++w) {
      const uint32_t key = s.h_out_key[size_t(p) * cap + i];
      out[w].harmonic = p + 1;
      out[w].template_num = int(key >> kKeyBinBits) - c;
      out[w].bin = key & ((1u << kKeyBinBits) - 1);
      out[w].power = s.h_out_pwr[size_t(p) * cap + i];
    }
  }
  if (n_written) *n_written = w;
  return overflow_status(ctx, s);
}

extern "C" fdas_status fdas_retrieve_harmonic_plane(fdas_ctx* ctx, uint32_t k, float* plane, fdas_part part,
                                                    int set) {
  CHECK_SET(ctx, set);
  DevSet& s = ctx->set[set];
  if (!plane) return fail(ctx, FDAS_EINVAL, "null plane pointer");
  if (k == 0 || k > ctx->cfg.n_harmonics) return fail(ctx, FDAS_EINVAL, "harmonic %u outside [1, %u]", k,
                                                      ctx->cfg.n_harmonics);
  if (!valid_part(part)) return fail(ctx, FDAS_EINVAL, "invalid FOP part %d", int(part));
  HIPCK(ctx, hipSetDevice(ctx->cfg.device));
  const size_t n = size_t(ctx->cfg.n_templates) * ctx->geo.n_bins;
  fdas_status st;
  if (!s.plane && (st = dev_alloc(ctx, &s.plane, n)) != FDAS_OK) return st;
  uint32_t rb, nr;
  part_rows(ctx, part, &rb, &nr);
  HIPCK(ctx, launch_hplane(s.fop, ctx->geo.n_bins, ctx->cfg.n_templates, rb, nr, k, s.plane, s.stream));
  HIPCK(ctx, hipMemcpyAsync(plane + size_t(rb) * ctx->geo.n_bins, s.plane + size_t(rb) * ctx->geo.n_bins,
                            size_t(nr) * ctx->geo.n_bins * sizeof(float), hipMemcpyDeviceToHost, s.stream));
  HIPCK(ctx, hipStreamSynchronize(s.stream));
  return FDAS_OK;
}

// hsum.py:125-131, the "mock-up" thresholds: thr_k = quantile(HP_k, 1 - n / M)
// over the M = rows * N values of the part, linear interpolation between the
// order statistics around q (M - 1), in FP64, rounded once to float32.  The two
// order statistics are the (n+1)-th and n-th largest values.  Instead of
// sorting M values: the (n+1)-th largest block maximum B_k bounds the (n+1)-th
// largest value from below (n+1 blocks hold a value >= B_k), so one summing
// pass with threshold just below B_k spills every value >= B_k, among them the
// n+1 largest, and those few are sorted on the host.
extern "C" fdas_status fdas_quantile_thresholds(fdas_ctx* ctx, uint32_t n_per_plane, fdas_part part, int set,
                                                float* thresholds) {
  CHECK_SET(ctx, set);
  DevSet& s = ctx->set[set];
  if (!thresholds) return fail(ctx, FDAS_EINVAL, "null thresholds pointer");
  if (!valid_part(part)) return fail(ctx, FDAS_EINVAL, "invalid FOP part %d", int(part));
  if (!s.fop_valid) return fail(ctx, FDAS_ESTATE, "no FOP in this buffer set (run the FT convolution first)");
  uint32_t rb, nr;
  part_rows(ctx, part, &rb, &nr);
  const uint32_t N = ctx->geo.n_bins, H = ctx->cfg.n_harmonics;
  const uint64_t M = uint64_t(nr) * N;
  if (n_per_plane < 1 || n_per_plane >= M)
    return fail(ctx, FDAS_EINVAL, "candidates per plane %u outside 1 .. %llu", n_per_plane,
                (unsigned long long)(M - 1));
  HIPCK(ctx, hipSetDevice(ctx->cfg.device));
  fdas_status st = wait_staging(ctx, s.staged_hsum, s.hsum_pending);
  if (st != FDAS_OK) return st;
  HIPCK(ctx, hipStreamSynchronize(s.stream));
  // 1. block maxima of every plane
  const uint32_t nb = hmax_blocks(N, nr);
  float* d_max = nullptr;
  if ((st = dev_alloc(ctx, &d_max, size_t(H) * nb)) != FDAS_OK) return st;
  std::vector<float> bmax(size_t(H) * nb);
  hipError_t e = launch_hmax(s.fop, N, ctx->cfg.n_templates, rb, nr, H, d_max, s.stream);
  if (e == hipSuccess) e = hipMemcpyAsync(bmax.data(), d_max, sizeof(float) * bmax.size(), hipMemcpyDeviceToHost,
                                          s.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s.stream);
  (void)hipFree(d_max);
  HIPCK(ctx, e);
  // 2. per plane a threshold t with count(HP_k > t) >= n + 2 that fits the spill
  //    lists: just below the (n+2)-th largest block maximum when there are that
  //    many blocks, refined (or, without a bound, found) by bisection over the
  //    bit patterns of non-negative floats, where each pass counts all planes.
  const size_t want = size_t(std::min<uint64_t>(uint64_t(n_per_plane) + 2, M));  // a spare rank for rounding
  auto value_of = [](int64_t b) { return b < 0 ? -INFINITY : __builtin_bit_cast(float, uint32_t(b)); };
  std::vector<int64_t> lo_b(H, -1), hi_b(H, 0x7f800000);  // count(> lo) >= want, count(> hi) < want
  for (uint32_t k = 0; k < H; ++k) {
    std::vector<float> v(bmax.begin() + size_t(k) * nb, bmax.begin() + size_t(k + 1) * nb);
    if (v.size() < want) continue;
    std::nth_element(v.begin(), v.begin() + (want - 1), v.end(), std::greater<float>());
    const float bound = std::nextafter(v[want - 1], -INFINITY);
    lo_b[k] = bound < 0.f ? -1 : int64_t(__builtin_bit_cast(uint32_t, bound));
  }
  std::vector<float> thr(H);
  std::vector<uint32_t> lo_count(H, 0);
  bool probing = false;  // the last pass used a probe threshold for some plane
  for (int pass = 0;; ++pass) {
    // pass 0 measures the starting bounds; later passes bisect the planes that overflow
    probing = false;
    std::vector<int64_t> at(H);
    for (uint32_t k = 0; k < H; ++k) {
      at[k] = lo_b[k];
      if (pass > 0 && lo_count[k] > ctx->spill) {
        at[k] = lo_b[k] + (hi_b[k] - lo_b[k]) / 2;
        probing = true;
      }
      thr[k] = value_of(at[k]);
    }
    if ((st = enqueue_hsum(ctx, s, part, thr.data())) != FDAS_OK) return st;
    HIPCK(ctx, hipMemcpyAsync(s.h_counts, s.counts, sizeof(uint32_t) * H, hipMemcpyDeviceToHost, s.stream));
    HIPCK(ctx, hipStreamSynchronize(s.stream));
    bool all_fit = true;
    for (uint32_t k = 0; k < H; ++k) {
      const uint32_t c = s.h_counts[k];
      if (pass == 0 || at[k] != lo_b[k]) {
        if (c >= want) {
          lo_b[k] = at[k];
          lo_count[k] = c;
        } else {
          hi_b[k] = at[k];
        }
      }
      if (lo_count[k] > ctx->spill) {
        all_fit = false;
        if (hi_b[k] - lo_b[k] <= 1)
          return fail(ctx, FDAS_EOVERFLOW, "plane %u: %u values tie at the quantile, beyond the spill capacity %u",
                      k + 1, lo_count[k], ctx->spill);
      }
    }
    if (all_fit) break;
  }
  if (probing) {  // the spill lists hold a probe's entries: spill once more at the final thresholds
    for (uint32_t k = 0; k < H; ++k) thr[k] = value_of(lo_b[k]);
    if ((st = enqueue_hsum(ctx, s, part, thr.data())) != FDAS_OK) return st;
    HIPCK(ctx, hipMemcpyAsync(s.h_counts, s.counts, sizeof(uint32_t) * H, hipMemcpyDeviceToHost, s.stream));
    HIPCK(ctx, hipStreamSynchronize(s.stream));
  }
  s.cands_valid = false;  // the lists now belong to these thresholds, not to a search
  const double q = 1.0 - double(n_per_plane) / double(M);
  const double vi = q * double(M - 1);
  const uint64_t below = uint64_t(std::floor(vi));  // ascending rank of the lower order statistic
  const double g = vi - double(below);
  for (uint32_t k = 0; k < H; ++k) {
    const uint32_t c = s.h_counts[k];
    if (c > ctx->spill)
      return fail(ctx, FDAS_EOVERFLOW, "plane %u: %u values above the bound exceed the spill capacity %u", k + 1, c,
                  ctx->spill);
    std::vector<float> p(c);
    if (c) HIPCK(ctx, hipMemcpy(p.data(), s.spill_pwr + size_t(k) * ctx->spill, sizeof(float) * c,
                                hipMemcpyDeviceToHost));
    // descending ranks of the two order statistics
    const uint64_t r_lo = M - 1 - below, r_hi = below + 1 < M ? M - 2 - below : r_lo;
    if (r_lo >= c) return fail(ctx, FDAS_EDEVICE, "plane %u: only %u values spilled, need %llu", k + 1, c,
                               (unsigned long long)(r_lo + 1));
    std::sort(p.begin(), p.end(), std::greater<float>());
    const double lo = p[r_lo], hi = p[r_hi];
    thresholds[k] = float(lo * (1.0 - g) + hi * g);
  }
  return FDAS_OK;
}

static float elapsed(const DevSet& s, int a, int b) {
  if (!s.ev_recorded[a] || !s.ev_recorded[b]) return 0.f;
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, s.ev[a], s.ev[b]) != hipSuccess) {
    (void)hipGetLastError();
    return 0.f;
  }
  return ms;
}

extern "C" fdas_status fdas_get_stage_times(fdas_ctx* ctx, int set, float* ms6) {
  CHECK_SET(ctx, set);
  if (!ms6) return fail(ctx, FDAS_EINVAL, "null output");
  const DevSet& s = ctx->set[set];
  ms6[0] = elapsed(s, EV_START, EV_INPUT);
  ms6[1] = elapsed(s, EV_FTC_K, EV_FTC);
  ms6[2] = elapsed(s, EV_HSUM_K0, EV_HSUM_K1);
  ms6[3] = elapsed(s, EV_HSUM_K1, EV_HSUM);
  ms6[4] = elapsed(s, EV_HSUM, EV_XFER);
  ms6[5] = elapsed(s, EV_START, EV_XFER);
  return FDAS_OK;
}

extern "C" void fdas_print_stats(fdas_ctx* ctx, int set, int reset) {
  if (!ctx || (set != FDAS_SET_A && set != FDAS_SET_B)) return;
  const DevSet& s = ctx->set[set];
  float t0 = 0.f, t1 = 0.f;
  (void)hipEventElapsedTime(&t0, ctx->epoch, s.ev[EV_START]);
  (void)hipEventElapsedTime(&t1, ctx->epoch, s.ev[EV_XFER]);
  if (reset) {
    ctx->pipeline_start = t0;
    ctx->pipeline_accum = 0.f;
    ctx->pipeline_launches = 1;
  } else {
    ++ctx->pipeline_launches;
  }
  const float acc = t1 - ctx->pipeline_start;
  info(ctx, "Launch  #%u%c complete @ %.3f ms II ~ %.3f ms", ctx->pipeline_launches, set == FDAS_SET_A ? 'A' : 'B',
       double(acc), double(acc - ctx->pipeline_accum));
  info(ctx, "  Preparation duration: %.3f ms", double(elapsed(s, EV_START, EV_INPUT)));
  info(ctx, "  FT convolution and IFFT duration: %.3f ms", double(elapsed(s, EV_FTC_K, EV_FTC)));
  info(ctx, "  Harmonic summing duration: %.3f ms", double(elapsed(s, EV_HSUM_K0, EV_HSUM)));
  info(ctx, "  Total duration: %.3f ms", double(elapsed(s, EV_START, EV_XFER)));
  ctx->pipeline_accum = acc;
}

extern "C" void fdas_print_events(fdas_ctx* ctx, int set) {
  if (!ctx || !ctx->log || (set != FDAS_SET_A && set != FDAS_SET_B)) return;
  const DevSet& s = ctx->set[set];
  float t[EV_COUNT] = {};
  for (int e = 0; e < EV_COUNT; ++e)
    if (s.ev_recorded[e]) (void)hipEventElapsedTime(&t[e], ctx->epoch, s.ev[e]);
  char buf[256];
  std::snprintf(buf, sizeof(buf), "@@@,%c,%.6f,%.6f,%.6f,%.6f,%.6f,%.6f,%.6f,%.6f", set == FDAS_SET_A ? 'A' : 'B',
                double(t[EV_START]), double(t[EV_INPUT]), double(t[EV_FTC_K]), double(t[EV_FTC]),
                double(t[EV_HSUM_K0]), double(t[EV_HSUM_K1]), double(t[EV_HSUM]), double(t[EV_XFER]));
  emit(ctx, FDAS_LOG_INFO, buf);
}

extern "C" int fdas_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

extern "C" fdas_status fdas_device_info(int device, char* name, size_t len, char* arch, size_t arch_len) {
  const int n = fdas_device_count();
  if (device < 0 || device >= n) return FDAS_EINVAL;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
    (void)hipGetLastError();
    return FDAS_EDEVICE;
  }
  if (name && len) {
    std::snprintf(name, len, "%s", prop.name[0] ? prop.name : prop.gcnArchName);
  }
  if (arch && arch_len) std::snprintf(arch, arch_len, "%s", prop.gcnArchName);
  return FDAS_OK;
}

extern "C" const char* fdas_platform_version(void) {
  static char buf[64] = {0};
  if (!buf[0]) {
    int v = 0;
    if (hipRuntimeGetVersion(&v) != hipSuccess) (void)hipGetLastError();
    std::snprintf(buf, sizeof(buf), "HIP %d.%d.%d", v / 10000000, (v / 100000) % 100, v % 100000);
  }
  return buf;
}
