// fdas_hsum_noalloc: harmonic summing, thresholding and candidate selection of
// a FOP already in device memory, with every piece of scratch (job
// descriptors, spill lists, counters, histograms, compaction lists) in one
// file-scope device array, so a call allocates nothing.  It runs the kernels
// of fdas_perform_harmonic_summing (FDAS.cpp:407-546 in the reference) and
// gives the same lists.  For callers that cannot allocate: code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <mutex>

#include "fdas_host.h"
#include "fdas_geometry.h"
#include "fdas_jobs.h"
#include "fdas_kernels.h"
#include "fdas_plan.h"

using namespace fdas;

// The scratch: one copy per GPU, alive until the process ends.  Host code
// never passes this name to a kernel (on the host it is only a stand-in for
// the symbol); it uses the per-device address from hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[FDAS_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;

// Job descriptors of one call, written into the scratch by a kernel that takes
// them as its argument (no host-to-device copy from pageable memory).
struct NoallocJobs {
  HsumJob hsum;
  SelectJob sel[kMaxHarmonics];
};
static_assert(sizeof(NoallocJobs) <= 3072, "job descriptors must fit a kernel argument block");
constexpr uint64_t kJobsBytes = 4096;

__global__ void put_jobs_kernel(NoallocJobs jobs, NoallocJobs* dst) {
  const uint32_t* src = reinterpret_cast<const uint32_t*>(&jobs);
  uint32_t* out = reinterpret_cast<uint32_t*>(dst);
  for (uint32_t i = threadIdx.x; i < sizeof(NoallocJobs) / 4; i += blockDim.x) out[i] = src[i];
}

std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

uint64_t align256(uint64_t v) { return (v + 255) / 256 * 256; }

fdas_status report(fdas_log_fn log, void* user, fdas_status st, const char* fmt, ...) {
  if (log) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    char msg[600];
    std::snprintf(msg, sizeof(msg), "[ERROR] fdas_hsum_noalloc: %s", buf);
    log(FDAS_LOG_ERROR, msg, user);
  }
  return st;
}

}  // namespace

extern "C" fdas_status fdas_hsum_noalloc(const float* d_fop, uint32_t n_bins, uint32_t n_templates, fdas_part part,
                                         uint32_t n_harmonics, const float* thresholds, uint32_t cands_per_harmonic,
                                         uint32_t spill_per_harmonic, uint32_t* d_key, float* d_power,
                                         uint32_t* d_counts, void* stream, fdas_log_fn log, void* user) {
  // Everything is checked before the device is touched.
  if (!d_fop || !thresholds || !d_key || !d_power || !d_counts)
    return report(log, user, FDAS_EINVAL, "null pointer argument");
  if (n_bins < 1 || n_bins > kMaxBins) return report(log, user, FDAS_EINVAL, "n_bins %u outside 1 .. %u", n_bins, kMaxBins);
  if (n_templates < 1 || n_templates > kMaxTemplates || n_templates % 2 == 0)
    return report(log, user, FDAS_EINVAL, "n_templates %u must be odd and <= %u", n_templates, kMaxTemplates);
  if (n_harmonics < 1 || n_harmonics > kMaxHarmonics)
    return report(log, user, FDAS_EINVAL, "n_harmonics %u outside 1 .. %u", n_harmonics, kMaxHarmonics);
  if (part != FDAS_NEGATIVE && part != FDAS_POSITIVE && part != FDAS_ALL)
    return report(log, user, FDAS_EINVAL, "invalid FOP part %d", int(part));
  if (cands_per_harmonic < 1 || cands_per_harmonic > kSelectMaxCap)
    return report(log, user, FDAS_EINVAL, "cands_per_harmonic %u outside 1 .. %u", cands_per_harmonic, kSelectMaxCap);
  for (uint32_t k = 0; k < n_harmonics; ++k)
    if (std::isnan(thresholds[k])) return report(log, user, FDAS_EINVAL, "threshold %u is NaN", k + 1);
  const uint32_t H = n_harmonics, cap = cands_per_harmonic, ccap = cap + kSelectPool;
  const uint32_t spill = spill_per_harmonic ? spill_per_harmonic : default_spill(n_templates, n_bins);
  if (spill < cap) return report(log, user, FDAS_EINVAL, "spill_per_harmonic %u below cands_per_harmonic %u", spill, cap);
  // Scratch layout: jobs | counters + histograms | compaction lists | spill lists.
  const uint64_t o_counts = kJobsBytes;
  const uint64_t o_compact = align256(o_counts + 4ull * H * (1 + kSelectBins));
  const uint64_t o_spill = align256(o_compact + 8ull * H * ccap);
  const uint64_t need = o_spill + 8ull * H * spill;
  if (need > FDAS_NOALLOC_SCRATCH_BYTES)
    return report(log, user, FDAS_EINVAL,
                  "%u planes x %u spilled detections (+ %u-entry lists) need %llu bytes of scratch; the maximum is "
                  "%llu (at most %llu spill entries per plane for %u planes and %u candidates)",
                  H, spill, cap, (unsigned long long)need, (unsigned long long)FDAS_NOALLOC_SCRATCH_BYTES,
                  (unsigned long long)((FDAS_NOALLOC_SCRATCH_BYTES - o_spill) / (8ull * H)), H, cap);

  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    (void)hipGetLastError();
    return report(log, user, FDAS_EDEVICE, "no current HIP device");
  }
  char* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
      void* p = nullptr;
      const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        return report(log, user, FDAS_EDEVICE, "scratch address of device %d unavailable (%s)", dev,
                      hipGetErrorString(e));
      }
      g_scratch[dev] = p;
    }
    base = static_cast<char*>(g_scratch[dev]);
  }
  uint32_t rb = 0, nr = n_templates;
  if (part == FDAS_NEGATIVE) nr = n_templates / 2 + 1;
  if (part == FDAS_POSITIVE) rb = n_templates / 2, nr = n_templates / 2 + 1;

  NoallocJobs* d_jobs = reinterpret_cast<NoallocJobs*>(base);
  uint32_t* compact_n = reinterpret_cast<uint32_t*>(base + o_counts);
  uint32_t* hist = compact_n + H;
  uint32_t* compact_key = reinterpret_cast<uint32_t*>(base + o_compact);
  float* compact_pwr = reinterpret_cast<float*>(base + o_compact + 4ull * H * ccap);
  uint32_t* spill_key = reinterpret_cast<uint32_t*>(base + o_spill);
  float* spill_pwr = reinterpret_cast<float*>(base + o_spill + 4ull * H * spill);

  NoallocJobs jobs{};
  HsumJob& j = jobs.hsum;
  j.fop = d_fop;
  j.n_bins = n_bins;
  j.n_templates = n_templates;
  j.row_begin = rb;
  j.n_rows = nr;
  j.n_harmonics = H;
  j.spill_cap = spill;
  j.block_begin = 0;
  hsum_row_chunks(nr, &j.rows_per_block, &j.n_row_chunks);
  j.spill_key = spill_key;
  j.spill_pwr = spill_pwr;
  j.counts = d_counts + H;  // detections per plane, before the cap
  for (uint32_t k = 0; k < H; ++k) j.thr[k] = thresholds[k];
  for (uint32_t p = 0; p < H; ++p) {
    SelectJob& q = jobs.sel[p];
    q.spill_key = spill_key + size_t(p) * spill;
    q.spill_pwr = spill_pwr + size_t(p) * spill;
    q.count = d_counts + H + p;
    q.spill_cap = spill;
    q.cap = cap;
    q.thr = thresholds[p];
    q.compact_cap = ccap;
    q.hist = hist + size_t(p) * kSelectBins;
    q.compact_key = compact_key + size_t(p) * ccap;
    q.compact_pwr = compact_pwr + size_t(p) * ccap;
    q.compact_n = compact_n + p;
    q.out_key = d_key + size_t(p) * cap;
    q.out_pwr = d_power + size_t(p) * cap;
    q.out_n = d_counts + p;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  // Every call shares this device's scratch and keeps no state between calls:
  // what must start at zero (counters, histograms, the caller's counts) is
  // cleared by fill kernels first.
  hipLaunchKernelGGL(put_jobs_kernel, dim3(1), dim3(256), 0, s, jobs, d_jobs);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = launch_zero_u32(d_counts, 2 * H, s);
  if (e == hipSuccess) e = launch_zero_u32(compact_n, H * (1 + kSelectBins), s);
  const uint32_t blocks = (n_bins + kHsumBinsPerBlock - 1) / kHsumBinsPerBlock * j.n_row_chunks;
  if (e == hipSuccess) e = launch_hsum(&d_jobs->hsum, 1, blocks, H, hsum_vec_ok(n_bins), s);
  if (e == hipSuccess) e = launch_select(d_jobs->sel, H, s);
  if (e != hipSuccess) return report(log, user, FDAS_EDEVICE, "kernel launch failed (%s)", hipGetErrorString(e));
  return FDAS_OK;
}
