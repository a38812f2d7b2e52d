// Batched and served FDAS behind the engine / plan entry points of include/fdas.h.
//
// The reference runs one trial per launch on two fixed buffer sets
// (FDAS.cpp:282-615).  Here an engine keeps banks, twiddles and a device arena
// bounded by an HBM budget; trials of any N, bank, harmonic count and part are
// cut into groups.  One group's FT convolution runs on the convolution stream
// while the previous group's harmonic summing, selection and read-back run on
// the summing stream, so the VALU-bound convolution overlaps the HBM-bound
// summing.  The same group machinery serves fixed batches (plans, optionally
// captured once as a HIP graph) and a continuous-batching request loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstdint>
#include <chrono>
#include <cmath>
#include <complex>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "fdas_geometry.h"
#include "fdas_host.h"
#include "fdas_jobs.h"
#include "fdas_kernels.h"
#include "fdas_plan.h"

using namespace fdas;

namespace {

constexpr uint32_t kTileSizes[3] = {1024, 2048, 4096};
constexpr uint32_t kMaxRowGroups = 64;  // same cap as the single-trial path, so both plan identical row groups
constexpr uint64_t kAlign = 256;
constexpr uint64_t kArenaChunk = 4ull << 30;

uint64_t align_up(uint64_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// First-fit sub-allocator over device chunks; the chunks are reserved lazily and
// never exceed the budget in total.
class Arena {
 public:
  explicit Arena(uint64_t budget) : budget_(budget) {}
  ~Arena() { release(); }

  void release() {
    for (auto& c : chunks_) (void)hipFree(c.base);
    chunks_.clear();
    live_.clear();
    reserved_ = in_use_ = 0;
  }

  // nullptr when `bytes` cannot be placed within the budget (err: the HIP error, if any).
  char* alloc(uint64_t bytes, hipError_t* err) {
    bytes = align_up(std::max<uint64_t>(bytes, 1));
    for (size_t ci = 0; ci < chunks_.size(); ++ci)
      if (char* p = take(ci, bytes)) return p;
    // Chunks reserved for earlier, smaller requests may together hold the budget
    // while each is too small: give back the wholly free ones and retry.
    if (reserved_ + bytes > budget_) release_free_chunks();
    if (reserved_ + bytes > budget_) return nullptr;
    const uint64_t size = std::min(budget_ - reserved_, std::max(bytes, kArenaChunk));
    void* base = nullptr;
    const hipError_t e = hipMalloc(&base, size);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      if (err) *err = e;
      return nullptr;
    }
    chunks_.push_back(Chunk{static_cast<char*>(base), size, {{0, size}}});
    reserved_ += size;
    return take(chunks_.size() - 1, bytes);
  }

  void free(void* p) {
    auto it = live_.find(p);
    if (it == live_.end()) return;
    Chunk& c = chunks_[it->second.first];
    uint64_t off = uint64_t(static_cast<char*>(p) - c.base), size = it->second.second;
    in_use_ -= size;
    live_.erase(it);
    auto nx = c.free.lower_bound(off);
    if (nx != c.free.end() && off + size == nx->first) {
      size += nx->second;
      nx = c.free.erase(nx);
    }
    if (nx != c.free.begin()) {
      auto pv = std::prev(nx);
      if (pv->first + pv->second == off) {
        pv->second += size;
        return;
      }
    }
    c.free.emplace(off, size);
  }

  // Returns every chunk that holds no live allocation to the device.
  void release_free_chunks() {
    std::vector<size_t> keep_index(chunks_.size(), SIZE_MAX);
    std::vector<Chunk> kept;
    for (size_t ci = 0; ci < chunks_.size(); ++ci) {
      Chunk& c = chunks_[ci];
      const bool empty = c.free.size() == 1 && c.free.begin()->first == 0 && c.free.begin()->second == c.size;
      if (empty) {
        (void)hipFree(c.base);
        reserved_ -= c.size;
      } else {
        keep_index[ci] = kept.size();
        kept.push_back(std::move(c));
      }
    }
    for (auto& kv : live_) kv.second.first = keep_index[kv.second.first];
    chunks_ = std::move(kept);
  }

  uint64_t budget() const { return budget_; }
  uint64_t reserved() const { return reserved_; }
  uint64_t in_use() const { return in_use_; }
  uint64_t peak() const { return peak_; }

 private:
  struct Chunk {
    char* base;
    uint64_t size;
    std::map<uint64_t, uint64_t> free;  // offset -> size
  };

  char* take(size_t ci, uint64_t bytes) {
    Chunk& c = chunks_[ci];
    for (auto it = c.free.begin(); it != c.free.end(); ++it) {
      if (it->second < bytes) continue;
      const uint64_t off = it->first, rest = it->second - bytes;
      c.free.erase(it);
      if (rest) c.free.emplace(off + bytes, rest);
      char* p = c.base + off;
      live_[p] = {ci, bytes};
      in_use_ += bytes;
      peak_ = std::max(peak_, in_use_);
      return p;
    }
    return nullptr;
  }

  uint64_t budget_, reserved_ = 0, in_use_ = 0, peak_ = 0;
  std::vector<Chunk> chunks_;
  std::unordered_map<void*, std::pair<size_t, uint64_t>> live_;
};

struct Bank {
  uint32_t n_templates = 0, n_taps = 0, max_extent = 1;
  std::vector<float> taps;        // T x n_taps complex
  std::vector<uint32_t> extent;   // non-zero taps per row
  std::map<uint32_t, float2*> spectra;  // tile -> [T][16][S/16] kernel layout, 1/S scaled
};

// Work shape of (N, bank, part), fixed once: the tile decides the FOP rounding,
// so a shape keeps it for the engine's lifetime.
struct Shape {
  uint32_t tile = 2048, rows_per_item = 16, hsum_rows = kHsumRowsPerBlock;
  uint32_t row_begin = 0, n_rows = 0;
  std::vector<FtcJob> ftc;  // row groups, pointers unset, item_begin relative
  uint32_t n_items = 0;
};

// Byte offsets of one trial's device state, in two parts.  The small part
// (lists, counters, selection histogram) is zeroed before every run and read
// back; the large part (selection pool, spill lists, input copy, FOP) is not.
// Plans keep the small parts of all their trials in one array, so a group
// zeroes and reads back its trials with one launch and one copy.
struct TrialLayout {
  // small part, from its base
  uint64_t out_n, counts, out_key, out_pwr, compact_n, hist, small_bytes;
  uint64_t readback_bytes;  // out_n .. out_pwr
  // large part, from its base
  uint64_t compact_key, compact_pwr, spill_key, spill_pwr, input, fop, large_bytes;
  uint64_t total;  // small_bytes + large_bytes
};

TrialLayout trial_layout(uint32_t n_bins, uint32_t T, uint32_t H, uint32_t cap, uint32_t spill, bool own_input) {
  TrialLayout l{};
  uint64_t o = 0;
  l.out_n = o;
  o += 4ull * H;
  l.counts = o;
  o += 4ull * H;
  l.out_key = o;
  o += 4ull * H * cap;
  l.out_pwr = o;
  o += 4ull * H * cap;
  l.readback_bytes = o;
  l.compact_n = o;
  o += 4ull * H;
  l.hist = o;
  o += 4ull * H * kSelectBins;
  l.small_bytes = align_up(o);
  const uint64_t ccap = uint64_t(cap) + kSelectPool;
  o = 0;
  l.compact_key = o;
  o = align_up(o + 4ull * H * ccap);
  l.compact_pwr = o;
  o = align_up(o + 4ull * H * ccap);
  l.spill_key = o;
  o = align_up(o + 4ull * H * spill);
  l.spill_pwr = o;
  o = align_up(o + 4ull * H * spill);
  l.input = o;
  if (own_input) o = align_up(o + 8ull * n_bins);
  l.fop = o;
  o = align_up(o + 4ull * T * n_bins);
  l.large_bytes = o;
  l.total = l.small_bytes + l.large_bytes;
  return l;
}

// One trial placed in a group.
struct Trial {
  fdas_request req{};
  fdas_result* res = nullptr;     // served trials: caller's result
  const Shape* shape = nullptr;
  const Bank* bank = nullptr;
  TrialLayout lay{};
  uint32_t spill = 0;             // spill capacity per plane
  char* small = nullptr;          // device small part (served trials: also the arena block)
  char* mem = nullptr;            // device large part
  const float2* input = nullptr;  // device spectrum
  char* h_read = nullptr;         // pinned mirror of the small part
  double t_submit = 0.0;
};

// Device and pinned job arrays of one group, and its launch list.
struct GroupBuf {
  FtcJob* d_ftc = nullptr;
  FtcJob* h_ftc = nullptr;
  uint32_t ftc_cap = 0;
  HsumJob* d_hsum = nullptr;
  HsumJob* h_hsum = nullptr;
  uint32_t hsum_cap = 0;
  SelectJob* d_sel = nullptr;
  SelectJob* h_sel = nullptr;
  uint32_t sel_cap = 0;
  // ev_ftc1 / ev_done order the streams; ev_ftc0 / ev_ftc_t1 / ev_hsum0 / ev_hsum1 only time the kernels.
  hipEvent_t ev_ftc0 = nullptr, ev_ftc1 = nullptr, ev_hsum0 = nullptr, ev_hsum1 = nullptr, ev_done = nullptr;
  hipEvent_t ev_ftc_t1 = nullptr;
  hipEvent_t ev_hsum_done = nullptr;  // harmonic summing finished (the next group's convolution waits on it)
  bool launch_fused = false;           // this group's convolution ran in a fused kernel (with the previous summing)
  bool hsum_fused = false;             // this group's summing ran in the next group's fused kernel
  struct FtcLaunch {
    uint32_t tile, first, n_jobs, n_items;
  };
  struct HsumLaunch {
    uint32_t H, first, n_jobs, n_blocks;
    bool vec;
  };
  std::vector<FtcLaunch> ftc;
  std::vector<HsumLaunch> hsum;
  uint32_t n_sel = 0;
  // Small parts of the group's trials when they are one contiguous span (plans).
  char* small_span = nullptr;
  char* h_span = nullptr;
  uint64_t span_bytes = 0;
};

}  // namespace

struct fdas_engine {
  fdas_engine_config cfg{};
  fdas_log_fn log = nullptr;
  void* user = nullptr;
  std::string last_error;
  std::mutex err_mu;  // last_error is written by API calls and by the scheduler thread
  std::mutex api_mu;  // banks, shapes, arena and streams (held by the scheduler while it launches)
  std::unique_ptr<Arena> arena;
  std::vector<std::unique_ptr<Bank>> banks;
  std::map<std::tuple<uint32_t, uint32_t, uint32_t>, std::unique_ptr<Shape>> shapes;
  std::map<uint32_t, std::pair<float2*, float2*>> twiddles;  // tile -> (tw1, tw2)
  hipStream_t s_ftc = nullptr, s_hsum = nullptr;
  float2* tune_input = nullptr;  // 2^22 bins of seeded complex noise, E|X|^2 = 1 (tuning input)
  uint32_t n_cu = 0;
  std::string device_key;        // device name, architecture and CU count (tuning-cache key)
  std::string tune_cache_path;   // tuned shapes are read from and appended to this file
  std::map<std::string, std::tuple<uint32_t, uint32_t, uint32_t>> tune_cache;  // key -> (tile, rows/item, hsum rows)

  // serving state (guarded by q_mu)
  std::mutex q_mu;
  std::condition_variable q_cv, done_cv;
  std::deque<Trial> queue;
  std::deque<fdas_result*> done;
  std::thread scheduler;
  bool stop = false, running = false;
  struct Ctx {
    GroupBuf buf;
    std::vector<Trial> trials;
    char* h_read = nullptr;
    uint64_t h_read_cap = 0;
    bool busy = false;
    uint64_t seq = 0;
  };
  std::vector<Ctx> ctxs;
  uint64_t launch_seq = 0;
  std::vector<fdas_plan*> plans;   // live plans, which hold arena memory (guarded by api_mu)
  hipEvent_t last_hsum = nullptr;  // summing-done event of the last served group
  fdas_engine_stats stats{};
};

struct fdas_plan {
  fdas_engine* eng = nullptr;
  std::vector<Trial> trials;
  std::vector<std::vector<uint32_t>> groups;
  std::vector<GroupBuf> bufs;
  std::vector<char*> regions;
  std::vector<char*> inputs;  // plan-owned device copies of host spectra (nullptr for device spectra)
  char* small_all = nullptr;  // small parts of all trials, in trial order
  char* h_read = nullptr;     // its pinned mirror
  uint64_t region_bytes = 0;
  hipEvent_t ev_start = nullptr, ev_end = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool use_graph = false, launched = false, eager_timed = false;
  uint32_t schedule = FDAS_SCHEDULE_LAGGED;
  uint32_t n_launches = 0;
};

namespace {

void emit(fdas_engine* e, int level, const std::string& msg) {
  if (e && e->log) e->log(level, msg.c_str(), e->user);
}

fdas_status fail(fdas_engine* e, fdas_status st, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (e) {
    const std::string msg = std::string("[ERROR] ") + buf;
    {
      std::lock_guard<std::mutex> lk(e->err_mu);
      e->last_error = msg;
    }
    emit(e, FDAS_LOG_ERROR, msg);
  }
  return st;
}

void info(fdas_engine* e, const char* fmt, ...) {
  if (!e || !e->log) return;
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  emit(e, FDAS_LOG_INFO, std::string("[INFO] ") + buf);
}

#define EHIP(eng, expr)                                                                                        \
  do {                                                                                                         \
    hipError_t e_ = (expr);                                                                                    \
    if (e_ != hipSuccess)                                                                                      \
      return fail(eng, FDAS_EDEVICE, "HIP call failed with status %d (%s):\n          %s [%s:%d]", int(e_), \
                  hipGetErrorString(e_), #expr, __FILE__, __LINE__);                                           \
  } while (0)

// Workgroups of an FT-convolution launch over n_items items.
uint32_t ftc_grid(const fdas_engine* e, uint32_t n_items) {
  const uint32_t cap = e->cfg.ftc_workgroups_per_cu * e->n_cu;
  return cap ? std::min(n_items, cap) : n_items;
}

void part_rows(uint32_t T, uint32_t part, uint32_t* begin, uint32_t* count) {
  const uint32_t c = T / 2;
  switch (part) {
    case FDAS_NEGATIVE: *begin = 0; *count = c + 1; break;
    case FDAS_POSITIVE: *begin = c; *count = c + 1; break;
    default: *begin = 0; *count = T; break;
  }
}

bool tile_usable(uint32_t tile, uint32_t max_extent) { return 2 * overlap_for(max_extent) <= tile; }

template <typename T>
fdas_status dev_alloc(fdas_engine* e, T** p, size_t count) {
  *p = nullptr;
  const hipError_t r = hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(count, 1) * sizeof(T));
  if (r != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return fail(e, FDAS_ENOMEM, "device allocation of %zu bytes failed (%s)", count * sizeof(T), hipGetErrorString(r));
  }
  return FDAS_OK;
}

template <typename T>
fdas_status host_alloc(fdas_engine* e, T** p, size_t count) {
  *p = nullptr;
  const hipError_t r = hipHostMalloc(reinterpret_cast<void**>(p), std::max<size_t>(count, 1) * sizeof(T), 0);
  if (r != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return fail(e, FDAS_ENOMEM, "pinned host allocation of %zu bytes failed (%s)", count * sizeof(T),
                hipGetErrorString(r));
  }
  return FDAS_OK;
}

void free_group(GroupBuf& g) {
  (void)hipFree(g.d_ftc);
  (void)hipFree(g.d_hsum);
  (void)hipFree(g.d_sel);
  (void)hipHostFree(g.h_ftc);
  (void)hipHostFree(g.h_hsum);
  (void)hipHostFree(g.h_sel);
  for (hipEvent_t ev : {g.ev_ftc0, g.ev_ftc1, g.ev_hsum0, g.ev_hsum1, g.ev_done, g.ev_ftc_t1, g.ev_hsum_done})
    if (ev) (void)hipEventDestroy(ev);
  g = GroupBuf{};
}

fdas_status group_events(fdas_engine* e, GroupBuf& g) {
  for (hipEvent_t* ev : {&g.ev_ftc0, &g.ev_ftc1, &g.ev_hsum0, &g.ev_hsum1, &g.ev_done, &g.ev_ftc_t1})
    if (!*ev) EHIP(e, hipEventCreate(ev));
  if (!g.ev_hsum_done) EHIP(e, hipEventCreateWithFlags(&g.ev_hsum_done, hipEventDisableTiming));
  return FDAS_OK;
}

// Grows the job arrays of g to hold at least the given counts.
fdas_status group_reserve(fdas_engine* e, GroupBuf& g, uint32_t n_ftc, uint32_t n_hsum, uint32_t n_sel) {
  fdas_status st;
  if (n_ftc > g.ftc_cap) {
    (void)hipFree(g.d_ftc);
    (void)hipHostFree(g.h_ftc);
    g.d_ftc = nullptr;
    g.h_ftc = nullptr;
    g.ftc_cap = 0;
    if ((st = dev_alloc(e, &g.d_ftc, n_ftc)) != FDAS_OK || (st = host_alloc(e, &g.h_ftc, n_ftc)) != FDAS_OK) return st;
    g.ftc_cap = n_ftc;
  }
  if (n_hsum > g.hsum_cap) {
    (void)hipFree(g.d_hsum);
    (void)hipHostFree(g.h_hsum);
    g.d_hsum = nullptr;
    g.h_hsum = nullptr;
    g.hsum_cap = 0;
    if ((st = dev_alloc(e, &g.d_hsum, n_hsum)) != FDAS_OK || (st = host_alloc(e, &g.h_hsum, n_hsum)) != FDAS_OK)
      return st;
    g.hsum_cap = n_hsum;
  }
  if (n_sel > g.sel_cap) {
    (void)hipFree(g.d_sel);
    (void)hipHostFree(g.h_sel);
    g.d_sel = nullptr;
    g.h_sel = nullptr;
    g.sel_cap = 0;
    if ((st = dev_alloc(e, &g.d_sel, n_sel)) != FDAS_OK || (st = host_alloc(e, &g.h_sel, n_sel)) != FDAS_OK) return st;
    g.sel_cap = n_sel;
  }
  return group_events(e, g);
}

// Fills the host job arrays of a group of placed trials (mem, input, h_read set)
// and its launch list: FT-convolution jobs grouped by tile, summing jobs by H.
fdas_status build_group(fdas_engine* e, GroupBuf& g, const std::vector<Trial*>& trials) {
  const uint32_t cap = e->cfg.cands_per_harmonic, ccap = cap + kSelectPool;
  uint32_t n_ftc = 0, n_sel = 0;
  for (const Trial* t : trials) {
    n_ftc += uint32_t(t->shape->ftc.size());
    n_sel += t->req.n_harmonics;
  }
  fdas_status st = group_reserve(e, g, n_ftc, uint32_t(trials.size()), n_sel);
  if (st != FDAS_OK) return st;
  g.ftc.clear();
  g.hsum.clear();
  uint32_t jf = 0;
  for (uint32_t tile : kTileSizes) {
    GroupBuf::FtcLaunch L{tile, jf, 0, 0};
    for (const Trial* t : trials) {
      if (t->shape->tile != tile) continue;
      float* fop = reinterpret_cast<float*>(t->mem + t->lay.fop);
      for (FtcJob j : t->shape->ftc) {
        j.spectrum = t->input;
        j.fop = fop;
        j.hperm = t->bank->spectra.at(tile);
        j.item_begin = L.n_items;
        L.n_items += j.n_tiles * j.n_chunks;
        g.h_ftc[jf++] = j;
      }
    }
    L.n_jobs = jf - L.first;
    if (L.n_jobs) g.ftc.push_back(L);
  }
  uint32_t jh = 0, js = 0;
  for (uint32_t H = 1; H <= kMaxHarmonics; ++H) {
    GroupBuf::HsumLaunch L{H, jh, 0, 0, true};
    for (const Trial* t : trials) {
      if (t->req.n_harmonics != H) continue;
      char* m = t->mem;
      char* sm = t->small;
      const TrialLayout& l = t->lay;
      const uint32_t spill = t->spill;
      HsumJob& j = g.h_hsum[jh++];
      j = HsumJob{};
      j.fop = reinterpret_cast<const float*>(m + l.fop);
      j.n_bins = t->req.n_bins;
      j.n_templates = t->bank->n_templates;
      j.row_begin = t->shape->row_begin;
      j.n_rows = t->shape->n_rows;
      j.n_harmonics = H;
      j.spill_cap = spill;
      j.block_begin = L.n_blocks;
      hsum_row_chunks(j.n_rows, &j.rows_per_block, &j.n_row_chunks, t->shape->hsum_rows);
      j.spill_key = reinterpret_cast<uint32_t*>(m + l.spill_key);
      j.spill_pwr = reinterpret_cast<float*>(m + l.spill_pwr);
      j.counts = reinterpret_cast<uint32_t*>(sm + l.counts);
      for (uint32_t k = 0; k < H; ++k) j.thr[k] = t->req.thresholds[k];
      L.n_blocks += (j.n_bins + kHsumBinsPerBlock - 1) / kHsumBinsPerBlock * j.n_row_chunks;
      L.vec = L.vec && hsum_vec_ok(j.n_bins);
      for (uint32_t p = 0; p < H; ++p) {
        SelectJob& q = g.h_sel[js++];
        q.spill_key = j.spill_key + size_t(p) * spill;
        q.spill_pwr = j.spill_pwr + size_t(p) * spill;
        q.count = j.counts + p;
        q.spill_cap = spill;
        q.cap = cap;
        q.thr = t->req.thresholds[p];
        q.compact_cap = ccap;
        q.hist = reinterpret_cast<uint32_t*>(sm + l.hist) + size_t(p) * kSelectBins;
        q.compact_key = reinterpret_cast<uint32_t*>(m + l.compact_key) + size_t(p) * ccap;
        q.compact_pwr = reinterpret_cast<float*>(m + l.compact_pwr) + size_t(p) * ccap;
        q.compact_n = reinterpret_cast<uint32_t*>(sm + l.compact_n) + p;
        q.out_key = reinterpret_cast<uint32_t*>(sm + l.out_key) + size_t(p) * cap;
        q.out_pwr = reinterpret_cast<float*>(sm + l.out_pwr) + size_t(p) * cap;
        q.out_n = reinterpret_cast<uint32_t*>(sm + l.out_n) + p;
      }
    }
    L.n_jobs = jh - L.first;
    if (L.n_jobs) g.hsum.push_back(L);
  }
  g.n_sel = js;
  (void)n_sel;
  // One zeroing launch and one read-back copy when the small parts are contiguous.
  g.small_span = nullptr;
  g.h_span = nullptr;
  g.span_bytes = 0;
  bool contiguous = !trials.empty();
  for (size_t i = 1; i < trials.size() && contiguous; ++i)
    contiguous = trials[i]->small == trials[i - 1]->small + trials[i - 1]->lay.small_bytes &&
                 trials[i]->h_read == trials[i - 1]->h_read + trials[i - 1]->lay.small_bytes;
  if (contiguous) {
    g.small_span = trials.front()->small;
    g.h_span = trials.front()->h_read;
    g.span_bytes = uint64_t(trials.back()->small - trials.front()->small) + trials.back()->lay.small_bytes;
  }
  return FDAS_OK;
}

fdas_status upload_group(fdas_engine* e, GroupBuf& g, uint32_t n_ftc, uint32_t n_hsum, hipStream_t s) {
  EHIP(e, hipMemcpyAsync(g.d_ftc, g.h_ftc, sizeof(FtcJob) * n_ftc, hipMemcpyHostToDevice, s));
  EHIP(e, hipMemcpyAsync(g.d_hsum, g.h_hsum, sizeof(HsumJob) * n_hsum, hipMemcpyHostToDevice, s));
  EHIP(e, hipMemcpyAsync(g.d_sel, g.h_sel, sizeof(SelectJob) * g.n_sel, hipMemcpyHostToDevice, s));
  return FDAS_OK;
}

uint32_t group_ftc_jobs(const GroupBuf& g) {
  uint32_t n = 0;
  for (const auto& L : g.ftc) n += L.n_jobs;
  return n;
}

uint32_t group_hsum_jobs(const GroupBuf& g) {
  uint32_t n = 0;
  for (const auto& L : g.hsum) n += L.n_jobs;
  return n;
}

// Enqueues one group: convolution on s_ftc, then (after its end event) zeroing,
// summing, selection and read-back on s_hsum.  wait_before: an event s_ftc waits
// on first (the group whose memory this one reuses), or nullptr.  prev_hsum: the
// previous group's summing-done event, or nullptr: the convolution starts after
// it, since the two kernels only slow each other down when they share the CUs
// (both VALU-heavy), while the selection and read-back that follow the summing
// still run beside the next convolution.  Timing events bracket the two kernels
// (not under graph capture, where this runtime rejects externally visible
// event records).
fdas_status enqueue_group(fdas_engine* e, GroupBuf& g, const std::vector<Trial*>& trials, hipEvent_t wait_before,
                          hipEvent_t prev_hsum, bool timing_events) {
  hipStream_t A = e->s_ftc, B = e->s_hsum;
  if (wait_before) EHIP(e, hipStreamWaitEvent(A, wait_before, 0));
  if (prev_hsum) EHIP(e, hipStreamWaitEvent(A, prev_hsum, 0));
  if (timing_events) EHIP(e, hipEventRecord(g.ev_ftc0, A));
  for (const auto& L : g.ftc) {
    const auto& tw = e->twiddles.at(L.tile);
    EHIP(e, launch_ftconv(int(L.tile), g.d_ftc + L.first, L.n_jobs, L.n_items, tw.first, tw.second,
                          ftc_grid(e, L.n_items), A));
  }
  if (timing_events) EHIP(e, hipEventRecord(g.ev_ftc_t1, A));
  EHIP(e, hipEventRecord(g.ev_ftc1, A));
  EHIP(e, hipStreamWaitEvent(B, g.ev_ftc1, 0));
  if (g.small_span) {
    EHIP(e, launch_zero_u32(reinterpret_cast<uint32_t*>(g.small_span), uint32_t(g.span_bytes / 4), B));
  } else {
    for (const Trial* t : trials)
      EHIP(e, launch_zero_u32(reinterpret_cast<uint32_t*>(t->small), uint32_t(t->lay.small_bytes / 4), B));
  }
  if (timing_events) EHIP(e, hipEventRecord(g.ev_hsum0, B));
  for (const auto& L : g.hsum) EHIP(e, launch_hsum(g.d_hsum + L.first, L.n_jobs, L.n_blocks, L.H, L.vec, B));
  if (timing_events) EHIP(e, hipEventRecord(g.ev_hsum1, B));
  EHIP(e, hipEventRecord(g.ev_hsum_done, B));
  EHIP(e, launch_select(g.d_sel, g.n_sel, B));
  if (g.small_span) {
    EHIP(e, hipMemcpyAsync(g.h_span, g.small_span, g.span_bytes, hipMemcpyDeviceToHost, B));
  } else {
    for (const Trial* t : trials)
      EHIP(e, hipMemcpyAsync(t->h_read, t->small, t->lay.readback_bytes, hipMemcpyDeviceToHost, B));
  }
  EHIP(e, hipEventRecord(g.ev_done, B));
  return FDAS_OK;
}

// Copies a completed trial's lists from its pinned read-back area to the result.
void finish_result(const fdas_engine* e, const Trial& t, fdas_result* r) {
  const uint32_t H = t.req.n_harmonics, cap = e->cfg.cands_per_harmonic;
  const uint32_t* out_n = reinterpret_cast<const uint32_t*>(t.h_read + t.lay.out_n);
  const uint32_t* counts = reinterpret_cast<const uint32_t*>(t.h_read + t.lay.counts);
  const uint32_t* key = reinterpret_cast<const uint32_t*>(t.h_read + t.lay.out_key);
  const float* pwr = reinterpret_cast<const float*>(t.h_read + t.lay.out_pwr);
  r->tag = t.req.tag;
  r->n_harmonics = H;
  r->tile = t.shape->tile;
  r->status = FDAS_OK;
  for (uint32_t p = 0; p < kMaxHarmonics; ++p) {
    r->n_detected[p] = p < H ? counts[p] : 0;
    r->n_kept[p] = p < H ? std::min(out_n[p], cap) : 0;
    if (p < H && counts[p] > t.spill) r->status = FDAS_EOVERFLOW;
  }
  if (r->key) std::memcpy(r->key, key, sizeof(uint32_t) * H * cap);
  if (r->power) std::memcpy(r->power, pwr, sizeof(float) * H * cap);
}

fdas_status validate_request(fdas_engine* e, const fdas_request* q) {
  if (!q) return fail(e, FDAS_EINVAL, "null request");
  if (!q->spectrum) return fail(e, FDAS_EINVAL, "request %llu: null spectrum", (unsigned long long)q->tag);
  if (q->n_bins < 1 || q->n_bins > kMaxBins)
    return fail(e, FDAS_EINVAL, "request %llu: n_bins %u outside 1 .. %u", (unsigned long long)q->tag, q->n_bins,
                kMaxBins);
  if (q->bank >= e->banks.size())
    return fail(e, FDAS_EINVAL, "request %llu: unknown bank %u", (unsigned long long)q->tag, q->bank);
  if (q->n_harmonics < 1 || q->n_harmonics > kMaxHarmonics)
    return fail(e, FDAS_EINVAL, "request %llu: n_harmonics %u outside 1 .. %u", (unsigned long long)q->tag,
                q->n_harmonics, kMaxHarmonics);
  if (q->part > FDAS_ALL) return fail(e, FDAS_EINVAL, "request %llu: invalid part %u", (unsigned long long)q->tag, q->part);
  for (uint32_t k = 0; k < q->n_harmonics; ++k)
    if (std::isnan(q->thresholds[k]))
      return fail(e, FDAS_EINVAL, "request %llu: threshold %u is NaN", (unsigned long long)q->tag, k + 1);
  return FDAS_OK;
}

// Times one launch list (best of `reps`, after one warm-up) on s_ftc.
template <typename F>
fdas_status time_best(fdas_engine* e, F&& enqueue, int reps, float* best) {
  hipEvent_t a, b;
  EHIP(e, hipEventCreate(&a));
  EHIP(e, hipEventCreate(&b));
  *best = FLT_MAX;
  fdas_status st = FDAS_OK;
  for (int r = 0; r <= reps && st == FDAS_OK; ++r) {
    if (hipEventRecord(a, e->s_ftc) != hipSuccess) st = FDAS_EDEVICE;
    if (st == FDAS_OK) st = enqueue();
    if (st == FDAS_OK && hipEventRecord(b, e->s_ftc) != hipSuccess) st = FDAS_EDEVICE;
    if (st == FDAS_OK && hipEventSynchronize(b) != hipSuccess) st = FDAS_EDEVICE;
    float ms = 0.f;
    if (st == FDAS_OK && r > 0 && hipEventElapsedTime(&ms, a, b) == hipSuccess) *best = std::min(*best, ms);
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (st == FDAS_EDEVICE) return fail(e, st, "HIP call failed while autotuning");
  return st;
}

std::vector<FtcJob> shape_jobs(const Bank& bank, uint32_t tile, uint32_t n_bins, uint32_t rb, uint32_t nr,
                               uint32_t rows_per_item, uint32_t* n_items) {
  std::vector<FtcJob> jobs =
      plan_ftc_rows(tile, n_bins, bank.n_taps, bank.extent, rb, nr, kMaxRowGroups, rows_per_item);
  uint32_t item = 0;
  for (FtcJob& j : jobs) {
    j.item_begin = item;
    item += j.n_tiles * j.n_chunks;
  }
  *n_items = item;
  return jobs;
}

// Seeded complex Gaussian noise, E|X|^2 = 1: SplitMix64 -> Box-Muller (FP64),
// the same for every engine, so tuning sees the detection density of
// noise-dominated trials at the usual thresholds.
std::vector<float2> tuning_noise(uint32_t n) {
  std::vector<float2> x(n);
  uint64_t state = 0x5eedf00dull;
  auto next = [&]() {
    uint64_t z = (state += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  };
  for (uint32_t i = 0; i < n; ++i) {
    const double u1 = (double(next() >> 11) + 1.0) * 0x1.0p-53, u2 = double(next() >> 11) * 0x1.0p-53;
    const double r = std::sqrt(-std::log(u1)), a = 2.0 * 3.14159265358979323846 * u2;  // variance 1/2 per part
    x[i] = make_float2(float(r * std::cos(a)), float(r * std::sin(a)));
  }
  return x;
}

fdas_status ensure_tune_input(fdas_engine* e) {
  if (e->tune_input) return FDAS_OK;
  const std::vector<float2> x = tuning_noise(kMaxBins);
  fdas_status st = dev_alloc(e, &e->tune_input, kMaxBins);
  if (st != FDAS_OK) return st;
  EHIP(e, hipMemcpy(e->tune_input, x.data(), sizeof(float2) * x.size(), hipMemcpyHostToDevice));
  return FDAS_OK;
}

// FNV-1a over the bank's taps and sizes.
uint64_t bank_digest(const Bank& b) {
  uint64_t h = 0xcbf29ce484222325ull;
  auto mix = [&](const void* p, size_t n) {
    const unsigned char* c = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) h = (h ^ c[i]) * 0x100000001b3ull;
  };
  mix(&b.n_templates, sizeof(b.n_templates));
  mix(&b.n_taps, sizeof(b.n_taps));
  mix(b.taps.data(), b.taps.size() * sizeof(float));
  return h;
}

// Cache key of a shape: device, N, bank contents, part and the per-plane cap.
std::string tune_key(const fdas_engine* e, uint32_t n_bins, const Bank& bank, uint32_t part) {
  char buf[320];
  std::snprintf(buf, sizeof(buf), "%s|N=%u|bank=%016llx|part=%u|cap=%u", e->device_key.c_str(), n_bins,
                (unsigned long long)bank_digest(bank), part, e->cfg.cands_per_harmonic);
  return buf;
}

// Deterministic shape without timing: the usable tile with the fewest
// transform points per FOP output over the bank's row groups (S log2 S per
// transform), 16 rows per item, the default summing rows.
void model_shape(const fdas_engine* e, const Bank& bank, uint32_t n_bins, uint32_t rb, uint32_t nr,
                 const std::vector<uint32_t>& tiles, Shape* out) {
  double best = 0.0;
  for (uint32_t tile : tiles) {
    const std::vector<FtcJob> jobs =
        plan_ftc_rows(tile, n_bins, bank.n_taps, bank.extent, rb, nr, kMaxRowGroups, 16);
    double cost = 0.0;
    for (const FtcJob& j : jobs) cost += double(j.n_rows) * j.n_tiles * tile * std::log2(double(tile));
    if (best == 0.0 || cost < best) {
      best = cost;
      out->tile = tile;
    }
  }
  out->rows_per_item = e->cfg.rows_per_item ? e->cfg.rows_per_item : 16;
  out->hsum_rows = e->cfg.hsum_rows ? e->cfg.hsum_rows : kHsumRowsPerBlock;
}

// Times FT convolution (tile x rows per item) and then harmonic summing +
// selection (rows per summing workgroup) for a shape, on the seeded noise
// spectrum with thresholds (k + 6 sqrt(k)) E|X|^2, and keeps the fastest.
// Caller holds api_mu; s_ftc is idle or may be waited on.
fdas_status time_shape(fdas_engine* e, uint32_t n_bins, uint32_t bank_id, uint32_t part,
                       const std::vector<uint32_t>& tiles, const std::vector<uint32_t>& rpis,
                       const std::vector<uint32_t>& hrows, Shape* out) {
  const Bank& bank = *e->banks[bank_id];
  const uint32_t T = bank.n_templates, H = 8;
  fdas_status st = ensure_tune_input(e);
  if (st != FDAS_OK) return st;
  const uint32_t spill = e->cfg.spill_per_harmonic ? e->cfg.spill_per_harmonic : default_spill(T, n_bins);
  const TrialLayout lay = trial_layout(n_bins, T, H, e->cfg.cands_per_harmonic, spill, false);
  // Candidates are timed the way plans launch them: G copies of the noise trial
  // in one launch, each with a FOP of its own (so the written and re-read bytes
  // are a real group's), sharing the first copy's counters and lists; fewer
  // copies when the HBM budget has no room for G.
  const uint64_t fop_bytes = align_up(uint64_t(T) * n_bins * sizeof(float));
  hipError_t herr = hipSuccess;
  uint32_t G = std::max(1u, std::min(e->cfg.group_trials, 8u));
  char* block = nullptr;
  for (; G >= 1; G /= 2) {
    if ((block = e->arena->alloc(align_up(lay.total) + uint64_t(G - 1) * fop_bytes, &herr))) break;
    if (G == 1) break;
  }
  if (!block) {
    info(e, "autotune of N=%u bank %u skipped: no room in the HBM budget; using tile %u", n_bins, bank_id, out->tile);
    return FDAS_OK;
  }
  char* mem = block + lay.small_bytes;
  auto fop_of = [&](uint32_t c) {
    return reinterpret_cast<float*>(c == 0 ? mem + lay.fop : block + align_up(lay.total) + uint64_t(c - 1) * fop_bytes);
  };
  GroupBuf g;
  st = group_reserve(e, g, kMaxRowGroups * G, G, H);
  auto run_ftc = [&](uint32_t tile, const std::vector<FtcJob>& jobs, uint32_t n_items) -> fdas_status {
    const auto& tw = e->twiddles.at(tile);
    EHIP(e, launch_ftconv(int(tile), g.d_ftc, uint32_t(jobs.size()), n_items, tw.first, tw.second,
                          ftc_grid(e, n_items), e->s_ftc));
    return FDAS_OK;
  };
  auto ftc_jobs = [&](uint32_t tile, uint32_t rpi, uint32_t* n_items) {
    uint32_t one = 0;
    const std::vector<FtcJob> base = shape_jobs(bank, tile, n_bins, out->row_begin, out->n_rows, rpi, &one);
    std::vector<FtcJob> jobs;
    for (uint32_t c = 0; c < G; ++c) {
      for (FtcJob j : base) {
        j.spectrum = e->tune_input;
        j.fop = fop_of(c);
        j.hperm = bank.spectra.at(tile);
        j.item_begin += c * one;
        jobs.push_back(j);
      }
    }
    *n_items = G * one;
    return jobs;
  };
  // Two sweeps over the candidates, timing only the second: the first brings
  // the device to the clock it sustains under load (an idle device ramps up
  // over the first milliseconds, which would penalise whichever runs first).
  float best = FLT_MAX;
  for (int sweep = 0; sweep < 2 && st == FDAS_OK; ++sweep) {
    for (uint32_t tile : tiles) {
      for (uint32_t rpi : rpis) {
        if (st != FDAS_OK) break;
        uint32_t n_items = 0;
        const std::vector<FtcJob> jobs = ftc_jobs(tile, rpi, &n_items);
        if (hipMemcpy(g.d_ftc, jobs.data(), sizeof(FtcJob) * jobs.size(), hipMemcpyHostToDevice) != hipSuccess) {
          st = fail(e, FDAS_EDEVICE, "job upload failed while autotuning");
          break;
        }
        float ms = FLT_MAX;
        st = time_best(e, [&]() { return run_ftc(tile, jobs, n_items); }, sweep ? 3 : 1, &ms);
        if (sweep == 0) continue;
        info(e, "autotune N=%u bank %u: tile %u, %u rows/item: %.4f ms for %u trials", n_bins, bank_id, tile, rpi,
             ms, G);
        // A later candidate (larger tile or more rows per item) must win by 2 %,
        // so that near-ties, which timing noise decides, keep one stable shape.
        if (st == FDAS_OK && ms < 0.98f * best) {
          best = ms;
          out->tile = tile;
          out->rows_per_item = rpi;
        }
      }
    }
  }
  if (st == FDAS_OK && hrows.size() > 1) {
    // The winner's FOP, then the summing shapes on it.
    uint32_t n_items = 0;
    const std::vector<FtcJob> jobs = ftc_jobs(out->tile, out->rows_per_item, &n_items);
    if (hipMemcpy(g.d_ftc, jobs.data(), sizeof(FtcJob) * jobs.size(), hipMemcpyHostToDevice) != hipSuccess)
      st = fail(e, FDAS_EDEVICE, "job upload failed while autotuning");
    if (st == FDAS_OK) st = run_ftc(out->tile, jobs, n_items);
    Trial t;
    t.req.n_bins = n_bins;
    t.req.bank = bank_id;
    t.req.n_harmonics = H;
    t.req.part = part;
    for (uint32_t k = 1; k <= H; ++k) t.req.thresholds[k - 1] = float(k + 6.0 * std::sqrt(double(k)));
    t.bank = &bank;
    t.lay = lay;
    t.spill = spill;
    t.small = block;
    t.mem = mem;
    t.input = e->tune_input;
    float hbest = FLT_MAX;
    uint32_t chosen = out->hsum_rows;
    Shape trial_shape;
    for (uint32_t rows : hrows) {
      if (st != FDAS_OK) break;
      trial_shape.row_begin = out->row_begin;
      trial_shape.n_rows = out->n_rows;
      trial_shape.tile = out->tile;
      trial_shape.hsum_rows = rows;
      t.shape = &trial_shape;
      const std::vector<Trial*> one{&t};
      if ((st = build_group(e, g, one)) != FDAS_OK) break;
      // G copies of the trial's summing job (block ranges one after another);
      // the selection does not depend on the summing shape and is not timed.
      const GroupBuf::HsumLaunch L1 = g.hsum.at(0);
      for (uint32_t c = 1; c < G; ++c) {
        g.h_hsum[c] = g.h_hsum[0];
        g.h_hsum[c].fop = fop_of(c);
        g.h_hsum[c].block_begin = c * L1.n_blocks;
      }
      if ((st = upload_group(e, g, 0, G, e->s_ftc)) != FDAS_OK) break;
      float ms = FLT_MAX;
      st = time_best(
          e,
          [&]() -> fdas_status {
            EHIP(e, launch_zero_u32(reinterpret_cast<uint32_t*>(t.small), uint32_t(lay.small_bytes / 4), e->s_ftc));
            EHIP(e, launch_hsum(g.d_hsum, G, G * L1.n_blocks, L1.H, L1.vec, e->s_ftc));
            return FDAS_OK;
          },
          3, &ms);
      info(e, "autotune N=%u bank %u: %u rows/summing workgroup: %.4f ms for %u trials", n_bins, bank_id, rows, ms,
           G);
      // Candidates come in increasing height; a taller chunk must win by 3 %: the
      // noise spectrum has few detections, and dense-detection rows (strong
      // signals) lengthen the tail of tall chunks more than of short ones.
      if (st == FDAS_OK && ms < 0.97f * hbest) {
        hbest = ms;
        chosen = rows;
      }
    }
    out->hsum_rows = chosen;
  }
  if (hipStreamSynchronize(e->s_ftc) != hipSuccess && st == FDAS_OK)
    st = fail(e, FDAS_EDEVICE, "HIP call failed while autotuning");
  free_group(g);
  e->arena->free(block);
  return st;
}

void save_tuned(fdas_engine* e, const std::string& key, const Shape& sh) {
  e->tune_cache[key] = std::make_tuple(sh.tile, sh.rows_per_item, sh.hsum_rows);
  if (e->tune_cache_path.empty()) return;
  if (FILE* f = std::fopen(e->tune_cache_path.c_str(), "a")) {
    std::fprintf(f, "%s tile=%u rows_per_item=%u hsum_rows=%u\n", key.c_str(), sh.tile, sh.rows_per_item,
                 sh.hsum_rows);
    std::fclose(f);
  }
}

// Chooses (tile, rows per item, rows per summing workgroup) of a shape.
// Configured values are kept; the rest comes from the tuning cache, else from
// timing (may_time), else from the cost model.  Caller holds api_mu.
fdas_status tune_shape(fdas_engine* e, uint32_t n_bins, uint32_t bank_id, uint32_t part, bool may_time,
                       Shape* out) {
  const Bank& bank = *e->banks[bank_id];
  const uint32_t T = bank.n_templates;
  uint32_t rb, nr;
  part_rows(T, part, &rb, &nr);
  out->row_begin = rb;
  out->n_rows = nr;
  std::vector<uint32_t> tiles;
  if (e->cfg.tile) {
    tiles.push_back(e->cfg.tile);
  } else {
    for (uint32_t s : kTileSizes)
      if (tile_usable(s, bank.max_extent)) tiles.push_back(s);
  }
  const std::vector<uint32_t> rpis = e->cfg.rows_per_item ? std::vector<uint32_t>{e->cfg.rows_per_item}
                                                          : std::vector<uint32_t>{8, 16, 32, 64};
  const std::vector<uint32_t> hrows = e->cfg.hsum_rows ? std::vector<uint32_t>{e->cfg.hsum_rows}
                                                       : std::vector<uint32_t>{11, 16, 22};
  model_shape(e, bank, n_bins, rb, nr, tiles, out);
  const std::string key = tune_key(e, n_bins, bank, part);
  const char* how = "cost model";
  auto cached = e->tune_cache.find(key);
  if (cached != e->tune_cache.end() &&
      std::find(tiles.begin(), tiles.end(), std::get<0>(cached->second)) != tiles.end()) {
    out->tile = std::get<0>(cached->second);
    out->rows_per_item = e->cfg.rows_per_item ? e->cfg.rows_per_item : std::get<1>(cached->second);
    out->hsum_rows = e->cfg.hsum_rows ? e->cfg.hsum_rows : std::get<2>(cached->second);
    how = "tuning cache";
  } else if (tiles.size() == 1 && rpis.size() == 1 && hrows.size() == 1) {
    out->tile = tiles[0];
    out->rows_per_item = rpis[0];
    out->hsum_rows = hrows[0];
    how = "configuration";
  } else if (may_time) {
    const fdas_status st = time_shape(e, n_bins, bank_id, part, tiles, rpis, hrows, out);
    if (st != FDAS_OK) return st;
    save_tuned(e, key, *out);
    how = "timing";
  }
  out->ftc = shape_jobs(bank, out->tile, n_bins, rb, nr, out->rows_per_item, &out->n_items);
  info(e, "shape N=%u bank %u part %u (%s): tile %u, %u rows/item, %u rows/summing workgroup, %zu row groups",
       n_bins, bank_id, part, how, out->tile, out->rows_per_item, out->hsum_rows, out->ftc.size());
  return FDAS_OK;
}

// Shape of a request, fixed on first use (caller holds api_mu).
fdas_status get_shape(fdas_engine* e, uint32_t n_bins, uint32_t bank, uint32_t part, bool may_time,
                      const Shape** out) {
  const auto key = std::make_tuple(n_bins, bank, part);
  auto it = e->shapes.find(key);
  if (it == e->shapes.end()) {
    auto sh = std::make_unique<Shape>();
    const fdas_status st = tune_shape(e, n_bins, bank, part, may_time, sh.get());
    if (st != FDAS_OK) return st;
    it = e->shapes.emplace(key, std::move(sh)).first;
  }
  *out = it->second.get();
  return FDAS_OK;
}

// own_input: a host spectrum gets its device copy inside the trial's block
// (served trials); plans keep such copies apart, since groups reuse blocks.
fdas_status make_trial(fdas_engine* e, const fdas_request& q, bool own_input, Trial* t) {
  fdas_status st = validate_request(e, &q);
  if (st != FDAS_OK) return st;
  const Shape* sh = nullptr;
  if ((st = get_shape(e, q.n_bins, q.bank, q.part, /*may_time=*/false, &sh)) != FDAS_OK) return st;
  t->req = q;
  t->shape = sh;
  t->bank = e->banks[q.bank].get();
  t->spill = e->cfg.spill_per_harmonic ? e->cfg.spill_per_harmonic : default_spill(t->bank->n_templates, q.n_bins);
  t->lay = trial_layout(q.n_bins, t->bank->n_templates, q.n_harmonics, e->cfg.cands_per_harmonic, t->spill,
                        own_input && (q.flags & FDAS_REQ_DEVICE_SPECTRUM) == 0);
  return FDAS_OK;
}

// ------------------------------------------------------------------ serving

void scheduler_loop(fdas_engine* e);

fdas_status start_scheduler(fdas_engine* e) {
  if (e->running) return FDAS_OK;
  e->ctxs.resize(e->cfg.groups_in_flight);
  e->stop = false;
  e->running = true;
  e->scheduler = std::thread(scheduler_loop, e);
  return FDAS_OK;
}

// Completes a finished context: results out, memory back to the arena.
void complete_ctx(fdas_engine* e, fdas_engine::Ctx& c) {
  const double t = now_ms();
  std::vector<fdas_result*> out;
  {
    std::lock_guard<std::mutex> lk(e->api_mu);
    for (Trial& tr : c.trials) {
      finish_result(e, tr, tr.res);
      tr.res->latency_ms = t - tr.t_submit;
      e->arena->free(tr.small);
      out.push_back(tr.res);
    }
  }
  c.trials.clear();
  c.busy = false;
  std::lock_guard<std::mutex> lk(e->q_mu);
  for (fdas_result* r : out) e->done.push_back(r);
  e->stats.completed += out.size();
  e->stats.in_flight -= uint32_t(out.size());
  e->done_cv.notify_all();
}

void fail_trial(fdas_engine* e, Trial& t, fdas_status st) {
  t.res->tag = t.req.tag;
  t.res->status = st;
  t.res->n_harmonics = t.req.n_harmonics;
  t.res->latency_ms = now_ms() - t.t_submit;
  std::memset(t.res->n_detected, 0, sizeof(t.res->n_detected));
  std::memset(t.res->n_kept, 0, sizeof(t.res->n_kept));
  std::lock_guard<std::mutex> lk(e->q_mu);
  e->done.push_back(t.res);
  e->stats.completed += 1;
  e->done_cv.notify_all();
}

// Admits queued trials into a free context and launches them.  Returns true if
// anything was admitted or failed (progress).
// mem_wait is set when the oldest queued trial waits for arena memory.
bool admit(fdas_engine* e, fdas_engine::Ctx& c, bool others_in_flight, bool& mem_wait) {
  std::vector<Trial> picked;
  {
    std::lock_guard<std::mutex> lk(e->q_mu);
    if (e->queue.empty()) return false;
  }
  bool progress = false;
  std::lock_guard<std::mutex> lk_api(e->api_mu);
  while (picked.size() < e->cfg.group_trials) {
    Trial t;
    {
      std::lock_guard<std::mutex> lk(e->q_mu);
      if (e->queue.empty()) break;
      t = e->queue.front();
    }
    hipError_t herr = hipSuccess;
    char* mem = e->arena->alloc(t.lay.total, &herr);
    if (!mem) {
      // Impossible when the trial alone exceeds the budget, or when nothing
      // else holds memory (no group in flight, no live plan) and it still
      // does not fit.  Otherwise it waits for memory to come back.
      const bool never = t.lay.total > e->arena->budget() || (picked.empty() && !others_in_flight && e->plans.empty());
      if (never) {
        {
          // fdas_engine_cancel may have withdrawn it meanwhile
          std::lock_guard<std::mutex> lk(e->q_mu);
          if (e->queue.empty() || e->queue.front().res != t.res) continue;
          e->queue.pop_front();
          e->stats.queued -= 1;
        }
        fail(e, FDAS_ENOMEM, "request %llu needs %llu bytes of device state, beyond the HBM budget",
             (unsigned long long)t.req.tag, (unsigned long long)t.lay.total);
        fail_trial(e, t, FDAS_ENOMEM);
        progress = true;
        continue;
      }
      mem_wait = true;
      break;  // wait for memory to come back
    }
    {
      std::lock_guard<std::mutex> lk(e->q_mu);
      if (e->queue.empty() || e->queue.front().res != t.res) {  // withdrawn by fdas_engine_cancel meanwhile
        e->arena->free(mem);
        continue;
      }
      e->queue.pop_front();
      e->stats.queued -= 1;
      e->stats.in_flight += 1;
    }
    t.small = mem;
    t.mem = mem + t.lay.small_bytes;
    picked.push_back(t);
  }
  if (picked.empty()) return progress;
  // Pinned read-back area of the context.
  uint64_t rb = 0;
  for (const Trial& t : picked) rb += t.lay.small_bytes;
  if (rb > c.h_read_cap) {
    (void)hipHostFree(c.h_read);
    c.h_read = nullptr;
    c.h_read_cap = 0;
    if (host_alloc(e, &c.h_read, rb) == FDAS_OK) c.h_read_cap = rb;
  }
  fdas_status st = c.h_read ? FDAS_OK : FDAS_ENOMEM;
  uint64_t off = 0;
  std::vector<Trial*> ptrs;
  for (Trial& t : picked) {
    t.h_read = c.h_read + off;
    off += t.lay.small_bytes;
    if (t.req.flags & FDAS_REQ_DEVICE_SPECTRUM) {
      t.input = static_cast<const float2*>(t.req.spectrum);
    } else {
      t.input = reinterpret_cast<const float2*>(t.mem + t.lay.input);
      if (st == FDAS_OK && hipMemcpyAsync(t.mem + t.lay.input, t.req.spectrum, 8ull * t.req.n_bins,
                                          hipMemcpyHostToDevice, e->s_ftc) != hipSuccess)
        st = fail(e, FDAS_EDEVICE, "spectrum upload of request %llu failed", (unsigned long long)t.req.tag);
    }
    ptrs.push_back(&t);
  }
  if (st == FDAS_OK) st = build_group(e, c.buf, ptrs);
  if (st == FDAS_OK) st = upload_group(e, c.buf, group_ftc_jobs(c.buf), group_hsum_jobs(c.buf), e->s_ftc);
  if (st == FDAS_OK) st = enqueue_group(e, c.buf, ptrs, nullptr, e->last_hsum, false);
  if (st == FDAS_OK) e->last_hsum = c.buf.ev_hsum_done;
  if (st != FDAS_OK) {
    // The streams may hold part of the group: drain them before the memory goes back.
    (void)hipStreamSynchronize(e->s_ftc);
    (void)hipStreamSynchronize(e->s_hsum);
    for (Trial& t : picked) {
      e->arena->free(t.small);
      {
        std::lock_guard<std::mutex> lk(e->q_mu);
        e->stats.in_flight -= 1;
      }
      fail_trial(e, t, st);
    }
    return true;
  }
  c.trials = std::move(picked);
  c.busy = true;
  c.seq = ++e->launch_seq;
  std::lock_guard<std::mutex> lk(e->q_mu);
  e->stats.groups_launched += 1;
  e->stats.max_group = std::max<uint32_t>(e->stats.max_group, uint32_t(c.trials.size()));
  return true;
}

void scheduler_loop(fdas_engine* e) {
  (void)hipSetDevice(e->cfg.device);
  for (;;) {
    bool progress = false;
    // Completions, oldest first.
    for (;;) {
      fdas_engine::Ctx* oldest = nullptr;
      for (auto& c : e->ctxs)
        if (c.busy && (!oldest || c.seq < oldest->seq)) oldest = &c;
      if (!oldest) break;
      if (hipEventQuery(oldest->buf.ev_done) != hipSuccess) {
        (void)hipGetLastError();  // "not ready" must not reach the next launch's error check
        break;
      }
      complete_ctx(e, *oldest);
      progress = true;
    }
    bool any_busy = false, any_free = false, mem_wait = false;
    for (auto& c : e->ctxs) (c.busy ? any_busy : any_free) = true;
    if (any_free) {
      for (auto& c : e->ctxs) {
        if (c.busy) continue;
        if (admit(e, c, any_busy, mem_wait)) {
          progress = true;
          any_busy = true;
        }
        break;
      }
    }
    if (progress) continue;
    std::unique_lock<std::mutex> lk(e->q_mu);
    if (e->stop && e->queue.empty() && !any_busy) break;
    if (!any_busy && !mem_wait) {
      e->q_cv.wait(lk, [&] { return e->stop || !e->queue.empty(); });
    } else if (!any_busy) {
      // Memory is held by live plans: look again when one is destroyed (or after 1 ms).
      e->q_cv.wait_for(lk, std::chrono::milliseconds(1));
    } else {
      e->q_cv.wait_for(lk, std::chrono::microseconds(50));
    }
  }
}

}  // namespace

// ------------------------------------------------------------------ C ABI

extern "C" {

void fdas_engine_default_config(fdas_engine_config* cfg) {
  if (!cfg) return;
  *cfg = fdas_engine_config{};
  cfg->cands_per_harmonic = 4096;
  cfg->spill_per_harmonic = 0;
  cfg->group_trials = 8;
  cfg->groups_in_flight = 3;
  cfg->hbm_budget_bytes = 200ull * 1000 * 1000 * 1000;
}

static fdas_status engine_init(fdas_engine* e) {
  const fdas_engine_config& c = e->cfg;
  int n_dev = 0;
  if (hipGetDeviceCount(&n_dev) != hipSuccess || c.device < 0 || c.device >= n_dev) {
    (void)hipGetLastError();
    return fail(e, FDAS_EDEVICE, "device %d not available (%d devices)", c.device, n_dev);
  }
  EHIP(e, hipSetDevice(c.device));
  int n_cu = 0;
  EHIP(e, hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, c.device));
  e->n_cu = uint32_t(n_cu);
  hipDeviceProp_t prop{};
  EHIP(e, hipGetDeviceProperties(&prop, c.device));
  e->device_key = std::string(prop.name) + "/" + prop.gcnArchName + "/" + std::to_string(n_cu) + "CU";
  for (char& ch : e->device_key)
    if (ch == ' ' || ch == '\t' || ch == '|') ch = '_';
  EHIP(e, hipStreamCreateWithFlags(&e->s_ftc, hipStreamNonBlocking));
  EHIP(e, hipStreamCreateWithFlags(&e->s_hsum, hipStreamNonBlocking));
  for (uint32_t tile : kTileSizes) {
    std::vector<float2> tw1, tw2;
    make_twiddles(tile, tw1, tw2);
    auto& tw = e->twiddles[tile];
    fdas_status st;
    if ((st = dev_alloc(e, &tw.first, tw1.size())) != FDAS_OK) return st;
    if ((st = dev_alloc(e, &tw.second, tw2.size())) != FDAS_OK) return st;
    EHIP(e, hipMemcpy(tw.first, tw1.data(), sizeof(float2) * tw1.size(), hipMemcpyHostToDevice));
    EHIP(e, hipMemcpy(tw.second, tw2.data(), sizeof(float2) * tw2.size(), hipMemcpyHostToDevice));
  }
  e->arena = std::make_unique<Arena>(c.hbm_budget_bytes);
  return FDAS_OK;
}

fdas_status fdas_engine_create(const fdas_engine_config* cfg_in, fdas_log_fn log, void* user, fdas_engine** out) {
  if (!out) return FDAS_EINVAL;
  *out = nullptr;
  auto e = std::make_unique<fdas_engine>();
  e->log = log;
  e->user = user;
  fdas_engine_default_config(&e->cfg);
  if (cfg_in) e->cfg = *cfg_in;
  fdas_engine_config& c = e->cfg;
  if (c.group_trials == 0) c.group_trials = 8;
  if (c.groups_in_flight == 0) c.groups_in_flight = 3;
  if (c.hbm_budget_bytes == 0) c.hbm_budget_bytes = 200ull * 1000 * 1000 * 1000;
  if (c.tile && c.tile != 1024 && c.tile != 2048 && c.tile != 4096)
    return fail(e.get(), FDAS_EINVAL, "tile %u is not 1024, 2048 or 4096", c.tile);
  if (c.rows_per_item > 256) return fail(e.get(), FDAS_EINVAL, "rows_per_item %u above 256", c.rows_per_item);
  if (c.hsum_rows > 256) return fail(e.get(), FDAS_EINVAL, "hsum_rows %u above 256", c.hsum_rows);
  if (c.plan_schedule > FDAS_SCHEDULE_STREAMS)
    return fail(e.get(), FDAS_EINVAL, "plan_schedule %u is not LAGGED, FUSED or STREAMS", c.plan_schedule);
  if (c.ftc_workgroups_per_cu > 64)
    return fail(e.get(), FDAS_EINVAL, "ftc_workgroups_per_cu %u above 64", c.ftc_workgroups_per_cu);
  if (c.cands_per_harmonic < 1 || c.cands_per_harmonic > kSelectMaxCap)
    return fail(e.get(), FDAS_EINVAL, "cands_per_harmonic %u outside 1 .. %u", c.cands_per_harmonic, kSelectMaxCap);
  if (c.spill_per_harmonic && c.spill_per_harmonic < c.cands_per_harmonic)
    return fail(e.get(), FDAS_EINVAL, "spill_per_harmonic %u below cands_per_harmonic %u", c.spill_per_harmonic,
                c.cands_per_harmonic);
  if (c.groups_in_flight < 2 || c.groups_in_flight > 8)
    return fail(e.get(), FDAS_EINVAL, "groups_in_flight %u outside 2 .. 8", c.groups_in_flight);
  if (c.group_trials > 256) return fail(e.get(), FDAS_EINVAL, "group_trials %u above 256", c.group_trials);
  const fdas_status st = engine_init(e.get());
  if (st != FDAS_OK) {
    fdas_engine_destroy(e.release());
    return st;
  }
  info(e.get(), "engine on device %d: HBM budget %.1f GB, groups of <= %u trials, %u groups in flight", c.device,
       double(c.hbm_budget_bytes) / 1e9, c.group_trials, c.groups_in_flight);
  *out = e.release();
  return FDAS_OK;
}

static void plan_release(fdas_engine* e, fdas_plan* p);

void fdas_engine_destroy(fdas_engine* e) {
  if (!e) return;
  if (e->running) {
    {
      std::lock_guard<std::mutex> lk(e->q_mu);
      e->stop = true;
    }
    e->q_cv.notify_all();
    e->scheduler.join();
  }
  if (e->s_ftc) (void)hipSetDevice(e->cfg.device);
  {
    // Plans the caller did not destroy go with the engine; their handles become invalid.
    std::lock_guard<std::mutex> lk(e->api_mu);
    while (!e->plans.empty()) plan_release(e, e->plans.back());
  }
  if (e->s_ftc) (void)hipStreamSynchronize(e->s_ftc);
  if (e->s_hsum) (void)hipStreamSynchronize(e->s_hsum);
  for (auto& c : e->ctxs) {
    free_group(c.buf);
    (void)hipHostFree(c.h_read);
  }
  for (auto& b : e->banks)
    for (auto& kv : b->spectra) (void)hipFree(kv.second);
  for (auto& kv : e->twiddles) {
    (void)hipFree(kv.second.first);
    (void)hipFree(kv.second.second);
  }
  (void)hipFree(e->tune_input);
  e->arena.reset();
  if (e->s_ftc) (void)hipStreamDestroy(e->s_ftc);
  if (e->s_hsum) (void)hipStreamDestroy(e->s_hsum);
  delete e;
}

const char* fdas_engine_last_error(const fdas_engine* e) {
  if (!e) return "null engine";
  thread_local std::string copy;
  std::lock_guard<std::mutex> lk(const_cast<fdas_engine*>(e)->err_mu);
  copy = e->last_error;
  return copy.c_str();
}

fdas_status fdas_engine_add_bank(fdas_engine* e, const float* taps_c64, uint32_t n_templates, uint32_t n_taps,
                                 uint32_t* bank_id) {
  if (!e) return FDAS_EINVAL;
  if (!taps_c64 || !bank_id) return fail(e, FDAS_EINVAL, "null bank argument");
  if (n_templates < 1 || n_templates > kMaxTemplates || n_templates % 2 == 0)
    return fail(e, FDAS_EINVAL, "n_templates %u must be odd and <= %u", n_templates, kMaxTemplates);
  if (n_taps < 1 || n_taps > 2048) return fail(e, FDAS_EINVAL, "n_taps %u outside 1 .. 2048", n_taps);
  std::lock_guard<std::mutex> lk(e->api_mu);
  EHIP(e, hipSetDevice(e->cfg.device));
  auto b = std::make_unique<Bank>();
  b->n_templates = n_templates;
  b->n_taps = n_taps;
  b->taps.assign(taps_c64, taps_c64 + size_t(2) * n_templates * n_taps);
  for (uint32_t tile : kTileSizes) {
    if (e->cfg.tile && tile != e->cfg.tile) continue;
    std::vector<uint32_t> extent;
    const std::vector<std::complex<double>> spec =
        taps_to_spectra(b->taps.data(), n_templates, n_taps, n_taps, tile, extent);
    b->extent = extent;
    b->max_extent = *std::max_element(extent.begin(), extent.end());
    if (!tile_usable(tile, b->max_extent)) continue;
    const std::vector<float2> perm = kernel_template_layout(spec, n_templates, tile);
    float2* d = nullptr;
    fdas_status st = dev_alloc(e, &d, perm.size());
    if (st != FDAS_OK) return st;
    b->spectra[tile] = d;
    EHIP(e, hipMemcpy(d, perm.data(), sizeof(float2) * perm.size(), hipMemcpyHostToDevice));
  }
  if (b->spectra.empty()) {
    const uint32_t need = 2 * overlap_for(b->max_extent);
    return fail(e, FDAS_EUNSUPPORTED, "templates of %u taps need a tile of at least %u points%s", b->max_extent, need,
                e->cfg.tile ? " (the configured tile is smaller)" : "");
  }
  *bank_id = uint32_t(e->banks.size());
  info(e, "bank %u: %u templates x %u taps (longest %u), %zu tile sizes", *bank_id, n_templates, n_taps, b->max_extent,
       b->spectra.size());
  e->banks.push_back(std::move(b));
  return FDAS_OK;
}

static fdas_status prepare_shape(fdas_engine* e, uint32_t n_bins, uint32_t bank, uint32_t part, bool may_time,
                                 uint32_t* tile) {
  std::lock_guard<std::mutex> lk(e->api_mu);
  EHIP(e, hipSetDevice(e->cfg.device));
  if (n_bins < 1 || n_bins > kMaxBins) return fail(e, FDAS_EINVAL, "n_bins %u outside 1 .. %u", n_bins, kMaxBins);
  if (bank >= e->banks.size()) return fail(e, FDAS_EINVAL, "unknown bank %u", bank);
  if (part > FDAS_ALL) return fail(e, FDAS_EINVAL, "invalid part %u", part);
  const Shape* sh = nullptr;
  const fdas_status st = get_shape(e, n_bins, bank, part, may_time, &sh);
  if (st == FDAS_OK && tile) *tile = sh->tile;
  return st;
}

fdas_status fdas_engine_prepare(fdas_engine* e, uint32_t n_bins, uint32_t bank, uint32_t part, uint32_t* tile) {
  if (!e) return FDAS_EINVAL;
  return prepare_shape(e, n_bins, bank, part, /*may_time=*/true, tile);
}

fdas_status fdas_engine_shape_info(fdas_engine* e, uint32_t n_bins, uint32_t bank, uint32_t part,
                                   fdas_shape_info* out) {
  if (!e) return FDAS_EINVAL;
  if (!out) return fail(e, FDAS_EINVAL, "null shape info");
  std::lock_guard<std::mutex> lk(e->api_mu);
  auto it = e->shapes.find(std::make_tuple(n_bins, bank, part));
  if (it == e->shapes.end())
    return fail(e, FDAS_EINVAL, "shape N=%u bank %u part %u has not been prepared", n_bins, bank, part);
  const Shape& sh = *it->second;
  *out = fdas_shape_info{};
  out->tile = sh.tile;
  out->rows_per_item = sh.rows_per_item;
  out->hsum_rows = sh.hsum_rows;
  out->n_row_groups = uint32_t(sh.ftc.size());
  for (const FtcJob& j : sh.ftc) {
    out->forward_transforms += uint64_t(j.n_tiles) * j.n_chunks;
    out->inverse_transforms += uint64_t(j.n_tiles) * j.n_rows;
    out->stored_outputs += uint64_t(j.n_rows) * n_bins;
  }
  return FDAS_OK;
}

fdas_status fdas_engine_tune_cache(fdas_engine* e, const char* path) {
  if (!e) return FDAS_EINVAL;
  if (!path || !*path) return fail(e, FDAS_EINVAL, "empty tuning-cache path");
  std::lock_guard<std::mutex> lk(e->api_mu);
  e->tune_cache_path = path;
  uint32_t n = 0;
  if (FILE* f = std::fopen(path, "r")) {
    char line[512];
    while (std::fgets(line, sizeof(line), f)) {
      char key[400];
      unsigned tile = 0, rpi = 0, rows = 0;
      if (std::sscanf(line, "%399s tile=%u rows_per_item=%u hsum_rows=%u", key, &tile, &rpi, &rows) != 4) continue;
      const bool ok = (tile == 1024 || tile == 2048 || tile == 4096) && rpi >= 1 && rpi <= 256 && rows >= 1 &&
                      rows <= 256;
      if (!ok) continue;
      e->tune_cache[key] = std::make_tuple(uint32_t(tile), uint32_t(rpi), uint32_t(rows));  // later lines win
      ++n;
    }
    std::fclose(f);
  }
  info(e, "tuning cache %s: %u tuned shapes", path, n);
  return FDAS_OK;
}

// ---------------------------------------------------------------- plans

fdas_status fdas_plan_create(fdas_engine* e, const fdas_request* reqs, uint32_t n, uint32_t flags, fdas_plan** out) {
  if (!e) return FDAS_EINVAL;
  if (!out || !reqs || n == 0) return fail(e, FDAS_EINVAL, "plan needs at least one request");
  *out = nullptr;
  // Shapes first (tuning takes the lock itself).
  for (uint32_t i = 0; i < n; ++i) {
    const fdas_status st = fdas_engine_prepare(e, reqs[i].n_bins, reqs[i].bank, reqs[i].part, nullptr);
    if (st != FDAS_OK) return st;
  }
  std::lock_guard<std::mutex> lk(e->api_mu);
  EHIP(e, hipSetDevice(e->cfg.device));
  auto p = std::make_unique<fdas_plan>();
  p->eng = e;
  p->use_graph = (flags & FDAS_PLAN_GRAPH) != 0;
  p->schedule = e->cfg.plan_schedule;
  p->trials.resize(n);
  for (uint32_t i = 0; i < n; ++i) {
    fdas_status st = make_trial(e, reqs[i], false, &p->trials[i]);
    if (st != FDAS_OK) return st;
  }
  // Groups of consecutive trials.
  const uint32_t G = e->cfg.group_trials, R = e->cfg.groups_in_flight;
  for (uint32_t i = 0; i < n; i += G) {
    std::vector<uint32_t> g;
    for (uint32_t j = i; j < std::min(n, i + G); ++j) g.push_back(j);
    p->groups.push_back(g);
  }
  const uint32_t n_groups = uint32_t(p->groups.size()), n_regions = std::min(R, n_groups);
  std::vector<uint64_t> region_size(n_regions, 0);
  for (uint32_t g = 0; g < n_groups; ++g) {
    uint64_t s = 0;
    for (uint32_t i : p->groups[g]) s += p->trials[i].lay.large_bytes;
    region_size[g % n_regions] = std::max(region_size[g % n_regions], s);
  }
  auto cleanup = [&](fdas_status st) {
    if (p->small_all) e->arena->free(p->small_all);
    for (char* r : p->regions) e->arena->free(r);
    for (char* m : p->inputs) e->arena->free(m);
    for (auto& b : p->bufs) free_group(b);
    (void)hipHostFree(p->h_read);
    if (p->ev_start) (void)hipEventDestroy(p->ev_start);
    if (p->ev_end) (void)hipEventDestroy(p->ev_end);
    return st;
  };
  for (uint32_t r = 0; r < n_regions; ++r) {
    hipError_t herr = hipSuccess;
    char* m = e->arena->alloc(region_size[r], &herr);
    if (!m)
      return cleanup(fail(e, FDAS_ENOMEM, "plan needs %llu bytes of device state for %u groups in flight; the HBM "
                                          "budget (%llu bytes, %llu reserved) cannot hold them",
                          (unsigned long long)region_size[r], n_regions,
                          (unsigned long long)e->cfg.hbm_budget_bytes, (unsigned long long)e->arena->reserved()));
    p->regions.push_back(m);
    p->region_bytes += region_size[r];
  }
  // Device copies of host spectra, one per trial, outside the reused regions.
  p->inputs.assign(n, nullptr);
  for (uint32_t i = 0; i < n; ++i) {
    if (p->trials[i].req.flags & FDAS_REQ_DEVICE_SPECTRUM) continue;
    hipError_t herr = hipSuccess;
    p->inputs[i] = e->arena->alloc(8ull * p->trials[i].req.n_bins, &herr);
    if (!p->inputs[i])
      return cleanup(fail(e, FDAS_ENOMEM, "no room in the HBM budget for the device copy of spectrum %u", i));
    p->region_bytes += align_up(8ull * p->trials[i].req.n_bins);
  }
  // Small parts of all trials in one array (and its pinned mirror), in trial order.
  uint64_t rb = 0;
  for (const Trial& t : p->trials) rb += t.lay.small_bytes;
  {
    hipError_t herr = hipSuccess;
    p->small_all = e->arena->alloc(rb, &herr);
    if (!p->small_all) return cleanup(fail(e, FDAS_ENOMEM, "no room in the HBM budget for the plan's lists"));
    p->region_bytes += align_up(rb);
  }
  fdas_status st = host_alloc(e, &p->h_read, rb);
  if (st != FDAS_OK) return cleanup(st);
  uint64_t roff = 0;
  for (Trial& t : p->trials) {
    t.small = p->small_all + roff;
    t.h_read = p->h_read + roff;
    roff += t.lay.small_bytes;
  }
  p->bufs.resize(n_groups);
  for (uint32_t g = 0; g < n_groups; ++g) {
    uint64_t off = 0;
    std::vector<Trial*> ptrs;
    for (uint32_t i : p->groups[g]) {
      Trial& t = p->trials[i];
      t.mem = p->regions[g % n_regions] + off;
      off += t.lay.large_bytes;
      t.input = (t.req.flags & FDAS_REQ_DEVICE_SPECTRUM) ? static_cast<const float2*>(t.req.spectrum)
                                                         : reinterpret_cast<const float2*>(p->inputs[i]);
      ptrs.push_back(&t);
    }
    if ((st = build_group(e, p->bufs[g], ptrs)) != FDAS_OK) return cleanup(st);
    if ((st = upload_group(e, p->bufs[g], group_ftc_jobs(p->bufs[g]), group_hsum_jobs(p->bufs[g]), e->s_ftc)) !=
        FDAS_OK)
      return cleanup(st);
  }
  if (hipStreamSynchronize(e->s_ftc) != hipSuccess) return cleanup(fail(e, FDAS_EDEVICE, "job upload failed"));
  if (hipEventCreate(&p->ev_start) != hipSuccess || hipEventCreate(&p->ev_end) != hipSuccess)
    return cleanup(fail(e, FDAS_EDEVICE, "event creation failed"));
  info(e, "plan: %u trials in %u groups, %u groups resident (%.2f GB)", n, n_groups, n_regions,
       double(p->region_bytes) / 1e9);
  e->plans.push_back(p.get());
  *out = p.release();
  return FDAS_OK;
}

fdas_status fdas_plan_set_request(fdas_plan* p, uint32_t i, const void* spectrum, const float* thresholds) {
  if (!p) return FDAS_EINVAL;
  fdas_engine* e = p->eng;
  if (p->launched) return fail(e, FDAS_ESTATE, "plan is running; wait for it before changing a request");
  if (i >= p->trials.size()) return fail(e, FDAS_EINVAL, "request index %u outside the plan's %zu", i, p->trials.size());
  std::lock_guard<std::mutex> lk(e->api_mu);
  EHIP(e, hipSetDevice(e->cfg.device));
  Trial& t = p->trials[i];
  if (spectrum) t.req.spectrum = spectrum;
  if (thresholds)
    for (uint32_t k = 0; k < t.req.n_harmonics; ++k) t.req.thresholds[k] = thresholds[k];
  if (t.req.flags & FDAS_REQ_DEVICE_SPECTRUM) t.input = static_cast<const float2*>(t.req.spectrum);
  // Rebuild the group's job arrays (pointers and thresholds live in device memory,
  // so a captured graph needs no update).
  for (uint32_t g = 0; g < p->groups.size(); ++g) {
    if (std::find(p->groups[g].begin(), p->groups[g].end(), i) == p->groups[g].end()) continue;
    std::vector<Trial*> ptrs;
    for (uint32_t j : p->groups[g]) ptrs.push_back(&p->trials[j]);
    fdas_status st = build_group(e, p->bufs[g], ptrs);
    if (st != FDAS_OK) return st;
    if ((st = upload_group(e, p->bufs[g], group_ftc_jobs(p->bufs[g]), group_hsum_jobs(p->bufs[g]), e->s_ftc)) !=
        FDAS_OK)
      return st;
    EHIP(e, hipStreamSynchronize(e->s_ftc));
  }
  return FDAS_OK;
}

// Lagged schedule of a plan's G groups on the convolution stream A: launch g
// (g = 0..G) runs group g's FT convolution together with group g-1's harmonic
// summing, as one fused kernel when the pair's shapes allow it (1024-point
// tiles, one summing shape), else as consecutive kernels.  Group g-1's
// selection and read-back follow on stream B beside launch g+1.
static fdas_status plan_enqueue_lagged(fdas_plan* p, bool capturing) {
  fdas_engine* e = p->eng;
  hipStream_t A = e->s_ftc, B = e->s_hsum;
  const uint32_t R = uint32_t(p->regions.size()), G = uint32_t(p->groups.size());
  std::vector<std::vector<Trial*>> ptrs(G);
  for (uint32_t g = 0; g < G; ++g)
    for (uint32_t i : p->groups[g]) ptrs[g].push_back(&p->trials[i]);
  for (uint32_t g = 0; g <= G; ++g) {
    GroupBuf* cur = g < G ? &p->bufs[g] : nullptr;    // convolved in this launch
    GroupBuf* prev = g > 0 ? &p->bufs[g - 1] : nullptr;  // summed in this launch
    if (cur && g >= R) EHIP(e, hipStreamWaitEvent(A, p->bufs[g - R].ev_done, 0));  // its region is free again
    if (prev) {
      if (prev->small_span) {
        EHIP(e, launch_zero_u32(reinterpret_cast<uint32_t*>(prev->small_span), uint32_t(prev->span_bytes / 4), A));
      } else {
        for (const Trial* t : ptrs[g - 1])
          EHIP(e, launch_zero_u32(reinterpret_cast<uint32_t*>(t->small), uint32_t(t->lay.small_bytes / 4), A));
      }
    }
    const bool fuse = p->schedule == FDAS_SCHEDULE_FUSED && cur && prev && cur->ftc.size() == 1 &&
                      prev->hsum.size() == 1 && cur->ftc[0].tile == 1024 &&
                      fused_supported(1024, prev->hsum[0].H, prev->hsum[0].vec);
    if (cur) cur->launch_fused = fuse;
    if (prev) prev->hsum_fused = fuse;
    // Timing events bracket each kernel: the convolution (or the fused kernel)
    // with ev_ftc0 / ev_ftc_t1 of its group, the summing with ev_hsum0 / ev_hsum1.
    if (cur && !capturing) EHIP(e, hipEventRecord(cur->ev_ftc0, A));
    if (fuse) {
      const auto& F = cur->ftc[0];
      const auto& S = prev->hsum[0];
      const auto& tw = e->twiddles.at(1024);
      EHIP(e, launch_fused(cur->d_ftc + F.first, F.n_jobs, F.n_items, tw.first, tw.second, prev->d_hsum + S.first,
                           S.n_jobs, S.n_blocks, S.H, S.vec, A));
      if (!capturing) EHIP(e, hipEventRecord(cur->ev_ftc_t1, A));
    } else {
      if (cur) {
        for (const auto& L : cur->ftc) {
          const auto& tw = e->twiddles.at(L.tile);
          EHIP(e, launch_ftconv(int(L.tile), cur->d_ftc + L.first, L.n_jobs, L.n_items, tw.first, tw.second,
                                ftc_grid(e, L.n_items), A));
        }
        if (!capturing) EHIP(e, hipEventRecord(cur->ev_ftc_t1, A));
      }
      if (prev) {
        if (!capturing) EHIP(e, hipEventRecord(prev->ev_hsum0, A));
        for (const auto& L : prev->hsum)
          EHIP(e, launch_hsum(prev->d_hsum + L.first, L.n_jobs, L.n_blocks, L.H, L.vec, A));
        if (!capturing) EHIP(e, hipEventRecord(prev->ev_hsum1, A));
      }
    }
    if (prev) {
      EHIP(e, hipEventRecord(prev->ev_hsum_done, A));
      EHIP(e, hipStreamWaitEvent(B, prev->ev_hsum_done, 0));
      EHIP(e, launch_select(prev->d_sel, prev->n_sel, B));
      if (prev->small_span) {
        EHIP(e, hipMemcpyAsync(prev->h_span, prev->small_span, prev->span_bytes, hipMemcpyDeviceToHost, B));
      } else {
        for (const Trial* t : ptrs[g - 1])
          EHIP(e, hipMemcpyAsync(t->h_read, t->small, t->lay.readback_bytes, hipMemcpyDeviceToHost, B));
      }
      EHIP(e, hipEventRecord(prev->ev_done, B));
    }
  }
  EHIP(e, hipStreamWaitEvent(A, p->bufs.back().ev_done, 0));
  return FDAS_OK;
}

// Two streams: group g's convolution (stream A) runs beside group g-1's
// summing, selection and read-back (stream B).
static fdas_status plan_enqueue_streams(fdas_plan* p, bool capturing) {
  fdas_engine* e = p->eng;
  const uint32_t R = uint32_t(p->regions.size());
  for (uint32_t g = 0; g < p->groups.size(); ++g) {
    std::vector<Trial*> ptrs;
    for (uint32_t i : p->groups[g]) ptrs.push_back(&p->trials[i]);
    const fdas_status st =
        enqueue_group(e, p->bufs[g], ptrs, g >= R ? p->bufs[g - R].ev_done : nullptr, nullptr, !capturing);
    if (st != FDAS_OK) return st;
  }
  EHIP(e, hipStreamWaitEvent(e->s_ftc, p->bufs.back().ev_done, 0));  // join the summing stream
  return FDAS_OK;
}

static fdas_status plan_enqueue(fdas_plan* p, bool capturing) {
  return p->schedule == FDAS_SCHEDULE_STREAMS ? plan_enqueue_streams(p, capturing) : plan_enqueue_lagged(p, capturing);
}

fdas_status fdas_plan_launch(fdas_plan* p) {
  if (!p) return FDAS_EINVAL;
  fdas_engine* e = p->eng;
  if (p->launched) return fail(e, FDAS_ESTATE, "plan already launched; wait for it first");
  std::lock_guard<std::mutex> lk(e->api_mu);
  EHIP(e, hipSetDevice(e->cfg.device));
  // Host spectra go to the plan's device copies first (outside any graph).
  for (Trial& t : p->trials)
    if (!(t.req.flags & FDAS_REQ_DEVICE_SPECTRUM))
      EHIP(e, hipMemcpyAsync(const_cast<float2*>(t.input), t.req.spectrum, 8ull * t.req.n_bins, hipMemcpyHostToDevice,
                             e->s_ftc));
  EHIP(e, hipEventRecord(p->ev_start, e->s_ftc));
  if (p->use_graph) {
    if (!p->exec) {
      EHIP(e, hipStreamBeginCapture(e->s_ftc, hipStreamCaptureModeThreadLocal));
      fdas_status st = plan_enqueue(p, true);
      hipGraph_t graph = nullptr;
      const hipError_t end = hipStreamEndCapture(e->s_ftc, &graph);
      if (st != FDAS_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
      }
      EHIP(e, end);
      p->graph = graph;
      EHIP(e, hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
      info(e, "plan captured as a HIP graph");
    }
    EHIP(e, hipGraphLaunch(p->exec, e->s_ftc));
    p->eager_timed = false;
  } else {
    const fdas_status st = plan_enqueue(p, false);
    if (st != FDAS_OK) return st;
    p->eager_timed = true;
  }
  EHIP(e, hipEventRecord(p->ev_end, e->s_ftc));
  p->launched = true;
  p->n_launches += 1;
  return FDAS_OK;
}

fdas_status fdas_plan_wait(fdas_plan* p, fdas_result* results) {
  if (!p) return FDAS_EINVAL;
  fdas_engine* e = p->eng;
  if (!p->launched) return fail(e, FDAS_ESTATE, "plan not launched");
  EHIP(e, hipSetDevice(e->cfg.device));
  EHIP(e, hipEventSynchronize(p->ev_end));
  p->launched = false;
  fdas_status worst = FDAS_OK;
  if (results) {
    for (size_t i = 0; i < p->trials.size(); ++i) {
      finish_result(e, p->trials[i], &results[i]);
      results[i].latency_ms = 0.0;
      if (results[i].status != FDAS_OK) worst = fdas_status(results[i].status);
    }
  }
  if (worst == FDAS_EOVERFLOW)
    return fail(e, FDAS_EOVERFLOW, "detections exceeded the spill capacity in at least one trial; its lists are "
                                   "incomplete (raise the threshold or spill_per_harmonic)");
  return FDAS_OK;
}

fdas_status fdas_plan_get_info(const fdas_plan* p, fdas_plan_info* out) {
  if (!p || !out) return FDAS_EINVAL;
  *out = fdas_plan_info{};
  out->n_trials = uint32_t(p->trials.size());
  out->n_groups = uint32_t(p->groups.size());
  out->n_launches = p->n_launches;
  out->graph = p->exec ? 1u : 0u;
  out->region_bytes = p->region_bytes;
  if (p->n_launches && !p->launched) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p->ev_start, p->ev_end) == hipSuccess) out->last_ms = ms;
    (void)hipGetLastError();
    out->lagged = p->schedule != FDAS_SCHEDULE_STREAMS;
    for (const GroupBuf& g : p->bufs) {
      if (!p->eager_timed) break;
      float a = 0.f, b = 0.f;
      if (hipEventElapsedTime(&a, g.ev_ftc0, g.ev_ftc_t1) == hipSuccess) {
        if (g.launch_fused) {
          out->fused_ms += a;
          out->n_fused += 1;
        } else {
          out->ftc_ms += a;
        }
      }
      (void)hipGetLastError();
      if (!g.hsum_fused && hipEventElapsedTime(&b, g.ev_hsum0, g.ev_hsum1) == hipSuccess) out->hsum_ms += b;
      (void)hipGetLastError();
    }
  }
  return FDAS_OK;
}

// Releases a plan's device state and removes it from its engine.  Caller holds api_mu.
static void plan_release(fdas_engine* e, fdas_plan* p) {
  (void)hipSetDevice(e->cfg.device);
  if (p->launched) (void)hipEventSynchronize(p->ev_end);
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  for (auto& b : p->bufs) free_group(b);
  for (char* r : p->regions) e->arena->free(r);
  for (char* m : p->inputs) e->arena->free(m);
  if (p->small_all) e->arena->free(p->small_all);
  (void)hipHostFree(p->h_read);
  if (p->ev_start) (void)hipEventDestroy(p->ev_start);
  if (p->ev_end) (void)hipEventDestroy(p->ev_end);
  e->plans.erase(std::find(e->plans.begin(), e->plans.end(), p));
  delete p;
}

void fdas_plan_destroy(fdas_plan* p) {
  if (!p) return;
  fdas_engine* e = p->eng;
  {
    std::lock_guard<std::mutex> lk(e->api_mu);
    plan_release(e, p);
  }
  e->q_cv.notify_all();  // a served trial may be waiting for the memory just returned
}

// ---------------------------------------------------------------- serving

fdas_status fdas_engine_submit(fdas_engine* e, const fdas_request* req, fdas_result* result) {
  if (!e) return FDAS_EINVAL;
  if (!result) return fail(e, FDAS_EINVAL, "null result");
  if (!req) return fail(e, FDAS_EINVAL, "null request");
  Trial t;
  {
    // A shape first seen while requests are being served is not timed (that
    // would stall the scheduler and skew its timings): it takes the tuning
    // cache's choice or the cost model's.  fdas_engine_prepare times it beforehand.
    bool serving;
    {
      std::lock_guard<std::mutex> lk(e->q_mu);
      serving = e->running;
    }
    fdas_status st = prepare_shape(e, req->n_bins, req->bank, req->part, !serving, nullptr);
    if (st != FDAS_OK) return st;
    std::lock_guard<std::mutex> lk(e->api_mu);
    if ((st = make_trial(e, *req, true, &t)) != FDAS_OK) return st;
  }
  t.res = result;
  t.t_submit = now_ms();
  std::lock_guard<std::mutex> lk(e->q_mu);
  if (e->stop) return fail(e, FDAS_ESTATE, "engine is shutting down");
  start_scheduler(e);
  e->queue.push_back(t);
  e->stats.submitted += 1;
  e->stats.queued += 1;
  e->q_cv.notify_all();
  return FDAS_OK;
}

fdas_status fdas_engine_poll(fdas_engine* e, fdas_result** done, uint32_t max, uint32_t* n_done, int32_t timeout_ms) {
  if (!e || !n_done || (!done && max)) return FDAS_EINVAL;
  *n_done = 0;
  std::unique_lock<std::mutex> lk(e->q_mu);
  if (timeout_ms < 0) {
    e->done_cv.wait(lk, [&] { return !e->done.empty() || e->stats.completed >= e->stats.submitted; });
  } else if (timeout_ms > 0) {
    e->done_cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), [&] { return !e->done.empty(); });
  }
  while (*n_done < max && !e->done.empty()) {
    done[(*n_done)++] = e->done.front();
    e->done.pop_front();
  }
  return FDAS_OK;
}

fdas_status fdas_engine_cancel(fdas_engine* e, uint64_t tag, uint32_t* n_canceled) {
  if (!e) return FDAS_EINVAL;
  uint32_t n = 0;
  {
    std::lock_guard<std::mutex> lk(e->q_mu);
    for (auto it = e->queue.begin(); it != e->queue.end();) {
      if (it->req.tag != tag) {
        ++it;
        continue;
      }
      fdas_result* r = it->res;
      r->tag = tag;
      r->status = FDAS_ECANCELED;
      r->n_harmonics = it->req.n_harmonics;
      r->tile = it->shape ? it->shape->tile : 0;
      r->latency_ms = now_ms() - it->t_submit;
      std::memset(r->n_detected, 0, sizeof(r->n_detected));
      std::memset(r->n_kept, 0, sizeof(r->n_kept));
      e->done.push_back(r);
      e->stats.queued -= 1;
      e->stats.completed += 1;
      it = e->queue.erase(it);
      ++n;
    }
    if (n) e->done_cv.notify_all();
  }
  if (n_canceled) *n_canceled = n;
  return FDAS_OK;
}

fdas_status fdas_engine_get_stats(fdas_engine* e, fdas_engine_stats* out) {
  if (!e || !out) return FDAS_EINVAL;
  fdas_engine_stats s;
  {
    std::lock_guard<std::mutex> lk(e->q_mu);
    s = e->stats;
  }
  {
    std::lock_guard<std::mutex> lk(e->api_mu);
    s.arena_reserved_bytes = e->arena->reserved();
    s.arena_in_use_bytes = e->arena->in_use();
    s.arena_peak_bytes = e->arena->peak();
  }
  *out = s;
  return FDAS_OK;
}

}  // extern "C"
