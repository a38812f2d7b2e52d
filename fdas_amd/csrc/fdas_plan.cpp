// Host-side planning shared by the single-trial context and the batch/serve engine.
#include "fdas_plan.h"

#include <algorithm>
#include <cmath>

#include "fdas_geometry.h"
#include "fdas_host.h"
#include "fdas_kernels.h"

namespace fdas {

namespace {
constexpr double kPi = 3.14159265358979323846264338327950288;
}

uint32_t ftc_rows_per_item(uint32_t n_rows, uint32_t target) {
  target = std::max<uint32_t>(1, target);
  const uint32_t chunks = std::max<uint32_t>(1, (n_rows + target - 1) / target);
  return (n_rows + chunks - 1) / chunks;
}

// Overlap-save returns the same linear convolution for any overlap >= L - 1,
// so the grouping changes the work (payload S - O per transform), not the
// result.  Minimises sum over groups of n_tiles(O) * (rows + forward
// transforms) by dynamic programming over the row boundaries.
std::vector<FtcJob> plan_ftc_rows(uint32_t tile, uint32_t n_bins, uint32_t max_taps,
                                  const std::vector<uint32_t>& row_taps, uint32_t row_begin, uint32_t n_rows,
                                  uint32_t max_groups, uint32_t rows_target) {
  const uint32_t S = tile, N = n_bins;
  auto taps = [&](uint32_t r) { return row_taps.empty() ? max_taps : std::max(1u, std::min(row_taps[r], max_taps)); };
  auto tiles_for = [&](uint32_t o) { return (N + (S - o) - 1) / (S - o); };
  std::vector<double> best(n_rows + 1, 1e300);
  std::vector<uint32_t> from(n_rows + 1, 0);
  best[0] = 0;
  for (uint32_t j = 1; j <= n_rows; ++j) {
    uint32_t lmax = 0;
    for (uint32_t i = j; i-- > 0;) {
      lmax = std::max(lmax, taps(row_begin + i));
      const uint32_t len = j - i;
      const uint32_t rpi = ftc_rows_per_item(len);
      const double cost = double(tiles_for(overlap_for(lmax))) * (double(len) + double((len + rpi - 1) / rpi));
      if (best[i] + cost < best[j]) {
        best[j] = best[i] + cost;
        from[j] = i;
      }
    }
  }
  std::vector<std::pair<uint32_t, uint32_t>> segs;
  for (uint32_t j = n_rows; j > 0; j = from[j]) segs.emplace_back(from[j], j);
  std::reverse(segs.begin(), segs.end());
  while (segs.size() > max_groups) {
    segs[segs.size() - 2].second = segs.back().second;
    segs.pop_back();
  }
  std::vector<FtcJob> jobs;
  for (const auto& sg : segs) {
    uint32_t lmax = 0;
    for (uint32_t r = sg.first; r < sg.second; ++r) lmax = std::max(lmax, taps(row_begin + r));
    FtcJob j{};
    j.n_bins = N;
    j.overlap = overlap_for(lmax);
    j.payload = S - j.overlap;
    j.n_tiles = tiles_for(j.overlap);
    j.row_begin = row_begin + sg.first;
    j.n_rows = sg.second - sg.first;
    j.rows_per_item = ftc_rows_per_item(j.n_rows, rows_target);
    j.n_chunks = (j.n_rows + j.rows_per_item - 1) / j.rows_per_item;
    jobs.push_back(j);
  }
  return jobs;
}

std::vector<float2> kernel_template_layout(const std::vector<std::complex<double>>& spec, uint32_t n_templates,
                                           uint32_t tile) {
  const uint32_t S = tile, NT = S / 16;
  std::vector<float2> perm(size_t(n_templates) * S);
  const double scale = 1.0 / double(S);
  for (uint32_t t = 0; t < n_templates; ++t)
    for (uint32_t slot = 0; slot < 16; ++slot)
      for (uint32_t tid = 0; tid < NT; ++tid) {
        const std::complex<double> v = spec[size_t(t) * S + tile_slot_bin(S, tid, slot)] * scale;
        perm[size_t(t) * S + size_t(slot) * NT + tid] = make_float2(float(v.real()), float(v.imag()));
      }
  return perm;
}

std::vector<std::complex<double>> taps_to_spectra(const float* taps_c64, uint32_t n_templates, uint32_t n_taps,
                                                  uint32_t max_taps, uint32_t tile, std::vector<uint32_t>& extent) {
  const uint32_t S = tile;
  std::vector<std::complex<double>> spec(size_t(n_templates) * S);
  extent.assign(n_templates, 1);
  for (uint32_t t = 0; t < n_templates; ++t) {
    std::complex<double>* row = spec.data() + size_t(t) * S;
    const uint32_t m_end = std::min(n_taps, max_taps);
    for (uint32_t m = 0; m < m_end; ++m) {
      const float* v = taps_c64 + 2 * (size_t(t) * n_taps + m);
      row[m] = std::complex<double>(v[0], v[1]);
      if (v[0] != 0.f || v[1] != 0.f) extent[t] = m + 1;
    }
    fft_inplace(row, S, -1);
  }
  return spec;
}

void make_twiddles(uint32_t tile, std::vector<float2>& tw1, std::vector<float2>& tw2) {
  const uint32_t S = tile, NT = S / 16, D = NT / 16;
  tw1.assign(16 * NT, make_float2(0.f, 0.f));
  tw2.assign(16 * D, make_float2(0.f, 0.f));
  for (uint32_t k = 0; k < 16; ++k) {
    for (uint32_t b = 0; b < NT; ++b) {
      const double a = -2.0 * kPi * double((uint64_t(b) * k) % S) / double(S);
      tw1[k * NT + b] = make_float2(float(std::cos(a)), float(std::sin(a)));
    }
    for (uint32_t d = 0; d < D; ++d) {
      const double a = -2.0 * kPi * double((uint64_t(d) * k) % NT) / double(NT);
      tw2[k * D + d] = make_float2(float(std::cos(a)), float(std::sin(a)));
    }
  }
}

void hsum_row_chunks(uint32_t n_rows, uint32_t* rows_per_block, uint32_t* n_row_chunks, uint32_t target) {
  target = std::max<uint32_t>(1, target);
  *n_row_chunks = (n_rows + target - 1) / target;
  *rows_per_block = (n_rows + *n_row_chunks - 1) / *n_row_chunks;
}

}  // namespace fdas
