/*
 * FDAS.hpp -- drop-in C++ host class over the C ABI in fdas.h.
 *
 * Same class name, enums, method names, argument meaning and bool returns as
 * the reference host class (fdas/src/host/FDAS.h:36-92), so the reference's
 * driver flow (fdas/src/host/main_gtest.cpp) compiles against it after
 * swapping the include.  Differences, all at initialise_accelerator:
 *   - the bitstream file name is accepted and ignored (kernels are built in);
 *   - the platform selector sees ("AMD HIP", "HIP x.y.z") and the device
 *     selector (ordinal, device_type_gpu, device name); the default
 *     choose_accelerator_devices accepts GPUs;
 *   - BufferMode (DDR bank placement, FDAS.cpp:188-219) has no GPU meaning and
 *     is accepted and ignored;
 *   - the compile-time GenInfo constants become an optional fdas_config.
 * Complex arrays may be of any 8-byte trivially copyable float-pair type:
 * cl_float2 from <CL/cl.hpp>, std::complex<float>, or float2_t below; so the
 * reference's callers keep their OpenCL types.
 * Header-only; link with libfdas_amd.so.
 */
#ifndef FDAS_AMD_FDAS_HPP
#define FDAS_AMD_FDAS_HPP

#include <cstdint>
#include <functional>
#include <ostream>
#include <string>
#include <type_traits>

#include "fdas.h"

namespace fdas_amd {

struct float2_t {
  float s[2];
};
static_assert(sizeof(float2_t) == 8, "float2_t must have the layout of cl_float2");

// Interleaved (re, im) floats of a complex array of any 8-byte float-pair type.
template <typename C2>
inline const float* complex_floats(const C2* p) {
  static_assert(sizeof(C2) == 2 * sizeof(float) && std::is_trivially_copyable<C2>::value,
                "complex elements must be two floats (re, im), e.g. cl_float2 or std::complex<float>");
  return reinterpret_cast<const float*>(p);
}
template <typename C2>
inline float* complex_floats(C2* p) {
  static_assert(sizeof(C2) == 2 * sizeof(float) && std::is_trivially_copyable<C2>::value,
                "complex elements must be two floats (re, im), e.g. cl_float2 or std::complex<float>");
  return reinterpret_cast<float*>(p);
}

class FDAS {
 public:
  enum FOPPart { NegativeAccelerations = FDAS_NEGATIVE, PositiveAccelerations = FDAS_POSITIVE,
                 AllAccelerations = FDAS_ALL };
  enum BufferSet { A = FDAS_SET_A, B = FDAS_SET_B };
  enum BufferMode { PerSet, PerStage, Crossed };

  static constexpr uint32_t device_type_gpu = 4;  // the value of CL_DEVICE_TYPE_GPU

  explicit FDAS(std::ostream& log) : log(log) { fdas_default_config(&cfg); }
  ~FDAS() { fdas_destroy(ctx); }
  FDAS(const FDAS&) = delete;
  FDAS& operator=(const FDAS&) = delete;

  // GenInfo replacement: call before initialise_accelerator to change T, taps, tile, H, cap.
  void configure(const fdas_config& c) { cfg = c; }
  const fdas_config& config() const { return cfg; }

  bool initialise_accelerator(std::string bitstream_file_name,
                              const std::function<bool(const std::string&, const std::string&)>& platform_selector,
                              const std::function<bool(uint32_t, uint32_t, const std::string&)>& device_selector,
                              uint32_t input_sz, BufferMode mode = PerSet, bool sync_pipeline = false) {
    (void)bitstream_file_name;
    (void)mode;
    if (!platform_selector("AMD HIP", fdas_platform_version())) {
      log << "[ERROR] No suitable platform found" << std::endl;
      return false;
    }
    int chosen = -1;
    const int n = fdas_device_count();
    for (int i = 0; i < n && chosen < 0; ++i) {
      char name[256] = {0};
      if (fdas_device_info(i, name, sizeof(name), nullptr, 0) != FDAS_OK) continue;
      if (device_selector(uint32_t(i), device_type_gpu, name)) {
        chosen = i;
        log << "[INFO] Device[" << i << "]: " << name << " (default)" << std::endl;
      }
    }
    if (chosen < 0) {
      log << "[ERROR] No suitable device found" << std::endl;
      return false;
    }
    fdas_destroy(ctx);
    ctx = nullptr;
    cfg.n_bins = input_sz;
    cfg.device = chosen;
    cfg.flags = (cfg.flags & ~FDAS_FLAG_SYNC_PIPELINE) | (sync_pipeline ? FDAS_FLAG_SYNC_PIPELINE : 0u);
    return ok(fdas_create(&cfg, &FDAS::forward_log, this, &ctx));
  }

  template <typename C2>
  bool upload_templates(const C2* templates, BufferSet ab) {
    return ok(fdas_upload_templates(ctx, complex_floats(templates), ab));
  }
  bool upload_thresholds(const float* thresholds, BufferSet ab) {
    return ok(fdas_upload_thresholds(ctx, thresholds, ab));
  }
  template <typename C2>
  bool perform_input_tiling(const C2* input, BufferSet ab) {
    return ok(fdas_perform_input_tiling(ctx, complex_floats(input), ab));
  }
  bool perform_ft_convolution(FOPPart which, BufferSet ab) {
    return ok(fdas_perform_ft_convolution(ctx, fdas_part(which), ab));
  }
  bool perform_harmonic_summing(const float* thresholds, FOPPart which, BufferSet ab) {
    return ok(fdas_perform_harmonic_summing(ctx, thresholds, fdas_part(which), ab));
  }
  template <typename C2>
  bool launch(const C2* input, const float* thresholds, uint32_t* detection_location, float* detection_power,
              FOPPart which, BufferSet ab) {
    return ok(fdas_launch(ctx, complex_floats(input), thresholds, detection_location, detection_power,
                          fdas_part(which), ab));
  }
  bool wait(BufferSet ab) { return ok(fdas_wait(ctx, ab)); }
  bool end_pipeline(BufferSet ab) { return ok(fdas_end_pipeline(ctx, ab)); }
  template <typename C2>
  bool retrieve_tiles(C2* tiles, BufferSet ab) {
    return ok(fdas_retrieve_tiles(ctx, complex_floats(tiles), ab));
  }
  bool retrieve_FOP(float* fop, BufferSet ab) { return ok(fdas_retrieve_fop(ctx, fop, ab)); }
  bool inject_FOP(const float* fop, BufferSet ab) { return ok(fdas_inject_fop(ctx, fop, ab)); }
  bool retrieve_candidates(uint32_t* detection_location, float* detection_power, BufferSet ab) {
    return ok(fdas_retrieve_candidates(ctx, detection_location, detection_power, ab));
  }
  void print_stats(BufferSet ab, bool reset = false) { fdas_print_stats(ctx, ab, reset ? 1 : 0); }
  void print_events(BufferSet ab) { fdas_print_events(ctx, ab); }

  uint32_t get_input_sz() const { return uint32_t(fdas_get_size(ctx, FDAS_SIZE_INPUT)); }
  uint32_t get_tiles_sz() const { return uint32_t(fdas_get_size(ctx, FDAS_SIZE_TILES)); }
  uint32_t get_templates_sz() const { return uint32_t(fdas_get_size(ctx, FDAS_SIZE_TEMPLATES)); }
  uint32_t get_thresholds_sz() const { return uint32_t(fdas_get_size(ctx, FDAS_SIZE_THRESHOLDS)); }
  uint32_t get_fop_sz() const { return uint32_t(fdas_get_size(ctx, FDAS_SIZE_FOP)); }
  uint32_t get_candidate_list_sz() const { return uint32_t(fdas_get_size(ctx, FDAS_SIZE_CANDIDATES)); }

  static bool choose_first_platform(const std::string&, const std::string&) { return true; }
  static bool choose_accelerator_devices(uint32_t, uint32_t device_type, const std::string&) {
    return device_type == device_type_gpu;
  }

  fdas_ctx* handle() const { return ctx; }

 private:
  static void forward_log(int, const char* msg, void* user) {
    static_cast<FDAS*>(user)->log << msg << std::endl;
  }
  bool ok(fdas_status st) { return st == FDAS_OK; }

  fdas_config cfg{};
  fdas_ctx* ctx = nullptr;
  std::ostream& log;
};

}  // namespace fdas_amd

#endif  // FDAS_AMD_FDAS_HPP
