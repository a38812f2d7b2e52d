/*
 * Stand-in for the reference's fdas/src/host/FDAS.h, so that the reference's
 * own test driver (fdas/src/host/main_gtest.cpp) compiles unchanged against
 * this library: the same include guard, the same GenInfo constants (the
 * reference's gen_info.h, found beside main_gtest.cpp), OpenCL's cl_float2 /
 * cl_float / cl_uint types, and a global `class FDAS` with the reference's
 * enums, method names and bool returns (FDAS.h:36-92), forwarding to
 * include/FDAS.hpp and through it to the C ABI.
 *
 * Differences from the reference class, all at construction: the template
 * spectra are taken in the reference's 4-parallel FFT order
 * (fdas_templates_*_ft_p4.npy), and GenInfo's sizes (85 templates, 421 taps,
 * 2048-point tiles, 8 planes, 4096-entry lists) configure the library.
 */
#ifndef FDAS_FDAS_H
#define FDAS_FDAS_H

#include <iostream>

#include <CL/cl.hpp>

#include "gen_info.h"
using namespace GenInfo;

#include "FDAS.hpp"

class FDAS : public fdas_amd::FDAS {
 public:
  explicit FDAS(std::ostream& log) : fdas_amd::FDAS(log) {
    fdas_config c = config();
    c.n_templates = Input::n_templates;
    c.max_taps = Input::max_template_len;
    c.tile = FTC::tile_sz;
    c.n_harmonics = HMS::n_planes;
    c.cands_per_harmonic = Output::n_candidates / HMS::n_planes;
    c.flags |= FDAS_FLAG_REFERENCE_FFT_ORDER;
    configure(c);
  }
};

#endif  // FDAS_FDAS_H
