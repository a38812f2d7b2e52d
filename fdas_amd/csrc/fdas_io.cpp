// SIGPROC time-series reader (ft_conv.py:39-67): the input of the front end.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "fdas_host.h"

namespace {

thread_local std::string g_io_error;

fdas_status io_fail(fdas_status st, const std::string& msg) {
  g_io_error = "[ERROR] " + msg;
  return st;
}

// Value that follows the first occurrence of `tag` in the header bytes.
template <typename T>
bool header_field(const std::vector<char>& hdr, const char* tag, T* out) {
  const size_t n = std::strlen(tag);
  for (size_t i = 0; i + n + sizeof(T) <= hdr.size(); ++i) {
    if (std::memcmp(hdr.data() + i, tag, n) == 0) {
      std::memcpy(out, hdr.data() + i + n, sizeof(T));
      return true;
    }
  }
  return false;
}

}  // namespace

extern "C" {

fdas_status fdas_read_tim(const char* path, int read_header, float* samples, uint64_t capacity,
                          uint64_t* n_samples, double* t_samp) {
  if (!path || !n_samples) return io_fail(FDAS_EINVAL, "null argument");
  *n_samples = 0;
  if (t_samp) *t_samp = 0.0;
  FILE* f = std::fopen(path, "rb");
  if (!f) return io_fail(FDAS_EINVAL, std::string("cannot open '") + path + "'");
  std::fseek(f, 0, SEEK_END);
  const long size = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  uint64_t offset = 0;
  if (read_header) {
    // The header lies within the first 512 bytes; it starts with a 4-byte
    // length followed by "HEADER_START" and ends with "HEADER_END".
    std::vector<char> hdr(size_t(std::min<long>(size, 512)));
    if (!hdr.empty() && std::fread(hdr.data(), 1, hdr.size(), f) != hdr.size()) {
      std::fclose(f);
      return io_fail(FDAS_EINVAL, "short read of the header");
    }
    static const char kStart[] = "HEADER_START", kEnd[] = "HEADER_END";
    const bool has_start = hdr.size() >= 4 + std::strlen(kStart) && std::memcmp(hdr.data() + 4, kStart, 12) == 0;
    if (has_start) {
      size_t end = 0;
      bool found = false;
      for (size_t i = 0; i + std::strlen(kEnd) <= hdr.size(); ++i)
        if (std::memcmp(hdr.data() + i, kEnd, std::strlen(kEnd)) == 0) {
          end = i + std::strlen(kEnd);
          found = true;
          break;
        }
      if (!found) {
        std::fclose(f);
        return io_fail(FDAS_EINVAL, "header end tag (HEADER_END) not found");
      }
      offset = end;
      int32_t data_type = 0, nbits = 0;
      double tsamp = 0.0;
      if (!header_field(hdr, "data_type", &data_type) || !header_field(hdr, "nbits", &nbits) ||
          !header_field(hdr, "tsamp", &tsamp)) {
        std::fclose(f);
        return io_fail(FDAS_EINVAL, "header field (data_type, nbits or tsamp) not found");
      }
      if (data_type != 2 || nbits != 32) {
        std::fclose(f);
        return io_fail(FDAS_EINVAL, "format mismatch: expecting 32-bit FP time series data");
      }
      if (t_samp) *t_samp = tsamp;
    }
  }
  const uint64_t n = size > long(offset) ? (uint64_t(size) - offset) / sizeof(float) : 0;
  *n_samples = n;
  if (samples) {
    if (capacity < n) {
      std::fclose(f);
      return io_fail(FDAS_EINVAL, "sample buffer smaller than the file's " + std::to_string(n) + " samples");
    }
    std::fseek(f, long(offset), SEEK_SET);
    if (std::fread(samples, sizeof(float), n, f) != n) {
      std::fclose(f);
      return io_fail(FDAS_EINVAL, "short read of the samples");
    }
  }
  std::fclose(f);
  return FDAS_OK;
}

const char* fdas_io_last_error(void) { return g_io_error.c_str(); }

}  // extern "C"
