// cpu_common.h -- error reporting and input checks of the host library
// (libcsamd_cpu.so).
#pragma once

#include "cosine_similarity_amd.h"

namespace cs {
namespace cpu {

// Record a formatted error for cs_cpu_last_error() and return `code`.
cs_status fail(cs_status code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
const char *last_error();

// Checks a spectrum set whose arrays are in host memory, all of it: sizes and
// pointers, then every spectrum with the device validator's checks
// (cg_check.h: offsets, length, m/z range and order).
cs_status check_host_spectra(const cs_spectra *s, const char *name);

}  // namespace cpu
}  // namespace cs
