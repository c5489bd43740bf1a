// cpu_common.h -- error reporting and input checks of the host library
// (libcsamd_cpu.so).
#pragma once

#include "cosine_similarity_amd.h"

namespace cs {
namespace cpu {

// Record a formatted error for cs_cpu_last_error() and return `code`.
cs_status fail(cs_status code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));

// Checks a spectrum set whose arrays are in host memory, all of it: sizes and
// pointers, offsets non-decreasing from offsets[0] >= 0 and spanning at most
// n_peaks, every spectrum at most max_peaks (<= 65535) long, m/z finite,
// ascending within each spectrum and within max_abs_mz.
cs_status check_host_spectra(const cs_spectra *s, const char *name);

}  // namespace cpu
}  // namespace cs
