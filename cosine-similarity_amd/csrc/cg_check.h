// cg_check.h -- the per-spectrum CSR checks, shared by the device validator
// (cs_validate_spectra, cosine_greedy.hip) and the host engine
// (check_host_spectra, cpu/cpu_common.cpp), so the CPU tests run the code the
// GPU runs.
#pragma once

#include <stdint.h>

#include "cg_lane.h"
#include "cosine_similarity_amd.h"

namespace cs {

// Fault bits (CS_FAULT_*) of spectrum s of a CSR set.  Peak arrays are read
// only inside [off0, off0 + n_peaks), the range the caller declares valid, and
// only after the spectrum's own offsets were found inside it.
CS_HD uint32_t spectrum_faults(const int64_t *offsets, int64_t s, int64_t n_peaks, int max_peaks,
                               float max_abs_mz, const float *mz, const double *mz_exact) {
    const int64_t off0 = offsets[0], a = offsets[s], b = offsets[s + 1];
    if (off0 < 0 || a < off0 || b < a || b - off0 > n_peaks) return CS_FAULT_OFFSETS;
    uint32_t f = 0;
    if (b - a > max_peaks) f |= CS_FAULT_LENGTH;
    float prev = -__builtin_huge_valf();
    double prevx = -__builtin_huge_val();
    for (int64_t g = a; g < b; ++g) {
        const float m = mz[g];
        // !(|m| <= max) also catches NaN
        if (!((m < 0.0f ? -m : m) <= max_abs_mz)) f |= CS_FAULT_MZ_RANGE;
        if (m < prev) f |= CS_FAULT_MZ_ORDER;
        prev = m;
        if (mz_exact) {
            const double x = mz_exact[g];
            if (!((x < 0.0 ? -x : x) <= (double)max_abs_mz)) f |= CS_FAULT_MZ_RANGE;
            if (x < prevx) f |= CS_FAULT_MZ_ORDER;
            prevx = x;
        }
    }
    return f;
}

}  // namespace cs
