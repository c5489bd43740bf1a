// cpu_common.cpp -- error reporting and input checks of libcsamd_cpu.so.
#include "cpu_common.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>

#include "cg_check.h"

namespace cs {
namespace cpu {
namespace {
thread_local char g_last_error[512] = "";
}

cs_status fail(cs_status code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof g_last_error, fmt, ap);
    va_end(ap);
    return code;
}

const char *last_error() { return g_last_error; }

cs_status check_host_spectra(const cs_spectra *s, const char *name) {
    if (!s) return fail(CS_ERR_INVALID, "%s is NULL", name);
    if (s->n_spectra < 0 || s->n_peaks < 0) return fail(CS_ERR_INVALID, "%s: negative size", name);
    if (s->max_peaks < 0 || s->max_peaks > 65535)
        return fail(CS_ERR_UNSUPPORTED, "%s: max_peaks %d outside [0, 65535]", name, s->max_peaks);
    if (s->n_spectra == 0) return CS_OK;
    if (!s->offsets || !s->norm) return fail(CS_ERR_INVALID, "%s: offsets/norm must not be NULL", name);
    if (s->n_peaks > 0 && (!s->mz || !s->int_pow))
        return fail(CS_ERR_INVALID, "%s: mz/int_pow must not be NULL", name);
    if (!(s->max_abs_mz >= 0.0f) || isinf(s->max_abs_mz))
        return fail(CS_ERR_INVALID, "%s: max_abs_mz must be finite", name);
    for (int64_t x = 0; x < s->n_spectra; ++x) {
        const uint32_t f = spectrum_faults(s->offsets, x, s->n_peaks, s->max_peaks, s->max_abs_mz,
                                           s->mz, s->mz_exact);
        if (f)
            return fail(CS_ERR_INVALID, "%s: spectrum %lld: %s%s%s%s", name, (long long)x,
                        (f & CS_FAULT_OFFSETS) ? "offsets decrease, are negative or pass n_peaks; " : "",
                        (f & CS_FAULT_LENGTH) ? "more peaks than max_peaks; " : "",
                        (f & CS_FAULT_MZ_RANGE) ? "m/z not finite or above max_abs_mz; " : "",
                        (f & CS_FAULT_MZ_ORDER) ? "m/z not ascending; " : "");
    }
    return CS_OK;
}

}  // namespace cpu
}  // namespace cs

extern "C" const char *cs_cpu_last_error(void) { return cs::cpu::last_error(); }
