// cpu_common.cpp -- error reporting and input checks of libcsamd_cpu.so.
#include "cpu_common.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>

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
    const int64_t *off = s->offsets;
    if (off[0] < 0) return fail(CS_ERR_INVALID, "%s: offsets[0] < 0", name);
    for (int64_t x = 0; x < s->n_spectra; ++x) {
        const int64_t a = off[x], b = off[x + 1];
        if (b < a) return fail(CS_ERR_INVALID, "%s: offsets decrease at spectrum %lld", name, (long long)x);
        if (b - off[0] > s->n_peaks)
            return fail(CS_ERR_INVALID, "%s: offsets[%lld] - offsets[0] exceeds n_peaks %lld", name,
                        (long long)(x + 1), (long long)s->n_peaks);
        if (b - a > s->max_peaks)
            return fail(CS_ERR_INVALID, "%s: spectrum %lld has %lld peaks > max_peaks %d", name,
                        (long long)x, (long long)(b - a), s->max_peaks);
        for (int64_t g = a; g < b; ++g) {
            const float m = s->mz[g];
            if (!isfinite(m) || fabsf(m) > s->max_abs_mz)
                return fail(CS_ERR_INVALID, "%s: m/z of peak %lld is not finite or exceeds max_abs_mz",
                            name, (long long)g);
            if (s->mz_exact && (!isfinite(s->mz_exact[g]) || fabs(s->mz_exact[g]) > s->max_abs_mz))
                return fail(CS_ERR_INVALID, "%s: m/z of peak %lld is not finite or exceeds max_abs_mz",
                            name, (long long)g);
            if (g > a && (m < s->mz[g - 1] || (s->mz_exact && s->mz_exact[g] < s->mz_exact[g - 1])))
                return fail(CS_ERR_INVALID, "%s: m/z not ascending in spectrum %lld", name, (long long)x);
        }
    }
    return CS_OK;
}

}  // namespace cpu
}  // namespace cs

extern "C" const char *cs_cpu_last_error(void) { return cs::cpu::g_last_error; }
