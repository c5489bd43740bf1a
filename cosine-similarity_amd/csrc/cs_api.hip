// cs_api.hip -- error reporting and version of the C ABI.
#include <stdarg.h>
#include <stdio.h>

#include "cs_internal.h"

namespace {
thread_local char g_last_error[1024] = "";
}

namespace cs {
cs_status fail(cs_status code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return code;
}
}  // namespace cs

extern "C" const char *cs_last_error(void) { return g_last_error; }

extern "C" int cs_version(void) { return 10000; }
