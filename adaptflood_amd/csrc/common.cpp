// common.cpp -- per-thread error text and version string of the C ABI.
#include <stdarg.h>
#include <stdio.h>

#include "../../include/adaptflood.h"

namespace af {

static thread_local char g_error[1024] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof(g_error), fmt, ap);
    va_end(ap);
}

}  // namespace af

extern "C" {

const char* af_last_error(void) { return af::g_error; }

const char* af_version(void) { return "adaptflood_amd 0.1 (gfx950)"; }

}  // extern "C"
