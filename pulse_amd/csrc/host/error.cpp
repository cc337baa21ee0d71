#include "common.h"

namespace pulse {

static thread_local std::string g_last_error;

void set_error(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

const char *last_error() { return g_last_error.c_str(); }

}  // namespace pulse

extern "C" const char *pulse_last_error(void) { return pulse::last_error(); }
