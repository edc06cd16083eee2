// nb_error.cpp -- the thread-local message of the last failing C-ABI call.
#include <cstdarg>
#include <cstdio>

#include "nb_internal.hpp"

namespace nb {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

void clear_error() { g_err[0] = 0; }

}  // namespace nb

extern "C" const char* nb_last_error(void) { return nb::g_err; }
