// Test harness (CPU only): the host block cache of igc_host.cpp compiled on its
// own, with the error recorder igc_internal.hpp declares.
#include <cstdarg>
#include <cstdint>
#include <cstdio>

#include "../indexgen_amd/csrc/igc_internal.hpp"

namespace igc {
static char g_msg[256];
int set_error(int code, const char *fmt, ...)
{
    va_list a;
    va_start(a, fmt);
    vsnprintf(g_msg, sizeof g_msg, fmt, a);
    va_end(a);
    return code;
}
const char *last_error() { return g_msg; }
} // namespace igc

extern "C" {
const char *igc_last_error(void) { return igc::last_error(); }
void *hc_alloc(size_t bytes) { return igc::host_alloc(bytes); }
void hc_free(void *p) { igc::host_free(p); }
void hc_trim(void) { igc::host_cache_trim(); }
}
