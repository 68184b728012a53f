// Library-wide C-ABI entry points: version and per-thread error text.
#include <stdio.h>

#include "gsr_internal.h"

static thread_local char g_err[512] = "";

void gsr_set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

extern "C" {
int gsr_version(void) { return 1; }
const char *gsr_last_error(void) { return g_err; }
}
