// phem_err.cpp -- error reporting (host only): the last error message of the
// calling thread, read back by phem_last_error().
#include <cstdarg>
#include <cstdio>

#include "phem_internal.h"

namespace phem {

static thread_local char g_err[512] = "";

int set_err(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

}  // namespace phem

extern "C" const char *phem_last_error(void) { return phem::g_err; }
