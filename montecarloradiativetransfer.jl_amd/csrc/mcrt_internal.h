// mcrt_internal.h — host-side helpers shared by the library's translation units.
#pragma once

#include "../../include/mcrt.h"

namespace mcrt {

// Record an error message on ctx (or on the calling thread when ctx is NULL)
// and return code.  printf-style.
int fail(mcrt_ctx* ctx, int code, const char* fmt, ...) __attribute__((format(printf, 3, 4)));

}  // namespace mcrt
