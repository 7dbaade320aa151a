// mcrt_internal.h — host-side helpers shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mcrt.h"

namespace mcrt {
struct TableCache;
}

// A device context: one HIP stream on one device, and the device tables of
// the scatterers its solves used (shared by content; handles pin them).
struct mcrt_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int cu_count = 0;
  size_t lds_per_block = 0;
  mcrt::TableCache* tables = nullptr;
  std::mutex table_mu;
  std::map<uint64_t, std::vector<const void*>> table_handles;  // handle -> pinned entries
  uint64_t next_table = 1;
};

namespace mcrt {

// Record an error message on ctx (or on the calling thread when ctx is NULL)
// and return code.  printf-style.
int fail(mcrt_ctx* ctx, int code, const char* fmt, ...) __attribute__((format(printf, 3, 4)));

// The reference's validate (plane.jl:52-62, sphere.jl:35-37) and the
// config's own consistency checks; MCRT_OK or MCRT_ERR_INVALID with a message.
int validate_geometry(mcrt_ctx* ctx, const mcrt_geometry* g);
int validate_config(mcrt_ctx* ctx, const mcrt_config* c);

}  // namespace mcrt
