// mcrt_coop.h — the persistent, workgroup-cooperative transport forms
// (FP64 and mixed precision).
#pragma once

#include <hip/hip_runtime.h>

#include <functional>
#include <string>

#include "../../include/mcrt.h"
#include "mcrt_kernels.h"

namespace mcrt {

struct CoopPlan {
  bool mixed = false;
  int blocks = 0;
  int threads = 0;
  size_t lds_bytes = 0;
  void* params_dev = nullptr;  // CoopParams<R> in device memory
  int nslab = 0;               // workgroup partial-bin slabs
  int64_t slab_stride = 0;     // doubles per slab (sum + compensation)
  double* slabs = nullptr;
  int64_t nout = 0;            // 5 + nrt + ncb
  double* dA = nullptr;
  double* dIrt = nullptr;
  double* dIcb = nullptr;
  int64_t nrt = 0, ncb = 0;
  int geo_kind = 0;
  bool mie = false;
  bool cb = false;
};

using DevAlloc = std::function<void*(size_t, hipError_t*)>;

// Validate, stage tables and parameters in device memory.  Returns an MCRT_*
// status and writes a message to *err on failure.
int coop_create(CoopPlan& cp, hipStream_t stream, int cu_count, size_t lds_max, const mcrt_config* cfg,
                const mcrt_geometry* geom, const double I0[4], uint64_t seed, uint64_t ray0,
                uint32_t instance, bool mixed,
                const DevAlloc& alloc, double* dA, double* dIrt, double* dIcb,
                unsigned long long* dcnt, std::string* err);

// Clear the workgroup slabs (before the timed transport launch).
hipError_t coop_prepare(CoopPlan& cp, hipStream_t stream);
// Enqueue the transport kernel.
hipError_t coop_launch(CoopPlan& cp, hipStream_t stream);
// Enqueue the reduction of the slabs into A, Irt, Icb.
hipError_t coop_reduce(CoopPlan& cp, hipStream_t stream);
// Enqueue the batch-means standard error of the N-normalised outputs
// (sig[nout] in the slab order: A[5], Irt, Icb).
hipError_t coop_sigma(CoopPlan& cp, hipStream_t stream, int64_t N, double* sig);

}  // namespace mcrt
