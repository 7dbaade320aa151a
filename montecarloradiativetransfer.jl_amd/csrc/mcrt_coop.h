// mcrt_coop.h — the persistent, workgroup-cooperative transport forms
// (FP64 and mixed precision).
#pragma once

#include <hip/hip_runtime.h>

#include <functional>
#include <string>
#include <vector>

#include "../../include/mcrt.h"
#include "mcrt_kernels.h"

namespace mcrt {

struct CoopPlan {
  bool mixed = false;
  int blocks = 0;
  int slots = 0;      // active ray slots per workgroup
  int wg_per_cu = 0;  // workgroups launched per CU
  int threads = 0;
  size_t lds_bytes = 0;
  size_t lds_max = 0;          // LDS per workgroup of the device
  void* params_dev = nullptr;  // CoopParams<R> in device memory
  int nslab = 0;               // workgroup partial-bin slabs
  int64_t slab_stride = 0;     // doubles per slab (sum + compensation)
  double* slabs = nullptr;
  int64_t nout = 0;            // outputs: 5 + nrt + ncb
  int64_t slab_words = 0;      // words of one slab half: 5 + nrt + cb_groups·ncb
  int cb_groups = 1;
  double* dA = nullptr;
  double* dIrt = nullptr;
  double* dIcb = nullptr;
  int64_t nrt = 0, ncb = 0;
  int geo_kind = 0;
  int sk = 0;       // kernel kind: 0 Rayleigh, 1 Mie, 3 any (per-layer tables, Spline)
  int sk0 = 0;      // kind of layer 1's scatterer (0 Rayleigh, 1 Mie, 2 Spline)
  int stage_s = 0;  // S table staged in LDS
  bool cb = false;
  std::vector<const void*> table_refs;  // cached tables in use (TableCache)
};

using DevAlloc = std::function<void*(size_t, hipError_t*)>;

// A ray range [rb, re) of one solve of a batched launch; the sums go to slab
// (one of the solve's batch slabs) with FP64 atomics.  cont: the chunk's own
// ray-slot block (NULL: the workgroup's block in the launch's slot memory);
// resume: cont holds rays in flight from a launch that stopped the chunk, and
// rb is the first ray not yet started.
struct BatchChunk {
  const void* params;  // the solve's parameter block (device)
  int64_t rb, re;
  double* slab;
  void* cont = nullptr;
  int32_t resume = 0;
  int32_t pad = 0;
};

// Device copies of scatterer tables keyed by their content, shared by the
// solves that use them: reference counted, and unreferenced tables are freed,
// least recently used first, when the cache holds more than its budget.
struct TableCache;
TableCache* table_cache_create(size_t budget_bytes);
void table_cache_destroy(TableCache* c);
void table_cache_release(TableCache* c, const std::vector<const void*>& refs);
void table_cache_stats(TableCache* c, size_t* bytes, uint64_t* hits, uint64_t* misses);

// One solve of a batched launch: parameters, tables and bins in device memory,
// its N rays split into chunks of chunk_rays, chunk k summed into batch slab
// k mod nbatch.
struct SolveDev {
  bool mixed = false;
  int geo_kind = 0;
  bool cb = false;
  int64_t N = 0, nout = 0, nrt = 0, ncb = 0;
  size_t lds_bytes = 0;
  void* params_dev = nullptr;
  int64_t chunk_rays = 0, nchunks = 0;
  int nbatch = 0;
  double* slabs = nullptr;                 // nbatch × nout raw sums
  int64_t* batch_rays = nullptr;           // nbatch (device)
  unsigned long long* counters = nullptr;  // kCntN
  double* out = nullptr;                   // nout raw sums: A[5], Irt, Icb
  double* sig = nullptr;                   // nout standard errors of out/N
  double cost = 0.0;                       // relative work of one chunk
  std::vector<const void*> table_refs;     // cached tables in use (TableCache)
};

// Rays per chunk when the caller does not choose: Throughput (batches) cuts
// by ray count, ⌈N/64⌉ within 512..8192; Latency (the server) keeps a chunk
// near a fixed number of bin evaluations from solve_work_per_ray, 64..8192
// rays.  At least two chunks when N ≥ 2.  A function of the solve alone, so
// that a solve is cut the same way alone and with any other solves.
enum class ChunkRule { Throughput, Latency };
double solve_work_per_ray(const mcrt_config* cfg, const mcrt_geometry* g);
int64_t default_chunk_rays(int64_t N, double work_per_ray, ChunkRule rule);
// Tables go to memory from alloc, or are shared through cache when given.
// slots, stage_s: the launch shape of the solve's chunks (mcrt_launch_shape;
// 0: the defaults).
int solve_create(SolveDev& sd, size_t lds_max, const mcrt_config* cfg, const mcrt_geometry* geom,
                 const double I0[4], uint64_t seed, uint64_t ray0, uint32_t instance, bool mixed,
                 int64_t chunk_rays, ChunkRule rule, const DevAlloc& alloc, std::string* err,
                 TableCache* cache = nullptr, int slots = 0, int stage_s = 0);
// Device bytes of a solve apart from cached tables (for admission).
size_t solve_bytes_estimate(const mcrt_config* cfg, const mcrt_geometry* g, bool mixed,
                            int64_t chunk_rays, ChunkRule rule);
void solve_chunk(const SolveDev& sd, int64_t k, BatchChunk* out);
// Zero the batch slabs and counters (kernels, so the sequence can be captured).
hipError_t solve_prepare(const SolveDev& sd, hipStream_t stream);
// Reduce the batch slabs into out and sig.
hipError_t solve_finish(const SolveDev& sd, hipStream_t stream);
size_t batch_slot_bytes(bool mixed, int blocks);
hipError_t batch_set_lds(bool mixed, size_t lds_bytes);
// One persistent launch over nchunks chunks (dnext: a device counter it zeroes
// first, unless zero_next is false and the caller has cleared it); blocks
// workgroups, each with its slot state in slot_mem (or the chunk's cont).
// With preempt (host memory the device reads), a workgroup that sees a
// non-zero flag stops its chunk at the next iteration boundary, after at
// least one iteration, and claims no further chunk; states[c] then receives
// chunk c's outcome: -2 finished, or the first ray not yet started (the
// rays in flight are kept in the chunk's cont), while unclaimed chunks keep
// what the caller stored there.
hipError_t batch_launch(bool mixed, const BatchChunk* dchunks, int nchunks, unsigned int* dnext,
                        void* slot_mem, int blocks, size_t lds_bytes, hipStream_t stream,
                        int64_t* states = nullptr, const unsigned int* preempt = nullptr,
                        bool zero_next = true);

// Find the address of the no-allocation solve's scratch array on `device`
// (current); mcrt_context_create calls it once per context.
hipError_t static_scratch_init(int device);

// Validate, stage tables and parameters in device memory.  Returns an MCRT_*
// status and writes a message to *err on failure.
int coop_create(CoopPlan& cp, hipStream_t stream, int cu_count, size_t lds_max, const mcrt_config* cfg,
                const mcrt_geometry* geom, const double I0[4], uint64_t seed, uint64_t ray0,
                uint32_t instance, bool mixed, const mcrt_launch_shape& shape,
                const DevAlloc& alloc, double* dA, double* dIrt, double* dIcb,
                unsigned long long* dcnt, std::string* err, TableCache* cache = nullptr);
// Pin a scatterer's FP32 and FP64 device tables in cache (references appended
// to refs; release them with table_cache_release).  Rayleigh has no tables.
int table_pin(TableCache* cache, const mcrt_scatterer* s, std::vector<const void*>* refs,
              std::string* err);

// Clear the workgroup slabs (before the timed transport launch).
hipError_t coop_prepare(CoopPlan& cp, hipStream_t stream);
// Enqueue the transport kernel.
hipError_t coop_launch(CoopPlan& cp, hipStream_t stream);
// Enqueue the reduction of the slabs into A, Irt, Icb.
hipError_t coop_reduce(CoopPlan& cp, hipStream_t stream);
// mcrt_plan_eval for a cooperative plan: the building blocks of its kernels
// on n host records (16 doubles in, 4 out); blocking on stream.
int coop_eval(const CoopPlan& cp, hipStream_t stream, int what, int64_t n, const double* in,
              double* out, std::string* err);
// Enqueue the batch-means standard error of the N-normalised outputs
// (sig[nout] in the slab order: A[5], Irt, Icb).
hipError_t coop_sigma(CoopPlan& cp, hipStream_t stream, int64_t N, double* sig);

}  // namespace mcrt
