// mcrt_serve.cpp — the request loop (C ABI mcrt_server_*, mcrt_submit, …).
//
// Continuous batching on one device: a host thread runs the scheduler of
// mcrt_sched.h (admission by device memory, launch windows of chunks taken
// round-robin from every admitted solve of one precision, preemption of the
// running window while work waits, resumption of the chunks it stopped) over
// the HIP backend below: one compute stream for the windows and the
// reductions that follow them, one set-up stream for clearing the slabs of
// newly admitted solves.  A solve is cut into chunks by a rule that depends on
// the solve alone, a chunk never shares an iteration with another, and a
// stopped chunk resumes exactly where it stopped, so every result is the same
// alone, in the mix and wherever the windows were cut, up to the order of FP64
// additions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/mcrt.h"
#include "mcrt_coop.h"
#include "mcrt_internal.h"
#include "mcrt_kernels.h"
#include "mcrt_sched.h"

using mcrt::fail;
using Clock = std::chrono::steady_clock;
namespace sched = mcrt::sched;

namespace {

// ---------------------------------------------------------- owned requests
struct OwnedScatterer {
  mcrt_scatterer s{};
  std::vector<double> csrn, F11, F12, F33, F34, S11, S22, F22, F44;
};

struct OwnedRequest {
  mcrt_config cfg{};
  std::vector<double> arr[12];
  mcrt_geometry geom{};
  std::vector<mcrt_layer> layers;
  std::vector<std::unique_ptr<OwnedScatterer>> sca;
  mcrt_request req{};
};

const double* keep(std::vector<double>& v, const double* p, int64_t n) {
  if (!p || n <= 0) return p;
  v.assign(p, p + n);
  return v.data();
}

const mcrt_scatterer* own_scatterer(OwnedRequest& o, const mcrt_scatterer* s,
                                    std::map<const mcrt_scatterer*, const mcrt_scatterer*>& seen) {
  if (!s) return nullptr;
  auto it = seen.find(s);
  if (it != seen.end()) return it->second;
  auto c = std::make_unique<OwnedScatterer>();
  c->s = *s;
  if (s->kind == MCRT_SCATTERER_MIE || s->kind == MCRT_SCATTERER_SPLINE) {
    const int64_t N = s->n_theta;
    c->s.csrn = keep(c->csrn, s->csrn, (int64_t)s->random_points + 1);
    c->s.F11 = keep(c->F11, s->F11, N);
    c->s.F12 = keep(c->F12, s->F12, N);
    c->s.F33 = keep(c->F33, s->F33, N);
    c->s.F34 = keep(c->F34, s->F34, N);
    c->s.S11 = keep(c->S11, s->S11, 2 * N);
    c->s.S22 = keep(c->S22, s->S22, 2 * N);
    if (s->kind == MCRT_SCATTERER_SPLINE) {
      c->s.F22 = keep(c->F22, s->F22, N);
      c->s.F44 = keep(c->F44, s->F44, N);
    }
  }
  const mcrt_scatterer* r = &c->s;
  o.sca.push_back(std::move(c));
  seen[s] = r;
  return r;
}

// Deep copy of a (validated) request, so that the caller's memory is free on return.
void own(const mcrt_request& q, OwnedRequest& o) {
  o.req = q;
  o.cfg = *q.cfg;
  mcrt_config& c = o.cfg;
  c.phi = keep(o.arr[0], c.phi, c.n_phi);
  c.cos_phi = keep(o.arr[1], c.cos_phi, c.n_phi);
  c.sin_phi = keep(o.arr[2], c.sin_phi, c.n_phi);
  c.cos_theta = keep(o.arr[3], c.cos_theta, c.n_theta);
  c.sin_theta = keep(o.arr[4], c.sin_theta, c.n_theta);
  c.inorm = keep(o.arr[5], c.inorm, c.n_theta);
  c.theta_b = keep(o.arr[6], c.theta_b, c.n_theta_b);
  c.cos_theta_b = keep(o.arr[7], c.cos_theta_b, c.n_theta_b);
  c.sin_theta_b = keep(o.arr[8], c.sin_theta_b, c.n_theta_b);
  c.phi_b = keep(o.arr[9], c.phi_b, c.n_phi_b);
  c.cos_phi_b = keep(o.arr[10], c.cos_phi_b, c.n_phi_b);
  c.sin_phi_b = keep(o.arr[11], c.sin_phi_b, c.n_phi_b);
  o.geom = *q.geom;
  std::map<const mcrt_scatterer*, const mcrt_scatterer*> seen;
  if (o.geom.kind == MCRT_GEOMETRY_PLANE) {
    o.layers.assign(q.geom->layers, q.geom->layers + q.geom->n_layers);
    for (auto& L : o.layers) L.scatterer = own_scatterer(o, L.scatterer, seen);
    o.geom.layers = o.layers.data();
    o.geom.scatterer = nullptr;
  } else {
    o.geom.scatterer = own_scatterer(o, q.geom->scatterer, seen);
    o.geom.layers = nullptr;
  }
  o.req.cfg = &o.cfg;
  o.req.geom = &o.geom;
}

// Memory kept for reuse in power-of-two size classes (device, or pinned host),
// so that admitting and retiring a solve calls no allocator in steady state —
// and no hipFree, which waits for the device.
class Pool {
 public:
  Pool(bool host, size_t keep) : host_(host), keep_(keep) {}
  ~Pool() {
    for (auto& kv : free_)
      for (void* p : kv.second) release_raw(p);
  }
  static size_t size_class(size_t n) {
    size_t c = 256;
    while (c < n) c <<= 1;
    return c;
  }
  void* get(size_t n, hipError_t* e) {
    const size_t c = size_class(n);
    {
      std::lock_guard<std::mutex> g(mu_);
      auto it = free_.find(c);
      if (it != free_.end() && !it->second.empty()) {
        void* p = it->second.back();
        it->second.pop_back();
        cached_ -= c;
        *e = hipSuccess;
        return p;
      }
    }
    void* p = nullptr;
    *e = host_ ? hipHostMalloc(&p, c, hipHostMallocDefault) : hipMalloc(&p, c);
    return *e == hipSuccess ? p : nullptr;
  }
  void put(void* p, size_t n) {
    if (!p) return;
    const size_t c = size_class(n);
    std::lock_guard<std::mutex> g(mu_);
    free_[c].push_back(p);
    cached_ += c;
    // beyond the limit, free the largest classes first
    for (auto it = free_.rbegin(); cached_ > keep_ && it != free_.rend(); ++it)
      while (cached_ > keep_ && !it->second.empty()) {
        release_raw(it->second.back());
        it->second.pop_back();
        cached_ -= it->first;
      }
  }

 private:
  void release_raw(void* p) { (void)(host_ ? hipHostFree(p) : hipFree(p)); }
  bool host_;
  size_t keep_;
  size_t cached_ = 0;
  std::mutex mu_;
  std::map<size_t, std::vector<void*>> free_;
};


struct Ticket {
  uint64_t id = 0;
  OwnedRequest r;
  int state = MCRT_TICKET_QUEUED;
  int err_code = 0;
  std::string err;
  bool mixed = false;
  size_t need = 0;
  mcrt::SolveDev sd;
  std::vector<std::pair<void*, size_t>> allocs;  // from the server's device pool
  bool released = true;
  bool admitted = false;  // set by the worker under the server lock
  // read-back of the results: one pinned block (host_pool) holding the sums,
  // the standard errors and the counters
  void* h_block = nullptr;
  size_t h_bytes = 0;
  double* h_out = nullptr;
  double* h_sig = nullptr;
  unsigned long long* h_cnt = nullptr;
  Clock::time_point t_submit, t_admit, t_start, t_done;
  bool started = false;
};

constexpr int kConcurrentFinishes = 64;

}  // namespace

struct mcrt_server;

namespace {

// The device side of the scheduler: windows on one compute stream.
class HipBackend final : public sched::Backend {
 public:
  explicit HipBackend(mcrt_server* S) : S_(S) {}
  hipError_t init();
  void shutdown();
  int create(uint64_t job, int64_t* nchunks, double* cost, std::string* err) override;
  void destroy(uint64_t job) override;
  int launch(bool mixed, const std::vector<sched::Task>& tasks, std::string* err) override;
  int poll(std::vector<int64_t>* states, sched::Time* start, std::string* err) override;
  void preempt() override { __atomic_store_n(flag_, 1u, __ATOMIC_RELEASE); }
  int64_t finish(const std::vector<uint64_t>& jobs, std::string* err) override;
  int finished(int64_t handle, std::string* err) override;

 private:
  Ticket* ticket(uint64_t id);
  mcrt_server* S_;
  hipStream_t lane_ = nullptr, setup_ = nullptr;
  hipEvent_t setup_ev_ = nullptr;   // after the last admitted solve's clear
  hipEvent_t ref_ev_ = nullptr;     // device time zero, taken at ref_t_
  Clock::time_point ref_t_;
  hipEvent_t start_ev_ = nullptr, done_ev_ = nullptr;  // of the window in flight
  // window staging: [next counter (64 B)] [chunks] [states], pinned and device
  unsigned char* h_win_ = nullptr;
  unsigned char* d_win_ = nullptr;
  size_t win_bytes_ = 0;
  int win_n_ = 0;
  unsigned int* flag_ = nullptr;   // preempt flag (coherent pinned host memory)
  unsigned int* d_flag_ = nullptr;
  unsigned char* conts_ = nullptr;  // continuation blocks
  size_t cont_bytes_ = 0;
  struct FinSlot {
    hipEvent_t ev = nullptr;
    bool used = false;
  } fins_[kConcurrentFinishes];
};

}  // namespace

struct mcrt_server {
  mcrt_ctx* ctx = nullptr;
  std::thread worker;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  bool stop = false;
  std::vector<Ticket*> incoming;
  std::vector<uint64_t> cancels;
  std::map<uint64_t, Ticket*> tickets;
  uint64_t next_id = 1;
  std::atomic<double> slice_ms{8.0};
  size_t budget = 0;
  int blocks = 0;
  mcrt::TableCache* cache = nullptr;
  Pool* dev_pool = nullptr;
  Pool* host_pool = nullptr;
  HipBackend* backend = nullptr;
  sched::Options opt;
  // statistics (the scheduler's, copied after every step)
  mcrt_server_stats st{};

  void run();
  void release(Ticket* t);
};

namespace {

Ticket* HipBackend::ticket(uint64_t id) {
  std::lock_guard<std::mutex> g(S_->mu);
  auto it = S_->tickets.find(id);
  return it == S_->tickets.end() ? nullptr : it->second;
}

hipError_t HipBackend::init() {
  const int cap = S_->opt.window_tasks;
  win_bytes_ = 64 + (sizeof(mcrt::BatchChunk) + sizeof(int64_t)) * (size_t)cap;
  cont_bytes_ = std::max(mcrt::batch_slot_bytes(true, 1), mcrt::batch_slot_bytes(false, 1));
  cont_bytes_ = (cont_bytes_ + 255) & ~size_t(255);
  hipError_t e = hipStreamCreateWithFlags(&lane_, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&setup_, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&setup_ev_, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreate(&ref_ev_);
  if (e == hipSuccess) e = hipEventCreate(&start_ev_);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&done_ev_, hipEventDisableTiming);
  for (auto& f : fins_)
    if (e == hipSuccess) e = hipEventCreateWithFlags(&f.ev, hipEventDisableTiming);
  if (e == hipSuccess) e = hipHostMalloc((void**)&h_win_, win_bytes_, hipHostMallocDefault);
  if (e == hipSuccess) e = hipMalloc((void**)&d_win_, win_bytes_);
  if (e == hipSuccess) e = hipHostMalloc((void**)&flag_, 256, hipHostMallocCoherent);
  if (e == hipSuccess) e = hipHostGetDevicePointer((void**)&d_flag_, flag_, 0);
  if (e == hipSuccess) e = hipMalloc((void**)&conts_, cont_bytes_ * (size_t)S_->opt.conts);
  if (e == hipSuccess) {
    std::memset(flag_, 0, 256);
    e = hipEventRecord(ref_ev_, lane_);
  }
  if (e == hipSuccess) e = hipEventSynchronize(ref_ev_);
  ref_t_ = Clock::now();
  return e;
}

void HipBackend::shutdown() {
  if (lane_) (void)hipStreamSynchronize(lane_);
  if (setup_) (void)hipStreamSynchronize(setup_);
  for (auto& f : fins_)
    if (f.ev) (void)hipEventDestroy(f.ev);
  if (setup_ev_) (void)hipEventDestroy(setup_ev_);
  if (ref_ev_) (void)hipEventDestroy(ref_ev_);
  if (start_ev_) (void)hipEventDestroy(start_ev_);
  if (done_ev_) (void)hipEventDestroy(done_ev_);
  if (h_win_) (void)hipHostFree(h_win_);
  if (d_win_) (void)hipFree(d_win_);
  if (flag_) (void)hipHostFree(flag_);
  if (conts_) (void)hipFree(conts_);
  if (lane_) (void)hipStreamDestroy(lane_);
  if (setup_) (void)hipStreamDestroy(setup_);
}

// Device state of a newly admitted solve (the clear runs on the set-up
// stream; windows wait for setup_ev_).
int HipBackend::create(uint64_t job, int64_t* nchunks, double* cost, std::string* err) {
  Ticket* t = ticket(job);
  if (!t) {
    *err = "unknown ticket";
    return MCRT_ERR_INVALID;
  }
  {
    std::lock_guard<std::mutex> g(S_->mu);
    t->admitted = true;
    t->state = MCRT_TICKET_RUNNING;
  }
  t->released = false;
  Pool* pool = S_->dev_pool;
  auto alloc = [t, pool](size_t n, hipError_t* e) -> void* {
    void* p = pool->get(n ? n : 8, e);
    if (!p) return nullptr;
    t->allocs.push_back({p, n ? n : 8});
    return p;
  };
  const mcrt_request& q = t->r.req;
  int rc = mcrt::solve_create(t->sd, S_->ctx->lds_per_block, q.cfg, q.geom, q.I0, q.seed, q.ray_begin,
                              q.instance, t->mixed, q.chunk_rays, mcrt::ChunkRule::Latency, alloc, err,
                              S_->cache, q.slots, q.stage_s);
  hipError_t e = hipSuccess;
  if (!rc) e = mcrt::solve_prepare(t->sd, setup_);
  if (!rc && e == hipSuccess) e = hipEventRecord(setup_ev_, setup_);
  if (!rc && e == hipSuccess) {
    const size_t n = (size_t)t->sd.nout;
    t->h_bytes = sizeof(double) * 2 * n + sizeof(unsigned long long) * mcrt::kCntN;
    t->h_block = S_->host_pool->get(t->h_bytes, &e);
    if (t->h_block) {
      t->h_out = (double*)t->h_block;
      t->h_sig = t->h_out + n;
      t->h_cnt = (unsigned long long*)(t->h_sig + n);
    }
  }
  if (!rc && e != hipSuccess) {
    rc = e == hipErrorOutOfMemory ? MCRT_ERR_NOMEM : MCRT_ERR_HIP;
    *err = hipGetErrorString(e);
  }
  *nchunks = rc ? 0 : t->sd.nchunks;
  *cost = rc ? 0.0 : t->sd.cost;
  return rc;
}

void HipBackend::destroy(uint64_t job) {
  Ticket* t = ticket(job);
  if (t) S_->release(t);
}

int HipBackend::launch(bool mixed, const std::vector<sched::Task>& tasks, std::string* err) {
  const int n = (int)tasks.size();
  if (n <= 0 || n > S_->opt.window_tasks) {
    *err = "window size out of range";
    return MCRT_ERR_INVALID;
  }
  // staging: next = 0, the chunk records, the outcome words (-1: unclaimed)
  std::memset(h_win_, 0, 64);
  auto* hc = (mcrt::BatchChunk*)(h_win_ + 64);
  auto* hs = (int64_t*)(h_win_ + 64 + sizeof(mcrt::BatchChunk) * (size_t)n);
  size_t lds = 0;
  for (int i = 0; i < n; ++i) {
    const sched::Task& k = tasks[i];
    Ticket* t = ticket(k.job);
    if (!t || k.cont < 0 || k.cont >= S_->opt.conts || k.chunk < 0 || k.chunk >= t->sd.nchunks ||
        t->sd.mixed != mixed) {
      *err = "malformed window task";
      return MCRT_ERR_INVALID;
    }
    mcrt::solve_chunk(t->sd, k.chunk, &hc[i]);
    if (k.resume) {
      if (k.rb < hc[i].rb || k.rb > hc[i].re) {
        *err = "resume point outside its chunk";
        return MCRT_ERR_INVALID;
      }
      hc[i].rb = k.rb;
      hc[i].resume = 1;
    }
    hc[i].cont = conts_ + cont_bytes_ * (size_t)k.cont;
    hs[i] = sched::kUnclaimed;
    lds = std::max(lds, t->sd.lds_bytes);
  }
  const size_t bytes = 64 + (sizeof(mcrt::BatchChunk) + sizeof(int64_t)) * (size_t)n;
  auto* dnext = (unsigned int*)d_win_;
  auto* dc = (const mcrt::BatchChunk*)(d_win_ + 64);
  auto* ds = (int64_t*)(d_win_ + 64 + sizeof(mcrt::BatchChunk) * (size_t)n);
  // the flag is clear before the launch is enqueued (no window is in flight)
  __atomic_store_n(flag_, 0u, __ATOMIC_RELEASE);
  hipError_t e = hipStreamWaitEvent(lane_, setup_ev_, 0);  // the solves' slabs are clear
  if (e == hipSuccess) e = hipMemcpyAsync(d_win_, h_win_, bytes, hipMemcpyHostToDevice, lane_);
  if (e == hipSuccess) e = hipEventRecord(start_ev_, lane_);
  if (e == hipSuccess)
    e = mcrt::batch_launch(mixed, dc, n, dnext, nullptr, S_->blocks, lds, lane_, ds, d_flag_, false);
  if (e == hipSuccess) e = hipMemcpyAsync(hs, ds, sizeof(int64_t) * (size_t)n, hipMemcpyDeviceToHost, lane_);
  if (e == hipSuccess) e = hipEventRecord(done_ev_, lane_);
  if (e != hipSuccess) {
    // an error after the first enqueue: wait for what was enqueued, then fail
    (void)hipStreamSynchronize(lane_);
    *err = hipGetErrorString(e);
    return MCRT_ERR_HIP;
  }
  win_n_ = n;
  return MCRT_OK;
}

int HipBackend::poll(std::vector<int64_t>* states, sched::Time* start, std::string* err) {
  const hipError_t q = hipEventQuery(done_ev_);
  if (q == hipErrorNotReady) return 0;
  if (q != hipSuccess) {
    *err = hipGetErrorString(q);
    return MCRT_ERR_HIP;
  }
  const auto* hs = (const int64_t*)(h_win_ + 64 + sizeof(mcrt::BatchChunk) * (size_t)win_n_);
  states->assign(hs, hs + win_n_);
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, ref_ev_, start_ev_) == hipSuccess)
    *start = ref_t_ + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double, std::milli>(ms));
  return 1;
}

int64_t HipBackend::finish(const std::vector<uint64_t>& jobs, std::string* err) {
  int slot = -1;
  for (int i = 0; i < kConcurrentFinishes; ++i)
    if (!fins_[i].used) {
      slot = i;
      break;
    }
  if (slot < 0) {
    *err = "too many read-backs in flight";
    return MCRT_ERR_HIP;
  }
  hipError_t e = hipSuccess;
  for (uint64_t id : jobs) {
    Ticket* t = ticket(id);
    if (!t) continue;
    const mcrt::SolveDev& sd = t->sd;
    const size_t n = (size_t)sd.nout;
    if (e == hipSuccess) e = mcrt::solve_finish(sd, lane_);
    if (e == hipSuccess) e = hipMemcpyAsync(t->h_out, sd.out, sizeof(double) * n, hipMemcpyDeviceToHost, lane_);
    if (e == hipSuccess) e = hipMemcpyAsync(t->h_sig, sd.sig, sizeof(double) * n, hipMemcpyDeviceToHost, lane_);
    if (e == hipSuccess)
      e = hipMemcpyAsync(t->h_cnt, sd.counters, sizeof(unsigned long long) * mcrt::kCntN,
                         hipMemcpyDeviceToHost, lane_);
  }
  if (e == hipSuccess) e = hipEventRecord(fins_[slot].ev, lane_);
  if (e != hipSuccess) {
    (void)hipStreamSynchronize(lane_);
    *err = hipGetErrorString(e);
    return MCRT_ERR_HIP;
  }
  fins_[slot].used = true;
  return slot;
}

int HipBackend::finished(int64_t h, std::string* err) {
  if (h < 0 || h >= kConcurrentFinishes || !fins_[h].used) {
    *err = "unknown read-back";
    return MCRT_ERR_INVALID;
  }
  const hipError_t q = hipEventQuery(fins_[h].ev);
  if (q == hipErrorNotReady) return 0;
  fins_[h].used = false;
  if (q != hipSuccess) {
    *err = hipGetErrorString(q);
    return MCRT_ERR_HIP;
  }
  return 1;
}

}  // namespace

void mcrt_server::release(Ticket* t) {
  if (t->released) return;
  mcrt::table_cache_release(cache, t->sd.table_refs);
  t->sd.table_refs.clear();
  for (auto& a : t->allocs) dev_pool->put(a.first, a.second);
  t->allocs.clear();
  t->released = true;
}

void mcrt_server::run() {
  (void)hipSetDevice(ctx->device);
  // a ticket is published once, after the scheduler has freed its device
  // state and nothing in flight holds it (its owner may then free it)
  auto publish = [this](const sched::Job& j, sched::Final f) {
    std::lock_guard<std::mutex> g(mu);
    auto it = tickets.find(j.id);
    if (it == tickets.end()) return;
    Ticket* t = it->second;
    t->t_admit = j.admitted ? j.t_admit : j.t_done;
    t->started = j.started;
    t->t_start = j.started ? j.t_start : j.t_done;
    t->t_done = j.t_done;
    if (f == sched::Final::Done) {
      t->state = MCRT_TICKET_DONE;
      st.rays += (uint64_t)t->sd.N;
    } else if (f == sched::Final::Cancelled) {
      t->state = MCRT_TICKET_CANCELLED;
    } else {
      t->state = MCRT_TICKET_FAILED;
      t->err_code = j.err_code ? j.err_code : MCRT_ERR_HIP;
      t->err = j.err;
    }
    cv_done.notify_all();
  };
  sched::Scheduler sc(backend, opt, publish);
  bool stopping = false;
  for (;;) {
    std::vector<Ticket*> in;
    std::vector<uint64_t> cn;
    {
      std::lock_guard<std::mutex> g(mu);
      in.swap(incoming);
      cn.swap(cancels);
      if (stop && !stopping) stopping = true;
    }
    const auto now = Clock::now();
    for (Ticket* t : in) sc.submit(t->id, t->mixed, t->need, t->t_submit);
    for (uint64_t id : cn) sc.cancel(id);
    if (stopping) sc.cancel_all();
    sc.set_slice(slice_ms.load());
    const bool progress = sc.step(now);
    {
      const sched::Stats s = sc.stats();
      std::lock_guard<std::mutex> g(mu);
      st.submitted = s.submitted;
      st.completed = s.completed;
      st.cancelled = s.cancelled;
      st.failed = s.failed;
      st.launches = s.launches;
      st.chunks = s.tasks;
      st.queued = s.queued + incoming.size();
      st.running = s.running;
      st.device_bytes_in_use = (int64_t)s.in_use;
      st.busy_ms = s.busy_ms;
      st.preemptions = s.preemptions;
      st.resumed = s.resumed;
    }
    if (stopping && sc.idle()) break;
    if (progress) continue;
    if (sc.in_flight()) {
      // a stopping window ends within an iteration: look again sooner
      std::this_thread::sleep_for(std::chrono::microseconds(sc.preempt_pending() ? 20 : 100));
    } else {
      std::unique_lock<std::mutex> lk(mu);
      cv_work.wait_for(lk, std::chrono::milliseconds(20),
                       [this] { return stop || !incoming.empty() || !cancels.empty(); });
    }
  }
}

extern "C" int mcrt_server_create(mcrt_ctx* ctx, int64_t device_budget_bytes,
                                  int64_t table_cache_bytes, mcrt_server** out) {
  if (!ctx) return fail(nullptr, MCRT_ERR_INVALID, "context is NULL");
  if (!out) return fail(ctx, MCRT_ERR_INVALID, "server_create: out is NULL");
  *out = nullptr;
  hipError_t e = hipSetDevice(ctx->device);
  size_t free_b = 0, total_b = 0;
  if (e == hipSuccess) e = hipMemGetInfo(&free_b, &total_b);
  if (e != hipSuccess) return fail(ctx, MCRT_ERR_HIP, "server_create: %s", hipGetErrorString(e));
  auto* S = new mcrt_server();
  S->ctx = ctx;
  S->budget = device_budget_bytes > 0 ? (size_t)device_budget_bytes : free_b / 2;
  S->cache = mcrt::table_cache_create(table_cache_bytes > 0 ? (size_t)table_cache_bytes : (size_t)1 << 30);
  S->blocks = std::max(1, ctx->cu_count);
  S->dev_pool = new Pool(false, (size_t)4 << 30);
  S->host_pool = new Pool(true, (size_t)1 << 30);
  S->opt.blocks = S->blocks;
  S->opt.window_tasks = 2 * S->blocks;
  S->opt.conts = 3 * S->blocks;
  S->opt.max_active = 1024;
  S->opt.budget = S->budget;
  S->opt.slice_ms = S->slice_ms.load();
  S->st.device_budget = (int64_t)S->budget;
  S->backend = new HipBackend(S);
  e = S->backend->init();
  // every solve's LDS fits in the largest request, so set it once
  if (e == hipSuccess) e = mcrt::batch_set_lds(true, ctx->lds_per_block - 16);
  if (e == hipSuccess) e = mcrt::batch_set_lds(false, ctx->lds_per_block - 16);
  if (e != hipSuccess) {
    int rc = fail(ctx, MCRT_ERR_HIP, "server_create: %s", hipGetErrorString(e));
    mcrt_server_destroy(S);
    return rc;
  }
  S->worker = std::thread([S] { S->run(); });
  *out = S;
  return MCRT_OK;
}

extern "C" int mcrt_submit(mcrt_server* S, const mcrt_request* q, uint64_t* ticket) {
  if (!S) return fail(nullptr, MCRT_ERR_INVALID, "server is NULL");
  if (!q || !ticket) return fail(S->ctx, MCRT_ERR_INVALID, "submit: NULL argument");
  // validation on the caller's thread (the reference's ArgumentError)
  mcrt_ctx tmp;  // messages of this call only
  tmp.device = S->ctx->device;
  int rc;
  if ((rc = mcrt::validate_config(&tmp, q->cfg)) || (rc = mcrt::validate_geometry(&tmp, q->geom)))
    return fail(nullptr, rc, "%s", tmp.err.c_str());
  if (q->precision != MCRT_PRECISION_FP64 && q->precision != MCRT_PRECISION_MIXED)
    return fail(nullptr, MCRT_ERR_UNSUPPORTED, "the server runs the fp64 and mixed forms");
  if (q->chunk_rays < 0) return fail(nullptr, MCRT_ERR_INVALID, "chunk_rays < 0");
  if (q->slots < 0 || q->slots % 64)
    return fail(nullptr, MCRT_ERR_INVALID, "slots must be 0 or a positive multiple of 64");
  auto* t = new Ticket();
  own(*q, t->r);
  t->mixed = q->precision == MCRT_PRECISION_MIXED;
  t->need = mcrt::solve_bytes_estimate(q->cfg, q->geom, t->mixed, q->chunk_rays,
                                       mcrt::ChunkRule::Latency);
  t->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> g(S->mu);
    if (S->stop) {
      delete t;
      return fail(nullptr, MCRT_ERR_CANCELLED, "the server is shutting down");
    }
    t->id = S->next_id++;
    S->tickets[t->id] = t;
    S->incoming.push_back(t);
  }
  S->cv_work.notify_one();
  *ticket = t->id;
  return MCRT_OK;
}

extern "C" int mcrt_poll(mcrt_server* S, uint64_t ticket, int32_t* state) {
  if (!S || !state) return fail(nullptr, MCRT_ERR_INVALID, "poll: NULL argument");
  std::lock_guard<std::mutex> g(S->mu);
  auto it = S->tickets.find(ticket);
  if (it == S->tickets.end()) return fail(nullptr, MCRT_ERR_INVALID, "unknown ticket %llu", (unsigned long long)ticket);
  *state = it->second->state;
  return MCRT_OK;
}

extern "C" int mcrt_wait(mcrt_server* S, uint64_t ticket, double timeout_ms, int32_t* state) {
  if (!S) return fail(nullptr, MCRT_ERR_INVALID, "server is NULL");
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->tickets.find(ticket);
  if (it == S->tickets.end()) return fail(nullptr, MCRT_ERR_INVALID, "unknown ticket %llu", (unsigned long long)ticket);
  Ticket* t = it->second;
  auto over = [t] { return t->state >= MCRT_TICKET_DONE; };
  if (timeout_ms < 0) S->cv_done.wait(lk, over);
  else S->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), over);
  if (state) *state = t->state;
  return over() ? MCRT_OK : MCRT_ERR_NOT_READY;
}

extern "C" int mcrt_result(mcrt_server* S, uint64_t ticket, double A[5], double* Irt, double* Icb,
                           double A_sigma[5], double* Irt_sigma, double* Icb_sigma, mcrt_stats* stats,
                           double* latency_ms) {
  if (!S) return fail(nullptr, MCRT_ERR_INVALID, "server is NULL");
  Ticket* t = nullptr;
  {
    std::lock_guard<std::mutex> g(S->mu);
    auto it = S->tickets.find(ticket);
    if (it == S->tickets.end()) return fail(nullptr, MCRT_ERR_INVALID, "unknown ticket %llu", (unsigned long long)ticket);
    t = it->second;
    if (t->state < MCRT_TICKET_DONE) return fail(nullptr, MCRT_ERR_NOT_READY, "ticket %llu is not finished", (unsigned long long)ticket);
    S->tickets.erase(it);
  }
  auto ms = [](Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
  };
  int rc = MCRT_OK;
  if (t->state == MCRT_TICKET_DONE) {
    const mcrt::SolveDev& sd = t->sd;
    const double N = (double)sd.N;
    if (A) for (int c = 0; c < 5; ++c) A[c] = t->h_out[c] / N;
    if (Irt) for (int64_t c = 0; c < sd.nrt; ++c) Irt[c] = t->h_out[5 + c] / N;
    if (Icb) for (int64_t c = 0; c < sd.ncb; ++c) Icb[c] = t->h_out[5 + sd.nrt + c] / N;
    if (A_sigma) for (int c = 0; c < 5; ++c) A_sigma[c] = t->h_sig[c];
    if (Irt_sigma) for (int64_t c = 0; c < sd.nrt; ++c) Irt_sigma[c] = t->h_sig[5 + c];
    if (Icb_sigma) for (int64_t c = 0; c < sd.ncb; ++c) Icb_sigma[c] = t->h_sig[5 + sd.nrt + c];
    if (stats) {
      const unsigned long long* c = t->h_cnt;
      stats->rays = c[mcrt::kCntRays];
      stats->events = c[mcrt::kCntEvents];
      stats->rt_evals = c[mcrt::kCntRt];
      stats->cb_evals = c[mcrt::kCntCb];
      stats->cb_evals_single = c[mcrt::kCntCb1];
      stats->rejections = c[mcrt::kCntRej];
      stats->nonfinite = c[mcrt::kCntNonfinite];
      stats->kernel_ms = ms(t->t_start, t->t_done);
      stats->total_ms = ms(t->t_submit, t->t_done);
      stats->batches = (uint64_t)sd.nbatch;
      stats->chunks = (uint64_t)sd.nchunks;
      stats->queue_ms = ms(t->t_submit, t->t_admit);
      stats->wait_ms = ms(t->t_admit, t->t_start);
    }
  } else if (t->state == MCRT_TICKET_CANCELLED) {
    rc = fail(nullptr, MCRT_ERR_CANCELLED, "ticket %llu was cancelled", (unsigned long long)ticket);
  } else {
    rc = fail(nullptr, t->err_code ? t->err_code : MCRT_ERR_HIP, "ticket %llu failed: %s",
              (unsigned long long)ticket, t->err.c_str());
  }
  if (latency_ms) *latency_ms = ms(t->t_submit, t->t_done);
  if (t->h_block) S->host_pool->put(t->h_block, t->h_bytes);
  delete t;
  return rc;
}

extern "C" int mcrt_cancel(mcrt_server* S, uint64_t ticket) {
  if (!S) return fail(nullptr, MCRT_ERR_INVALID, "server is NULL");
  {
    std::lock_guard<std::mutex> g(S->mu);
    auto it = S->tickets.find(ticket);
    if (it == S->tickets.end()) return fail(nullptr, MCRT_ERR_INVALID, "unknown ticket %llu", (unsigned long long)ticket);
    if (it->second->state < MCRT_TICKET_DONE) S->cancels.push_back(ticket);
  }
  S->cv_work.notify_one();
  return MCRT_OK;
}

extern "C" int mcrt_server_stats_get(mcrt_server* S, mcrt_server_stats* out) {
  if (!S || !out) return fail(nullptr, MCRT_ERR_INVALID, "server_stats: NULL argument");
  std::lock_guard<std::mutex> g(S->mu);
  *out = S->st;
  out->device_budget = (int64_t)S->budget;
  size_t cb = 0;
  mcrt::table_cache_stats(S->cache, &cb, &out->table_hits, &out->table_misses);
  out->table_cache_bytes = (int64_t)cb;
  return MCRT_OK;
}

extern "C" int mcrt_server_set_slice(mcrt_server* S, double slice_ms) {
  if (!S) return fail(nullptr, MCRT_ERR_INVALID, "server is NULL");
  if (!(slice_ms >= 0.0 && slice_ms <= 1e6)) return fail(S->ctx, MCRT_ERR_INVALID, "slice_ms out of range");
  S->slice_ms.store(slice_ms);
  return MCRT_OK;
}

extern "C" void mcrt_server_destroy(mcrt_server* S) {
  if (!S) return;
  {
    std::lock_guard<std::mutex> g(S->mu);
    S->stop = true;
  }
  S->cv_work.notify_all();
  if (S->worker.joinable()) S->worker.join();
  (void)hipSetDevice(S->ctx->device);
  if (S->backend) S->backend->shutdown();
  for (auto& kv : S->tickets) {
    Ticket* t = kv.second;
    S->release(t);
    if (t->h_block) S->host_pool->put(t->h_block, t->h_bytes);
    delete t;
  }
  delete S->backend;
  mcrt::table_cache_destroy(S->cache);
  delete S->dev_pool;
  delete S->host_pool;
  delete S;
}
