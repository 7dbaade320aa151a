// mcrt_serve.cpp — the request loop (C ABI mcrt_server_*, mcrt_submit, …).
//
// Continuous batching on one device: a host thread admits submitted solves
// while their device memory (bins, slabs, parameters; Mie tables are shared
// through a content-keyed cache) fits a budget, and keeps two launch lanes
// (HIP streams) busy with persistent batch_kernel launches.  Each launch takes
// a window of chunks round-robin from every admitted solve of one precision,
// so a solve that arrives while others run joins the next window; with two
// lanes the next window is queued on the device while the current one runs
// and its workgroups start as the current one's finish.  A solve is cut into
// chunks by a rule that depends on the solve alone, and a chunk never shares
// an iteration with another, so every result is the same alone and in the mix
// up to the order of FP64 additions.
#include <hip/hip_runtime.h>

#include <algorithm>
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

using mcrt::fail;
using Clock = std::chrono::steady_clock;

namespace {

// ---------------------------------------------------------- owned requests
struct OwnedScatterer {
  mcrt_scatterer s{};
  std::vector<double> csrn, F11, F12, F33, F34, S11, S22;
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
  if (s->kind == MCRT_SCATTERER_MIE) {
    const int64_t N = s->n_theta;
    c->s.csrn = keep(c->csrn, s->csrn, (int64_t)s->random_points + 1);
    c->s.F11 = keep(c->F11, s->F11, N);
    c->s.F12 = keep(c->F12, s->F12, N);
    c->s.F33 = keep(c->F33, s->F33, N);
    c->s.F34 = keep(c->F34, s->F34, N);
    c->s.S11 = keep(c->S11, s->S11, 2 * N);
    c->s.S22 = keep(c->S22, s->S22, 2 * N);
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
  bool cancel = false;
  size_t need = 0;
  bool admitted = false;  // holds `need` bytes of the budget
  mcrt::SolveDev sd;
  std::vector<std::pair<void*, size_t>> allocs;  // from the server's device pool
  int64_t next_chunk = 0;
  int inflight = 0;           // launches still running with chunks of this solve
  bool finished = false;      // the launch with its reduction and read-back completed
  bool released = false;      // device memory freed
  // read-back of the results: one pinned block (host_pool) holding the sums,
  // the standard errors and the counters
  void* h_block = nullptr;
  size_t h_bytes = 0;
  double* h_out = nullptr;
  double* h_sig = nullptr;
  unsigned long long* h_cnt = nullptr;
  Clock::time_point t_submit, t_start, t_done;
};

constexpr int kLanes = 2;
constexpr int kLaneChunks = 1 << 16;  // chunk-list capacity of one launch
// Admitted solves at a time.  An admitted solve holds device and pinned
// memory; under overload, admitting everything the budget allows (tens of
// thousands of small solves) only multiplies live allocations beyond what the
// next launch windows can use.  The rest wait in the queue.
constexpr size_t kMaxActive = 1024;

}  // namespace

struct mcrt_server {
  mcrt_ctx* ctx = nullptr;
  std::thread worker;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  bool stop = false;
  std::deque<Ticket*> pending;
  std::vector<Ticket*> active;
  std::map<uint64_t, Ticket*> tickets;
  uint64_t next_id = 1;
  size_t budget = 0, in_use = 0;
  mcrt::TableCache* cache = nullptr;
  Pool* dev_pool = nullptr;
  Pool* host_pool = nullptr;
  hipStream_t setup = nullptr;
  hipEvent_t setup_ev = nullptr;  // after the last admitted solve's clear
  struct Lane {
    hipStream_t s = nullptr;
    void* slot_mem = nullptr;
    unsigned int* dnext = nullptr;
    mcrt::BatchChunk* dchunks = nullptr;
    mcrt::BatchChunk* hchunks = nullptr;  // pinned
    hipEvent_t ev = nullptr;
    bool busy = false;
    std::vector<Ticket*> members, finishing;
  } lanes[kLanes];
  int blocks = 0;
  bool prefer_mixed = true;
  // statistics
  mcrt_server_stats st{};
  double busy_ms = 0.0;

  void run();
  void admit(std::vector<Ticket*>& batch);
  bool launch(Lane& L);
  void release(Ticket* t);
};

void mcrt_server::release(Ticket* t) {
  if (t->released) return;
  mcrt::table_cache_release(cache, t->sd.table_refs);
  t->sd.table_refs.clear();
  for (auto& a : t->allocs) dev_pool->put(a.first, a.second);
  t->allocs.clear();
  t->released = true;
  if (t->admitted) {
    std::lock_guard<std::mutex> g(mu);
    in_use -= t->need;
    t->admitted = false;
  }
}

// Create the device state of newly admitted solves (outside the lock).
void mcrt_server::admit(std::vector<Ticket*>& batch) {
  for (Ticket* t : batch) {
    Pool* pool = dev_pool;
    auto alloc = [t, pool](size_t n, hipError_t* e) -> void* {
      void* p = pool->get(n ? n : 8, e);
      if (!p) return nullptr;
      t->allocs.push_back({p, n ? n : 8});
      return p;
    };
    std::string err;
    const mcrt_request& q = t->r.req;
    int rc = mcrt::solve_create(t->sd, ctx->lds_per_block, q.cfg, q.geom, q.I0, q.seed, q.ray_begin,
                                q.instance, t->mixed, q.chunk_rays, mcrt::ChunkRule::Latency,
                                alloc, &err, cache);
    hipError_t e = hipSuccess;
    // the clear runs on the set-up stream; launches wait for setup_ev
    if (!rc) e = mcrt::solve_prepare(t->sd, setup);
    if (!rc && e == hipSuccess) {
      const size_t n = (size_t)t->sd.nout;
      t->h_bytes = sizeof(double) * 2 * n + sizeof(unsigned long long) * mcrt::kCntN;
      t->h_block = host_pool->get(t->h_bytes, &e);
      if (t->h_block) {
        t->h_out = (double*)t->h_block;
        t->h_sig = t->h_out + n;
        t->h_cnt = (unsigned long long*)(t->h_sig + n);
      }
    }
    if (!rc && e != hipSuccess) {
      rc = e == hipErrorOutOfMemory ? MCRT_ERR_NOMEM : MCRT_ERR_HIP;
      err = hipGetErrorString(e);
    }
    if (rc) {
      release(t);
      std::lock_guard<std::mutex> g(mu);
      t->state = MCRT_TICKET_FAILED;
      t->err_code = rc;
      t->err = err;
      t->t_done = Clock::now();
      ++st.failed;
      cv_done.notify_all();
      continue;
    }
    std::lock_guard<std::mutex> g(mu);
    t->state = MCRT_TICKET_RUNNING;
    t->t_start = Clock::now();
    active.push_back(t);
  }
  (void)hipEventRecord(setup_ev, setup);
}

// Fill and launch one window on lane L; false when there is nothing to run.
bool mcrt_server::launch(Lane& L) {
  std::vector<Ticket*> elig;
  bool mixed;
  {
    std::lock_guard<std::mutex> g(mu);
    bool any[2] = {false, false};
    for (Ticket* t : active)
      if (!t->cancel && t->next_chunk < t->sd.nchunks) any[t->mixed ? 1 : 0] = true;
    if (!any[0] && !any[1]) return false;
    mixed = any[1] && (prefer_mixed || !any[0]);
    prefer_mixed = !mixed;  // alternate when both precisions wait
    for (Ticket* t : active)
      if (!t->cancel && t->mixed == mixed && t->next_chunk < t->sd.nchunks) elig.push_back(t);
  }
  // round-robin over the solves (oldest first) up to ~2 chunks per workgroup
  const int target = std::min(kLaneChunks, std::max(2 * blocks, (int)elig.size()));
  struct Item { double cost; Ticket* t; int64_t k; };
  std::vector<Item> items;
  std::vector<int64_t> take(elig.size(), 0);
  for (bool more = true; more && (int)items.size() < target;) {
    more = false;
    for (size_t i = 0; i < elig.size() && (int)items.size() < target; ++i) {
      Ticket* t = elig[i];
      const int64_t k = t->next_chunk + take[i];
      if (k >= t->sd.nchunks) continue;
      items.push_back({t->sd.cost, t, k});
      ++take[i];
      more = true;
    }
  }
  if (items.empty()) return false;
  std::stable_sort(items.begin(), items.end(), [](const Item& a, const Item& b) { return a.cost > b.cost; });
  size_t lds = 0;
  L.members.clear();
  L.finishing.clear();
  {
    std::lock_guard<std::mutex> g(mu);  // mcrt_cancel reads next_chunk
    for (size_t i = 0; i < elig.size(); ++i) {
      if (!take[i]) continue;
      Ticket* t = elig[i];
      t->next_chunk += take[i];
      ++t->inflight;
      L.members.push_back(t);
      if (t->next_chunk >= t->sd.nchunks) L.finishing.push_back(t);
      lds = std::max(lds, t->sd.lds_bytes);
    }
  }
  for (size_t j = 0; j < items.size(); ++j) mcrt::solve_chunk(items[j].t->sd, items[j].k, &L.hchunks[j]);
  hipError_t e = hipMemcpyAsync(L.dchunks, L.hchunks, sizeof(mcrt::BatchChunk) * items.size(),
                                hipMemcpyHostToDevice, L.s);
  if (e == hipSuccess) e = hipStreamWaitEvent(L.s, setup_ev, 0);  // the solves' slabs are clear
  if (e == hipSuccess)
    e = mcrt::batch_launch(mixed, L.dchunks, (int)items.size(), L.dnext, L.slot_mem, blocks, lds, L.s);
  // a solve finishing here may still have chunks in the other lane's launch:
  // its reduction waits for that launch
  for (Lane& O : lanes) {
    if (&O == &L || !O.busy || e != hipSuccess) continue;
    bool shared = false;
    for (Ticket* t : L.finishing)
      if (std::find(O.members.begin(), O.members.end(), t) != O.members.end()) shared = true;
    if (shared) e = hipStreamWaitEvent(L.s, O.ev, 0);
  }
  for (Ticket* t : L.finishing) {
    if (e != hipSuccess) break;
    e = mcrt::solve_finish(t->sd, L.s);
    const size_t n = (size_t)t->sd.nout;
    if (e == hipSuccess) e = hipMemcpyAsync(t->h_out, t->sd.out, sizeof(double) * n, hipMemcpyDeviceToHost, L.s);
    if (e == hipSuccess) e = hipMemcpyAsync(t->h_sig, t->sd.sig, sizeof(double) * n, hipMemcpyDeviceToHost, L.s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(t->h_cnt, t->sd.counters, sizeof(unsigned long long) * mcrt::kCntN,
                         hipMemcpyDeviceToHost, L.s);
  }
  if (e == hipSuccess) e = hipEventRecord(L.ev, L.s);
  if (e != hipSuccess) {
    // the lane is unusable: its solves launch nothing more, and the sweep
    // publishes them as failed once no launch holds them (a ticket is never
    // published while the server still lists it: its owner may free it)
    (void)hipStreamSynchronize(L.s);
    std::lock_guard<std::mutex> g(mu);
    for (Ticket* t : L.members) {
      --t->inflight;
      t->cancel = true;
      if (!t->err_code) {
        t->err_code = MCRT_ERR_HIP;
        t->err = hipGetErrorString(e);
      }
    }
    L.members.clear();
    L.finishing.clear();
    return false;
  }
  L.busy = true;
  ++st.launches;
  st.chunks += items.size();
  return true;
}

void mcrt_server::run() {
  (void)hipSetDevice(ctx->device);
  auto t_busy = Clock::now();
  bool was_busy = false;
  for (;;) {
    bool progress = false;
    // ---- admission
    std::vector<Ticket*> batch;
    {
      std::lock_guard<std::mutex> g(mu);
      while (!pending.empty()) {
        Ticket* t = pending.front();
        if (t->cancel) {
          pending.pop_front();
          t->state = MCRT_TICKET_CANCELLED;
          t->t_done = Clock::now();
          ++st.cancelled;
          cv_done.notify_all();
          continue;
        }
        if (t->need > budget) {
          pending.pop_front();
          t->state = MCRT_TICKET_FAILED;
          t->err_code = MCRT_ERR_NOMEM;
          t->err = "the solve needs more device memory than the server's budget";
          t->t_done = Clock::now();
          ++st.failed;
          cv_done.notify_all();
          continue;
        }
        if (in_use + t->need > budget) break;  // wait for memory to free up
        if (active.size() + batch.size() >= kMaxActive) break;
        pending.pop_front();
        in_use += t->need;
        t->admitted = true;
        batch.push_back(t);
      }
    }
    if (!batch.empty()) {
      admit(batch);
      progress = true;
    }
    // ---- completions
    for (Lane& L : lanes) {
      if (!L.busy || hipEventQuery(L.ev) != hipSuccess) continue;
      L.busy = false;
      progress = true;
      std::lock_guard<std::mutex> g(mu);
      for (Ticket* t : L.members) --t->inflight;
      for (Ticket* t : L.finishing) t->finished = true;
      L.members.clear();
      L.finishing.clear();
    }
    // ---- solves that will launch nothing more: free their device memory,
    // then publish the final state
    {
      std::vector<Ticket*> over;
      {
        std::lock_guard<std::mutex> g(mu);
        for (auto it = active.begin(); it != active.end();) {
          Ticket* t = *it;
          if (t->inflight == 0 && (t->finished || t->cancel)) {
            over.push_back(t);
            it = active.erase(it);
          } else {
            ++it;
          }
        }
      }
      for (Ticket* t : over) release(t);
      if (!over.empty()) {
        progress = true;
        std::lock_guard<std::mutex> g(mu);
        for (Ticket* t : over) {
          t->t_done = Clock::now();
          if (t->err_code) {  // a launch holding its chunks failed
            t->state = MCRT_TICKET_FAILED;
            ++st.failed;
          } else if (t->finished) {
            t->state = MCRT_TICKET_DONE;
            ++st.completed;
            st.rays += (uint64_t)t->sd.N;
          } else {
            t->state = MCRT_TICKET_CANCELLED;
            ++st.cancelled;
          }
        }
        cv_done.notify_all();
      }
    }
    // ---- launches
    for (Lane& L : lanes)
      if (!L.busy && launch(L)) progress = true;
    const bool busy = lanes[0].busy || lanes[1].busy;
    if (busy && !was_busy) t_busy = Clock::now();
    if (!busy && was_busy) busy_ms += std::chrono::duration<double, std::milli>(Clock::now() - t_busy).count();
    was_busy = busy;
    if (progress) continue;
    std::unique_lock<std::mutex> lk(mu);
    if (stop && pending.empty() && active.empty() && !busy) break;
    if (busy) {
      lk.unlock();
      std::this_thread::sleep_for(std::chrono::microseconds(200));
    } else {
      cv_work.wait_for(lk, std::chrono::milliseconds(20),
                       [this] { return stop || !pending.empty(); });
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
  const size_t slot = std::max(mcrt::batch_slot_bytes(true, S->blocks), mcrt::batch_slot_bytes(false, S->blocks));
  e = hipStreamCreateWithFlags(&S->setup, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&S->setup_ev, hipEventDisableTiming);
  for (auto& L : S->lanes) {
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&L.s, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&L.slot_mem, slot);
    if (e == hipSuccess) e = hipMalloc((void**)&L.dnext, sizeof(unsigned int));
    if (e == hipSuccess) e = hipMalloc((void**)&L.dchunks, sizeof(mcrt::BatchChunk) * kLaneChunks);
    if (e == hipSuccess)
      e = hipHostMalloc((void**)&L.hchunks, sizeof(mcrt::BatchChunk) * kLaneChunks, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&L.ev, hipEventDisableTiming);
  }
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
    S->pending.push_back(t);
    ++S->st.submitted;
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
      stats->kernel_ms = std::chrono::duration<double, std::milli>(t->t_done - t->t_start).count();
      stats->total_ms = std::chrono::duration<double, std::milli>(t->t_done - t->t_submit).count();
      stats->batches = (uint64_t)sd.nbatch;
      stats->chunks = (uint64_t)sd.nchunks;
    }
  } else if (t->state == MCRT_TICKET_CANCELLED) {
    rc = fail(nullptr, MCRT_ERR_CANCELLED, "ticket %llu was cancelled", (unsigned long long)ticket);
  } else {
    rc = fail(nullptr, t->err_code ? t->err_code : MCRT_ERR_HIP, "ticket %llu failed: %s",
              (unsigned long long)ticket, t->err.c_str());
  }
  if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(t->t_done - t->t_submit).count();
  S->host_pool->put(t->h_block, t->h_bytes);
  delete t;
  return rc;
}

extern "C" int mcrt_cancel(mcrt_server* S, uint64_t ticket) {
  if (!S) return fail(nullptr, MCRT_ERR_INVALID, "server is NULL");
  {
    std::lock_guard<std::mutex> g(S->mu);
    auto it = S->tickets.find(ticket);
    if (it == S->tickets.end()) return fail(nullptr, MCRT_ERR_INVALID, "unknown ticket %llu", (unsigned long long)ticket);
    Ticket* t = it->second;
    if (t->state < MCRT_TICKET_DONE && t->next_chunk < (t->sd.nchunks ? t->sd.nchunks : 1)) t->cancel = true;
  }
  S->cv_work.notify_one();
  return MCRT_OK;
}

extern "C" int mcrt_server_stats_get(mcrt_server* S, mcrt_server_stats* out) {
  if (!S || !out) return fail(nullptr, MCRT_ERR_INVALID, "server_stats: NULL argument");
  std::lock_guard<std::mutex> g(S->mu);
  *out = S->st;
  out->queued = S->pending.size();
  out->running = S->active.size();
  out->device_bytes_in_use = (int64_t)S->in_use;
  out->device_budget = (int64_t)S->budget;
  out->busy_ms = S->busy_ms;
  size_t cb = 0;
  mcrt::table_cache_stats(S->cache, &cb, &out->table_hits, &out->table_misses);
  out->table_cache_bytes = (int64_t)cb;
  return MCRT_OK;
}

extern "C" void mcrt_server_destroy(mcrt_server* S) {
  if (!S) return;
  {
    std::lock_guard<std::mutex> g(S->mu);
    S->stop = true;
    for (Ticket* t : S->pending) t->cancel = true;
    for (Ticket* t : S->active) t->cancel = true;
  }
  S->cv_work.notify_all();
  if (S->worker.joinable()) S->worker.join();
  (void)hipSetDevice(S->ctx->device);
  for (auto& kv : S->tickets) {
    Ticket* t = kv.second;
    S->release(t);
    S->host_pool->put(t->h_block, t->h_bytes);
    delete t;
  }
  for (auto& L : S->lanes) {
    if (L.s) (void)hipStreamSynchronize(L.s);
    if (L.ev) (void)hipEventDestroy(L.ev);
    if (L.hchunks) (void)hipHostFree(L.hchunks);
    if (L.dchunks) (void)hipFree(L.dchunks);
    if (L.dnext) (void)hipFree(L.dnext);
    if (L.slot_mem) (void)hipFree(L.slot_mem);
    if (L.s) (void)hipStreamDestroy(L.s);
  }
  if (S->setup) (void)hipStreamSynchronize(S->setup);
  if (S->setup_ev) (void)hipEventDestroy(S->setup_ev);
  if (S->setup) (void)hipStreamDestroy(S->setup);
  mcrt::table_cache_destroy(S->cache);
  delete S->dev_pool;
  delete S->host_pool;
  delete S;
}
