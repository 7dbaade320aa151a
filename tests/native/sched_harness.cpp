// sched_harness.cpp — the server's scheduler (mcrt_sched.cpp) over a
// simulated device, for host-only tests.
//
// The simulated device runs a launch window as list scheduling of its tasks
// over `blocks` workgroups in virtual time: each workgroup claims the next
// task when free; a preempted window's workgroups stop their task at the
// first iteration boundary after the request (and after at least one
// iteration of it) and claim no further task.  Every call of the scheduler is
// checked against what the real device relies on: a resumed task carries the
// continuation block its chunk was stopped in, no chunk finishes twice, no
// window or read-back touches a job whose device state was freed, a job is
// published once, after its state was freed and with nothing in flight.
// Injected faults: launch() refusing before anything is enqueued, a window
// that the device fails, a create() that fails.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "../../include/mcrt.h"
#include "../../montecarloradiativetransfer.jl_amd/csrc/mcrt_sched.h"

namespace {

using mcrt::sched::Task;
using mcrt::sched::Time;

Time at(double ms) {
  return Time(std::chrono::duration_cast<mcrt::sched::Clock::duration>(
      std::chrono::duration<double, std::milli>(ms)));
}
double ms_of(Time t) {
  return std::chrono::duration<double, std::milli>(t.time_since_epoch()).count();
}

struct FakeJob {
  int64_t nchunks = 0;
  double chunk_ms = 1.0;
  bool fail_create = false;
  bool created = false, destroyed = false;
  std::vector<double> remaining;
  std::vector<int> cont;        // block holding the chunk's stopped rays, or -1
  std::vector<int64_t> rb;      // resume point given out with it
  std::vector<int> finished;    // times finished
};

struct Harness;

struct Fake final : mcrt::sched::Backend {
  Harness* H = nullptr;
  int blocks = 4;
  double iter_ms = 0.05;
  double now = 0.0;
  std::map<uint64_t, FakeJob> jobs;
  bool live = false, preempted = false;
  double t0 = 0.0, tp = 0.0;
  std::vector<Task> tasks;
  int fail_launches = 0, fail_windows = 0;
  std::map<int64_t, std::pair<double, std::vector<uint64_t>>> fins;
  int64_t next_fin = 0;
  std::vector<std::string> violations;

  void violate(const std::string& m) {
    if (violations.size() < 64) violations.push_back(m);
    else violations.back() = "(more)";
  }

  int create(uint64_t job, int64_t* nchunks, double* cost, std::string* err) override {
    FakeJob& j = jobs[job];
    j.created = true;
    if (j.fail_create) {
      *err = "simulated allocation failure";
      *nchunks = 0;
      return MCRT_ERR_NOMEM;
    }
    j.remaining.assign(j.nchunks, j.chunk_ms);
    j.cont.assign(j.nchunks, -1);
    j.rb.assign(j.nchunks, -1);
    j.finished.assign(j.nchunks, 0);
    *nchunks = j.nchunks;
    *cost = j.chunk_ms;
    return MCRT_OK;
  }

  void destroy(uint64_t job) override {
    FakeJob& j = jobs[job];
    if (j.destroyed) violate("destroy twice: job " + std::to_string(job));
    if (live)
      for (const Task& t : tasks)
        if (t.job == job) violate("destroy while a window holds job " + std::to_string(job));
    for (auto& f : fins)
      for (uint64_t id : f.second.second)
        if (id == job) violate("destroy while a read-back holds job " + std::to_string(job));
    j.destroyed = true;
  }

  int launch(bool, const std::vector<Task>& ts, std::string* err) override {
    if (live) violate("launch while a window is in flight");
    if (fail_launches > 0) {
      --fail_launches;
      *err = "simulated launch failure";
      return MCRT_ERR_HIP;
    }
    std::set<std::pair<uint64_t, int64_t>> seen;
    std::set<int> conts;
    for (const Task& t : ts) {
      auto it = jobs.find(t.job);
      if (it == jobs.end() || !it->second.created || it->second.destroyed) {
        violate("window task of a job without device state: " + std::to_string(t.job));
        continue;
      }
      FakeJob& j = it->second;
      if (t.chunk < 0 || t.chunk >= j.nchunks) {
        violate("chunk out of range");
        continue;
      }
      if (!seen.insert({t.job, t.chunk}).second) violate("chunk twice in one window");
      if (t.cont < 0 || !conts.insert(t.cont).second) violate("continuation block missing or shared");
      if (j.finished[t.chunk]) violate("finished chunk launched again");
      const bool saved = j.cont[t.chunk] >= 0;
      if (t.resume != saved) violate("resume flag does not match the chunk's saved state");
      if (saved && (t.cont != j.cont[t.chunk] || t.rb != j.rb[t.chunk]))
        violate("chunk resumed from another block or point");
    }
    tasks = ts;
    live = true;
    preempted = false;
    t0 = now;
    return MCRT_OK;
  }

  // The simulated schedule of the window: outcome per task and the end time.
  double simulate(std::vector<int64_t>* st, std::vector<double>* ran) const {
    const size_t n = tasks.size();
    st->assign(n, mcrt::sched::kUnclaimed);
    ran->assign(n, 0.0);
    std::vector<double> free_at(blocks, t0);
    std::vector<int> claims(blocks, 0);
    double end = t0;
    size_t next = 0;
    for (;;) {
      int w = -1;
      for (int k = 0; k < blocks; ++k)
        if (w < 0 || free_at[k] < free_at[w]) w = k;
      if (next >= n) break;
      const double ts = free_at[w];
      if (preempted && claims[w] > 0 && ts >= tp) {
        free_at[w] = INFINITY;  // this workgroup has exited
        bool any = false;
        for (int k = 0; k < blocks; ++k) any |= std::isfinite(free_at[k]);
        if (!any) break;
        continue;
      }
      const size_t i = next++;
      ++claims[w];
      const Task& t = tasks[i];
      auto it = jobs.find(t.job);
      const double r = it == jobs.end() || t.chunk >= (int64_t)it->second.remaining.size()
                           ? 0.0
                           : it->second.remaining[t.chunk];
      double stop = INFINITY;
      if (preempted) stop = std::max(tp, ts + iter_ms);
      if (ts + r <= stop) {
        (*st)[i] = mcrt::sched::kFinished;
        (*ran)[i] = r;
        free_at[w] = ts + r;
      } else {
        (*st)[i] = 1000 * t.chunk + (int64_t)(ts * 10) % 997;  // a resume point ≥ 0
        (*ran)[i] = stop - ts;
        free_at[w] = stop;
      }
      end = std::max(end, free_at[w]);
    }
    for (int k = 0; k < blocks; ++k)
      if (std::isfinite(free_at[k])) end = std::max(end, free_at[k]);
    return end;
  }

  int poll(std::vector<int64_t>* states, Time* start, std::string* err) override {
    if (!live) {
      violate("poll with no window");
      return 0;
    }
    std::vector<int64_t> st;
    std::vector<double> ran;
    const double end = simulate(&st, &ran);
    if (now < end) return 0;
    live = false;
    *start = at(t0);
    if (fail_windows > 0) {
      --fail_windows;
      tasks.clear();
      *err = "simulated device failure";
      return MCRT_ERR_HIP;
    }
    for (size_t i = 0; i < tasks.size(); ++i) {
      const Task& t = tasks[i];
      FakeJob& j = jobs[t.job];
      if (t.chunk < 0 || t.chunk >= j.nchunks) continue;
      if (st[i] == mcrt::sched::kFinished) {
        ++j.finished[t.chunk];
        if (j.finished[t.chunk] > 1) violate("chunk finished twice");
        j.remaining[t.chunk] = 0.0;
        j.cont[t.chunk] = -1;
        j.rb[t.chunk] = -1;
      } else if (st[i] >= 0) {
        j.remaining[t.chunk] -= ran[i];
        j.cont[t.chunk] = t.cont;
        j.rb[t.chunk] = st[i];
      }
    }
    *states = st;
    tasks.clear();
    return 1;
  }

  void preempt() override {
    if (!live) violate("preempt with no window");
    if (!preempted) {
      preempted = true;
      tp = now;
    }
  }

  int64_t finish(const std::vector<uint64_t>& js, std::string*) override {
    for (uint64_t id : js) {
      FakeJob& j = jobs[id];
      if (j.destroyed) violate("read-back of a freed job");
      for (int64_t c = 0; c < j.nchunks; ++c)
        if (j.finished[c] != 1) violate("read-back before every chunk finished once");
    }
    const int64_t h = next_fin++;
    fins[h] = {now + 0.02, js};
    return h;
  }

  int finished(int64_t h, std::string* err) override {
    auto it = fins.find(h);
    if (it == fins.end()) {
      *err = "unknown read-back";
      violate("unknown read-back handle");
      return MCRT_ERR_INVALID;
    }
    if (now < it->second.first) return 0;
    fins.erase(it);
    return 1;
  }

  bool holds(uint64_t job) const {
    if (live)
      for (const Task& t : tasks)
        if (t.job == job) return true;
    for (auto& f : fins)
      for (uint64_t id : f.second.second)
        if (id == job) return true;
    return false;
  }
};

struct Published {
  int count = 0;
  int final = -1;
  int err = 0;
  double t[4] = {0, 0, 0, 0};
};

struct Harness {
  Fake fake;
  mcrt::sched::Scheduler* sc = nullptr;
  std::map<uint64_t, Published> pub;
  size_t max_running = 0, max_in_use = 0;
};

}  // namespace

extern "C" {

void* sh_create(int blocks, int window_tasks, int conts, int max_active, double budget,
                double slice_ms, double iter_ms) {
  auto* H = new Harness();
  H->fake.H = H;
  H->fake.blocks = blocks;
  H->fake.iter_ms = iter_ms;
  mcrt::sched::Options o;
  o.blocks = blocks;
  o.window_tasks = window_tasks;
  o.conts = conts;
  o.max_active = (size_t)max_active;
  o.budget = (size_t)budget;
  o.slice_ms = slice_ms;
  H->sc = new mcrt::sched::Scheduler(&H->fake, o, [H](const mcrt::sched::Job& j, mcrt::sched::Final f) {
    Published& p = H->pub[j.id];
    ++p.count;
    if (p.count > 1) H->fake.violate("published twice: job " + std::to_string(j.id));
    if (H->fake.holds(j.id)) H->fake.violate("published while held: job " + std::to_string(j.id));
    auto it = H->fake.jobs.find(j.id);
    if (it != H->fake.jobs.end() && it->second.created && !it->second.destroyed)
      H->fake.violate("published before its device state was freed: job " + std::to_string(j.id));
    p.final = (int)f;
    p.err = j.err_code;
    p.t[0] = ms_of(j.t_submit);
    p.t[1] = j.admitted ? ms_of(j.t_admit) : ms_of(j.t_done);
    p.t[2] = j.started ? ms_of(j.t_start) : ms_of(j.t_done);
    p.t[3] = ms_of(j.t_done);
  });
  return H;
}

void sh_destroy(void* h) {
  auto* H = (Harness*)h;
  delete H->sc;
  delete H;
}

int sh_submit(void* h, uint64_t id, int mixed, double need, int64_t nchunks, double chunk_ms,
              int fail_create, double now) {
  auto* H = (Harness*)h;
  FakeJob& j = H->fake.jobs[id];
  j.nchunks = nchunks;
  j.chunk_ms = chunk_ms;
  j.fail_create = fail_create != 0;
  H->sc->submit(id, mixed != 0, (size_t)need, at(now));
  return 0;
}

void sh_cancel(void* h, uint64_t id) { ((Harness*)h)->sc->cancel(id); }
void sh_cancel_all(void* h) { ((Harness*)h)->sc->cancel_all(); }
void sh_fail_launches(void* h, int n) { ((Harness*)h)->fake.fail_launches = n; }
void sh_fail_windows(void* h, int n) { ((Harness*)h)->fake.fail_windows = n; }
void sh_set_slice(void* h, double ms) { ((Harness*)h)->sc->set_slice(ms); }

int sh_step(void* h, double now) {
  auto* H = (Harness*)h;
  H->fake.now = now;
  const bool p = H->sc->step(at(now));
  const auto s = H->sc->stats();
  H->max_running = std::max(H->max_running, s.running);
  H->max_in_use = std::max(H->max_in_use, s.in_use);
  return p ? 1 : 0;
}

int sh_idle(void* h) { return ((Harness*)h)->sc->idle() ? 1 : 0; }
int sh_window_live(void* h) { return ((Harness*)h)->fake.live ? 1 : 0; }

int sh_final(void* h, uint64_t id) {
  auto* H = (Harness*)h;
  auto it = H->pub.find(id);
  return it == H->pub.end() ? -1 : it->second.final;
}
int sh_publish_count(void* h, uint64_t id) {
  auto* H = (Harness*)h;
  auto it = H->pub.find(id);
  return it == H->pub.end() ? 0 : it->second.count;
}
int sh_err(void* h, uint64_t id) {
  auto* H = (Harness*)h;
  auto it = H->pub.find(id);
  return it == H->pub.end() ? 0 : it->second.err;
}
double sh_time(void* h, uint64_t id, int which) {
  auto* H = (Harness*)h;
  auto it = H->pub.find(id);
  return (it == H->pub.end() || which < 0 || which > 3) ? NAN : it->second.t[which];
}
// Was job id in the window in flight when the window was last launched?
int sh_in_window(void* h, uint64_t id) {
  auto* H = (Harness*)h;
  if (!H->fake.live) return 0;
  for (const Task& t : H->fake.tasks)
    if (t.job == id) return 1;
  return 0;
}

// submitted, completed, cancelled, failed, launches, tasks, resumed,
// preemptions, finishes, queued, running, in_use, max_running, max_in_use,
// violations
void sh_stats(void* h, double* out) {
  auto* H = (Harness*)h;
  const auto s = H->sc->stats();
  const double v[15] = {(double)s.submitted, (double)s.completed, (double)s.cancelled, (double)s.failed,
                        (double)s.launches, (double)s.tasks, (double)s.resumed, (double)s.preemptions,
                        (double)s.finishes, (double)s.queued, (double)s.running, (double)s.in_use,
                        (double)H->max_running, (double)H->max_in_use,
                        (double)H->fake.violations.size()};
  for (int i = 0; i < 15; ++i) out[i] = v[i];
}

const char* sh_violation(void* h, int i) {
  auto* H = (Harness*)h;
  return (i >= 0 && i < (int)H->fake.violations.size()) ? H->fake.violations[i].c_str() : "";
}

}  // extern "C"

#ifdef SCHED_HARNESS_MAIN
// A randomised run under the sanitizers: many small jobs, cancellations and
// injected faults; exits non-zero on any violated invariant.
#include <random>
int main() {
  void* h = sh_create(16, 32, 48, 64, 1e9, 2.0, 0.05);
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  double now = 0.0;
  uint64_t id = 1;
  for (int step = 0; step < 200000; ++step) {
    now += 0.01;
    if (U(rng) < 0.02 && id < 6000) {
      const int64_t nch = 1 + (int64_t)(U(rng) * 20);
      sh_submit(h, id, U(rng) < 0.5, 1e6 * (1 + U(rng)), nch, 0.05 + 5 * U(rng) * U(rng),
                U(rng) < 0.01, now);
      ++id;
    }
    if (U(rng) < 0.002) sh_cancel(h, 1 + (uint64_t)(U(rng) * (double)id));
    if (U(rng) < 0.0001) sh_fail_launches(h, 1 + (int)(U(rng) * 3));
    if (U(rng) < 0.00005) sh_fail_windows(h, 1);
    sh_step(h, now);
  }
  for (int k = 0; k < 2000000 && !sh_idle(h); ++k) sh_step(h, now += 0.01);
  double s[15];
  sh_stats(h, s);
  int bad = 0;
  for (uint64_t j = 1; j < id; ++j)
    if (sh_publish_count(h, j) != 1) ++bad;
  std::printf("submitted %.0f completed %.0f cancelled %.0f failed %.0f launches %.0f preemptions %.0f "
              "resumed %.0f max_running %.0f violations %.0f unpublished %d idle %d\n",
              s[0], s[1], s[2], s[3], s[4], s[7], s[6], s[12], s[14], bad, sh_idle(h));
  for (int i = 0; i < (int)s[14]; ++i) std::printf("violation: %s\n", sh_violation(h, i));
  const int rc = (s[14] == 0 && bad == 0 && sh_idle(h) && s[12] <= 64) ? 0 : 1;
  sh_destroy(h);
  return rc;
}
#endif
