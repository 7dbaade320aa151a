// mcrt_sched.h — the request scheduler of the server, free of device code.
//
// A job is one submitted solve, cut into chunks (ray ranges).  Chunks run in
// launch windows: one persistent launch at a time takes a list of tasks (a
// chunk that starts, or a chunk that resumes from the ray-slot state a previous
// window left in its continuation block) that the device's workgroups claim in
// order.  While other work waits, the running window is asked to stop after
// `slice_ms` (at once when it holds fewer tasks than there are workgroups, so
// that part of the device is idle): every workgroup then finishes the
// iteration it is in, keeps the state of its chunk's rays in the chunk's
// continuation block and exits, and the next window resumes those chunks
// first-come in round-robin with the waiting ones.  A chunk's rays iterate
// together with no other chunk's, and a resumed chunk continues from exactly
// where it stopped, so a solve's sums do not depend on where the windows were
// cut beyond the order of FP64 additions.
//
// The scheduler is single-threaded: the server's worker calls step(); the
// device side sits behind Backend, which a test replaces with a simulation.
#pragma once

#include <chrono>
#include <cstdint>
#include <deque>
#include <functional>
#include <map>
#include <string>
#include <vector>

namespace mcrt {
namespace sched {

using Clock = std::chrono::steady_clock;
using Time = Clock::time_point;

// Outcome of a task in a finished window (else: the first ray not yet
// started, ≥ 0, of a chunk that stopped with rays in flight).
constexpr int64_t kUnclaimed = -1;
constexpr int64_t kFinished = -2;

struct Task {
  uint64_t job = 0;
  int64_t chunk = 0;
  int64_t rb = -1;     // resume point (-1: the chunk starts from its first ray)
  int cont = -1;       // continuation block
  bool resume = false;  // the block holds the chunk's rays in flight
};

// The device side.  Every call is made from the scheduler's thread.
class Backend {
 public:
  virtual ~Backend() = default;
  // Device state of a job: *nchunks chunks of relative work *cost each.
  // MCRT_OK, or an MCRT_ERR_* code with a message.
  virtual int create(uint64_t job, int64_t* nchunks, double* cost, std::string* err) = 0;
  // Free a job's device state (nothing in flight holds it).
  virtual void destroy(uint64_t job) = 0;
  // Enqueue a window of tasks of one precision.  Non-zero: an MCRT_ERR_*
  // code, and nothing was enqueued.
  virtual int launch(bool mixed, const std::vector<Task>& tasks, std::string* err) = 0;
  // The window in flight: 0 running; 1 finished, with one outcome per task in
  // *states and the host time its kernel started in *start; < 0 (an MCRT_ERR_*
  // code) the device failed it.
  virtual int poll(std::vector<int64_t>* states, Time* start, std::string* err) = 0;
  // Ask the window in flight to stop at the next iteration boundary.
  virtual void preempt() = 0;
  // Enqueue the reduction and read-back of jobs whose chunks all finished;
  // returns a handle for finished(), or < 0 (MCRT_ERR_*: nothing enqueued).
  virtual int64_t finish(const std::vector<uint64_t>& jobs, std::string* err) = 0;
  // 0 running, 1 done, < 0 failed (MCRT_ERR_*).
  virtual int finished(int64_t handle, std::string* err) = 0;
};

struct Options {
  int blocks = 256;           // workgroups of a window
  int window_tasks = 512;     // tasks per window at most
  int conts = 768;            // continuation blocks
  size_t max_active = 1024;   // admitted jobs at a time
  size_t budget = 0;          // device bytes of admitted jobs at a time
  double slice_ms = 8.0;      // time a window keeps the device while work waits
};

enum class Final { Done, Cancelled, Failed };

struct Job {
  uint64_t id = 0;
  bool mixed = false;
  size_t need = 0;
  bool admitted = false, created = false;
  int64_t nchunks = 0;
  double cost = 0.0;
  int64_t next_chunk = 0;        // chunks never launched: [next_chunk, nchunks)
  std::vector<int64_t> returned;  // launched but never claimed
  std::deque<Task> suspended;     // stopped with rays in flight (resume first)
  int64_t done_chunks = 0;
  int held = 0;                   // tasks in the window in flight, plus a finish
  bool cancel = false;
  bool done = false;              // reduction and read-back completed
  int err_code = 0;
  std::string err;
  Time t_submit{}, t_admit{}, t_start{}, t_done{};
  bool started = false;
  bool waiting() const {
    return !cancel && !err_code && (!suspended.empty() || !returned.empty() || next_chunk < nchunks);
  }
};

struct Stats {
  uint64_t submitted = 0, completed = 0, cancelled = 0, failed = 0;
  uint64_t launches = 0, tasks = 0, resumed = 0, preemptions = 0;
  uint64_t finishes = 0;
  size_t queued = 0, running = 0, in_use = 0;
  double busy_ms = 0.0;  // time with a window or a finish in flight
};

// Called once per job when it ends, after its device state was freed and
// nothing in flight holds it.
using Publish = std::function<void(const Job&, Final)>;

class Scheduler {
 public:
  Scheduler(Backend* be, const Options& o, Publish pub);
  // Queue a job (its id is the caller's ticket).
  void submit(uint64_t id, bool mixed, size_t need, Time now);
  // Stop a job: nothing more of it launches; it ends Cancelled once nothing
  // in flight holds it (Done if its last chunk already finished).
  void cancel(uint64_t id);
  void set_slice(double ms) { o_.slice_ms = ms; }
  // One pass: admission, completions, publication, launch, preemption.
  // Returns true when anything changed.
  bool step(Time now);
  bool in_flight() const { return win_.live || !fin_.empty(); }
  bool preempt_pending() const { return win_.live && win_.preempted; }
  bool idle() const { return queue_.empty() && active_.empty() && !in_flight(); }
  // Every job, queued or admitted, ends Cancelled (what is in flight is
  // waited for by further steps).
  void cancel_all();
  Stats stats() const;
  const Options& options() const { return o_; }

 private:
  struct Window {
    bool live = false, mixed = false, preempted = false;
    std::vector<Task> tasks;
    Time t_launch{};
  };
  struct Fin {
    int64_t handle;
    std::vector<uint64_t> jobs;
  };
  bool admit(Time now);
  bool complete(Time now);
  bool start_finishes();
  bool sweep(Time now);
  bool launch(Time now);
  void maybe_preempt(Time now);
  void fail_job(Job& j, int code, const std::string& msg);
  void drop_waiting(Job& j);
  int take_cont();
  void give_cont(int c);

  Backend* be_;
  Options o_;
  Publish pub_;
  std::deque<Job*> queue_;
  std::vector<Job*> active_;  // admitted, in admission order
  std::map<uint64_t, Job*> jobs_;
  std::vector<int> free_conts_;
  Window win_;
  std::vector<Fin> fin_;
  bool last_mixed_ = false;
  Stats st_;
  Time t_busy_{};
  bool was_busy_ = false;
};

}  // namespace sched
}  // namespace mcrt
