// mcrt_sched.cpp — the request scheduler of the server (host only).
#include "mcrt_sched.h"

#include <algorithm>

#include "../../include/mcrt.h"

namespace mcrt {
namespace sched {

Scheduler::Scheduler(Backend* be, const Options& o, Publish pub) : be_(be), o_(o), pub_(std::move(pub)) {
  o_.blocks = std::max(1, o_.blocks);
  o_.window_tasks = std::max(1, o_.window_tasks);
  o_.conts = std::max(1, o_.conts);
  for (int c = o_.conts - 1; c >= 0; --c) free_conts_.push_back(c);
}

void Scheduler::submit(uint64_t id, bool mixed, size_t need, Time now) {
  Job* j = new Job();
  j->id = id;
  j->mixed = mixed;
  j->need = need;
  j->t_submit = now;
  jobs_[id] = j;
  queue_.push_back(j);
  ++st_.submitted;
}

void Scheduler::cancel(uint64_t id) {
  auto it = jobs_.find(id);
  if (it == jobs_.end()) return;
  Job& j = *it->second;
  if (j.done) return;
  j.cancel = true;
  drop_waiting(j);
}

void Scheduler::cancel_all() {
  for (auto& kv : jobs_) cancel(kv.first);
}

int Scheduler::take_cont() {
  if (free_conts_.empty()) return -1;
  const int c = free_conts_.back();
  free_conts_.pop_back();
  return c;
}

void Scheduler::give_cont(int c) {
  if (c >= 0) free_conts_.push_back(c);
}

// A job that launches nothing more gives back the blocks of its stopped chunks.
void Scheduler::drop_waiting(Job& j) {
  for (const Task& t : j.suspended) give_cont(t.cont);
  j.suspended.clear();
  j.returned.clear();
  j.next_chunk = j.nchunks;
}

void Scheduler::fail_job(Job& j, int code, const std::string& msg) {
  if (!j.err_code) {
    j.err_code = code;
    j.err = msg;
  }
  drop_waiting(j);
}

// Queued jobs in order while their device memory fits the budget and fewer
// than max_active are admitted.  An admitted job holds device and pinned
// memory; under overload, admitting everything the budget allows (tens of
// thousands of small solves) only multiplies live allocations beyond what the
// next windows can use, so the rest wait in the queue.
bool Scheduler::admit(Time now) {
  bool any = false;
  while (!queue_.empty()) {
    Job* j = queue_.front();
    if (j->cancel) {
      queue_.pop_front();
      any = true;
      j->t_done = now;
      ++st_.cancelled;
      jobs_.erase(j->id);
      pub_(*j, Final::Cancelled);
      delete j;
      continue;
    }
    if (j->need > o_.budget) {
      queue_.pop_front();
      any = true;
      j->err_code = MCRT_ERR_NOMEM;
      j->err = "the solve needs more device memory than the server's budget";
      j->t_done = now;
      ++st_.failed;
      jobs_.erase(j->id);
      pub_(*j, Final::Failed);
      delete j;
      continue;
    }
    if (st_.in_use + j->need > o_.budget || active_.size() >= o_.max_active) break;
    queue_.pop_front();
    any = true;
    st_.in_use += j->need;
    j->admitted = true;
    j->t_admit = now;
    std::string err;
    const int rc = be_->create(j->id, &j->nchunks, &j->cost, &err);
    j->created = true;  // destroy() frees what create() took, also on failure
    if (rc) fail_job(*j, rc, err);
    active_.push_back(j);
  }
  return any;
}

// Outcomes of the window and of the finishes in flight.
bool Scheduler::complete(Time now) {
  bool any = false;
  if (win_.live) {
    std::vector<int64_t> states;
    Time start = now;
    std::string err;
    const int r = be_->poll(&states, &start, &err);
    if (r != 0) {
      any = true;
      win_.live = false;
      if (r > 0 && states.size() != win_.tasks.size()) {
        err = "window outcome count mismatch";
        states.clear();
      }
      for (size_t i = 0; i < win_.tasks.size(); ++i) {
        const Task& t = win_.tasks[i];
        Job& j = *jobs_.at(t.job);
        --j.held;
        const int64_t s = (r > 0 && !states.empty()) ? states[i] : kUnclaimed;
        if (r < 0 || states.empty()) {
          give_cont(t.cont);
          fail_job(j, r < 0 ? r : MCRT_ERR_HIP, err);
          continue;
        }
        if (s != kUnclaimed && !j.started) {
          j.started = true;
          j.t_start = start;
        }
        if (j.cancel || j.err_code) {
          give_cont(t.cont);
          continue;
        }
        if (s == kFinished) {
          ++j.done_chunks;
          give_cont(t.cont);
        } else if (s >= 0) {
          Task n = t;
          n.rb = s;
          n.resume = true;
          j.suspended.push_back(n);
        } else if (t.resume) {
          j.suspended.push_back(t);
        } else {
          j.returned.push_back(t.chunk);
          give_cont(t.cont);
        }
      }
      win_.tasks.clear();
    }
  }
  for (auto it = fin_.begin(); it != fin_.end();) {
    std::string err;
    const int r = be_->finished(it->handle, &err);
    if (r == 0) {
      ++it;
      continue;
    }
    any = true;
    for (uint64_t id : it->jobs) {
      Job& j = *jobs_.at(id);
      --j.held;
      if (r > 0) j.done = true;
      else fail_job(j, r, err);
    }
    it = fin_.erase(it);
  }
  return any;
}

// Jobs whose chunks all finished (or that have none): their reduction and
// read-back, enqueued before the next window.
bool Scheduler::start_finishes() {
  std::vector<uint64_t> fin;
  for (Job* j : active_)
    if (!j->cancel && !j->err_code && j->created && j->held == 0 && !j->done &&
        j->done_chunks == j->nchunks)
      fin.push_back(j->id);
  if (fin.empty()) return false;
  std::string err;
  const int64_t h = be_->finish(fin, &err);
  if (h < 0) {
    for (uint64_t id : fin) fail_job(*jobs_.at(id), (int)h, err);
  } else {
    for (uint64_t id : fin) ++jobs_.at(id)->held;
    fin_.push_back({h, fin});
    ++st_.finishes;
  }
  return true;
}

// Jobs that nothing in flight holds and that launch nothing more: free their
// device state, then publish (a job is published once, and never while a
// window or a finish still holds it: its owner may free it).
bool Scheduler::sweep(Time now) {
  bool any = false;
  for (auto it = active_.begin(); it != active_.end();) {
    Job* j = *it;
    const bool over = j->held == 0 && (j->done || j->cancel || j->err_code);
    if (!over) {
      ++it;
      continue;
    }
    any = true;
    it = active_.erase(it);
    drop_waiting(*j);
    if (j->created) be_->destroy(j->id);
    st_.in_use -= j->need;
    j->t_done = now;
    Final f;
    if (j->done && !j->err_code) {
      f = Final::Done;
      ++st_.completed;
    } else if (j->err_code) {
      f = Final::Failed;
      ++st_.failed;
    } else {
      f = Final::Cancelled;
      ++st_.cancelled;
    }
    jobs_.erase(j->id);
    pub_(*j, f);
    delete j;
  }
  return any;
}

// The next window: one precision (alternating when both wait), tasks taken
// round-robin over the admitted jobs in admission order, each job giving its
// stopped chunks first, then chunks never claimed, then new ones (which need
// a free continuation block).
bool Scheduler::launch(Time now) {
  if (win_.live) return false;
  bool any[2] = {false, false};
  for (Job* j : active_)
    if (j->waiting()) any[j->mixed ? 1 : 0] = true;
  if (!any[0] && !any[1]) return false;
  const bool mixed = any[1] && (!any[0] || !last_mixed_);
  std::vector<Task> tasks;
  std::vector<Job*> el;
  for (Job* j : active_)
    if (j->mixed == mixed && j->waiting()) el.push_back(j);
  bool conts_left = true;
  for (bool more = true; more && (int)tasks.size() < o_.window_tasks;) {
    more = false;
    for (Job* j : el) {
      if ((int)tasks.size() >= o_.window_tasks) break;
      Task t;
      t.job = j->id;
      if (!j->suspended.empty()) {
        t = j->suspended.front();
        j->suspended.pop_front();
      } else if (conts_left && (!j->returned.empty() || j->next_chunk < j->nchunks)) {
        const int c = take_cont();
        if (c < 0) {
          conts_left = false;
          continue;
        }
        if (!j->returned.empty()) {
          t.chunk = j->returned.back();
          j->returned.pop_back();
        } else {
          t.chunk = j->next_chunk++;
        }
        t.cont = c;
        t.rb = -1;
        t.resume = false;
      } else {
        continue;
      }
      tasks.push_back(t);
      ++j->held;
      more = true;
    }
  }
  if (tasks.empty()) return false;
  std::string err;
  const int rc = be_->launch(mixed, tasks, &err);
  if (rc) {
    // nothing was enqueued: the tasks are back with their jobs, which fail
    for (const Task& t : tasks) {
      Job& j = *jobs_.at(t.job);
      --j.held;
      if (t.resume) j.suspended.push_back(t);
      else {
        j.returned.push_back(t.chunk);
        give_cont(t.cont);
      }
    }
    for (const Task& t : tasks) fail_job(*jobs_.at(t.job), rc, err);
    return true;
  }
  for (const Task& t : tasks)
    if (t.resume) ++st_.resumed;
  win_.live = true;
  win_.mixed = mixed;
  win_.preempted = false;
  win_.tasks = std::move(tasks);
  win_.t_launch = now;
  last_mixed_ = mixed;
  ++st_.launches;
  st_.tasks += win_.tasks.size();
  return true;
}

void Scheduler::maybe_preempt(Time now) {
  if (!win_.live || win_.preempted) return;
  bool waiting = false;
  for (Job* j : active_)
    if (j->waiting()) {
      waiting = true;
      break;
    }
  if (!waiting) return;
  const double ms = std::chrono::duration<double, std::milli>(now - win_.t_launch).count();
  if (ms >= o_.slice_ms || (int)win_.tasks.size() < o_.blocks) {
    be_->preempt();
    win_.preempted = true;
    ++st_.preemptions;
  }
}

bool Scheduler::step(Time now) {
  bool p = false;
  p |= admit(now);
  p |= complete(now);
  p |= start_finishes();
  p |= sweep(now);
  p |= launch(now);
  maybe_preempt(now);
  const bool busy = in_flight();
  if (busy && !was_busy_) t_busy_ = now;
  if (!busy && was_busy_) st_.busy_ms += std::chrono::duration<double, std::milli>(now - t_busy_).count();
  was_busy_ = busy;
  return p;
}

Stats Scheduler::stats() const {
  Stats s = st_;
  s.queued = queue_.size();
  s.running = active_.size();
  return s;
}

}  // namespace sched
}  // namespace mcrt
