// fp_serve.hip -- the serving worker of a context (fp_submit / fp_poll).
//
// Continuous batching: requests that arrive while others are on the GPU join
// the next round; every round advances each active projection by one stage
// (Projector::serve_round), so new work overlaps running work on the device
// and a grid's result does not depend on what it shares rounds with.
#include "fp_tall.hpp"
#include "fp_uniform.hpp"

#include <chrono>
#include <cstring>

namespace fp {

static double wall_seconds() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

Server::Server(hipStream_t s, int device, Projector* proj, TallProjector* tall, std::mutex* gpu_mutex)
    : stream_(s), device_(device), proj_(proj), tall_(tall), gpu_mutex_(gpu_mutex) {
  thread_ = std::thread(&Server::loop, this);
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  work_cv_.notify_all();
  if (thread_.joinable()) thread_.join();
}

int Server::submit(UniformGrid* g, TallGrid* t, const fp_inputs* in, const fp_outputs* out, const fp_options* opt,
                   uint64_t* ticket) {
  std::unique_ptr<Job> j(new Job());
  j->grid = g;
  j->tall = t;
  j->in = *in;
  j->out = *out;
  j->opt = *opt;
  j->t_submit = wall_seconds();
  const void* key = g ? (const void*)g : (const void*)t;
  {
    std::lock_guard<std::mutex> lk(mu_);
    // one projection in flight per grid: jobs share the grid's buffers
    if (jobs_on_locked(key) > 0) {
      set_error("fp_submit: the grid has a submitted projection in flight (fp_poll it first)");
      return FP_ERR_STATE;
    }
    if (in->stream_order) {
      // the worker's launches for this job wait for the caller's queued work
      (void)hipSetDevice(device_);
      hipError_t e = hipEventCreateWithFlags(&j->ready, hipEventDisableTiming);
      if (e == hipSuccess) e = hipEventRecord(j->ready, (hipStream_t)in->stream);
      if (e != hipSuccess) {
        if (j->ready) (void)hipEventDestroy(j->ready);
        set_error("fp_submit: ordering after the caller's stream: %s", hipGetErrorString(e));
        return FP_ERR_DEVICE;
      }
    }
    j->ticket = next_ticket_++;
    *ticket = j->ticket;
    incoming_.push_back(std::move(j));
  }
  work_cv_.notify_one();
  return FP_OK;
}

int Server::jobs_on(const void* g) {
  std::lock_guard<std::mutex> lk(mu_);
  return jobs_on_locked(g);
}

int Server::jobs_on_locked(const void* g) {
  int n = 0;
  for (const auto& j : incoming_) n += (const void*)j->grid == g || (const void*)j->tall == g;
  for (const auto& j : active_) n += (const void*)j->grid == g || (const void*)j->tall == g;
  return n;
}

int Server::pending() {
  std::lock_guard<std::mutex> lk(mu_);
  return (int)(incoming_.size() + active_.size());
}

int Server::poll(uint64_t ticket, int timeout_ms, fp_report* rep, double* latency_ms) {
  std::unique_lock<std::mutex> lk(mu_);
  auto ready = [&] { return done_.count(ticket) != 0; };
  if (!ready()) {
    if (timeout_ms != 0) {
      if (timeout_ms < 0) done_cv_.wait(lk, ready);
      else done_cv_.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready);
    }
    if (!ready()) {
      bool known = false;
      for (const auto& j : incoming_) known |= j->ticket == ticket;
      for (const auto& j : active_) known |= j->ticket == ticket;
      if (!known) {
        set_error("fp_poll: unknown ticket %llu", (unsigned long long)ticket);
        return FP_ERR_INVALID_ARGUMENT;
      }
      return 1;
    }
  }
  auto it = done_.find(ticket);
  std::unique_ptr<Job> j = std::move(it->second);
  done_.erase(it);
  if (rep) *rep = j->report;
  if (latency_ms) *latency_ms = (j->t_done - j->t_submit) * 1e3;
  return j->rc;
}

void Server::loop() {
  (void)hipSetDevice(device_);
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(mu_);
      work_cv_.wait(lk, [&] { return stop_ || !incoming_.empty() || !active_.empty(); });
      if (stop_) return;
      while (!incoming_.empty()) {
        Job* j = incoming_.front().get();
        if (j->ready) (void)hipStreamWaitEvent(stream_, j->ready, 0);
        active_.push_back(std::move(incoming_.front()));
        incoming_.pop_front();
      }
    }
    std::vector<Job*> jobs, tjobs;
    for (auto& j : active_) (j->tall ? tjobs : jobs).push_back(j.get());
    int rc = FP_OK;
    {
      // uniform and tall-cell jobs advance in the same round, each kind in
      // its own batched launch set
      std::lock_guard<std::mutex> g(*gpu_mutex_);
      if (!jobs.empty()) rc = proj_->serve_round(stream_, jobs);
      if (rc == FP_OK && !tjobs.empty()) rc = tall_->serve_round(stream_, tjobs);
    }
    {
      std::lock_guard<std::mutex> lk(mu_);
      for (auto& j : active_) {
        if (rc != FP_OK && j->stage != Job::DONE) {
          j->rc = rc;
          j->t_done = wall_seconds();
          j->stage = Job::DONE;
        }
      }
      for (auto it = active_.begin(); it != active_.end();) {
        if ((*it)->stage == Job::DONE) {
          if ((*it)->ready) {
            (void)hipEventDestroy((*it)->ready);
            (*it)->ready = nullptr;
          }
          const uint64_t t = (*it)->ticket;
          done_[t] = std::move(*it);
          it = active_.erase(it);
        } else {
          ++it;
        }
      }
    }
    done_cv_.notify_all();
  }
}

}  // namespace fp
