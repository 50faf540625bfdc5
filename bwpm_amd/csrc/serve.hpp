// Request queue with continuous batching on one device.
//
// Requests (a solve spec, its input field and iteration count) arrive at any time. Admission
// workers take them in arrival order while their device footprint fits the HBM budget, build the
// engine (host set-up of boundaries and tables) and upload the field; several workers build
// engines concurrently. Admitted solves of one (precision, grid) join a continuous
// lane (run_continuous_lane) that steps all its members together, one column and one row launch
// per slab step; a member joins a free slot between chunks of slab steps and leaves when its
// program ends, so requests that arrive while others are on the GPU overlap with them. Lanes of
// different grids run concurrently on their own streams; solves with index fluctuations run one per
// worker on a few workers of their own (their GMRES synchronises with the host every iteration, so
// several overlap on the device). A completion thread waits for each finished solve, copies its
// result to the host, frees its device memory and records its latency (submit → result on host).
#pragma once

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "engine.hpp"

namespace bwpm {

enum TicketState : int { kTicketPending = 0, kTicketAdmitted = 1, kTicketDone = 2, kTicketFailed = 3 };

struct QueueStats {
  double p50_ms = 0, p99_ms = 0, mean_ms = 0;  // latency of completed requests
  double updates_per_s = 0;                     // grid-point·slab updates of completed requests / span
  long long completed = 0, failed = 0, pending = 0, running = 0;
  double used_bytes = 0, budget_bytes = 0, peak_bytes = 0;
};

class Queue {
 public:
  // budget_bytes <= 0: 90 % of the device memory free at creation. slots: solves per lane.
  Queue(int device, double budget_bytes, int slots, int chunk);
  ~Queue();
  Queue(const Queue&) = delete;
  Queue& operator=(const Queue&) = delete;

  long long submit(Spec spec, std::vector<c128> field, int iterations);
  int state(long long ticket);
  // Waits until the ticket is done or failed, or the timeout (seconds, < 0: none) passes.
  int wait(long long ticket, double timeout_s);
  // Result of a done ticket (the ticket is forgotten); a failed ticket throws its error.
  RunOutput take(long long ticket);
  QueueStats stats();

 private:
  using Clock = std::chrono::steady_clock;
  struct Ticket {
    long long id = 0;
    Spec spec;
    std::vector<c128> field;
    int iterations = 0;
    long long updates = 0;
    size_t bytes = 0;  // reserved device bytes
    std::unique_ptr<EngineBase> engine;
    RunOutput out;
    int state = kTicketPending;
    int code = kOk;
    std::string error;
    Clock::time_point t_submit, t_done;
  };
  using Key = std::tuple<int, long long, long long>;  // precision, ny, nx
  struct Lane;

  void admit_loop();
  void complete_loop();
  void born_loop();
  void lane_loop(Lane* lane);
  void finish(Ticket& t, std::unique_lock<std::mutex>& lock);  // lock held on entry and exit
  void fail(Ticket& t, int code, const std::string& msg);  // caller holds mu_
  void release(Ticket& t);                                  // caller holds mu_
  Lane* lane_of(const Key& k);                              // caller holds mu_

  int device_;
  int slots_, chunk_;
  double budget_ = 0, used_ = 0, peak_ = 0;
  long long next_id_ = 1;
  bool stop_ = false;           // no new requests; lanes end when empty
  bool complete_stop_ = false;  // every producer of finished solves has ended

  std::mutex mu_;
  std::condition_variable cv_admit_, cv_done_, cv_complete_, cv_born_;
  std::map<long long, std::unique_ptr<Ticket>> tickets_;
  std::deque<long long> pending_;                      // arrival order
  std::deque<std::pair<long long, hipEvent_t>> finished_;
  std::deque<long long> born_ready_;
  std::map<Key, std::unique_ptr<Lane>> lanes_;
  std::vector<double> latencies_ms_;
  long long completed_ = 0, failed_ = 0, running_ = 0;
  long long done_updates_ = 0;
  Clock::time_point first_submit_{}, last_done_{};
  bool any_submit_ = false;

  static constexpr int kAdmitWorkers = 4;
  static constexpr int kBornWorkers = 8;
  std::thread complete_thread_;
  std::vector<std::thread> admit_threads_, born_threads_;
};

}  // namespace bwpm
