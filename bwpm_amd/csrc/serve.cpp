// Request queue with continuous batching on one device (see serve.hpp).
#include "serve.hpp"

#include <cstdlib>

#include <algorithm>
#include <set>

namespace bwpm {

struct Queue::Lane final : LaneFeed {
  Queue* q = nullptr;
  Key key;
  hipStream_t stream{};
  std::deque<long long> ready;  // admitted, waiting for a slot
  std::set<long long> members;  // in the lane
  std::condition_variable cv;
  std::thread thread;

  LaneMember next() override {
    std::lock_guard<std::mutex> lock(q->mu_);
    if (ready.empty()) return {};
    const long long id = ready.front();
    ready.pop_front();
    members.insert(id);
    Ticket& t = *q->tickets_.at(id);
    return LaneMember{t.engine.get(), t.iterations, id};
  }
  void finished(const LaneMember& m, hipEvent_t done) override {
    std::lock_guard<std::mutex> lock(q->mu_);
    members.erase(m.tag);
    q->finished_.emplace_back(m.tag, done);
    q->cv_complete_.notify_one();
  }
  bool wait() override {
    std::unique_lock<std::mutex> lock(q->mu_);
    cv.wait(lock, [&] { return q->stop_ || !ready.empty(); });
    return !ready.empty();
  }
};

Queue::Queue(int device, double budget_bytes, int slots, int chunk)
    : device_(device), slots_(slots), chunk_(chunk) {
  require(slots >= 1 && slots <= 1024, "lane slots must be in 1 … 1024");
  require(chunk >= 1, "chunk must be >= 1");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) throw Error(kEdevice, "no GPU device available");
  require(device >= 0 && device < count, "device index out of range");
  HIP_CHECK(hipSetDevice(device));
  size_t free_b = 0, total_b = 0;
  HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
  budget_ = budget_bytes > 0 ? std::min(budget_bytes, static_cast<double>(total_b)) : 0.9 * static_cast<double>(free_b);
  for (int i = 0; i < kAdmitWorkers; ++i) admit_threads_.emplace_back([this] { admit_loop(); });
  complete_thread_ = std::thread([this] { complete_loop(); });
  // Born solves run one per worker thread and stream (BWPM_BORN_WORKERS, 1 … 16, overrides the count)
  int born_workers = kBornWorkers;
  if (const char* env = std::getenv("BWPM_BORN_WORKERS")) {
    const int v = std::atoi(env);
    if (v >= 1 && v <= 16) born_workers = v;
  }
  for (int i = 0; i < born_workers; ++i) born_threads_.emplace_back([this] { born_loop(); });
}

Queue::~Queue() {
  {
    std::lock_guard<std::mutex> lock(mu_);
    stop_ = true;
    for (auto& kv : lanes_) kv.second->cv.notify_all();
  }
  cv_admit_.notify_all();
  cv_born_.notify_all();
  for (auto& t : admit_threads_)
    if (t.joinable()) t.join();
  for (auto& kv : lanes_)
    if (kv.second->thread.joinable()) kv.second->thread.join();
  for (auto& t : born_threads_)
    if (t.joinable()) t.join();
  {
    std::lock_guard<std::mutex> lock(mu_);
    complete_stop_ = true;
  }
  cv_complete_.notify_all();
  if (complete_thread_.joinable()) complete_thread_.join();
  (void)hipSetDevice(device_);
  for (auto& kv : lanes_) (void)hipStreamDestroy(kv.second->stream);
}

long long Queue::submit(Spec spec, std::vector<c128> field, int iterations) {
  require(iterations >= 0, "iterations must be >= 0");
  require(field.size() == static_cast<size_t>(2 * spec.ny * spec.nx), "field must have shape (ny, nx, 2)");
  spec.device = device_;
  auto t = std::make_unique<Ticket>();
  t->spec = std::move(spec);
  t->field = std::move(field);
  t->iterations = iterations;
  const long long passes = iterations > 0 ? 2LL * iterations : 1LL;
  t->updates = t->spec.ny * t->spec.nx * t->spec.nz * passes;
  t->t_submit = Clock::now();
  std::lock_guard<std::mutex> lock(mu_);
  require(!stop_, "queue is closed");
  t->id = next_id_++;
  if (!any_submit_) {
    first_submit_ = t->t_submit;
    any_submit_ = true;
  }
  const long long id = t->id;
  pending_.push_back(id);
  tickets_.emplace(id, std::move(t));
  cv_admit_.notify_one();
  return id;
}

int Queue::state(long long ticket) {
  std::lock_guard<std::mutex> lock(mu_);
  auto it = tickets_.find(ticket);
  require(it != tickets_.end(), "unknown ticket");
  return it->second->state;
}

int Queue::wait(long long ticket, double timeout_s) {
  std::unique_lock<std::mutex> lock(mu_);
  auto it = tickets_.find(ticket);
  require(it != tickets_.end(), "unknown ticket");
  Ticket* t = it->second.get();
  auto ready = [&] { return t->state == kTicketDone || t->state == kTicketFailed; };
  if (timeout_s < 0)
    cv_done_.wait(lock, ready);
  else
    cv_done_.wait_for(lock, std::chrono::duration<double>(timeout_s), ready);
  return t->state;
}

RunOutput Queue::take(long long ticket) {
  std::lock_guard<std::mutex> lock(mu_);
  auto it = tickets_.find(ticket);
  require(it != tickets_.end(), "unknown ticket");
  Ticket& t = *it->second;
  if (t.state == kTicketFailed) {
    const Error err(t.code, t.error);
    tickets_.erase(it);
    throw err;
  }
  require(t.state == kTicketDone, "ticket has not finished");
  RunOutput out = std::move(t.out);
  tickets_.erase(it);
  return out;
}

QueueStats Queue::stats() {
  std::lock_guard<std::mutex> lock(mu_);
  QueueStats s;
  std::vector<double> v = latencies_ms_;
  if (!v.empty()) {
    std::sort(v.begin(), v.end());
    auto pct = [&](double p) {
      const double r = p * static_cast<double>(v.size() - 1);
      const size_t lo = static_cast<size_t>(r);
      const size_t hi = std::min(lo + 1, v.size() - 1);
      return v[lo] + (v[hi] - v[lo]) * (r - static_cast<double>(lo));
    };
    s.p50_ms = pct(0.50);
    s.p99_ms = pct(0.99);
    double sum = 0;
    for (double x : v) sum += x;
    s.mean_ms = sum / static_cast<double>(v.size());
  }
  if (completed_ > 0) {
    const double span = std::chrono::duration<double>(last_done_ - first_submit_).count();
    s.updates_per_s = span > 0 ? static_cast<double>(done_updates_) / span : 0;
  }
  s.completed = completed_;
  s.failed = failed_;
  s.pending = static_cast<long long>(pending_.size());
  s.running = running_;
  s.used_bytes = used_;
  s.budget_bytes = budget_;
  s.peak_bytes = peak_;
  return s;
}

void Queue::fail(Ticket& t, int code, const std::string& msg) {
  t.state = kTicketFailed;
  t.code = code;
  t.error = msg;
  t.t_done = Clock::now();
  ++failed_;
  cv_done_.notify_all();
}

void Queue::release(Ticket& t) {
  used_ -= static_cast<double>(t.bytes);
  t.bytes = 0;
  --running_;
  cv_admit_.notify_all();
}

Queue::Lane* Queue::lane_of(const Key& k) {
  auto it = lanes_.find(k);
  if (it != lanes_.end()) return it->second.get();
  auto lane = std::make_unique<Lane>();
  lane->q = this;
  lane->key = k;
  HIP_CHECK(hipStreamCreateWithFlags(&lane->stream, hipStreamNonBlocking));
  Lane* raw = lane.get();
  lanes_.emplace(k, std::move(lane));
  raw->thread = std::thread([this, raw] { lane_loop(raw); });
  return raw;
}

void Queue::admit_loop() {
  (void)hipSetDevice(device_);
  std::unique_lock<std::mutex> lock(mu_);
  for (;;) {
    cv_admit_.wait(lock, [&] { return stop_ || !pending_.empty(); });
    if (stop_) {
      for (long long id : pending_) fail(*tickets_.at(id), kEinval, "queue closed before the request was admitted");
      pending_.clear();
      return;
    }
    const long long id = pending_.front();
    Ticket& t = *tickets_.at(id);
    const size_t need = estimate_device_bytes(t.spec, t.iterations);
    if (used_ + static_cast<double>(need) > budget_) {
      if (running_ == 0) {
        pending_.pop_front();
        fail(t, kEnomem,
             "request needs about " + std::to_string(need) + " device bytes, more than the HBM budget of " +
                 std::to_string(static_cast<long long>(budget_)));
      } else {
        cv_admit_.wait(lock);  // a running solve releases its memory
      }
      continue;
    }
    pending_.pop_front();
    t.bytes = need;
    used_ += static_cast<double>(need);
    ++running_;
    lock.unlock();
    std::unique_ptr<EngineBase> e;
    int code = kOk;
    std::string msg;
    try {
      e = make_engine(std::move(t.spec));
      e->upload(t.field.data());
    } catch (const Error& err) {
      code = err.code;
      msg = err.what();
      e.reset();
    } catch (const std::exception& err) {
      code = kEinternal;
      msg = err.what();
      e.reset();
    }
    lock.lock();
    if (!e) {
      release(t);
      fail(t, code, msg);
      continue;
    }
    const double actual = static_cast<double>(e->device_bytes());
    used_ += actual - static_cast<double>(t.bytes);
    t.bytes = static_cast<size_t>(actual);
    peak_ = std::max(peak_, used_);
    t.engine = std::move(e);
    t.state = kTicketAdmitted;
    t.field = std::vector<c128>();  // the inputs now live on the device
    t.spec.labels = std::vector<int32_t>();
    t.spec.fluct = std::vector<c128>();
    if (t.engine->has_born()) {
      born_ready_.push_back(id);
      cv_born_.notify_one();
    } else {
      Lane* lane = lane_of(Key{t.engine->precision(), t.engine->ny(), t.engine->nx()});
      lane->ready.push_back(id);
      lane->cv.notify_one();
    }
  }
}

void Queue::lane_loop(Lane* lane) {
  (void)hipSetDevice(device_);
  try {
    run_continuous_lane(std::get<0>(lane->key), *lane, slots_, chunk_, lane->stream);
  } catch (const std::exception& err) {
    const Error* e = dynamic_cast<const Error*>(&err);
    std::lock_guard<std::mutex> lock(mu_);
    for (long long id : lane->members) {
      Ticket& t = *tickets_.at(id);
      t.engine.reset();
      release(t);
      fail(t, e ? e->code : kEinternal, err.what());
    }
    lane->members.clear();
    for (long long id : lane->ready) {
      Ticket& t = *tickets_.at(id);
      t.engine.reset();
      release(t);
      fail(t, e ? e->code : kEinternal, err.what());
    }
    lane->ready.clear();
  }
}

// Result of a finished solve to the host, its device memory released, its latency recorded.
void Queue::finish(Ticket& t, std::unique_lock<std::mutex>& lock) {
  RunOutput out;
  int code = kOk;
  std::string msg;
  lock.unlock();
  try {
    t.engine->download(out, t.iterations, {});
  } catch (const Error& err) {
    code = err.code;
    msg = err.what();
  } catch (const std::exception& err) {
    code = kEinternal;
    msg = err.what();
  }
  t.engine.reset();
  lock.lock();
  release(t);
  if (code != kOk) {
    fail(t, code, msg);
    return;
  }
  t.out = std::move(out);
  t.state = kTicketDone;
  t.t_done = Clock::now();
  latencies_ms_.push_back(std::chrono::duration<double, std::milli>(t.t_done - t.t_submit).count());
  ++completed_;
  done_updates_ += t.updates;
  last_done_ = t.t_done;
  cv_done_.notify_all();
}

void Queue::complete_loop() {
  (void)hipSetDevice(device_);
  std::unique_lock<std::mutex> lock(mu_);
  for (;;) {
    cv_complete_.wait(lock, [&] { return complete_stop_ || !finished_.empty(); });
    if (finished_.empty()) return;
    const auto [id, ev] = finished_.front();
    finished_.pop_front();
    lock.unlock();
    const hipError_t st = hipEventSynchronize(ev);
    (void)hipEventDestroy(ev);
    lock.lock();
    Ticket& t = *tickets_.at(id);
    if (st != hipSuccess) {
      lock.unlock();
      t.engine.reset();
      lock.lock();
      release(t);
      fail(t, kEdevice, std::string("device error in a lane: ") + hipGetErrorString(st));
      continue;
    }
    finish(t, lock);
  }
}

void Queue::born_loop() {
  (void)hipSetDevice(device_);
  std::unique_lock<std::mutex> lock(mu_);
  for (;;) {
    cv_born_.wait(lock, [&] { return stop_ || !born_ready_.empty(); });
    if (born_ready_.empty()) return;
    const long long id = born_ready_.front();
    born_ready_.pop_front();
    Ticket& t = *tickets_.at(id);
    lock.unlock();  // each engine runs on its own stream
    int code = kOk;
    std::string msg;
    try {
      t.engine->run_resident(t.iterations, false, nullptr);
    } catch (const Error& err) {
      code = err.code;
      msg = err.what();
    } catch (const std::exception& err) {
      code = kEinternal;
      msg = err.what();
    }
    lock.lock();
    if (code != kOk) {
      lock.unlock();
      t.engine.reset();
      lock.lock();
      release(t);
      fail(t, code, msg);
      continue;
    }
    finish(t, lock);
  }
}

}  // namespace bwpm
