// Host-side helpers shared by the engine and the analysis module: device and pinned buffers,
// uploads, and FFT plans with their twiddle tables.
#pragma once

#include <cmath>
#include <string>
#include <vector>

#include "common.hpp"
#include "fft.hpp"

namespace bwpm {
namespace detail {

// ---- device memory -------------------------------------------------------------------------

class DevBuf {
 public:
  DevBuf() = default;
  explicit DevBuf(size_t bytes) { alloc(bytes); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    release();
    p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0;
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t bytes) {
    release();
    if (bytes == 0) return;
    const hipError_t e = hipMalloc(&p_, bytes);
    if (e != hipSuccess) {
      p_ = nullptr;
      throw Error(kEnomem, "device allocation of " + std::to_string(bytes) + " bytes failed: " +
                               hipGetErrorString(e));
    }
    n_ = bytes;
  }
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  template <typename U>
  U* as() const { return static_cast<U*>(p_); }
  size_t bytes() const { return n_; }

 private:
  void* p_ = nullptr;
  size_t n_ = 0;
};

// Page-locked host memory (allocated pinned; existing host memory is never registered).
class PinnedBuf {
 public:
  explicit PinnedBuf(size_t bytes) {
    if (bytes == 0) return;
    const hipError_t e = hipHostMalloc(&p_, bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
      p_ = nullptr;
      throw Error(kEnomem, "pinned host allocation of " + std::to_string(bytes) + " bytes failed: " +
                               hipGetErrorString(e));
    }
  }
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
  PinnedBuf(PinnedBuf&& o) noexcept : p_(o.p_) { o.p_ = nullptr; }
  ~PinnedBuf() {
    if (p_) (void)hipHostFree(p_);
  }
  template <typename U>
  U* as() const { return static_cast<U*>(p_); }

 private:
  void* p_ = nullptr;
};

// The copy has completed when this returns. Callers pass short-lived host vectors in pageable
// memory, freed (and their blocks reused) as soon as the set-up step that built them returns; an
// asynchronous copy still in flight would then move whatever the allocator put there.
template <typename U>
inline DevBuf upload_vec(const std::vector<U>& v, hipStream_t s) {
  DevBuf b(v.size() * sizeof(U));
  if (!v.empty()) {
    HIP_CHECK(hipMemcpyAsync(b.as<void>(), v.data(), v.size() * sizeof(U), hipMemcpyHostToDevice, s));
    HIP_CHECK(hipStreamSynchronize(s));
  }
  return b;
}

// ---- FFT plans ---------------------------------------------------------------------------------

// Radices of the plan of an n-point line into plan.n, plan.nst, plan.rad (no allocation).
inline void plan_radices(long long n, FftPlan& plan) {
  plan.n = static_cast<int>(n);
  plan.nst = 0;
  long long m = n;
  for (int R : {8, 4, 2, 3, 5, 7, 11, 13})
    while (m % R == 0) {
      if (plan.nst == kMaxStages) throw Error(kEinval, "grid size " + std::to_string(n) + " needs too many FFT stages");
      plan.rad[plan.nst++] = R;
      m /= R;
    }
  if (m != 1)
    throw Error(kEinval, "grid size " + std::to_string(n) +
                             " has a prime factor above 13; supported sizes are 2^a·3^b·5^c·7^d·11^e·13^f");
}

inline std::vector<int> factorize(long long n) {
  FftPlan plan{};
  plan_radices(n, plan);
  return std::vector<int>(plan.rad, plan.rad + plan.nst);
}

// The twiddle W_n^t = exp(−2πi t/n), from the mathematics in extended precision, then rounded to
// FP64 (c, s).
inline void twiddle_value(long long t, long long n, double& c, double& s) {
  const long double a = -2.0L * 3.141592653589793238462643383279502884L * static_cast<long double>(t) /
                        static_cast<long double>(n);
  c = static_cast<double>(cosl(a));
  s = static_cast<double>(sinl(a));
}

template <typename T>
struct Twiddles {
  DevBuf hi, lo;
  FftPlan plan{};
  void build(long long n, hipStream_t s) {
    plan.n = static_cast<int>(n);
    const std::vector<int> r = n > 1 ? factorize(n) : std::vector<int>{};
    plan.nst = static_cast<int>(r.size());
    for (size_t i = 0; i < r.size(); ++i) plan.rad[i] = r[i];
    std::vector<cx<T>> h(n), l(n);
    for (long long t = 0; t < n; ++t) {
      double c = 0, sn = 0;
      twiddle_value(t, n, c, sn);
      const T ch = static_cast<T>(c), sh = static_cast<T>(sn);
      h[t] = cx<T>{ch, sh};
      l[t] = cx<T>{static_cast<T>(c - static_cast<double>(ch)), static_cast<T>(sn - static_cast<double>(sh))};
    }
    hi = upload_vec(h, s);
    lo = upload_vec(l, s);
    plan.tw_hi = hi.as<void>();
    plan.tw_lo = lo.as<void>();
  }
};

}  // namespace detail
}  // namespace bwpm
