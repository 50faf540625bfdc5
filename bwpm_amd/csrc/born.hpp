// Born scattering at weak index fluctuations δn (placeholder until the device Born path lands).
#pragma once

#include "engine.hpp"
#include "fft.hpp"

namespace bwpm {

template <typename T, bool COMP>
class Born {
 public:
  Born(const Spec&, const std::vector<char>&, const std::vector<c128>&, const double*, const double*, hipStream_t) {
    throw Error(kEinval, "index fluctuations are not supported by this build");
  }
  void forward(const px2<T>*, long long, bool) {}
  void backward(const px2<T>*, long long) {}
  const px2<T>* increment() const { return nullptr; }
  const px2<T>* fwd_plane(long long) const { return nullptr; }
  const px2<T>* bwd_plane(long long) const { return nullptr; }
  long long iterations() const { return 0; }
  size_t device_bytes() const { return 0; }
};

}  // namespace bwpm
