// Common device/host types of the MI355X BWPM engine: complex numbers, the (E_x, E_y) pixel
// vector, compensated ("hi + lo") constants for the FP32 bulk path, and error plumbing.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>

namespace bwpm {

#define BWPM_HD __host__ __device__ __forceinline__

// Complex number of precision T (float for the bulk path, double for the FP64 path).
template <typename T>
struct alignas(2 * sizeof(T)) cx {
  T x, y;
};

// One transverse pixel (E_x, E_y); 16 B in FP32 so one dwordx4 load moves a whole pixel.
template <typename T>
struct alignas(4 * sizeof(T)) px2 {
  cx<T> a, b;
};

template <typename T>
BWPM_HD cx<T> mk(T x, T y) {
  return cx<T>{x, y};
}
template <typename T>
BWPM_HD cx<T> operator+(cx<T> a, cx<T> b) {
  return {a.x + b.x, a.y + b.y};
}
template <typename T>
BWPM_HD cx<T> operator-(cx<T> a, cx<T> b) {
  return {a.x - b.x, a.y - b.y};
}
template <typename T>
BWPM_HD cx<T> operator-(cx<T> a) {
  return {-a.x, -a.y};
}
template <typename T>
BWPM_HD cx<T> operator*(cx<T> a, cx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
BWPM_HD cx<T> operator*(cx<T> a, T s) {
  return {a.x * s, a.y * s};
}
template <typename T>
BWPM_HD cx<T> operator*(T s, cx<T> a) {
  return {a.x * s, a.y * s};
}
template <typename T>
BWPM_HD cx<T> conj(cx<T> a) {
  return {a.x, -a.y};
}
// multiply by ±i
template <typename T>
BWPM_HD cx<T> mul_i(cx<T> a) {
  return {-a.y, a.x};
}
template <typename T>
BWPM_HD cx<T> mul_mi(cx<T> a) {
  return {a.y, -a.x};
}

BWPM_HD float fmat(float a, float b, float c) { return fmaf(a, b, c); }
BWPM_HD double fmat(double a, double b, double c) { return fma(a, b, c); }

// a·(w_hi + w_lo) for a compensated complex constant: the lo part removes the systematic rounding
// of w, so the per-slab error of the FP32 path is rounding noise (random sign) and does not grow
// linearly with the number of slabs. In FP64 (COMP = false) lo is ignored.
template <bool COMP, typename T>
BWPM_HD cx<T> cmul2(cx<T> a, cx<T> hi, cx<T> lo) {
  if constexpr (COMP) {
    const T rl = fmat(a.x, lo.x, -a.y * lo.y);
    const T il = fmat(a.x, lo.y, a.y * lo.x);
    return {fmat(a.x, hi.x, fmat(-a.y, hi.y, rl)), fmat(a.x, hi.y, fmat(a.y, hi.x, il))};
  } else {
    return a * hi;
  }
}
// a·(c_hi + c_lo) for a real compensated constant.
template <bool COMP, typename T>
BWPM_HD cx<T> rmul2(cx<T> a, T hi, T lo) {
  if constexpr (COMP) {
    return {fmat(a.x, hi, a.x * lo), fmat(a.y, hi, a.y * lo)};
  } else {
    return {a.x * hi, a.y * hi};
  }
}

// ---- device global-memory access ------------------------------------------------------------
// Loads and stores through pointers in the global address space compile to global_load/store
// rather than flat ones, so the LDS waits of the FFT passes do not also wait for HBM traffic.
#define BWPM_GLOBAL __attribute__((address_space(1)))
template <typename T>
struct Vec2;
template <>
struct Vec2<float> {
  typedef float type __attribute__((ext_vector_type(2)));
};
template <>
struct Vec2<double> {
  typedef double type __attribute__((ext_vector_type(2)));
};

// Element i of an array whose byte size is below 4 GiB: the byte offset is formed in 32 bits, so
// the access uses the scalar base + 32-bit vector offset form when the base is uniform.
template <typename U, typename E>
__device__ __forceinline__ const BWPM_GLOBAL E* gaddr(const U* p, unsigned i) {
  return (const BWPM_GLOBAL E*)((const BWPM_GLOBAL char*)p + i * static_cast<unsigned>(sizeof(E)));
}
template <typename U, typename E>
__device__ __forceinline__ BWPM_GLOBAL E* gaddr(U* p, unsigned i) {
  return (BWPM_GLOBAL E*)((BWPM_GLOBAL char*)p + i * static_cast<unsigned>(sizeof(E)));
}

template <typename T>
__device__ __forceinline__ cx<T> gload(const cx<T>* p, unsigned i) {
  using V = typename Vec2<T>::type;
  const V v = *gaddr<cx<T>, V>(p, i);
  return cx<T>{v.x, v.y};
}
template <typename T>
__device__ __forceinline__ void gstore(cx<T>* p, unsigned i, cx<T> z) {
  using V = typename Vec2<T>::type;
  V v;
  v.x = z.x;
  v.y = z.y;
  *gaddr<cx<T>, V>(p, i) = v;
}
template <typename U>
__device__ __forceinline__ U gload1(const U* p, unsigned i) {
  return *gaddr<U, U>(p, i);
}

// Values every lane of a wavefront holds equally, moved to scalar registers (so branches on them
// are scalar and addresses derived from them use the scalar base of global loads).
__device__ __forceinline__ int uniform(int v) {
  return __builtin_amdgcn_readfirstlane(v);
}
template <typename P>
__device__ __forceinline__ P* uniform(P* p) {
  const unsigned long long u = reinterpret_cast<unsigned long long>(p);
  const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<int>(u & 0xffffffffULL)));
  const unsigned hi = static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<int>(u >> 32)));
  return reinterpret_cast<P*>((static_cast<unsigned long long>(hi) << 32) | lo);
}

// Status codes of the C ABI (include/bwpm_hip.h).
enum Status : int { kOk = 0, kEinval = 1, kEnoconv = 2, kEdevice = 3, kEnomem = 4, kEinternal = 5 };

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess)
    throw Error(kEdevice, std::string(what) + ": " + hipGetErrorString(e));
}
#define HIP_CHECK(x) ::bwpm::hip_check((x), #x)

inline void require(bool ok, const std::string& msg) {
  if (!ok) throw Error(kEinval, msg);
}

}  // namespace bwpm
