// Plain 2D FFTs of (E_x, E_y)-shaped planes for the Born machinery: a row pass and a column pass,
// each serving up to kMaxPlanes planes per launch (blockIdx.y), with an optional per-pixel
// complex multiplier applied after the transform (fuses "F⁻¹ then ×V") and an optional real scale
// (the 1/N of the inverse). Register-resident kernels for power-of-two lines, LDS Stockham kernels
// otherwise.
#pragma once

#include "fft.hpp"
#include "fft_reg.hpp"
#include "kernels.hpp"

namespace bwpm {

inline constexpr int kMaxPlanes = 4;

template <typename T>
struct PlaneSet {
  const px2<T>* in[kMaxPlanes];
  px2<T>* out[kMaxPlanes];
  const cx<T>* mul[kMaxPlanes];  // per-pixel multiplier of both components (nullptr: none)
  T scale;                       // applied to every output value
};

template <typename T>
__device__ __forceinline__ px2<T> post(px2<T> v, const cx<T>* mul, long long q, T scale) {
  if (mul) {
    const cx<T> m = mul[q];
    v.a = v.a * m;
    v.b = v.b * m;
  }
  if (scale != static_cast<T>(1)) {
    v.a = v.a * scale;
    v.b = v.b * scale;
  }
  return v;
}

// ---- generic lengths -----------------------------------------------------------------------------

template <typename T, bool COMP, bool INV>
__global__ __launch_bounds__(kThreads) void k_fft2_rows(PlaneSet<T> ps, int ny, int nx, int rows, int ls, FftPlan plan) {
  const px2<T>* in = ps.in[blockIdx.y];
  px2<T>* out = ps.out[blockIdx.y];
  const cx<T>* mul = ps.mul[blockIdx.y];
  cx<T>* buf = lds<T>();
  const int y0 = blockIdx.x * rows;
  const int nrow = min(rows, ny - y0);
  const int count = nrow * nx;
  load_rows(buf, in, y0, nrow, nx, rows, ls);
  __syncthreads();
  fft_lines<COMP, INV>(buf, ls, 2 * rows, plan);
  for (int i = threadIdx.x; i < count; i += kThreads) {
    const int r = i / nx, x = i - r * nx;
    const long long q = static_cast<long long>(y0 + r) * nx + x;
    out[q] = post(px2<T>{buf[x * ls + r], buf[x * ls + rows + r]}, mul, q, ps.scale);
  }
}

template <typename T, bool COMP, bool INV>
__global__ __launch_bounds__(kThreads) void k_fft2_cols(PlaneSet<T> ps, int ny, int nx, int cols, int ls, FftPlan plan) {
  const px2<T>* in = ps.in[blockIdx.y];
  px2<T>* out = ps.out[blockIdx.y];
  const cx<T>* mul = ps.mul[blockIdx.y];
  cx<T>* buf = lds<T>();
  const int W = cols;
  const int x0 = blockIdx.x * W;
  const int count = ny * W;
  for (int i = threadIdx.x; i < count; i += kThreads) {
    const int y = i / W, c = i - y * W, x = x0 + c;
    px2<T> v{};
    if (x < nx) v = in[static_cast<long long>(y) * nx + x];
    buf[y * ls + c] = v.a;
    buf[y * ls + W + c] = v.b;
  }
  __syncthreads();
  fft_lines<COMP, INV>(buf, ls, 2 * W, plan);
  for (int i = threadIdx.x; i < count; i += kThreads) {
    const int y = i / W, c = i - y * W, x = x0 + c;
    if (x >= nx) continue;
    const long long q = static_cast<long long>(y) * nx + x;
    out[q] = post(px2<T>{buf[y * ls + c], buf[y * ls + W + c]}, mul, q, ps.scale);
  }
}

// ---- power-of-two lengths (register-resident) -----------------------------------------------------

template <int N, int RW, typename T, bool COMP, bool INV>
__global__ __launch_bounds__(N / 8 * RW) void k_fft2_rows_fast(PlaneSet<T> ps, const cx<T>* __restrict__ twh,
                                                               const cx<T>* __restrict__ twl) {
  constexpr int TT = N / 8;
  constexpr int LL = reg_lds_len<N>();
  __shared__ px2<T> sh[RW * LL];
  const px2<T>* in = ps.in[blockIdx.y];
  px2<T>* out = ps.out[blockIdx.y];
  const cx<T>* mul = ps.mul[blockIdx.y];
  const int r = threadIdx.x / TT, t = threadIdx.x % TT;
  const long long row = static_cast<long long>(blockIdx.x * RW + r) * N;
  px2<T> v[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) v[m] = in[row + t + TT * m];
  reg_fft<N, COMP, INV>(v, sh + r * LL, t, twh, twl);
#pragma unroll
  for (int m = 0; m < 8; ++m) out[row + t + TT * m] = post(v[m], mul, row + t + TT * m, ps.scale);
}

template <int N, int W, typename T, bool COMP, bool INV>
__global__ __launch_bounds__(N / 8 * W) void k_fft2_cols_fast(PlaneSet<T> ps, int nx, const cx<T>* __restrict__ twh,
                                                              const cx<T>* __restrict__ twl) {
  constexpr int TT = N / 8;
  constexpr int LL = reg_lds_len<N>();
  __shared__ px2<T> sh[W * LL];
  const px2<T>* in = ps.in[blockIdx.y];
  px2<T>* out = ps.out[blockIdx.y];
  const cx<T>* mul = ps.mul[blockIdx.y];
  const int c = threadIdx.x % W, t = threadIdx.x / W;
  const int x = blockIdx.x * W + c;
  px2<T> v[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) v[m] = in[static_cast<long long>(t + TT * m) * nx + x];
  reg_fft<N, COMP, INV>(v, sh + c * LL, t, twh, twl);
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const long long q = static_cast<long long>(t + TT * m) * nx + x;
    out[q] = post(v[m], mul, q, ps.scale);
  }
}

}  // namespace bwpm
