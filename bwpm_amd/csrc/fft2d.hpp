// Plain 2D FFTs of (E_x, E_y)-shaped planes for the Born machinery: a row pass and a column pass,
// each serving up to kMaxPlanes planes per launch (blockIdx.y), with an optional per-pixel
// complex multiplier applied after the transform (fuses "F⁻¹ then ×V") and an optional real scale
// (the 1/N of the inverse). Register-resident kernels for power-of-two lines, LDS Stockham kernels
// otherwise.
#pragma once

#include "fft.hpp"
#include "fft_reg.hpp"
#include "kernels.hpp"

#include <algorithm>

namespace bwpm {

inline constexpr int kMaxPlanes = 4;

template <typename T>
struct PlaneSet {
  const px2<T>* in[kMaxPlanes];
  px2<T>* out[kMaxPlanes];
  const cx<T>* mul[kMaxPlanes];  // per-pixel multiplier of both components (nullptr: none)
  const cx<T>* pre[kMaxPlanes];  // per-pixel multiplier of the input (register row kernels only)
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

template <typename T, bool COMP, bool INV, bool LEAN = false>
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
  fft_lines<COMP, INV, LEAN>(buf, ls, 2 * rows, plan);
  for (int i = threadIdx.x; i < count; i += kThreads) {
    const int r = i / nx, x = i - r * nx;
    const long long q = static_cast<long long>(y0 + r) * nx + x;
    out[q] = post(px2<T>{buf[x * ls + r], buf[x * ls + rows + r]}, mul, q, ps.scale);
  }
}

template <typename T, bool COMP, bool INV, bool LEAN = false>
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
  fft_lines<COMP, INV, LEAN>(buf, ls, 2 * W, plan);
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
  constexpr int LQ = lds_stride(2 * reg_lds_len<N>(), 1, sizeof(cx<T>));  // a line: E_x and E_y planes
  __shared__ cx<T> sh[RW * LQ];
  const px2<T>* in = ps.in[blockIdx.y];
  px2<T>* out = ps.out[blockIdx.y];
  const cx<T>* mul = ps.mul[blockIdx.y];
  const cx<T>* pre = ps.pre[blockIdx.y];
  const int r = threadIdx.x / TT, t = threadIdx.x % TT;
  const long long row = static_cast<long long>(blockIdx.x * RW + r) * N;
  px2<T> v[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) v[m] = in[row + t + TT * m];
  if (pre) {
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const cx<T> f = pre[row + t + TT * m];
      v[m] = px2<T>{v[m].a * f, v[m].b * f};
    }
  }
  reg_fft<N, COMP, INV>(v, sh + r * LQ, t, twh, twl);
#pragma unroll
  for (int m = 0; m < 8; ++m) out[row + t + TT * m] = post(v[m], mul, row + t + TT * m, ps.scale);
}

template <int N, int W, typename T, bool COMP, bool INV>
__global__ __launch_bounds__(N / 8 * W) void k_fft2_cols_fast(PlaneSet<T> ps, int nx, const cx<T>* __restrict__ twh,
                                                              const cx<T>* __restrict__ twl) {
  constexpr int TT = N / 8;
  constexpr int LQ = lds_stride(2 * reg_lds_len<N>(), W, sizeof(cx<T>));  // a line: E_x and E_y planes
  __shared__ cx<T> sh[W * LQ];
  const px2<T>* in = ps.in[blockIdx.y];
  px2<T>* out = ps.out[blockIdx.y];
  const cx<T>* mul = ps.mul[blockIdx.y];
  const int c = threadIdx.x % W, t = threadIdx.x / W;
  const int x = xcd_group(blockIdx.x, gridDim.x) * W + c;
  px2<T> v[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) v[m] = in[static_cast<long long>(t + TT * m) * nx + x];
  reg_fft<N, COMP, INV>(v, sh + c * LQ, t, twh, twl);
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const long long q = static_cast<long long>(t + TT * m) * nx + x;
    out[q] = post(v[m], mul, q, ps.scale);
  }
}

// ---- host driver ---------------------------------------------------------------------------------

// In-place or out-of-place 2D transforms of up to kMaxPlanes (ny, nx, 2) planes on one stream: a
// row pass then a column pass. Register kernels for power-of-two lines up to 2048 where the
// precision mode matches them (FP64 plain, FP32 compensated), LDS Stockham kernels otherwise.
template <typename T, bool COMP>
class Fft2D {
 public:
  void init(long long ny, long long nx, const FftPlan& px, const FftPlan& py, hipStream_t stream) {
    ny_ = ny;
    nx_ = nx;
    npix_ = ny * nx;
    px_ = px;
    py_ = py;
    stream_ = stream;
    auto ls_of = [](int lines) { return lines <= 2 ? lines : (lines | 1); };
    const int elems = Budget<T>::elems;
    rows_fast_ = fast_n(nx_) && COMP == (sizeof(T) == 4);
    cols_fast_ = fast_n(ny_) && COMP == (sizeof(T) == 4);
    int rw = static_cast<int>(std::min<long long>(ny_, 64));
    while (rw > 1 && nx_ * ls_of(2 * rw) > elems) rw /= 2;
    rows_ = rw;
    ls_row_ = ls_of(2 * rw);
    int w = 16;
    while (w > 1 && ny_ * ls_of(2 * w) > elems) w /= 2;
    while (w > 1 && w > nx_) w /= 2;
    cols_ = w;
    ls_col_ = ls_of(2 * w);
    require(nx_ * ls_row_ <= elems && ny_ * ls_col_ <= elems, "grid too large for the 2D transforms");
    // narrow geometry for launches of few workgroups (kNarrowLds of LDS per workgroup)
    while (rw > 1 && nx_ * ls_of(2 * rw) * static_cast<long long>(sizeof(cx<T>)) > kNarrowLds) rw /= 2;
    rows_n_ = rw;
    ls_row_n_ = ls_of(2 * rw);
    while (w > 1 && ny_ * ls_of(2 * w) * static_cast<long long>(sizeof(cx<T>)) > kNarrowLds) w /= 2;
    cols_n_ = w;
    ls_col_n_ = ls_of(2 * w);
    lean_rows_ = lean_plan(px_);
    lean_cols_ = lean_plan(py_);
    for (const void* k : {reinterpret_cast<const void*>(&k_fft2_rows<T, COMP, false>),
                          reinterpret_cast<const void*>(&k_fft2_rows<T, COMP, true>),
                          reinterpret_cast<const void*>(&k_fft2_cols<T, COMP, false>),
                          reinterpret_cast<const void*>(&k_fft2_cols<T, COMP, true>),
                          reinterpret_cast<const void*>(&k_fft2_rows<T, COMP, false, true>),
                          reinterpret_cast<const void*>(&k_fft2_rows<T, COMP, true, true>),
                          reinterpret_cast<const void*>(&k_fft2_cols<T, COMP, false, true>),
                          reinterpret_cast<const void*>(&k_fft2_cols<T, COMP, true, true>)})
      HIP_CHECK(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
  }

  // In-place transform of np planes; the inverse is scaled by 1/N and optionally multiplied by `mul`.
  void planes(px2<T>* const* pl, int np, bool inverse, const cx<T>* mul) {
    PlaneSet<T> a{}, b{};
    for (int i = 0; i < np; ++i) {
      a.in[i] = b.in[i] = pl[i];
      a.out[i] = b.out[i] = pl[i];
      b.mul[i] = mul;
    }
    run(a, b, np, inverse);
  }
  // Both passes run the register kernels (power-of-two lines in the precision they support).
  bool fast() const { return rows_fast_ && cols_fast_; }
  const FftPlan& row_plan() const { return px_; }
  const FftPlan& col_plan() const { return py_; }

  // The column pass alone, for callers that made the row pass themselves: out = F_y(in) or
  // F_y⁻¹(in)/N_pixels (the scale of the whole inverse 2D transform).
  void cols(const px2<T>* in, px2<T>* out, bool inverse) {
    PlaneSet<T> b{};
    b.in[0] = in;
    b.out[0] = out;
    b.scale = inverse ? static_cast<T>(1.0 / static_cast<double>(npix_)) : static_cast<T>(1);
    if (inverse)
      cols_pass<true>(b, 1);
    else
      cols_pass<false>(b, 1);
  }

  // out = F(in·pre) or F⁻¹(in·pre) (1/N scaled; pre: per-pixel multiplier of the input, nullptr:
  // none, only where the row pass runs the register kernels); `in` is left unchanged unless it is
  // `out`.
  void one(const px2<T>* in, px2<T>* out, bool inverse, const cx<T>* pre = nullptr) {
    require(pre == nullptr || rows_fast_, "an input multiplier needs the register row kernels");
    PlaneSet<T> a{}, b{};
    a.pre[0] = pre;
    a.in[0] = in;
    a.out[0] = out;
    b.in[0] = out;
    b.out[0] = out;
    run(a, b, 1, inverse);
  }

 private:
  static bool fast_n(long long n) {
    return n == 64 || n == 128 || n == 256 || n == 512 || n == 1024 || n == 2048 || n == 192 || n == 384 ||
           n == 768 || n == 1536;
  }
  // Lines per workgroup of the register kernels: a power of two with at most `threads` threads
  // (at least one line).
  template <int N>
  static constexpr int fit_lines(int threads) {
    int l = 1;
    while (2 * l * (N / 8) <= threads) l *= 2;
    return l;
  }
  void run(PlaneSet<T>& a, PlaneSet<T>& b, int np, bool inverse) {
    a.scale = 1;
    b.scale = inverse ? static_cast<T>(1.0 / static_cast<double>(npix_)) : static_cast<T>(1);
    if (inverse) {
      rows_pass<true>(a, np);
      cols_pass<true>(b, np);
    } else {
      rows_pass<false>(a, np);
      cols_pass<false>(b, np);
    }
  }
  template <bool INV>
  void rows_pass(const PlaneSet<T>& ps, int np) {
    if (rows_fast_) {
      switch (nx_) {
        case 64: rows_fast_n<64, INV>(ps, np); return;
        case 128: rows_fast_n<128, INV>(ps, np); return;
        case 256: rows_fast_n<256, INV>(ps, np); return;
        case 512: rows_fast_n<512, INV>(ps, np); return;
        case 1024: rows_fast_n<1024, INV>(ps, np); return;
        case 2048: rows_fast_n<2048, INV>(ps, np); return;
        case 192: rows_fast_n<192, INV>(ps, np); return;
        case 384: rows_fast_n<384, INV>(ps, np); return;
        case 768: rows_fast_n<768, INV>(ps, np); return;
        default: rows_fast_n<1536, INV>(ps, np); return;
      }
    }
    const bool narrow = (ny_ + rows_ - 1) / rows_ * np < kNarrowBelow;
    const int rows = narrow ? rows_n_ : rows_, ls = narrow ? ls_row_n_ : ls_row_;
    const size_t lds = static_cast<size_t>(nx_) * ls * sizeof(cx<T>);
    const dim3 grid(static_cast<unsigned>((ny_ + rows - 1) / rows), np);
    if (lean_rows_)
      hipLaunchKernelGGL((k_fft2_rows<T, COMP, INV, true>), grid, dim3(kThreads), lds, stream_, ps, static_cast<int>(ny_),
                         static_cast<int>(nx_), rows, ls, px_);
    else
      hipLaunchKernelGGL((k_fft2_rows<T, COMP, INV>), grid, dim3(kThreads), lds, stream_, ps, static_cast<int>(ny_),
                         static_cast<int>(nx_), rows, ls, px_);
    HIP_CHECK(hipGetLastError());
  }
  template <int N, bool INV>
  void rows_fast_n(const PlaneSet<T>& ps, int np) {
    constexpr int RW = fit_lines<N>(sizeof(T) == 4 ? 512 : 256);
    const auto* th = static_cast<const cx<T>*>(px_.tw_hi);
    const auto* tl = static_cast<const cx<T>*>(px_.tw_lo);
    if (ny_ % RW == 0)
      hipLaunchKernelGGL((k_fft2_rows_fast<N, RW, T, COMP, INV>), dim3(static_cast<unsigned>(ny_ / RW), np),
                         dim3(N / 8 * RW), 0, stream_, ps, th, tl);
    else
      hipLaunchKernelGGL((k_fft2_rows_fast<N, 1, T, COMP, INV>), dim3(static_cast<unsigned>(ny_), np), dim3(N / 8), 0,
                         stream_, ps, th, tl);
    HIP_CHECK(hipGetLastError());
  }
  template <bool INV>
  void cols_pass(const PlaneSet<T>& ps, int np) {
    if (cols_fast_) {
      switch (ny_) {
        case 64: cols_fast_n<64, INV>(ps, np); return;
        case 128: cols_fast_n<128, INV>(ps, np); return;
        case 256: cols_fast_n<256, INV>(ps, np); return;
        case 512: cols_fast_n<512, INV>(ps, np); return;
        case 1024: cols_fast_n<1024, INV>(ps, np); return;
        case 2048: cols_fast_n<2048, INV>(ps, np); return;
        case 192: cols_fast_n<192, INV>(ps, np); return;
        case 384: cols_fast_n<384, INV>(ps, np); return;
        case 768: cols_fast_n<768, INV>(ps, np); return;
        default: cols_fast_n<1536, INV>(ps, np); return;
      }
    }
    const bool narrow = (nx_ + cols_ - 1) / cols_ * np < kNarrowBelow;
    const int cols = narrow ? cols_n_ : cols_, ls = narrow ? ls_col_n_ : ls_col_;
    const size_t lds = static_cast<size_t>(ny_) * ls * sizeof(cx<T>);
    const dim3 grid(static_cast<unsigned>((nx_ + cols - 1) / cols), np);
    if (lean_cols_)
      hipLaunchKernelGGL((k_fft2_cols<T, COMP, INV, true>), grid, dim3(kThreads), lds, stream_, ps, static_cast<int>(ny_),
                         static_cast<int>(nx_), cols, ls, py_);
    else
      hipLaunchKernelGGL((k_fft2_cols<T, COMP, INV>), grid, dim3(kThreads), lds, stream_, ps, static_cast<int>(ny_),
                         static_cast<int>(nx_), cols, ls, py_);
    HIP_CHECK(hipGetLastError());
  }
  template <int N, bool INV>
  void cols_fast_n(const PlaneSet<T>& ps, int np) {
    constexpr int W0 = fit_lines<N>(sizeof(T) == 4 ? 512 : 256);
    constexpr int W = W0 > 16 ? 16 : W0;
    const auto* th = static_cast<const cx<T>*>(py_.tw_hi);
    const auto* tl = static_cast<const cx<T>*>(py_.tw_lo);
    if (nx_ % W == 0)
      hipLaunchKernelGGL((k_fft2_cols_fast<N, W, T, COMP, INV>), dim3(static_cast<unsigned>(nx_ / W), np),
                         dim3(N / 8 * W), 0, stream_, ps, static_cast<int>(nx_), th, tl);
    else
      hipLaunchKernelGGL((k_fft2_cols_fast<N, 1, T, COMP, INV>), dim3(static_cast<unsigned>(nx_), np), dim3(N / 8), 0,
                         stream_, ps, static_cast<int>(nx_), th, tl);
    HIP_CHECK(hipGetLastError());
  }

  long long ny_ = 0, nx_ = 0, npix_ = 0;
  FftPlan px_{}, py_{};
  hipStream_t stream_ = nullptr;
  bool rows_fast_ = false, cols_fast_ = false;
  bool lean_rows_ = false, lean_cols_ = false;  // generic plans of radices 2, 3, 4 and 8 only
  int rows_n_ = 1, ls_row_n_ = 1, cols_n_ = 1, ls_col_n_ = 1;  // narrow geometry
  int rows_ = 1, ls_row_ = 2, cols_ = 1, ls_col_ = 2;
};

}  // namespace bwpm
