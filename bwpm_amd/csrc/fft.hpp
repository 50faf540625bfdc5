// Mixed-radix Stockham FFT of many short lines held in LDS, for 256-thread workgroups (four
// wave64 wavefronts).
//
// A line of length n is transformed in place in LDS by stages of radix R (the product of the
// radices is n). Stage s with Ns = product of the previous radices reads the R inputs
// x_r = a[j + r·n/R] of butterfly j, multiplies x_r by the twiddle W_n^{r·(j mod Ns)·n/(Ns·R)},
// applies the size-R DFT and writes y_r to a[(j div Ns)·Ns·R + (j mod Ns) + r·Ns]. Every thread
// first loads all of its butterflies into registers, the workgroup synchronises, and only then
// writes, so one LDS buffer suffices (out-of-place in registers, in-place in LDS).
//
// Twiddles come from a per-length table W_n^t = exp(−2πi t/n) computed in FP64 on the host and
// split into hi + lo parts for the FP32 path (COMP = true); constant multipliers inside the
// butterflies (1/√2, √3/2, cos/sin 2π/5) are compensated the same way. The inverse transform
// (INV = true) uses conjugated twiddles and is unnormalised (the 1/N lives in the operator tables).
#pragma once

#include "common.hpp"

namespace bwpm {

inline constexpr int kThreads = 512;  // threads per workgroup of every FFT kernel
// LDS per workgroup of the generic kernels' narrow geometry, and the workgroup count (of the wide
// geometry) below which a launch takes it. Wide workgroups are the more efficient ones as long as
// a launch has enough of them: a lane of 16 192² solves (192 workgroups) runs 1.5× faster wide,
// while a single 384² solve (48) runs 1.4× and a single 192² solve (12) 1.5× faster narrow.
inline constexpr long long kNarrowLds = 65536;
inline constexpr long long kNarrowBelow = 128;
inline constexpr int kMaxStages = 12;

struct FftPlan {
  int n;                    // line length
  int nst;                  // number of stages
  int rad[kMaxStages];      // radices, applied in order
  const void* tw_hi;        // cx<T>[n], W_n^t
  const void* tw_lo;        // cx<T>[n], rounding residual (FP32 path only)
};

// Compensated real constants: hi = the constant rounded to T, lo = the residual.
template <typename T>
struct KC {
  static constexpr double r2 = 0.70710678118654752440084436210485;   // 1/√2
  static constexpr double s3 = 0.86602540378443864676372317075294;   // √3/2
  static constexpr double c1 = 0.30901699437494742410229341718282;   // cos 2π/5
  static constexpr double c2 = -0.80901699437494742410229341718282;  // cos 4π/5
  static constexpr double s1 = 0.95105651629515357211643933337938;   // sin 2π/5
  static constexpr double s2 = 0.58778525229247312916870595463907;   // sin 4π/5
  static constexpr T hi(double v) { return static_cast<T>(v); }
  static constexpr T lo(double v) { return static_cast<T>(v - static_cast<double>(static_cast<T>(v))); }
};

// size-R DFTs on registers (forward sign −, INV: +)
template <bool COMP, bool INV, typename T>
__device__ __forceinline__ void dft2(cx<T>* x) {
  const cx<T> a = x[0], b = x[1];
  x[0] = a + b;
  x[1] = a - b;
}

template <bool COMP, bool INV, typename T>
__device__ __forceinline__ void dft4(cx<T>* x) {
  const cx<T> t0 = x[0] + x[2], t1 = x[0] - x[2];
  const cx<T> t2 = x[1] + x[3];
  const cx<T> d = x[1] - x[3];
  const cx<T> t3 = INV ? mul_i(d) : mul_mi(d);
  x[0] = t0 + t2;
  x[2] = t0 - t2;
  x[1] = t1 + t3;
  x[3] = t1 - t3;
}

template <bool COMP, bool INV, typename T>
__device__ __forceinline__ void dft8(cx<T>* x) {
  cx<T> e[4] = {x[0], x[2], x[4], x[6]};
  cx<T> o[4] = {x[1], x[3], x[5], x[7]};
  dft4<COMP, INV>(e);
  dft4<COMP, INV>(o);
  constexpr T h = KC<T>::hi(KC<T>::r2), l = KC<T>::lo(KC<T>::r2);
  // W8 = (1 ∓ i)/√2, W8² = ∓i, W8³ = (−1 ∓ i)/√2
  cx<T> w1, w3;
  if constexpr (INV) {
    w1 = cx<T>{o[1].x - o[1].y, o[1].x + o[1].y};
    w3 = cx<T>{-o[3].x - o[3].y, o[3].x - o[3].y};
  } else {
    w1 = cx<T>{o[1].x + o[1].y, o[1].y - o[1].x};
    w3 = cx<T>{o[3].y - o[3].x, -o[3].x - o[3].y};
  }
  w1 = rmul2<COMP>(w1, h, l);
  w3 = rmul2<COMP>(w3, h, l);
  const cx<T> w2 = INV ? mul_i(o[2]) : mul_mi(o[2]);
  x[0] = e[0] + o[0];
  x[4] = e[0] - o[0];
  x[1] = e[1] + w1;
  x[5] = e[1] - w1;
  x[2] = e[2] + w2;
  x[6] = e[2] - w2;
  x[3] = e[3] + w3;
  x[7] = e[3] - w3;
}

template <bool COMP, bool INV, typename T>
__device__ __forceinline__ void dft3(cx<T>* x) {
  constexpr T h = KC<T>::hi(KC<T>::s3), l = KC<T>::lo(KC<T>::s3);
  const cx<T> t = x[1] + x[2];
  const cx<T> d = rmul2<COMP>(x[1] - x[2], h, l);
  const cx<T> a = x[0] - t * static_cast<T>(0.5);
  x[0] = x[0] + t;
  const cx<T> id = INV ? mul_i(d) : mul_mi(d);
  x[1] = a + id;
  x[2] = a - id;
}

template <bool COMP, bool INV, typename T>
__device__ __forceinline__ void dft5(cx<T>* x) {
  using K = KC<T>;
  const cx<T> t1 = x[1] + x[4], t2 = x[2] + x[3];
  const cx<T> d1 = x[1] - x[4], d2 = x[2] - x[3];
  const cx<T> c1t1 = rmul2<COMP>(t1, K::hi(K::c1), K::lo(K::c1));
  const cx<T> c2t1 = rmul2<COMP>(t1, K::hi(K::c2), K::lo(K::c2));
  const cx<T> c1t2 = rmul2<COMP>(t2, K::hi(K::c1), K::lo(K::c1));
  const cx<T> c2t2 = rmul2<COMP>(t2, K::hi(K::c2), K::lo(K::c2));
  const cx<T> s1d1 = rmul2<COMP>(d1, K::hi(K::s1), K::lo(K::s1));
  const cx<T> s2d1 = rmul2<COMP>(d1, K::hi(K::s2), K::lo(K::s2));
  const cx<T> s1d2 = rmul2<COMP>(d2, K::hi(K::s1), K::lo(K::s1));
  const cx<T> s2d2 = rmul2<COMP>(d2, K::hi(K::s2), K::lo(K::s2));
  const cx<T> a1 = x[0] + c1t1 + c2t2, a2 = x[0] + c2t1 + c1t2;
  const cx<T> b1 = s1d1 + s2d2, b2 = s2d1 - s1d2;
  x[0] = x[0] + t1 + t2;
  const cx<T> ib1 = INV ? mul_i(b1) : mul_mi(b1);
  const cx<T> ib2 = INV ? mul_i(b2) : mul_mi(b2);
  x[1] = a1 + ib1;
  x[4] = a1 - ib1;
  x[2] = a2 + ib2;
  x[3] = a2 - ib2;
}

// Odd prime R (7, 11, 13): direct DFT with the length-n twiddle table, W_R^t = W_n^{t·n/R}.
template <int R, bool COMP, bool INV, typename T>
__device__ __forceinline__ void dft_generic(cx<T>* x, const cx<T>* twh, const cx<T>* twl, int n) {
  cx<T> y[R];
  const int step = n / R;
#pragma unroll
  for (int q = 0; q < R; ++q) {
    cx<T> s = x[0];
#pragma unroll
    for (int r = 1; r < R; ++r) {
      const int t = ((r * q) % R) * step;
      cx<T> wh = twh[t], wl = twl[t];
      if constexpr (INV) {
        wh = conj(wh);
        wl = conj(wl);
      }
      s = s + cmul2<COMP>(x[r], wh, wl);
    }
    y[q] = s;
  }
#pragma unroll
  for (int q = 0; q < R; ++q) x[q] = y[q];
}

template <int R, bool COMP, bool INV, typename T>
__device__ __forceinline__ void dft(cx<T>* x, const cx<T>* twh, const cx<T>* twl, int n) {
  if constexpr (R == 2) dft2<COMP, INV>(x);
  else if constexpr (R == 3) dft3<COMP, INV>(x);
  else if constexpr (R == 4) dft4<COMP, INV>(x);
  else if constexpr (R == 5) dft5<COMP, INV>(x);
  else if constexpr (R == 8) dft8<COMP, INV>(x);
  else dft_generic<R, COMP, INV>(x, twh, twl, n);
}

// Elements (line length × line stride) one workgroup may hold in one LDS buffer.
template <typename T>
struct Budget {
  static constexpr int elems = 8192;
};

// One Stockham stage over `lines` interleaved lines of length n: element j of line l lives at
// buf[j·ls + l] (ls ≥ lines, odd to spread LDS banks). Consecutive threads take consecutive lines,
// so the reads of a stage are contiguous in LDS.
template <int R, bool COMP, bool INV, typename T>
__device__ __forceinline__ void fft_stage(cx<T>* buf, int n, int ls, int lines, int ns,
                                          const cx<T>* twh, const cx<T>* twl) {
  constexpr int BPT = (Budget<T>::elems / R + kThreads - 1) / kThreads;
  const int m = n / R;
  const int total = m * lines;
  const int tstride = n / (ns * R);
  cx<T> v[BPT][R];
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int idx = threadIdx.x + b * kThreads;
    if (idx < total) {
      const int j = idx / lines, line = idx - j * lines;
      const cx<T>* src = buf + j * ls + line;
#pragma unroll
      for (int r = 0; r < R; ++r) v[b][r] = src[r * m * ls];
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < BPT; ++b) {
    const int idx = threadIdx.x + b * kThreads;
    if (idx < total) {
      const int j = idx / lines, line = idx - j * lines;
      const int k = j % ns;
      if (ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          const int t = r * k * tstride;
          cx<T> wh = twh[t], wl = twl[t];
          if constexpr (INV) {
            wh = conj(wh);
            wl = conj(wl);
          }
          v[b][r] = cmul2<COMP>(v[b][r], wh, wl);
        }
      }
      dft<R, COMP, INV>(v[b], twh, twl, n);
      cx<T>* dst = buf + ((j - k) * R + k) * ls + line;
#pragma unroll
      for (int r = 0; r < R; ++r) dst[r * ns * ls] = v[b][r];
    }
  }
  __syncthreads();
}

// A plan whose radices are all 2, 3, 4 or 8 (sizes 2^a·3^b): kernels for it compile only those
// stages, which keeps them out of register spills.
inline bool lean_plan(const FftPlan& p) {
  for (int s = 0; s < p.nst; ++s)
    if (p.rad[s] != 2 && p.rad[s] != 3 && p.rad[s] != 4 && p.rad[s] != 8) return false;
  return true;
}

// Full transform of `lines` interleaved lines in LDS (element j of line l at buf[j·ls + l]).
// Must be called by all threads of the workgroup. LEAN: the plan is lean_plan's.
template <bool COMP, bool INV, bool LEAN = false, typename T>
__device__ void fft_lines(cx<T>* buf, int ls, int lines, const FftPlan& p) {
  const cx<T>* twh = static_cast<const cx<T>*>(p.tw_hi);
  const cx<T>* twl = static_cast<const cx<T>*>(p.tw_lo);
  int ns = 1;
  for (int s = 0; s < p.nst; ++s) {
    const int R = p.rad[s];
    if constexpr (LEAN) {
      switch (R) {
        case 2: fft_stage<2, COMP, INV>(buf, p.n, ls, lines, ns, twh, twl); break;
        case 3: fft_stage<3, COMP, INV>(buf, p.n, ls, lines, ns, twh, twl); break;
        case 4: fft_stage<4, COMP, INV>(buf, p.n, ls, lines, ns, twh, twl); break;
        default: fft_stage<8, COMP, INV>(buf, p.n, ls, lines, ns, twh, twl); break;
      }
      ns *= R;
      continue;
    }
    switch (R) {
      case 2: fft_stage<2, COMP, INV>(buf, p.n, ls, lines, ns, twh, twl); break;
      case 3: fft_stage<3, COMP, INV>(buf, p.n, ls, lines, ns, twh, twl); break;
      case 4: fft_stage<4, COMP, INV>(buf, p.n, ls, lines, ns, twh, twl); break;
      case 5: fft_stage<5, COMP, INV>(buf, p.n, ls, lines, ns, twh, twl); break;
      case 7: fft_stage<7, COMP, INV>(buf, p.n, ls, lines, ns, twh, twl); break;
      case 8: fft_stage<8, COMP, INV>(buf, p.n, ls, lines, ns, twh, twl); break;
      case 11: fft_stage<11, COMP, INV>(buf, p.n, ls, lines, ns, twh, twl); break;
      default: fft_stage<13, COMP, INV>(buf, p.n, ls, lines, ns, twh, twl); break;
    }
    ns *= R;
  }
}

}  // namespace bwpm
