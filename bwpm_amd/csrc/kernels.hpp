// Slab-step kernels of the bidirectional engine (templated on precision T and on COMP, the
// compensated FP32 arithmetic of the bulk path).
//
// One slab step  out = Σ_pairs I_{r,s} ⊙ F⁻¹{M_{kind,r,s}·F{in}}  (reference: Solver::apply,
// cpp/src/solver.cpp:384-417) is split over two kernels so that each moves the (E_x, E_y) plane
// through HBM once:
//
//   column kernel  X (rows already FFT'd along x) → column FFT → for each product o (one per label
//                  pair, T kinds first, then R kinds): M_o·Ŝ or phase_o·Ŝ → column IFFT → plane P_o
//   row kernel     for each T product: row IFFT of P_o, keep the pixels of its pair; mask; Born
//                  increment; sparse reflection sources and dense Born planes added; the plane is
//                  written out only where needed (kept planes, pass ends, Born inputs); R products:
//                  row IFFT and gather of the interface pixels into the sparse store; finally the
//                  row FFT of the result for the next slab step (fused, so the field never makes
//                  an extra trip).
//
// Every launch serves a batch of solves of equal grid: blockIdx.y selects the solve and its
// per-step descriptor (a null descriptor is a no-op), so many planes share one launch.
#pragma once

#include "fft.hpp"
#include "fft_reg.hpp"
#include "interface_math.hpp"

namespace bwpm {

extern __shared__ __attribute__((aligned(16))) unsigned char g_lds[];

template <typename T>
__device__ __forceinline__ cx<T>* lds() {
  return reinterpret_cast<cx<T>*>(g_lds);
}

// Operator product: scalar phase table (type 0, one complex per k) or symmetric 2×2 matrix table
// (type 1, three complex per k: m00, m01, m11), each as hi and lo parts.
template <typename T>
struct Prod {
  int type;
  const cx<T>* hi;
  const cx<T>* lo;
};

// Storage of the two streaming planes (X and the product planes) may be narrower than the
// arithmetic: ST = float with T = double keeps them as complex64 in HBM and computes in FP64.
template <typename T, typename U>
__device__ __forceinline__ cx<T> widen(cx<U> z) {
  return cx<T>{static_cast<T>(z.x), static_cast<T>(z.y)};
}
template <typename T, typename U>
__device__ __forceinline__ px2<T> widen(px2<U> v) {
  return px2<T>{widen<T>(v.a), widen<T>(v.b)};
}

// Block-floating-point storage (ST = bfp, the mixed mode's streaming planes): one complex value in
// 64 bits, its two parts as 28-bit two's-complement mantissas that share one exponent. With
// m = max(|re|, |im|) = f·2^e (f in [0.5, 1)), each part is stored as q = rint(v·2^(27 − e)),
// |q| ≤ 2^27 − 1, in bits 4 … 31 of its word; the biased exponent E = e + 126 sits in the two low
// nibbles (real part's word: bits 0 … 3 of E, imaginary part's word: bits 4 … 7). The rounding error
// is at most 2^−27·m, against 2^−24 per part for complex64, in the same 8 bytes, over complex64's
// normal range: E = 0 is a zero (m below 2^−126 flushes to it), E = 255 a value with a part that is
// not finite or m ≥ 2^128 (it reads back as NaN).
struct bfp {
  uint32_t w;
};

BWPM_HD cx<bfp> bfp_encode(cx<double> z) {
  const double ax = fabs(z.x), ay = fabs(z.y);
  const double m = ax > ay ? ax : ay;
  int e = 0;
  (void)frexp(m, &e);
  // comparisons with NaN are false, so a NaN in either part is not finite
  const bool finite = ax <= 1.7976931348623157e308 && ay <= 1.7976931348623157e308;
  const bool over = !finite || e > 128;
  const bool flush = !(m > 0) || e < -125;  // zero or below the range
  const int s = 27 - e;
  const double r0 = rint(ldexp(z.x, s)), r1 = rint(ldexp(z.y, s));
  int q0 = static_cast<int>(fmin(r0, 134217727.0)), q1 = static_cast<int>(fmin(r1, 134217727.0));
  unsigned E = static_cast<unsigned>(e + 126);
  if (flush || over) {
    q0 = 0;
    q1 = 0;
    E = over ? 255u : 0u;
  }
  return cx<bfp>{bfp{(static_cast<unsigned>(q0) << 4) | (E & 15u)}, bfp{(static_cast<unsigned>(q1) << 4) | (E >> 4)}};
}

BWPM_HD cx<double> bfp_decode(unsigned w0, unsigned w1) {
  const int E = static_cast<int>((w0 & 15u) | ((w1 & 15u) << 4));
  const int sh = E - 153;  // 2^(e − 27)
  cx<double> z{ldexp(static_cast<double>(static_cast<int>(w0) >> 4), sh),
               ldexp(static_cast<double>(static_cast<int>(w1) >> 4), sh)};
  if (E == 255) z = cx<double>{__builtin_nan(""), __builtin_nan("")};
  return z;
}

template <>
__device__ __forceinline__ cx<bfp> widen<bfp, double>(cx<double> z) {
  return bfp_encode(z);
}
template <>
__device__ __forceinline__ cx<double> widen<double, bfp>(cx<bfp> z) {
  return bfp_decode(z.x.w, z.y.w);
}

// Streaming-plane access: element i (a complex index) of a plane stored as ST, computed as T.
template <typename T, typename ST>
__device__ __forceinline__ cx<T> sload(const px2<ST>* p, unsigned i) {
  return widen<T>(gload(reinterpret_cast<const cx<ST>*>(p), i));
}
template <typename T, typename ST>
__device__ __forceinline__ void sstore(px2<ST>* p, unsigned i, cx<T> v) {
  gstore(reinterpret_cast<cx<ST>*>(p), i, widen<ST>(v));
}
template <>
__device__ __forceinline__ cx<double> sload<double, bfp>(const px2<bfp>* p, unsigned i) {
  typedef unsigned u2 __attribute__((ext_vector_type(2)));
  const u2 v = *gaddr<px2<bfp>, u2>(p, i);
  return bfp_decode(v.x, v.y);
}
template <>
__device__ __forceinline__ void sstore<double, bfp>(px2<bfp>* p, unsigned i, cx<double> z) {
  typedef unsigned u2 __attribute__((ext_vector_type(2)));
  const cx<bfp> b = bfp_encode(z);
  u2 v;
  v.x = b.x.w;
  v.y = b.y.w;
  *gaddr<px2<bfp>, u2>(p, i) = v;
}

// Column-kernel work of one solve for one step.
template <typename T, typename ST = T>
struct ColDesc {
  const px2<ST>* in;     // X plane; nullptr: no column work for this solve at this step
  px2<ST>* out;          // product planes, `plane` pixels apart
  const Prod<T>* prods;  // nprod descriptors (T kinds first)
  int nprod;
  int nt;                   // T products (the first nt); the rest are R products
  const uint32_t* rowuse;   // per-row product use (kRowUse bits; nullptr: every row uses every product)
};

// Per-row product use at a boundary with several label pairs: bit min(o, 30) is set if a pixel of
// the row selects T pair o, bit 31 if the row holds interface pixels (it uses the R products).
inline constexpr int kRowUseR = 31;
inline constexpr unsigned kRowUseT = (1u << 30) - 1u;  // the bits of single T pairs 0 … 29
__host__ __device__ inline constexpr unsigned row_use_bit(int o) { return static_cast<unsigned>(o < 30 ? o : 30); }
// The rows of a use mask (−1: unknown) select exactly one T pair below 30.
__host__ __device__ inline constexpr bool single_t(int use) {
  return use != -1 && (static_cast<unsigned>(use) & ((1u << 31) - 1u)) != 0 &&
         (static_cast<unsigned>(use) & ((1u << 31) - 1u) & ~kRowUseT) == 0 &&
         ((static_cast<unsigned>(use) & kRowUseT) & ((static_cast<unsigned>(use) & kRowUseT) - 1u)) == 0;
}

// Row-kernel work of one solve for one step.
template <typename T, typename ST = T>
struct RowDesc {
  int active;                 // 0: no row work at this step
  int nt, nr;                 // T products (0: start a pass from e_in), R products
  const px2<ST>* prod;        // product planes
  const uint8_t* tsel;        // T pair of each pixel (nt > 1)
  const uint32_t* rowuse;     // per-row product use (kRowUse bits; nullptr: every row uses every product)
  const T* mask;              // transmission mask plane or nullptr
  const double* mask_y;       // or the mask as the product of two profiles, mask_y[y]·mask_x[x] (FP64)
  const double* mask_x;
  const px2<T>* e_in;         // nt == 0: start plane (nullptr: zero)
  const px2<T>* inc;          // Born increment, times mask if mask != nullptr
  const int* rowptr;          // ny + 1 offsets into the sparse list (nullptr: none)
  const int* sp_pix;          // pixel index of each sparse entry
  const uint8_t* sp_rid;      // R product of each sparse entry
  const px2<T>* sp_add;       // source values added at the sparse pixels (nullptr: none)
  px2<T>* sp_store;           // reflection values stored at the sparse pixels (nullptr: none)
  const px2<T>* dense;        // dense Born plane added (nullptr: none)
  px2<T>* e_out;              // real-space result (nullptr: not written)
  px2<T>* e_out2;             // second copy (kept plane)
  px2<ST>* x_out;             // row FFT of the result for the next step (nullptr: none)
};

// The mask factor of pixel (y, x) (plane index q) of a row descriptor: the two profiles multiplied
// in FP64 (the host checked that this product gives every value of the mask plane bit for bit), or
// the mask plane itself.
template <typename T, typename ST>
__device__ __forceinline__ T mask_factor(const RowDesc<T, ST>& a, unsigned y, unsigned x, unsigned q) {
  if (a.mask_y) return static_cast<T>(a.mask_y[y] * a.mask_x[x]);
  return a.mask[q];
}

template <typename T, bool COMP>
__device__ __forceinline__ px2<T> apply_prod(const Prod<T>& pd, long long k, px2<T> e) {
  if (pd.type == 0) {
    const cx<T> ph = pd.hi[k], pl = pd.lo[k];
    return px2<T>{cmul2<COMP>(e.a, ph, pl), cmul2<COMP>(e.b, ph, pl)};
  }
  const cx<T>* h = pd.hi + 3 * k;
  const cx<T>* l = pd.lo + 3 * k;
  const cx<T> h00 = h[0], h01 = h[1], h11 = h[2];
  const cx<T> l00 = l[0], l01 = l[1], l11 = l[2];
  return px2<T>{cmul2<COMP>(e.a, h00, l00) + cmul2<COMP>(e.b, h01, l01),
                cmul2<COMP>(e.a, h01, l01) + cmul2<COMP>(e.b, h11, l11)};
}

// ================================ generic sizes (LDS Stockham) ================================

struct ColGeom {
  int ny, nx, cols, ls;
  long long plane;
  FftPlan plan;  // length ny
};

template <typename T, bool COMP, typename ST = T, bool LEAN = false>
__global__ __launch_bounds__(kThreads) void k_col_step(const ColDesc<T, ST>* __restrict__ descs, ColGeom g) {
  const ColDesc<T, ST> d = descs[blockIdx.y];
  if (!d.in) return;
  cx<T>* A = lds<T>();
  const int W = g.cols, ny = g.ny, nx = g.nx, ls = g.ls;
  const int lines = 2 * W;
  const int x0 = blockIdx.x * W;
  const int count = ny * W;
  for (int i = threadIdx.x; i < count; i += kThreads) {
    const int y = i / W, c = i - y * W, x = x0 + c;
    px2<T> v{};
    if (x < nx) v = widen<T>(d.in[static_cast<long long>(y) * nx + x]);
    A[y * ls + c] = v.a;
    A[y * ls + W + c] = v.b;
  }
  __syncthreads();
  fft_lines<COMP, false, LEAN>(A, ls, lines, g.plan);
  // several products: keep the spectrum in this block's own positions of X and reload it per
  // product, so one LDS buffer suffices (in the storage precision of X)
  px2<ST>* xio = const_cast<px2<ST>*>(d.in);
  if (d.nprod > 1) {
    for (int i = threadIdx.x; i < count; i += kThreads) {
      const int y = i / W, c = i - y * W, x = x0 + c;
      if (x < nx) xio[static_cast<long long>(y) * nx + x] = widen<ST>(px2<T>{A[y * ls + c], A[y * ls + W + c]});
    }
    __syncthreads();
  }
  for (int o = 0; o < d.nprod; ++o) {
    const Prod<T> pd = d.prods[o];
    for (int i = threadIdx.x; i < count; i += kThreads) {
      const int y = i / W, c = i - y * W, x = x0 + c;
      if (x >= nx) continue;
      const long long k = static_cast<long long>(y) * nx + x;
      const px2<T> e = o == 0 ? px2<T>{A[y * ls + c], A[y * ls + W + c]} : widen<T>(xio[k]);
      const px2<T> r = apply_prod<T, COMP>(pd, k, e);
      A[y * ls + c] = r.a;
      A[y * ls + W + c] = r.b;
    }
    __syncthreads();
    fft_lines<COMP, true, LEAN>(A, ls, lines, g.plan);
    px2<ST>* out = d.out + o * g.plane;
    for (int i = threadIdx.x; i < count; i += kThreads) {
      const int y = i / W, c = i - y * W, x = x0 + c;
      if (x < nx) out[static_cast<long long>(y) * nx + x] = widen<ST>(px2<T>{A[y * ls + c], A[y * ls + W + c]});
    }
    __syncthreads();
  }
}

struct RowGeom {
  int ny, nx, rows, ls;
  long long plane;
  FftPlan plan;  // length nx
};

template <typename T, typename U>
__device__ __forceinline__ void load_rows(cx<T>* buf, const px2<U>* src, int y0, int nrow, int nx, int rows,
                                          int ls) {
  const int count = nrow * nx;
  for (int i = threadIdx.x; i < count; i += kThreads) {
    const int r = i / nx, x = i - r * nx;
    const px2<T> v = widen<T>(src[static_cast<long long>(y0 + r) * nx + x]);
    buf[x * ls + r] = v.a;
    buf[x * ls + rows + r] = v.b;
  }
}

template <typename T, bool COMP, typename ST = T, bool LEAN = false>
__global__ __launch_bounds__(kThreads) void k_row_step(const RowDesc<T, ST>* __restrict__ descs, RowGeom g) {
  const RowDesc<T, ST> a = descs[blockIdx.y];
  if (!a.active) return;
  const int nx = g.nx, rows = g.rows, ls = g.ls;
  const int y0 = blockIdx.x * rows;
  const int nrow = min(rows, g.ny - y0);
  const int lines = 2 * rows;
  const int count = nrow * nx;
  cx<T>* acc = lds<T>();
  cx<T>* work = acc + static_cast<long long>(nx) * ls;

  // 1. transmitted field (or the start plane)
  if (a.nt == 0) {
    if (a.e_in) {
      load_rows(acc, a.e_in, y0, nrow, nx, rows, ls);
    } else {
      for (int i = threadIdx.x; i < count; i += kThreads) {
        const int r = i / nx, x = i - r * nx;
        acc[x * ls + r] = cx<T>{0, 0};
        acc[x * ls + rows + r] = cx<T>{0, 0};
      }
    }
    __syncthreads();
  } else if (a.nt == 1) {
    load_rows(acc, a.prod, y0, nrow, nx, rows, ls);
    __syncthreads();
    fft_lines<COMP, true, LEAN>(acc, ls, lines, g.plan);
  } else {
    for (int o = 0; o < a.nt; ++o) {
      load_rows(work, a.prod + o * g.plane, y0, nrow, nx, rows, ls);
      __syncthreads();
      fft_lines<COMP, true, LEAN>(work, ls, lines, g.plan);
      for (int i = threadIdx.x; i < count; i += kThreads) {
        const int r = i / nx, x = i - r * nx;
        if (a.tsel[static_cast<long long>(y0 + r) * nx + x] == o) {
          acc[x * ls + r] = work[x * ls + r];
          acc[x * ls + rows + r] = work[x * ls + rows + r];
        }
      }
      __syncthreads();
    }
  }

  // 2. epilogue in the order of the reference: mask, Born increment, sparse sources, Born plane
  for (int i = threadIdx.x; i < count; i += kThreads) {
    const int r = i / nx, x = i - r * nx;
    const long long q = static_cast<long long>(y0 + r) * nx + x;
    cx<T> ea = acc[x * ls + r], eb = acc[x * ls + rows + r];
    if (a.mask || a.mask_y) {
      const T m = mask_factor(a, static_cast<unsigned>(y0 + r), static_cast<unsigned>(x), static_cast<unsigned>(q));
      ea = ea * m;
      eb = eb * m;
      if (a.inc) {
        const px2<T> dd = a.inc[q];
        ea = ea + dd.a * m;
        eb = eb + dd.b * m;
      }
    } else if (a.inc) {
      const px2<T> dd = a.inc[q];
      ea = ea + dd.a;
      eb = eb + dd.b;
    }
    acc[x * ls + r] = ea;
    acc[x * ls + rows + r] = eb;
  }
  __syncthreads();
  if (a.rowptr && a.sp_add) {
    const int e0 = a.rowptr[y0], e1 = a.rowptr[y0 + nrow];
    for (int e = e0 + threadIdx.x; e < e1; e += kThreads) {
      const int pix = a.sp_pix[e];
      const int r = pix / nx - y0, x = pix - (pix / nx) * nx;
      const px2<T> v = a.sp_add[e];
      acc[x * ls + r] = acc[x * ls + r] + v.a;
      acc[x * ls + rows + r] = acc[x * ls + rows + r] + v.b;
    }
    __syncthreads();
  }
  if (a.dense) {
    for (int i = threadIdx.x; i < count; i += kThreads) {
      const int r = i / nx, x = i - r * nx;
      const px2<T> dd = a.dense[static_cast<long long>(y0 + r) * nx + x];
      acc[x * ls + r] = acc[x * ls + r] + dd.a;
      acc[x * ls + rows + r] = acc[x * ls + rows + r] + dd.b;
    }
    __syncthreads();
  }

  // 3. reflection products: row IFFT and gather of the interface pixels
  if (a.nr > 0 && a.rowptr && a.sp_store) {
    const int e0 = a.rowptr[y0], e1 = a.rowptr[y0 + nrow];
    for (int o = 0; o < a.nr; ++o) {
      load_rows(work, a.prod + (a.nt + o) * g.plane, y0, nrow, nx, rows, ls);
      __syncthreads();
      fft_lines<COMP, true, LEAN>(work, ls, lines, g.plan);
      for (int e = e0 + threadIdx.x; e < e1; e += kThreads) {
        if (a.sp_rid[e] != o) continue;
        const int pix = a.sp_pix[e];
        const int r = pix / nx - y0, x = pix - (pix / nx) * nx;
        a.sp_store[e] = px2<T>{work[x * ls + r], work[x * ls + rows + r]};
      }
      __syncthreads();
    }
  }

  // 4. real-space outputs
  if (a.e_out || a.e_out2) {
    for (int i = threadIdx.x; i < count; i += kThreads) {
      const int r = i / nx, x = i - r * nx;
      const long long q = static_cast<long long>(y0 + r) * nx + x;
      const px2<T> v{acc[x * ls + r], acc[x * ls + rows + r]};
      if (a.e_out) a.e_out[q] = v;
      if (a.e_out2) a.e_out2[q] = v;
    }
  }

  // 5. row FFT for the next slab step
  if (a.x_out) {
    fft_lines<COMP, false, LEAN>(acc, ls, lines, g.plan);
    for (int i = threadIdx.x; i < count; i += kThreads) {
      const int r = i / nx, x = i - r * nx;
      a.x_out[static_cast<long long>(y0 + r) * nx + x] = widen<ST>(px2<T>{acc[x * ls + r], acc[x * ls + rows + r]});
    }
  }
}

// ---- operator tables (FP64 math, stored as T hi + lo) ---------------------------------------

template <typename T>
__device__ __forceinline__ void split(zd v, cx<T>* hi, cx<T>* lo) {
  const T hx = static_cast<T>(v.x), hy = static_cast<T>(v.y);
  *hi = cx<T>{hx, hy};
  *lo = cx<T>{static_cast<T>(v.x - static_cast<double>(hx)), static_cast<T>(v.y - static_cast<double>(hy))};
}

// Homogeneous slab (index(r) = index(s) = n): exp(i k_z dz)/N.
template <typename T>
__global__ void k_build_phase(const double* kx, const double* ky, int ny, int nx, zd n, double k0,
                              double dz, double scale, cx<T>* hi, cx<T>* lo) {
  const long long k = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
  if (k >= static_cast<long long>(ny) * nx) return;
  const int y = static_cast<int>(k / nx), x = static_cast<int>(k - static_cast<long long>(y) * nx);
  const double kp2 = kx[x] * kx[x] + ky[y] * ky[y];
  const zd p = zmul(propagation(kz_of(n, k0, kp2), dz, 1.0), scale);
  split(p, hi + k, lo + k);
}

// Interface pair: M_{kind, r, s}(k)/N of table 1.
template <typename T>
__global__ void k_build_matrix(const double* kx, const double* ky, int ny, int nx, int kind, zd nr,
                               zd ns, double k0, double dz, double scale, cx<T>* hi, cx<T>* lo) {
  const long long k = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
  if (k >= static_cast<long long>(ny) * nx) return;
  const int y = static_cast<int>(k / nx), x = static_cast<int>(k - static_cast<long long>(y) * nx);
  const Mat2 m = operator_matrix(kind, kx[x], ky[y], nr, ns, k0, dz);
  split(zmul(m.m00, scale), hi + 3 * k, lo + 3 * k);
  split(zmul(m.m01, scale), hi + 3 * k + 1, lo + 3 * k + 1);
  split(zmul(m.m11, scale), hi + 3 * k + 2, lo + 3 * k + 2);
}

// Workgroups are dispatched to the eight XCDs round-robin. Column groups that share cache lines
// (W·sizeof(px2) < 128 B) are mapped so that neighbouring groups run on the same XCD and its L2:
// XCD j takes the j-th contiguous eighth of the groups.
__device__ __forceinline__ int xcd_group(int b, int n) {
  return (n & 7) ? b : (b & 7) * (n >> 3) + (b >> 3);
}

// ---- small elementwise helpers ---------------------------------------------------------------

template <typename T>
__global__ void k_from_c128(const zd* src, px2<T>* dst, long long npix) {
  const long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
  if (i >= npix) return;
  const zd a = src[2 * i], b = src[2 * i + 1];
  dst[i] = px2<T>{cx<T>{static_cast<T>(a.x), static_cast<T>(a.y)}, cx<T>{static_cast<T>(b.x), static_cast<T>(b.y)}};
}

template <typename T>
__global__ void k_to_c128(const px2<T>* src, zd* dst, long long npix) {
  const long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
  if (i >= npix) return;
  const px2<T> v = src[i];
  dst[2 * i] = zd{static_cast<double>(v.a.x), static_cast<double>(v.a.y)};
  dst[2 * i + 1] = zd{static_cast<double>(v.b.x), static_cast<double>(v.b.y)};
}

template <typename T>
__global__ void k_fill_zero(px2<T>* dst, long long npix) {
  const long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
  if (i < npix) dst[i] = px2<T>{};
}

}  // namespace bwpm
