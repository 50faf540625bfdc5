// Slab-step kernels for power-of-two grids: register-resident FFT (fft_reg.hpp) with one field
// component per thread. See kernels.hpp for how one slab step is split over the column and the
// row kernel.
//
// Lanes 2i and 2i + 1 hold E_x and E_y of the same pixels, so a wavefront reads and writes whole
// (E_x, E_y) pixel vectors contiguously. Each lane owns a complex line in LDS for the FFT
// exchanges. The 2×2 interface operator needs the other component of the pixel, and gets it
// from the partner lane with a DPP quad permutation. Register economy: a thread keeps eight
// complex values. A column step with several products stores the column spectrum back over its
// own input positions of X and reloads it per product. A row step with several transmitted
// products merges the selected pixels of each product into the positions of product plane 0 (each
// thread touches only what it owns). Reflection products are handled before the transmitted ones.
#pragma once

#include "fft_reg.hpp"
#include "kernels.hpp"

namespace bwpm {

// ---- geometry shared by the kernels and their launches --------------------------------------

// Lines of one LDS bank group (lds_stride): column kernels with W columns (lanes
// (t·W + c)·2 + comp), row kernels (the two components of a row).
constexpr int col_group_lines(int w) { return 2 * w; }
constexpr int kRowGroupLines = 2;

// Static LDS of a kernel with `lines` component lines of n points (complex size csize), `group`
// of them side by side in a bank group, with room for the staged twiddles (n − rad[0] entries:
// plan_tw_len, at most n).
constexpr long long fast_lds_bytes(int n, int lines, int group, int csize, bool comp) {
  return static_cast<long long>(lines) * lds_stride(n + n / 8, group, csize) * csize +
         static_cast<long long>(n) * csize * (comp ? 2 : 1);
}

// Lines (columns or rows) per workgroup: a power of two, 256 threads (n/4 per line) where the line
// allows (192 for lines of 3·2^k points), and at most 52 KB of LDS, so that three workgroups share
// a CU and hide each other's memory and barrier waits.
constexpr int fast_lines(int n, bool columns, int csize, bool comp) {
  int w = 1;
  while (2 * w * n <= 1024) w *= 2;
  while (w > 1 && fast_lds_bytes(n, 2 * w, columns ? col_group_lines(w) : kRowGroupLines, csize, comp) > 53248)
    w /= 2;
  return w;
}

// Twiddles are staged in LDS when the line buffers leave room for them (160 KB per workgroup);
// the 4096-point FP64 lines read them from global memory.
template <int N, int LINES, int GROUP, typename T, bool COMP>
constexpr bool fast_tw_in_lds() {
  return fast_lds_bytes(N, LINES, GROUP, static_cast<int>(sizeof(cx<T>)), COMP) <= 163840;
}

// ---- lane-pair exchange (lanes 2i ↔ 2i + 1): DPP quad permutation [1, 0, 3, 2] ---------------

__device__ __forceinline__ int pair_swap(int v) {
  return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);
}
__device__ __forceinline__ float pair_swap(float v) {
  return __int_as_float(pair_swap(__float_as_int(v)));
}
__device__ __forceinline__ double pair_swap(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = pair_swap(static_cast<int>(b & 0xffffffffLL));
  const int hi = pair_swap(static_cast<int>(b >> 32));
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}
template <typename T>
__device__ __forceinline__ cx<T> pair_swap(cx<T> z) {
  return cx<T>{pair_swap(z.x), pair_swap(z.y)};
}

// Operator product on one component (comp 0: E_x, 1: E_y) of spectrum entry k. The phase type is
// the same for both components; the symmetric matrix type takes the partner lane's component:
// E_x' = m00·E_x + m01·E_y, E_y' = m01·E_x + m11·E_y (terms in this order). Every lane of the
// wavefront must call it with the same product (the swap reads the partner's register).
template <typename T, bool COMP>
__device__ __forceinline__ cx<T> apply_prod1(const Prod<T>& pd, unsigned k, int comp, cx<T> e) {
  if (pd.type == 0) {
    cx<T> pl{};
    if constexpr (COMP) pl = gload(pd.lo, k);
    return cmul2<COMP>(e, gload(pd.hi, k), pl);
  }
  const cx<T> o = pair_swap(e);
  const cx<T> ea = comp ? o : e;
  const cx<T> eb = comp ? e : o;
  const unsigned i = 3u * k + static_cast<unsigned>(comp);
  cx<T> l0{}, l1{};
  if constexpr (COMP) {
    l0 = gload(pd.lo, i);
    l1 = gload(pd.lo, i + 1);
  }
  return cmul2<COMP>(ea, gload(pd.hi, i), l0) + cmul2<COMP>(eb, gload(pd.hi, i + 1), l1);
}

// Column kernel: W adjacent columns per workgroup, N/8 threads per column and component (thread
// index (t·W + c)·2 + comp, so a wavefront reads W·sizeof(px2) contiguous bytes per row).
// MULTI (launched for steps with several products): each thread keeps its eight spectrum values in
// an LDS stash across the products instead of writing them back to X and reloading them. The
// arithmetic is the same, so both forms give the same bits where X is stored in the arithmetic
// precision (ST = T); with narrower storage only the LDS stash keeps the spectrum unrounded.
template <typename PL, int W, typename T, bool COMP, bool MULTI, typename ST = T>
__global__ __launch_bounds__(PL::N / 4 * W) void k_col_fast(const ColDesc<T, ST>* __restrict__ descs, int nx,
                                                        long long plane, const cx<T>* __restrict__ g_twh,
                                                        const cx<T>* __restrict__ g_twl) {
  constexpr int N = PL::N;
  constexpr int TT = N / 8;
  constexpr int CS = static_cast<int>(sizeof(cx<T>));
  constexpr int LS = lds_stride(reg_lds_len<N>(), col_group_lines(W), CS);
  constexpr int SS = lds_stride(N, col_group_lines(W), CS);  // stash line stride
  constexpr bool TWL = fast_tw_in_lds<N, 2 * W, col_group_lines(W), T, COMP>();
  __shared__ cx<T> sh[2 * W * LS];
  __shared__ cx<T> tw_hi_s[TWL ? plan_tw_len<PL>() : 1];
  __shared__ cx<T> tw_lo_s[TWL && COMP ? plan_tw_len<PL>() : 1];
  __shared__ cx<T> stash[MULTI ? 2 * W * SS : 1];
  const ColDesc<T, ST> d = descs[blockIdx.y];
  if (!d.in) return;
  if constexpr (TWL) stage_twiddles<PL, COMP>(tw_hi_s, tw_lo_s, g_twh, g_twl);
  const cx<T>* tw_hi = TWL ? tw_hi_s : g_twh;
  const cx<T>* tw_lo = TWL ? tw_lo_s : g_twl;
  const int comp = threadIdx.x & 1;
  const int q = threadIdx.x >> 1;
  const int c = q % W, t = q / W;
  const unsigned x = xcd_group(blockIdx.x, gridDim.x) * W + c;
  cx<T>* line = sh + (2 * c + comp) * LS;
  px2<ST>* xio = const_cast<px2<ST>*>(d.in);
  // pixel index of row t + T·m and complex index of this lane's component (planes < 2^31 values)
  auto pix = [&](int m) { return static_cast<unsigned>(t + TT * m) * static_cast<unsigned>(nx) + x; };
  auto at = [&](int m) { return 2u * pix(m) + static_cast<unsigned>(comp); };
  const int nprod = uniform(d.nprod);
  // the rows' product use, for the interface variant's stores (all ones: store every row)
  unsigned use[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) use[m] = ~0u;
  if constexpr (MULTI) {
    if (d.rowuse) {
#pragma unroll
      for (int m = 0; m < 8; ++m) use[m] = gload1(d.rowuse, t + TT * m);
    }
  }
  cx<T> s[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) s[m] = sload<T>(xio, at(m));
  reg_fft<PL, COMP, false, TWL>(s, line, t, tw_hi, tw_lo);
  // several products: keep the spectrum (in the LDS stash, or in the thread's own positions of X);
  // stash entry m of this lane is own[m·T] on its component line
  cx<T>* own = stash;
  if constexpr (MULTI) own = stash + (2 * c + comp) * SS + t;
  if (nprod > 1) {
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      if constexpr (MULTI)
        own[m * TT] = s[m];
      else
        sstore<T>(xio, at(m), s[m]);
    }
  }
  for (int o = 0; o < nprod; ++o) {
    if (o > 0) {
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        if constexpr (MULTI)
          s[m] = own[m * TT];
        else
          s[m] = sload<T>(xio, at(m));
      }
    }
    const Prod<T>* pp = d.prods + o;
    const Prod<T> pd{uniform(gload1(&pp->type, 0)), uniform(gload1(&pp->hi, 0)), uniform(gload1(&pp->lo, 0))};
#pragma unroll
    for (int m = 0; m < 8; ++m) s[m] = apply_prod1<T, COMP>(pd, pix(m), comp, s[m]);
    reg_fft<PL, COMP, true, TWL>(s, line, t, tw_hi, tw_lo);
    px2<ST>* out = d.out + o * plane;
    if constexpr (MULTI) {
      // a product is stored only in the rows whose row transforms use it: rows with a pixel that
      // selects T product o, rows with interface pixels for the R products (the row kernel
      // transforms each row on its own, so values of the other rows never reach a used pixel).
      // Only the interface variant does this; the plain variant keeps its registers.
      const unsigned bit = o >= d.nt ? static_cast<unsigned>(kRowUseR) : row_use_bit(o);
#pragma unroll
      for (int m = 0; m < 8; ++m)
        if ((use[m] >> bit) & 1u) sstore<T>(out, at(m), s[m]);
    } else {
#pragma unroll
      for (int m = 0; m < 8; ++m) sstore<T>(out, at(m), s[m]);
    }
  }
}

// Row kernel: RW rows per workgroup, N/8 threads per row and component (thread index
// (r·N/8 + t)·2 + comp). MULTI (launched for steps with several transmitted products): the
// selected pixels of each product are merged in an LDS stash instead of product plane 0.
template <typename PL, int RW, typename T, bool COMP, bool MULTI, typename ST = T>
__global__ __launch_bounds__(PL::N / 4 * RW) void k_row_fast(const RowDesc<T, ST>* __restrict__ descs, long long plane,
                                                         const cx<T>* __restrict__ g_twh,
                                                         const cx<T>* __restrict__ g_twl) {
  constexpr int N = PL::N;
  constexpr int TT = N / 8;
  constexpr int CS = static_cast<int>(sizeof(cx<T>));
  constexpr int LS = lds_stride(reg_lds_len<N>(), kRowGroupLines, CS);
  constexpr int SS = lds_stride(N, kRowGroupLines, CS);  // stash line stride
  constexpr bool TWL = fast_tw_in_lds<N, 2 * RW, kRowGroupLines, T, COMP>();
  __shared__ cx<T> sh[2 * RW * LS];
  __shared__ cx<T> tw_hi_s[TWL ? plan_tw_len<PL>() : 1];
  __shared__ cx<T> tw_lo_s[TWL && COMP ? plan_tw_len<PL>() : 1];
  __shared__ cx<T> stash[MULTI ? 2 * RW * SS : 1];
  const RowDesc<T, ST> a = descs[blockIdx.y];
  if (!a.active) return;
  if constexpr (TWL) stage_twiddles<PL, COMP>(tw_hi_s, tw_lo_s, g_twh, g_twl);
  const cx<T>* tw_hi = TWL ? tw_hi_s : g_twh;
  const cx<T>* tw_lo = TWL ? tw_lo_s : g_twl;
  const int comp = threadIdx.x & 1;
  const int q = threadIdx.x >> 1;
  const int r = q / TT, t = q % TT;
  const int y = blockIdx.x * RW + r;
  const unsigned row = static_cast<unsigned>(y) * N;
  cx<T>* line = sh + (2 * r + comp) * LS;
  // complex index of (pixel, component) of this lane (planes < 2^31 values)
  auto at = [&](int m) { return 2u * (row + t + TT * m) + static_cast<unsigned>(comp); };
  auto cplane = [](const px2<T>* p) { return reinterpret_cast<const cx<T>*>(p); };
  cx<T> acc[8];
  // Work a product does only for rows that use it is skipped for the whole workgroup when none of
  // its rows does (the conditions are read from memory the kernel does not write, so every thread
  // evaluates them equally and the barriers inside the transforms stay uniform): transmitted
  // products no pixel of these rows selects, and the reflection products when these rows hold no
  // interface pixel. Those products' values would only be stored at pixels of other rows, so the
  // results are the same bits.
  int use = -1;
  if (a.rowuse) {
    unsigned u = 0;
#pragma unroll
    for (int i = 0; i < RW; ++i) u |= gload1(a.rowuse, blockIdx.x * RW + i);
    use = uniform(static_cast<int>(u));
  }

  // 1. reflection products: row IFFT and gather of the interface pixels
  if (a.nr > 0 && a.rowptr && a.sp_store && ((static_cast<unsigned>(use) >> kRowUseR) & 1u)) {
    const int e0 = gload1(a.rowptr, y), e1 = gload1(a.rowptr, y + 1);
    cx<T>* store = reinterpret_cast<cx<T>*>(a.sp_store);
    for (int o = 0; o < a.nr; ++o) {
      const px2<ST>* src = a.prod + (a.nt + o) * plane;
#pragma unroll
      for (int m = 0; m < 8; ++m) acc[m] = sload<T>(src, at(m));
      reg_fft<PL, COMP, true, TWL>(acc, line, t, tw_hi, tw_lo);
#pragma unroll
      for (int m = 0; m < 8; ++m) line[lpad(t + TT * m)] = acc[m];
      __syncthreads();
      for (int e = e0 + t; e < e1; e += TT)
        if (gload1(a.sp_rid, e) == o) gstore(store, 2u * e + comp, line[lpad(gload1(a.sp_pix, e) - y * N)]);
      __syncthreads();
    }
  }

  // 2. transmitted field (or the start plane)
  if (a.nt == 0) {
    const cx<T>* ein = cplane(a.e_in);
#pragma unroll
    for (int m = 0; m < 8; ++m) acc[m] = ein ? gload(ein, at(m)) : cx<T>{};
  } else if (a.nt == 1 || single_t(use)) {
    // one transmitted product serves every pixel of these rows: no merge
    const int o = a.nt == 1 ? 0 : __builtin_ctz(static_cast<unsigned>(use) & kRowUseT);
#pragma unroll
    for (int m = 0; m < 8; ++m) acc[m] = sload<T>(a.prod + o * plane, at(m));
    reg_fft<PL, COMP, true, TWL>(acc, line, t, tw_hi, tw_lo);
  } else {
    // the selected pixels are collected in the LDS stash or in product plane 0
    px2<ST>* merged = const_cast<px2<ST>*>(a.prod);
    cx<T>* own = stash;  // stash entry m of this lane is own[m·T]
    if constexpr (MULTI) own = stash + (2 * r + comp) * SS + t;
    for (int o = 0; o < a.nt; ++o) {
      if (!((static_cast<unsigned>(use) >> row_use_bit(o)) & 1u)) continue;
      const px2<ST>* src = a.prod + o * plane;
#pragma unroll
      for (int m = 0; m < 8; ++m) acc[m] = sload<T>(src, at(m));
      reg_fft<PL, COMP, true, TWL>(acc, line, t, tw_hi, tw_lo);
#pragma unroll
      for (int m = 0; m < 8; ++m)
        if (gload1(a.tsel, row + t + TT * m) == o) {
          if constexpr (MULTI)
            own[m * TT] = acc[m];
          else
            sstore<T>(merged, at(m), acc[m]);
        }
    }
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      if constexpr (MULTI)
        acc[m] = own[m * TT];
      else
        acc[m] = sload<T>(merged, at(m));
    }
  }

  // 3. epilogue in the order of the reference: mask, Born increment, sparse sources, Born plane
  const cx<T>* inc = cplane(a.inc);
  if (a.mask || a.mask_y) {
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const T mk = mask_factor(a, static_cast<unsigned>(y), static_cast<unsigned>(t + TT * m), row + t + TT * m);
      acc[m] = acc[m] * mk;
      if (inc) acc[m] = acc[m] + gload(inc, at(m)) * mk;
    }
  } else if (inc) {
#pragma unroll
    for (int m = 0; m < 8; ++m) acc[m] = acc[m] + gload(inc, at(m));
  }
  if (a.rowptr && a.sp_add) {
    const cx<T>* add = cplane(a.sp_add);
#pragma unroll
    for (int m = 0; m < 8; ++m) line[lpad(t + TT * m)] = acc[m];
    __syncthreads();
    const int e1 = gload1(a.rowptr, y + 1);
    for (int e = gload1(a.rowptr, y) + t; e < e1; e += TT) {
      const int xx = lpad(gload1(a.sp_pix, e) - y * N);
      line[xx] = line[xx] + gload(add, 2u * e + comp);
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 8; ++m) acc[m] = line[lpad(t + TT * m)];
    __syncthreads();
  }
  if (a.dense) {
    const cx<T>* dense = cplane(a.dense);
#pragma unroll
    for (int m = 0; m < 8; ++m) acc[m] = acc[m] + gload(dense, at(m));
  }

  // 4. real-space outputs
  if (a.e_out) {
    cx<T>* o = reinterpret_cast<cx<T>*>(a.e_out);
#pragma unroll
    for (int m = 0; m < 8; ++m) gstore(o, at(m), acc[m]);
  }
  if (a.e_out2) {
    cx<T>* o = reinterpret_cast<cx<T>*>(a.e_out2);
#pragma unroll
    for (int m = 0; m < 8; ++m) gstore(o, at(m), acc[m]);
  }

  // 5. row FFT for the next slab step
  if (a.x_out) {
    reg_fft<PL, COMP, false, TWL>(acc, line, t, tw_hi, tw_lo);
#pragma unroll
    for (int m = 0; m < 8; ++m) sstore<T>(a.x_out, at(m), acc[m]);
  }
}

}  // namespace bwpm
