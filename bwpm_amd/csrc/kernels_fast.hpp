// Slab-step kernels for power-of-two grids (register-resident FFT; see kernels.hpp for the split
// of one slab step over the column and the row kernel, and fft_reg.hpp for the FFT).
//
// Register economy: each thread keeps one array of eight pixel vectors. A column step with several
// products stores the column spectrum back over its own input positions of X and reloads it per
// product; a row step with several transmitted products merges the selected pixels of each product
// into the positions of product plane 0 (each thread touches only the pixels it owns); reflection
// products are handled before the transmitted ones. Twiddles are staged in LDS once per workgroup.
#pragma once

#include "fft_reg.hpp"
#include "kernels.hpp"

namespace bwpm {

// Column kernel: W adjacent columns per workgroup, N/8 threads per column (thread index t·W + c,
// so a wavefront reads W·sizeof(px2) contiguous bytes per row).
template <int N, int W, typename T, bool COMP>
__global__ __launch_bounds__(N / 8 * W) void k_col_fast(const ColDesc<T>* __restrict__ descs, int nx,
                                                        long long plane, const cx<T>* __restrict__ g_twh,
                                                        const cx<T>* __restrict__ g_twl) {
  constexpr int TT = N / 8;
  constexpr int LL = reg_lds_len<N>();
  __shared__ px2<T> sh[W * LL];
  __shared__ cx<T> tw_hi[N];
  __shared__ cx<T> tw_lo[COMP ? N : 1];
  const ColDesc<T> d = descs[blockIdx.y];
  if (!d.in) return;
  stage_twiddles<N, COMP>(tw_hi, tw_lo, g_twh, g_twl);
  const int c = threadIdx.x % W, t = threadIdx.x / W;
  const int x = blockIdx.x * W + c;
  px2<T>* line = sh + c * LL;
  px2<T>* xin = const_cast<px2<T>*>(d.in);
  px2<T> s[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) s[m] = d.in[static_cast<long long>(t + TT * m) * nx + x];
  reg_fft<N, COMP, false>(s, line, t, tw_hi, tw_lo);
  if (d.nprod > 1) {  // keep the spectrum in the thread's own positions of X
#pragma unroll
    for (int m = 0; m < 8; ++m) xin[static_cast<long long>(t + TT * m) * nx + x] = s[m];
  }
  for (int o = 0; o < d.nprod; ++o) {
    if (o > 0) {
#pragma unroll
      for (int m = 0; m < 8; ++m) s[m] = xin[static_cast<long long>(t + TT * m) * nx + x];
    }
    const Prod<T> pd = d.prods[o];
#pragma unroll
    for (int m = 0; m < 8; ++m) s[m] = apply_prod<T, COMP>(pd, static_cast<long long>(t + TT * m) * nx + x, s[m]);
    reg_fft<N, COMP, true>(s, line, t, tw_hi, tw_lo);
    px2<T>* out = d.out + o * plane;
#pragma unroll
    for (int m = 0; m < 8; ++m) out[static_cast<long long>(t + TT * m) * nx + x] = s[m];
  }
}

// Row kernel: RW rows per workgroup, N/8 threads per row (thread index r·N/8 + t).
template <int N, int RW, typename T, bool COMP>
__global__ __launch_bounds__(N / 8 * RW) void k_row_fast(const RowDesc<T>* __restrict__ descs, long long plane,
                                                         const cx<T>* __restrict__ g_twh,
                                                         const cx<T>* __restrict__ g_twl) {
  constexpr int TT = N / 8;
  constexpr int LL = reg_lds_len<N>();
  __shared__ px2<T> sh[RW * LL];
  __shared__ cx<T> tw_hi[N];
  __shared__ cx<T> tw_lo[COMP ? N : 1];
  const RowDesc<T> a = descs[blockIdx.y];
  if (!a.active) return;
  stage_twiddles<N, COMP>(tw_hi, tw_lo, g_twh, g_twl);
  const int r = threadIdx.x / TT, t = threadIdx.x % TT;
  const int y = blockIdx.x * RW + r;
  const long long row = static_cast<long long>(y) * N;
  px2<T>* line = sh + r * LL;
  px2<T> acc[8];

  // 1. reflection products: row IFFT and gather of the interface pixels
  if (a.nr > 0 && a.rowptr && a.sp_store) {
    const int e0 = a.rowptr[y], e1 = a.rowptr[y + 1];
    for (int o = 0; o < a.nr; ++o) {
      const px2<T>* src = a.prod + (a.nt + o) * plane;
#pragma unroll
      for (int m = 0; m < 8; ++m) acc[m] = src[row + t + TT * m];
      reg_fft<N, COMP, true>(acc, line, t, tw_hi, tw_lo);
#pragma unroll
      for (int m = 0; m < 8; ++m) line[lpad(t + TT * m)] = acc[m];
      __syncthreads();
      for (int e = e0 + t; e < e1; e += TT)
        if (a.sp_rid[e] == o) a.sp_store[e] = line[lpad(a.sp_pix[e] - y * N)];
      __syncthreads();
    }
  }

  // 2. transmitted field (or the start plane)
  if (a.nt == 0) {
#pragma unroll
    for (int m = 0; m < 8; ++m) acc[m] = a.e_in ? a.e_in[row + t + TT * m] : px2<T>{};
  } else if (a.nt == 1) {
#pragma unroll
    for (int m = 0; m < 8; ++m) acc[m] = a.prod[row + t + TT * m];
    reg_fft<N, COMP, true>(acc, line, t, tw_hi, tw_lo);
  } else {
    px2<T>* merged = const_cast<px2<T>*>(a.prod);  // product plane 0 collects the selected pixels
    for (int o = 0; o < a.nt; ++o) {
      const px2<T>* src = a.prod + o * plane;
#pragma unroll
      for (int m = 0; m < 8; ++m) acc[m] = src[row + t + TT * m];
      reg_fft<N, COMP, true>(acc, line, t, tw_hi, tw_lo);
#pragma unroll
      for (int m = 0; m < 8; ++m)
        if (a.tsel[row + t + TT * m] == o) merged[row + t + TT * m] = acc[m];
    }
#pragma unroll
    for (int m = 0; m < 8; ++m) acc[m] = merged[row + t + TT * m];
  }

  // 3. epilogue in the order of the reference: mask, Born increment, sparse sources, Born plane
  if (a.mask) {
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const T mk = a.mask[row + t + TT * m];
      acc[m].a = acc[m].a * mk;
      acc[m].b = acc[m].b * mk;
      if (a.inc) {
        const px2<T> dd = a.inc[row + t + TT * m];
        acc[m].a = acc[m].a + dd.a * mk;
        acc[m].b = acc[m].b + dd.b * mk;
      }
    }
  } else if (a.inc) {
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const px2<T> dd = a.inc[row + t + TT * m];
      acc[m].a = acc[m].a + dd.a;
      acc[m].b = acc[m].b + dd.b;
    }
  }
  if (a.rowptr && a.sp_add) {
#pragma unroll
    for (int m = 0; m < 8; ++m) line[lpad(t + TT * m)] = acc[m];
    __syncthreads();
    const int e1 = a.rowptr[y + 1];
    for (int e = a.rowptr[y] + t; e < e1; e += TT) {
      const int x = a.sp_pix[e] - y * N;
      const px2<T> v = a.sp_add[e];
      px2<T> u = line[lpad(x)];
      u.a = u.a + v.a;
      u.b = u.b + v.b;
      line[lpad(x)] = u;
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 8; ++m) acc[m] = line[lpad(t + TT * m)];
    __syncthreads();
  }
  if (a.dense) {
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const px2<T> dd = a.dense[row + t + TT * m];
      acc[m].a = acc[m].a + dd.a;
      acc[m].b = acc[m].b + dd.b;
    }
  }

  // 4. real-space outputs
  if (a.e_out) {
#pragma unroll
    for (int m = 0; m < 8; ++m) a.e_out[row + t + TT * m] = acc[m];
  }
  if (a.e_out2) {
#pragma unroll
    for (int m = 0; m < 8; ++m) a.e_out2[row + t + TT * m] = acc[m];
  }

  // 5. row FFT for the next slab step
  if (a.x_out) {
    reg_fft<N, COMP, false>(acc, line, t, tw_hi, tw_lo);
#pragma unroll
    for (int m = 0; m < 8; ++m) a.x_out[row + t + TT * m] = acc[m];
  }
}

}  // namespace bwpm
