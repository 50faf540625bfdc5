// Persistent slab-step kernels for the plain steps of a lane: steps where every solve of the launch
// has one product, a homogeneous phase, in the column kernel, and one transmitted product with no
// reflections, sparse sources or Born planes in the row kernel (the mask, kept planes and pass-end
// outputs are allowed). A single 2×2 operator stays with the general column kernel: its partner
// swap and two table entries per value would cost this kernel a wave of occupancy. These are most steps of a solve; the steps at
// interfaces, pass starts and Born slabs run the general kernels of kernels_fast.hpp.
//
// A launch has about as many workgroups as the device holds at once. Workgroup b takes the line
// tiles b, b + P, b + 2P, … of all solves (P workgroups), and issues the global loads of its next
// tile before it transforms the current one: the HBM latency of the next tile overlaps the FFT
// passes, the operator and the stores of this one. The twiddles are staged in LDS once per
// workgroup instead of once per tile. Per tile the arithmetic is exactly that of the general
// kernels (same transforms, same operator products in the same order), so both give the same bits.
//
// Column tiles are W adjacent columns of one solve, G column groups per solve. Workgroups are
// dispatched to the eight XCDs in turn (b mod 8), and the host makes the grid a multiple of 8, so
// every tile of a workgroup has the workgroup's residue. Tile u goes to XCD x = u mod 8, which takes
// the column groups [x·G/8, (x+1)·G/8) of every solve in turn: neighbouring groups, which share the
// 128-byte lines of X, meet in that XCD's L2, and so does the XCD's slice of the phase table, which
// the solves of a lane with one structure share (one table in HBM), from one solve to the next.
// (The mapping only affects speed; any dispatch order gives the same results.)
#pragma once

#include "kernels_fast.hpp"

namespace bwpm {

// One complex value of a streaming plane as stored (ST), loaded without conversion, so a prefetch
// holds storage words only; decoded with widen<T>.
template <typename ST>
__device__ __forceinline__ cx<ST> sraw(const px2<ST>* p, unsigned i) {
  return gload(reinterpret_cast<const cx<ST>*>(p), i);
}
template <>
__device__ __forceinline__ cx<bfp> sraw<bfp>(const px2<bfp>* p, unsigned i) {
  typedef unsigned u2 __attribute__((ext_vector_type(2)));
  const u2 v = *gaddr<px2<bfp>, u2>(p, i);
  return cx<bfp>{bfp{v.x}, bfp{v.y}};
}

// XCDs of the device; workgroup b of a launch runs on XCD b mod kXcds.
constexpr int kXcds = 8;

// Tile order of the column kernel (see the file comment): linear tile (solve·G + group) of
// dispatch index u, for G column groups per solve. A bijection of [0, S·G) when G is a multiple of
// 8 (the identity otherwise).
__device__ __host__ __forceinline__ int xcd_tile(int u, int groups) {
  if (groups & 7) return u;
  const int g8 = groups >> 3, k = u >> 3;
  return (k / g8) * groups + (u & 7) * g8 + k % g8;
}

template <typename PL, int W, typename T, bool COMP, typename ST>
__global__ __launch_bounds__(PL::N / 4 * W) void k_col_plain(const ColDesc<T, ST>* __restrict__ descs, int S,
                                                         int groups, int nx, const cx<T>* __restrict__ g_twh,
                                                         const cx<T>* __restrict__ g_twl) {
  constexpr int N = PL::N;
  constexpr int TT = N / 8;
  constexpr int CS = static_cast<int>(sizeof(cx<T>));
  constexpr int LS = lds_stride(reg_lds_len<N>(), col_group_lines(W), CS);
  constexpr bool TWL = fast_tw_in_lds<N, 2 * W, col_group_lines(W), T, COMP>();
  __shared__ cx<T> sh[2 * W * LS];
  __shared__ cx<T> tw_hi_s[TWL ? plan_tw_len<PL>() : 1];
  __shared__ cx<T> tw_lo_s[TWL && COMP ? plan_tw_len<PL>() : 1];
  const int ntiles = S * groups;
  // the next tile (dispatch index) of this workgroup whose solve has column work; tile indices are
  // uniform, so descriptors are read with scalar loads (which do not wait behind vector loads)
  auto advance = [&](int u) {
    for (u = uniform(u); u < ntiles; u = uniform(u + static_cast<int>(gridDim.x)))
      if (descs[xcd_tile(u, groups) / groups].in != nullptr) break;
    return u;
  };
  int u = advance(static_cast<int>(blockIdx.x));
  if (u >= ntiles) return;
  if constexpr (TWL) stage_twiddles<PL, COMP>(tw_hi_s, tw_lo_s, g_twh, g_twl);
  const cx<T>* tw_hi = TWL ? tw_hi_s : g_twh;
  const cx<T>* tw_lo = TWL ? tw_lo_s : g_twl;
  const int comp = threadIdx.x & 1;
  const int q = threadIdx.x >> 1;
  const int c = q % W, t = q / W;
  cx<T>* line = sh + (2 * c + comp) * LS;
  // pixel index of row t + T·m in column x, and the complex index of this lane's component
  auto pix = [&](unsigned x, int m) { return static_cast<unsigned>(t + TT * m) * static_cast<unsigned>(nx) + x; };
  auto at = [&](unsigned x, int m) { return 2u * pix(x, m) + static_cast<unsigned>(comp); };
  auto load = [&](int uu, cx<ST> (&r)[8]) {
    const int lin = xcd_tile(uu, groups);
    const px2<ST>* in = descs[lin / groups].in;
    const unsigned x = static_cast<unsigned>((lin % groups) * W + c);
#pragma unroll
    for (int m = 0; m < 8; ++m) r[m] = sraw(in, at(x, m));
  };
  // one tile from `cur`, with the next tile's loads issued into `nxt`; the loop alternates the two
  // buffers (no register copy, which would wait for the loads, and for the stores behind them)
  auto step = [&](cx<ST> (&cur)[8], cx<ST> (&nxt)[8]) {
    const int lin = xcd_tile(u, groups);
    const ColDesc<T, ST>& d = descs[lin / groups];
    // the product as scalars (table loads then address as base + offset)
    const Prod<T>* pp = d.prods;
    const Prod<T> pd{uniform(gload1(&pp->type, 0)), uniform(gload1(&pp->hi, 0)), uniform(gload1(&pp->lo, 0))};
    px2<ST>* out = d.out;
    const unsigned x = static_cast<unsigned>((lin % groups) * W + c);
    cx<T> v[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) v[m] = widen<T>(cur[m]);
    reg_fft<PL, COMP, false, TWL>(v, line, t, tw_hi, tw_lo);
    // the phase values are loaded before the next tile's prefetch is issued, so that waiting for
    // them does not wait for the prefetch (loads complete in order); the prefetch then overlaps the
    // operator, the inverse transform and the stores. The host launches this kernel for phase
    // products only (type 0); the test of the type also keeps the table loads and the product in
    // blocks of their own, which the register allocation needs (160 instead of 192 VGPRs for FP64
    // arithmetic, i.e. three waves per SIMD instead of two).
    cx<T> ph[8], pl[8];
    if (pd.type == 0) {
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        ph[m] = gload(pd.hi, pix(x, m));
        pl[m] = cx<T>{};
        if constexpr (COMP) pl[m] = gload(pd.lo, pix(x, m));
      }
    }
    const int un = advance(u + static_cast<int>(gridDim.x));
    if (un < ntiles) load(un, nxt);
    if (pd.type == 0) {
#pragma unroll
      for (int m = 0; m < 8; ++m) v[m] = cmul2<COMP>(v[m], ph[m], pl[m]);
    }
    reg_fft<PL, COMP, true, TWL>(v, line, t, tw_hi, tw_lo);
#pragma unroll
    for (int m = 0; m < 8; ++m) sstore<T>(out, at(x, m), v[m]);
    u = un;
  };
  cx<ST> a[8], b[8];
  load(u, a);
  for (;;) {
    step(a, b);
    if (u >= ntiles) break;
    step(b, a);
    if (u >= ntiles) break;
  }
}

template <typename PL, int RW, typename T, bool COMP, typename ST>
__global__ __launch_bounds__(PL::N / 4 * RW) void k_row_plain(const RowDesc<T, ST>* __restrict__ descs, int S,
                                                          int groups, const cx<T>* __restrict__ g_twh,
                                                          const cx<T>* __restrict__ g_twl) {
  constexpr int N = PL::N;
  constexpr int TT = N / 8;
  constexpr int CS = static_cast<int>(sizeof(cx<T>));
  constexpr int LS = lds_stride(reg_lds_len<N>(), kRowGroupLines, CS);
  constexpr bool TWL = fast_tw_in_lds<N, 2 * RW, kRowGroupLines, T, COMP>();
  __shared__ cx<T> sh[2 * RW * LS];
  __shared__ cx<T> tw_hi_s[TWL ? plan_tw_len<PL>() : 1];
  __shared__ cx<T> tw_lo_s[TWL && COMP ? plan_tw_len<PL>() : 1];
  const int ntiles = S * groups;
  auto advance = [&](int u) {
    for (u = uniform(u); u < ntiles; u = uniform(u + static_cast<int>(gridDim.x)))
      if (descs[u / groups].active != 0) break;
    return u;
  };
  int u = advance(static_cast<int>(blockIdx.x));
  if (u >= ntiles) return;
  if constexpr (TWL) stage_twiddles<PL, COMP>(tw_hi_s, tw_lo_s, g_twh, g_twl);
  const cx<T>* tw_hi = TWL ? tw_hi_s : g_twh;
  const cx<T>* tw_lo = TWL ? tw_lo_s : g_twl;
  const int comp = threadIdx.x & 1;
  const int q = threadIdx.x >> 1;
  const int r = q / TT, t = q % TT;
  cx<T>* line = sh + (2 * r + comp) * LS;
  auto at = [&](unsigned y, int m) {
    return 2u * (y * N + static_cast<unsigned>(t + TT * m)) + static_cast<unsigned>(comp);
  };
  auto load = [&](int uu, cx<ST> (&rr)[8]) {
    const px2<ST>* src = descs[uu / groups].prod;
    const unsigned y = static_cast<unsigned>((uu % groups) * RW + r);
#pragma unroll
    for (int m = 0; m < 8; ++m) rr[m] = sraw(src, at(y, m));
  };
  auto step = [&](cx<ST> (&cur)[8], cx<ST> (&nxt)[8]) {
    const RowDesc<T, ST>& a = descs[u / groups];
    const unsigned y = static_cast<unsigned>((u % groups) * RW + r);
    const double* mask_y = a.mask_y;
    const double* mask_x = a.mask_x;
    const T* mask = a.mask;
    px2<T>* e_out = a.e_out;
    px2<T>* e_out2 = a.e_out2;
    px2<ST>* x_out = a.x_out;
    // the mask factors (as mask_factor) are loaded before the next tile's prefetch (loads
    // complete in order)
    T fac[8];
    if (mask_y) {
      const double my = gload1(mask_y, y);
#pragma unroll
      for (int m = 0; m < 8; ++m) fac[m] = static_cast<T>(my * gload1(mask_x, static_cast<unsigned>(t + TT * m)));
    } else if (mask) {
#pragma unroll
      for (int m = 0; m < 8; ++m) fac[m] = gload1(mask, y * N + static_cast<unsigned>(t + TT * m));
    }
    const int un = advance(u + static_cast<int>(gridDim.x));
    if (un < ntiles) load(un, nxt);
    cx<T> acc[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) acc[m] = widen<T>(cur[m]);
    reg_fft<PL, COMP, true, TWL>(acc, line, t, tw_hi, tw_lo);
    if (mask_y || mask) {
#pragma unroll
      for (int m = 0; m < 8; ++m) acc[m] = acc[m] * fac[m];
    }
    if (e_out) {
      cx<T>* o = reinterpret_cast<cx<T>*>(e_out);
#pragma unroll
      for (int m = 0; m < 8; ++m) gstore(o, at(y, m), acc[m]);
    }
    if (e_out2) {
      cx<T>* o = reinterpret_cast<cx<T>*>(e_out2);
#pragma unroll
      for (int m = 0; m < 8; ++m) gstore(o, at(y, m), acc[m]);
    }
    if (x_out) {
      reg_fft<PL, COMP, false, TWL>(acc, line, t, tw_hi, tw_lo);
#pragma unroll
      for (int m = 0; m < 8; ++m) sstore<T>(x_out, at(y, m), acc[m]);
    }
    u = un;
  };
  cx<ST> a[8], b[8];
  load(u, a);
  for (;;) {
    step(a, b);
    if (u >= ntiles) break;
    step(b, a);
    if (u >= ntiles) break;
  }
}

}  // namespace bwpm
