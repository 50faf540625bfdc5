// Register-resident radix-8 Stockham FFT of power-of-two lines (16 … 4096 points) for the hot
// slab-step kernels.
//
// A line of N points is owned by T = N/8 threads. Thread t holds the eight pixel vectors
// a[t + T·m] (m = 0 … 7, both field components), which is exactly the read pattern of every
// Stockham pass of this plan: a pass of radix R lets each thread run 8/R butterflies
// j = t + T·b on v[b + r·8/R], multiplies by W_N^{r·(j mod Ns)·N/(Ns·R)}, applies the size-R DFT
// and writes y_r to LDS at (j − k)·R + k + r·Ns; the next pass reads a[t + T·m] back. The first
// and the last pass are radix 8, so the input is taken in natural order from registers and the
// last pass leaves A[t + T·r] in v[r] (natural order) without touching LDS: P − 1 LDS exchanges.
//
// In the px2 form the two field components are transformed one after the other inside each
// butterfly, which halves the live temporaries; the cx form transforms a single component. Twiddles
// are read from the table W_N^e in global memory, or (cx form) from a copy the kernel stages in LDS
// once, laid out per pass so that the lanes of a wavefront read consecutive entries.
// LDS index i is padded to i + i/8 to spread the stride-8 writes of the first pass over banks.
#pragma once

#include "fft.hpp"

namespace bwpm {

// LDS stride (complex entries of csize bytes) of lines of `len` entries. A 128-bit (64-bit) LDS
// access serves a wavefront's lanes in groups of g = 128 / csize (8 for complex128, 16 for
// complex64), and a group covers g / L consecutive positions of L lines side by side (L = 1 where
// the lanes of a group share a line). A stride ≡ g / L (mod g) puts the g lanes of a group in
// distinct bank groups.
constexpr int lds_stride(int len, int group_lines, int csize) {
  const int g = 128 / csize;
  const int target = g / (group_lines < g ? group_lines : g);
  return len + ((target - len % g) % g + g) % g;
}

template <int N>
struct RegPlan;
template <> struct RegPlan<16> { static constexpr int P = 2; static constexpr int rad[P] = {8, 2}; };
template <> struct RegPlan<32> { static constexpr int P = 2; static constexpr int rad[P] = {8, 4}; };
template <> struct RegPlan<64> { static constexpr int P = 2; static constexpr int rad[P] = {8, 8}; };
template <> struct RegPlan<128> { static constexpr int P = 3; static constexpr int rad[P] = {8, 2, 8}; };
template <> struct RegPlan<256> { static constexpr int P = 3; static constexpr int rad[P] = {8, 4, 8}; };
template <> struct RegPlan<512> { static constexpr int P = 3; static constexpr int rad[P] = {8, 8, 8}; };
template <> struct RegPlan<1024> { static constexpr int P = 4; static constexpr int rad[P] = {8, 2, 8, 8}; };
template <> struct RegPlan<2048> { static constexpr int P = 4; static constexpr int rad[P] = {8, 4, 8, 8}; };
template <> struct RegPlan<4096> { static constexpr int P = 4; static constexpr int rad[P] = {8, 8, 8, 8}; };

// The size-16 and size-32 plans end with a radix below 8; their last pass writes through LDS too.

template <int N>
constexpr int reg_lds_len() {
  return N + N / 8;
}
__device__ __forceinline__ int lpad(int i) { return i + (i >> 3); }

// Staged twiddle layout: pass p ≥ 1 (stride Ns, radix R) keeps W_N^{r·k·N/(Ns·R)} at
// reg_tw_off(p) + (r − 1)·Ns + k for r in [1, R), k in [0, Ns). In the table W_N^e the lanes of a
// pass read entries r·N/(Ns·R) apart (all in one LDS bank group in the second pass); here
// consecutive k are consecutive entries. Pass 0 (Ns = 1) takes no twiddles.
template <int N>
constexpr int reg_tw_off(int pass) {
  int off = 0, ns = RegPlan<N>::rad[0];
  for (int p = 1; p < pass; ++p) {
    off += (RegPlan<N>::rad[p] - 1) * ns;
    ns *= RegPlan<N>::rad[p];
  }
  return off;
}
template <int N>
constexpr int reg_tw_len() {
  return reg_tw_off<N>(RegPlan<N>::P);
}

template <int N, int PASS, int NS, bool COMP, typename T>
__device__ __forceinline__ void stage_tw_pass(cx<T>* sh_hi, cx<T>* sh_lo, const cx<T>* __restrict__ g_hi,
                                              const cx<T>* __restrict__ g_lo) {
  if constexpr (PASS < RegPlan<N>::P) {
    constexpr int R = RegPlan<N>::rad[PASS];
    constexpr int off = reg_tw_off<N>(PASS);
    for (int i = threadIdx.x; i < (R - 1) * NS; i += blockDim.x) {
      const int r = i / NS + 1, k = i % NS;
      const int e = r * k * (N / (NS * R));
      sh_hi[off + i] = g_hi[e];
      if constexpr (COMP) sh_lo[off + i] = g_lo[e];
    }
    stage_tw_pass<N, PASS + 1, NS * R, COMP>(sh_hi, sh_lo, g_hi, g_lo);
  }
}

// Copies the twiddles of every pass from the table W_N^e (g_hi, and g_lo for COMP) into the staged
// layout in LDS (reg_tw_len<N>() entries each); reg_fft reads them with STAGED.
template <int N, bool COMP, typename T>
__device__ __forceinline__ void stage_twiddles(cx<T>* sh_hi, cx<T>* sh_lo, const cx<T>* __restrict__ g_hi,
                                               const cx<T>* __restrict__ g_lo) {
  stage_tw_pass<N, 1, RegPlan<N>::rad[0], COMP>(sh_hi, sh_lo, g_hi, g_lo);
  __syncthreads();
}

template <int R, bool COMP, bool INV, typename T>
__device__ __forceinline__ void dft_r(cx<T>* x) {
  if constexpr (R == 2) dft2<COMP, INV>(x);
  else if constexpr (R == 4) dft4<COMP, INV>(x);
  else dft8<COMP, INV>(x);
}

template <int N, int PASS, int NS, bool COMP, bool INV, bool STAGED, typename T>
__device__ __forceinline__ void reg_pass(px2<T> (&v)[8], cx<T>* lds, int t, const cx<T>* twh, const cx<T>* twl) {
  using PL = RegPlan<N>;
  if constexpr (PASS < PL::P) {
    constexpr int R = PL::rad[PASS];
    constexpr int B = 8 / R;
    constexpr int TT = N / 8;
    constexpr bool last = PASS == PL::P - 1;
    constexpr bool in_regs = last && R == 8;  // natural-order result stays in registers
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const int j = t + TT * b;
      const int k = j & (NS - 1);
      cx<T> wh[R], wl[R];
      if constexpr (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          const int e = STAGED ? reg_tw_off<N>(PASS) + (r - 1) * NS + k : r * k * (N / (NS * R));
          wh[r] = INV ? conj(twh[e]) : twh[e];
          if constexpr (COMP) wl[r] = INV ? conj(twl[e]) : twl[e];
        }
      }
      const int base = (j - k) * R + k;
#pragma unroll
      for (int comp = 0; comp < 2; ++comp) {
        cx<T> x[R];
#pragma unroll
        for (int r = 0; r < R; ++r) x[r] = comp == 0 ? v[b + r * B].a : v[b + r * B].b;
        if constexpr (NS > 1) {
#pragma unroll
          for (int r = 1; r < R; ++r) x[r] = cmul2<COMP>(x[r], wh[r], wl[r]);
        }
        dft_r<R, COMP, INV>(x);
        if constexpr (in_regs) {
#pragma unroll
          for (int r = 0; r < R; ++r) {
            if (comp == 0) v[r].a = x[r];
            else v[r].b = x[r];
          }
        } else {
#pragma unroll
          for (int r = 0; r < R; ++r) lds[comp * reg_lds_len<N>() + lpad(base + r * NS)] = x[r];
        }
      }
    }
    if constexpr (!in_regs) {
      __syncthreads();
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        v[m].a = lds[lpad(t + TT * m)];
        v[m].b = lds[reg_lds_len<N>() + lpad(t + TT * m)];
      }
      __syncthreads();
    }
    reg_pass<N, PASS + 1, NS * R, COMP, INV, STAGED>(v, lds, t, twh, twl);
  }
}

// One field component per thread: the same passes on v[m] = a[t + T·m] of a single complex line
// (the slab-step kernels give the two components of a pixel to two adjacent lanes).
template <int N, int PASS, int NS, bool COMP, bool INV, bool STAGED, typename T>
__device__ __forceinline__ void reg_pass1(cx<T> (&v)[8], cx<T>* lds, int t, const cx<T>* twh, const cx<T>* twl) {
  using PL = RegPlan<N>;
  if constexpr (PASS < PL::P) {
    constexpr int R = PL::rad[PASS];
    constexpr int B = 8 / R;
    constexpr int TT = N / 8;
    constexpr bool last = PASS == PL::P - 1;
    constexpr bool in_regs = last && R == 8;
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const int j = t + TT * b;
      const int k = j & (NS - 1);
      cx<T> x[R];
#pragma unroll
      for (int r = 0; r < R; ++r) x[r] = v[b + r * B];
      if constexpr (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          const int e = STAGED ? reg_tw_off<N>(PASS) + (r - 1) * NS + k : r * k * (N / (NS * R));
          const cx<T> wh = INV ? conj(twh[e]) : twh[e];
          cx<T> wl{};
          if constexpr (COMP) wl = INV ? conj(twl[e]) : twl[e];
          x[r] = cmul2<COMP>(x[r], wh, wl);
        }
      }
      dft_r<R, COMP, INV>(x);
      if constexpr (in_regs) {
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = x[r];
      } else {
        const int base = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) lds[lpad(base + r * NS)] = x[r];
      }
    }
    if constexpr (!in_regs) {
      __syncthreads();
#pragma unroll
      for (int m = 0; m < 8; ++m) v[m] = lds[lpad(t + TT * m)];
      __syncthreads();
    }
    reg_pass1<N, PASS + 1, NS * R, COMP, INV, STAGED>(v, lds, t, twh, twl);
  }
}

// Transform one line held as v[m] = a[t + T·m]; on return v[m] = A[t + T·m]. All threads of the
// workgroup must call it together (it synchronises the workgroup between passes). The px2 form
// exchanges through 2·reg_lds_len<N>() LDS entries, the E_x plane and then the E_y plane, so that
// each LDS instruction of a wavefront touches consecutive 16-byte entries of one plane; the cx form
// through reg_lds_len<N>() entries. twh/twl point
// to the twiddle table W_N^e (LDS or global), or with STAGED to the copy that stage_twiddles
// made. The px2 form transforms both components of a pixel vector in one thread,
// the cx form one component.
template <int N, bool COMP, bool INV, bool STAGED = false, typename T>
__device__ __forceinline__ void reg_fft(px2<T> (&v)[8], cx<T>* lds, int t, const cx<T>* twh, const cx<T>* twl) {
  reg_pass<N, 0, 1, COMP, INV, STAGED>(v, lds, t, twh, twl);
}
template <int N, bool COMP, bool INV, bool STAGED = false, typename T>
__device__ __forceinline__ void reg_fft(cx<T> (&v)[8], cx<T>* lds, int t, const cx<T>* twh, const cx<T>* twl) {
  reg_pass1<N, 0, 1, COMP, INV, STAGED>(v, lds, t, twh, twl);
}

}  // namespace bwpm
