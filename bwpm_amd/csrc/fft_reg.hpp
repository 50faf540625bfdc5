// Register-resident mixed-radix Stockham FFT for the hot slab-step kernels: power-of-two lines
// (16 … 4096 points) and lines of 3·2^k and 9·2^k points (192 … 3072).
//
// A line of N points is owned by T = N/8 threads. Thread t holds the eight pixel vectors
// a[t + T·m] (m = 0 … 7, both field components), which is exactly the read pattern of every
// Stockham pass whose radix R divides 8: such a pass lets each thread run 8/R butterflies
// j = t + T·b on v[b + r·8/R], multiplies by W_N^{r·(j mod Ns)·N/(Ns·R)}, applies the size-R DFT
// and writes y_r to LDS at (j − k)·R + k + r·Ns; the next pass reads a[t + T·m] back. The first
// and the last pass are radix 8, so the input is taken in natural order from registers and the
// last pass leaves A[t + T·r] in v[r] (natural order) without touching LDS: P − 1 LDS exchanges.
//
// A pass of radix 3, 6, 9 or 12 (an "irregular" pass: R does not divide the eight values of a
// thread) has N/R butterflies, which the T threads take as j = t, t + T, … (the last round only
// partly): it reads its inputs x_r = a[j + r·N/R] straight from the LDS the previous pass wrote,
// synchronises, and writes its outputs back to LDS in the same Stockham order. A 384-point line is
// 8·6·8, so it has two exchanges like a 512-point one.
//
// In the px2 form the two field components are transformed one after the other inside each
// butterfly, which halves the live temporaries; the cx form transforms a single component. Twiddles
// are read from the table W_N^e in global memory, or (cx form) from a copy the kernel stages in LDS
// once, laid out per pass so that the lanes of a wavefront read consecutive entries.
// LDS index i is padded to i + i/8 to spread the stride-8 writes of the first pass over banks.
#pragma once

#include <type_traits>

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

// A radix plan: the passes' radices in order; N is their product.
template <int... R>
struct Plan {
  static constexpr int P = sizeof...(R);
  static constexpr int rad[P] = {R...};
  static constexpr int N = (R * ...);
};

// Plans per line length: the default, and an alternative that the plan tuner times against it
// (alt = type where a length has only one plan worth running). Every plan starts with a radix that
// divides 8, so the first pass reads natural order from registers; every 3·2^k and 9·2^k plan has
// one irregular pass.
template <int N>
struct RegPlans;
template <> struct RegPlans<16> { using type = Plan<8, 2>; using alt = type; };
template <> struct RegPlans<32> { using type = Plan<8, 4>; using alt = type; };
template <> struct RegPlans<64> { using type = Plan<8, 8>; using alt = type; };
template <> struct RegPlans<128> { using type = Plan<8, 2, 8>; using alt = Plan<4, 4, 8>; };
template <> struct RegPlans<256> { using type = Plan<8, 4, 8>; using alt = Plan<4, 8, 8>; };
template <> struct RegPlans<512> { using type = Plan<8, 8, 8>; using alt = Plan<8, 4, 2, 8>; };
template <> struct RegPlans<1024> { using type = Plan<8, 2, 8, 8>; using alt = Plan<4, 4, 8, 8>; };
template <> struct RegPlans<2048> { using type = Plan<8, 4, 8, 8>; using alt = Plan<4, 8, 8, 8>; };
template <> struct RegPlans<4096> { using type = Plan<8, 8, 8, 8>; using alt = type; };
template <> struct RegPlans<192> { using type = Plan<8, 3, 8>; using alt = Plan<4, 6, 8>; };
template <> struct RegPlans<384> { using type = Plan<8, 6, 8>; using alt = Plan<4, 12, 8>; };
template <> struct RegPlans<576> { using type = Plan<8, 9, 8>; using alt = Plan<8, 3, 3, 8>; };
template <> struct RegPlans<768> { using type = Plan<8, 12, 8>; using alt = Plan<8, 6, 2, 8>; };
template <> struct RegPlans<1152> { using type = Plan<8, 9, 2, 8>; using alt = Plan<4, 4, 9, 8>; };
template <> struct RegPlans<1536> { using type = Plan<8, 8, 3, 8>; using alt = Plan<4, 8, 6, 8>; };
template <> struct RegPlans<3072> { using type = Plan<8, 8, 6, 8>; using alt = Plan<4, 8, 12, 8>; };

template <int N>
using RegPlan = typename RegPlans<N>::type;
// Plan v (0: the default, 1: the alternative) of line length N.
template <int N, int V>
using RegPlanV = std::conditional_t<V == 0, typename RegPlans<N>::type, typename RegPlans<N>::alt>;
template <int N>
constexpr int reg_plan_count() {
  return std::is_same_v<typename RegPlans<N>::type, typename RegPlans<N>::alt> ? 1 : 2;
}

// The size-16 and size-32 plans end with a radix below 8; their last pass writes through LDS too.

// A pass whose radix divides the eight values a thread holds takes its inputs from registers.
constexpr bool reg_regular(int radix) { return 8 % radix == 0; }
template <typename PL>
constexpr bool plan_all_regular() {
  for (int p = 0; p < PL::P; ++p)
    if (!reg_regular(PL::rad[p])) return false;
  return true;
}

template <int N>
constexpr int reg_lds_len() {
  return N + N / 8;
}
__device__ __forceinline__ int lpad(int i) { return i + (i >> 3); }

// Staged twiddle layout: pass p ≥ 1 (stride Ns, radix R) keeps W_N^{r·k·N/(Ns·R)} at
// plan_tw_off(p) + (r − 1)·Ns + k for r in [1, R), k in [0, Ns). In the table W_N^e the lanes of a
// pass read entries r·N/(Ns·R) apart (all in one LDS bank group in the second pass); here
// consecutive k are consecutive entries. Pass 0 (Ns = 1) takes no twiddles; the layout holds
// N − rad[0] entries.
template <typename PL>
constexpr int plan_tw_off(int pass) {
  int off = 0, ns = PL::rad[0];
  for (int p = 1; p < pass; ++p) {
    off += (PL::rad[p] - 1) * ns;
    ns *= PL::rad[p];
  }
  return off;
}
template <typename PL>
constexpr int plan_tw_len() {
  return plan_tw_off<PL>(PL::P);
}
template <int N>
constexpr int reg_tw_off(int pass) {
  return plan_tw_off<RegPlan<N>>(pass);
}
template <int N>
constexpr int reg_tw_len() {
  return plan_tw_len<RegPlan<N>>();
}

template <typename PL, int PASS, int NS, bool COMP, typename T>
__device__ __forceinline__ void stage_tw_pass(cx<T>* sh_hi, cx<T>* sh_lo, const cx<T>* __restrict__ g_hi,
                                              const cx<T>* __restrict__ g_lo) {
  if constexpr (PASS < PL::P) {
    constexpr int N = PL::N;
    constexpr int R = PL::rad[PASS];
    constexpr int off = plan_tw_off<PL>(PASS);
    for (int i = threadIdx.x; i < (R - 1) * NS; i += blockDim.x) {
      const int r = i / NS + 1, k = i % NS;
      const int e = r * k * (N / (NS * R));
      sh_hi[off + i] = g_hi[e];
      if constexpr (COMP) sh_lo[off + i] = g_lo[e];
    }
    stage_tw_pass<PL, PASS + 1, NS * R, COMP>(sh_hi, sh_lo, g_hi, g_lo);
  }
}

// Copies the twiddles of every pass of plan PL from the table W_N^e (g_hi, and g_lo for COMP) into
// the staged layout in LDS (plan_tw_len<PL>() entries each); reg_fft reads them with STAGED.
template <typename PL, bool COMP, typename T>
__device__ __forceinline__ void stage_twiddles(cx<T>* sh_hi, cx<T>* sh_lo, const cx<T>* __restrict__ g_hi,
                                               const cx<T>* __restrict__ g_lo) {
  stage_tw_pass<PL, 1, PL::rad[0], COMP>(sh_hi, sh_lo, g_hi, g_lo);
  __syncthreads();
}
template <int N, bool COMP, typename T>
__device__ __forceinline__ void stage_twiddles(cx<T>* sh_hi, cx<T>* sh_lo, const cx<T>* __restrict__ g_hi,
                                               const cx<T>* __restrict__ g_lo) {
  stage_twiddles<RegPlan<N>, COMP>(sh_hi, sh_lo, g_hi, g_lo);
}

// cos(2π·m/36) for any integer m, from cos(10°·i), i = 0 … 9 (correctly rounded literals).
constexpr double cos_36th(int m) {
  constexpr double c[10] = {1.0,
                            0.98480775301220805936674302458952,
                            0.93969262078590838405410927732473,
                            0.86602540378443864676372317075294,
                            0.76604444311897803520239265055542,
                            0.64278760968653932632264340990726,
                            0.5,
                            0.34202014332566873304409961468226,
                            0.17364817766693034885171662676931,
                            0.0};
  m %= 36;
  if (m < 0) m += 36;
  if (m <= 9) return c[m];
  if (m <= 18) return -c[18 - m];
  if (m <= 27) return -c[m - 18];
  return c[36 - m];
}

template <int R, bool COMP, bool INV, typename T>
__device__ __forceinline__ void dft_r(cx<T>* x);

// x ← W_R^e·x for the size-R DFT's internal twiddles (W_R = e^{∓2πi/R}, R dividing 36): exact for
// multiples of a quarter turn, else a (compensated) complex product with the constant.
template <int R, int E, bool COMP, bool INV, typename T>
__device__ __forceinline__ cx<T> dft_twiddle(cx<T> x) {
  constexpr int m = (E * (36 / R)) % 36;
  if constexpr (m == 0) {
    return x;
  } else if constexpr (m == 9) {
    return INV ? mul_i(x) : mul_mi(x);
  } else if constexpr (m == 18) {
    return cx<T>{-x.x, -x.y};
  } else if constexpr (m == 27) {
    return INV ? mul_mi(x) : mul_i(x);
  } else {
    constexpr double c = cos_36th(m), sn = (INV ? 1.0 : -1.0) * cos_36th(9 - m);
    constexpr cx<T> hi{static_cast<T>(c), static_cast<T>(sn)};
    constexpr cx<T> lo{static_cast<T>(c - static_cast<double>(static_cast<T>(c))),
                       static_cast<T>(sn - static_cast<double>(static_cast<T>(sn)))};
    return cmul2<COMP>(x, hi, lo);
  }
}

template <int R1, int R2, int N2, int K1, bool COMP, bool INV, typename T>
__device__ __forceinline__ void dft_ct_tw(cx<T> (&y)[R2][R1], const cx<T>* a) {
  if constexpr (K1 < R1) {
    y[N2][K1] = dft_twiddle<R1 * R2, N2 * K1, COMP, INV>(a[K1]);
    dft_ct_tw<R1, R2, N2, K1 + 1, COMP, INV>(y, a);
  }
}
// Size R1·R2 DFT by one Cooley–Tukey split: input n = R2·n1 + n2, output k = k1 + R1·k2. R2 DFTs of
// size R1 over n1, the twiddles W_R^{n2·k1}, then R1 DFTs of size R2 over n2.
template <int R1, int R2, bool COMP, bool INV, typename T, int N2 = 0>
__device__ __forceinline__ void dft_ct_rows(cx<T> (&y)[R2][R1], const cx<T>* x) {
  if constexpr (N2 < R2) {
    cx<T> a[R1];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) a[n1] = x[R2 * n1 + N2];
    dft_r<R1, COMP, INV>(a);
    y[N2][0] = a[0];
    dft_ct_tw<R1, R2, N2, 1, COMP, INV>(y, a);
    dft_ct_rows<R1, R2, COMP, INV, T, N2 + 1>(y, x);
  }
}
template <int R1, int R2, bool COMP, bool INV, typename T>
__device__ __forceinline__ void dft_ct(cx<T>* x) {
  cx<T> y[R2][R1];
  dft_ct_rows<R1, R2, COMP, INV>(y, x);
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) {
    cx<T> b[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) b[n2] = y[n2][k1];
    dft_r<R2, COMP, INV>(b);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) x[k1 + R1 * k2] = b[k2];
  }
}

template <int R, bool COMP, bool INV, typename T>
__device__ __forceinline__ void dft_r(cx<T>* x) {
  if constexpr (R == 2) dft2<COMP, INV>(x);
  else if constexpr (R == 3) dft3<COMP, INV>(x);
  else if constexpr (R == 4) dft4<COMP, INV>(x);
  else if constexpr (R == 6) dft_ct<3, 2, COMP, INV>(x);
  else if constexpr (R == 8) dft8<COMP, INV>(x);
  else if constexpr (R == 9) dft_ct<3, 3, COMP, INV>(x);
  else if constexpr (R == 12) dft_ct<4, 3, COMP, INV>(x);
  else static_assert(R == 2, "register FFT radices: 2, 3, 4, 6, 8, 9, 12");
}

template <int N, int PASS, int NS, bool COMP, bool INV, bool STAGED, typename T>
__device__ __forceinline__ void reg_pass(px2<T> (&v)[8], cx<T>* lds, int t, const cx<T>* twh, const cx<T>* twl) {
  using PL = RegPlan<N>;
  static_assert(plan_all_regular<RegPlan<N>>(), "the two-component form runs plans of radices 2, 4 and 8");
  if constexpr (PASS < PL::P) {
    constexpr int R = PL::rad[PASS];
    constexpr int B = 8 / R;
    constexpr int TT = N / 8;
    constexpr bool last = PASS == PL::P - 1;
    constexpr bool in_regs = last && R == 8;  // natural-order result stays in registers
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const int j = t + TT * b;
      const int k = j % NS;
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
// (the slab-step kernels give the two components of a pixel to two adjacent lanes). Regular passes
// run from registers; an irregular pass runs from the LDS the previous pass left (see the file
// comment), so a regular pass followed by an irregular one does not read its output back.
template <typename PL, int PASS, int NS, bool COMP, bool INV, bool STAGED, typename T>
__device__ __forceinline__ void reg_pass1(cx<T> (&v)[8], cx<T>* lds, int t, const cx<T>* twh, const cx<T>* twl) {
  constexpr int N = PL::N;
  if constexpr (PASS < PL::P) {
    constexpr int R = PL::rad[PASS];
    constexpr int TT = N / 8;
    constexpr int M = N / R;  // butterflies of the pass
    constexpr bool last = PASS == PL::P - 1;
    constexpr bool regular = reg_regular(R);
    constexpr bool in_regs = last && R == 8;
    // the next pass reads its inputs from LDS itself
    constexpr bool next_lds = !last && !reg_regular(PL::rad[PASS + 1 < PL::P ? PASS + 1 : 0]);
    static_assert(PASS > 0 || (regular && NS == 1), "the first pass takes natural order from registers");
    // twiddled DFT of butterfly j on x, written to LDS (or kept in registers by the last pass)
    auto butterfly = [&](int j, cx<T>* x) {
      const int k = j % NS;
      if constexpr (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          const int e = STAGED ? plan_tw_off<PL>(PASS) + (r - 1) * NS + k : r * k * (N / (NS * R));
          const cx<T> wh = INV ? conj(twh[e]) : twh[e];
          cx<T> wl{};
          if constexpr (COMP) wl = INV ? conj(twl[e]) : twl[e];
          x[r] = cmul2<COMP>(x[r], wh, wl);
        }
      }
      dft_r<R, COMP, INV>(x);
      if constexpr (!in_regs) {
        const int base = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) lds[lpad(base + r * NS)] = x[r];
      }
    };
    if constexpr (regular) {
      constexpr int B = 8 / R;
#pragma unroll
      for (int b = 0; b < B; ++b) {
        cx<T> x[R];
#pragma unroll
        for (int r = 0; r < R; ++r) x[r] = v[b + r * B];
        butterfly(t + TT * b, x);
        if constexpr (in_regs) {
#pragma unroll
          for (int r = 0; r < R; ++r) v[r] = x[r];
        }
      }
    } else {
      // butterflies j = t + T·b; the last round covers j < M only
      constexpr int B = (M + TT - 1) / TT;
      cx<T> x[B][R];
#pragma unroll
      for (int b = 0; b < B; ++b) {
        const int j = t + TT * b;
        if (b * TT + TT <= M || j < M) {
#pragma unroll
          for (int r = 0; r < R; ++r) x[b][r] = lds[lpad(j + r * M)];
        }
      }
      __syncthreads();  // every input is read before the outputs overwrite the line
#pragma unroll
      for (int b = 0; b < B; ++b) {
        const int j = t + TT * b;
        if (b * TT + TT <= M || j < M) butterfly(j, x[b]);
      }
    }
    if constexpr (!in_regs) {
      __syncthreads();
      if constexpr (!next_lds) {
#pragma unroll
        for (int m = 0; m < 8; ++m) v[m] = lds[lpad(t + TT * m)];
        __syncthreads();
      }
    }
    reg_pass1<PL, PASS + 1, NS * R, COMP, INV, STAGED>(v, lds, t, twh, twl);
  }
}

// Both components per thread, for plans with an irregular pass (lines of 3·2^k and 9·2^k points in
// the 2D transforms and the Born operator): the passes of reg_pass1 on the two LDS planes, the E_x
// plane and then the E_y plane (reg_pass keeps serving the all-regular plans).
template <typename PL, int PASS, int NS, bool COMP, bool INV, bool STAGED, typename T>
__device__ __forceinline__ void reg_pass2(px2<T> (&v)[8], cx<T>* lds, int t, const cx<T>* twh, const cx<T>* twl) {
  constexpr int N = PL::N;
  constexpr int L = reg_lds_len<N>();
  if constexpr (PASS < PL::P) {
    constexpr int R = PL::rad[PASS];
    constexpr int TT = N / 8;
    constexpr int M = N / R;
    constexpr bool last = PASS == PL::P - 1;
    constexpr bool regular = reg_regular(R);
    constexpr bool in_regs = last && R == 8;
    constexpr bool next_lds = !last && !reg_regular(PL::rad[PASS + 1 < PL::P ? PASS + 1 : 0]);
    static_assert(PASS > 0 || (regular && NS == 1), "the first pass takes natural order from registers");
    // the twiddles of butterfly j (shared by both components)
    auto twiddles = [&](int j, cx<T>* wh, cx<T>* wl) {
      const int k = j % NS;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const int e = STAGED ? plan_tw_off<PL>(PASS) + (r - 1) * NS + k : r * k * (N / (NS * R));
        wh[r] = INV ? conj(twh[e]) : twh[e];
        wl[r] = cx<T>{};
        if constexpr (COMP) wl[r] = INV ? conj(twl[e]) : twl[e];
      }
    };
    auto butterfly = [&](int j, int comp, cx<T>* x, const cx<T>* wh, const cx<T>* wl) {
      const int k = j % NS;
      if constexpr (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) x[r] = cmul2<COMP>(x[r], wh[r], wl[r]);
      }
      dft_r<R, COMP, INV>(x);
      if constexpr (!in_regs) {
        const int base = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) lds[comp * L + lpad(base + r * NS)] = x[r];
      }
    };
    if constexpr (regular) {
      constexpr int B = 8 / R;
#pragma unroll
      for (int b = 0; b < B; ++b) {
        const int j = t + TT * b;
        cx<T> wh[R], wl[R];
        if constexpr (NS > 1) twiddles(j, wh, wl);
#pragma unroll
        for (int comp = 0; comp < 2; ++comp) {
          cx<T> x[R];
#pragma unroll
          for (int r = 0; r < R; ++r) x[r] = comp == 0 ? v[b + r * B].a : v[b + r * B].b;
          butterfly(j, comp, x, wh, wl);
          if constexpr (in_regs) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
              if (comp == 0) v[r].a = x[r];
              else v[r].b = x[r];
            }
          }
        }
      }
    } else {
      // butterflies j = t + T·b of both planes; the last round covers j < M only
      constexpr int B = (M + TT - 1) / TT;
      cx<T> x[2][B][R];
#pragma unroll
      for (int b = 0; b < B; ++b) {
        const int j = t + TT * b;
        if (b * TT + TT <= M || j < M) {
#pragma unroll
          for (int comp = 0; comp < 2; ++comp)
#pragma unroll
            for (int r = 0; r < R; ++r) x[comp][b][r] = lds[comp * L + lpad(j + r * M)];
        }
      }
      __syncthreads();  // every input is read before the outputs overwrite the lines
#pragma unroll
      for (int b = 0; b < B; ++b) {
        const int j = t + TT * b;
        if (b * TT + TT <= M || j < M) {
          cx<T> wh[R], wl[R];
          if constexpr (NS > 1) twiddles(j, wh, wl);
          butterfly(j, 0, x[0][b], wh, wl);
          butterfly(j, 1, x[1][b], wh, wl);
        }
      }
    }
    if constexpr (!in_regs) {
      __syncthreads();
      if constexpr (!next_lds) {
#pragma unroll
        for (int m = 0; m < 8; ++m) {
          v[m].a = lds[lpad(t + TT * m)];
          v[m].b = lds[L + lpad(t + TT * m)];
        }
        __syncthreads();
      }
    }
    reg_pass2<PL, PASS + 1, NS * R, COMP, INV, STAGED>(v, lds, t, twh, twl);
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
  if constexpr (plan_all_regular<RegPlan<N>>())
    reg_pass<N, 0, 1, COMP, INV, STAGED>(v, lds, t, twh, twl);
  else
    reg_pass2<RegPlan<N>, 0, 1, COMP, INV, STAGED>(v, lds, t, twh, twl);
}
template <int N, bool COMP, bool INV, bool STAGED = false, typename T>
__device__ __forceinline__ void reg_fft(cx<T> (&v)[8], cx<T>* lds, int t, const cx<T>* twh, const cx<T>* twl) {
  reg_pass1<RegPlan<N>, 0, 1, COMP, INV, STAGED>(v, lds, t, twh, twl);
}
// The same with an explicit plan PL (one component per thread; STAGED twiddles from
// stage_twiddles<PL>).
template <typename PL, bool COMP, bool INV, bool STAGED = false, typename T>
__device__ __forceinline__ void reg_fft(cx<T> (&v)[8], cx<T>* lds, int t, const cx<T>* twh, const cx<T>* twl) {
  reg_pass1<PL, 0, 1, COMP, INV, STAGED>(v, lds, t, twh, twl);
}

}  // namespace bwpm
