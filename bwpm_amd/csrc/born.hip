// Non-templated Born kernels (one definition for the whole library).
#include "born.hpp"

namespace bwpm {

// out[b] = Σ_i partial[b·kRedBlocks + i] in a fixed tree order (deterministic).
__global__ __launch_bounds__(kRedThreads) void k_dot_final(const double2* __restrict__ partial,
                                                           double2* __restrict__ out, const int* __restrict__ skip) {
  if (skip && *skip) return;
  __shared__ double sr[kRedThreads], si[kRedThreads];
  const double2* p = partial + static_cast<long long>(blockIdx.x) * kRedBlocks;
  double re = 0, im = 0;
  for (int i = threadIdx.x; i < kRedBlocks; i += kRedThreads) {
    re += p[i].x;
    im += p[i].y;
  }
  sr[threadIdx.x] = re;
  si[threadIdx.x] = im;
  __syncthreads();
  for (int s = kRedThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sr[threadIdx.x] += sr[threadIdx.x + s];
      si[threadIdx.x] += si[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = double2{sr[0], si[0]};
}

// ---- GMRES cycle state on the device ----------------------------------------------------------

namespace {
__device__ inline double2 zmul(double2 a, double2 b) { return double2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ inline double2 zadd(double2 a, double2 b) { return double2{a.x + b.x, a.y + b.y}; }
__device__ inline double2 zconj(double2 a) { return double2{a.x, -a.y}; }
__device__ inline double2 zneg(double2 a) { return double2{-a.x, -a.y}; }
__device__ inline double2 zscale(double2 a, double s) { return double2{a.x * s, a.y * s}; }
__device__ inline double zabs(double2 a) { return hypot(a.x, a.y); }
}  // namespace

__global__ void k_gm_givens(GmState* st, const double2* __restrict__ res, int j, double tol_abs) {
  if (threadIdx.x != 0 || blockIdx.x != 0 || st->done) return;
  constexpr int M = kGmRestart;
  double2 hv[M + 1];
  for (int i = 0; i <= j; ++i) hv[i] = zadd(res[i], res[kGmH2 + i]);
  const double hnext = sqrt(res[kGmN2].x);
  for (int i = 0; i < j; ++i) {  // the previous rotations on column j
    const double2 a = hv[i], c = hv[i + 1];
    hv[i] = zadd(zmul(st->cs[i], a), zmul(st->sn[i], c));
    hv[i + 1] = zadd(zmul(zneg(zconj(st->sn[i])), a), zmul(st->cs[i], c));
  }
  const double2 a = hv[j];
  const double aa = zabs(a);
  const double r = sqrt(a.x * a.x + a.y * a.y + hnext * hnext);
  double2 c, sn;
  if (aa == 0.0) {
    c = double2{0.0, 0.0};
    sn = double2{1.0, 0.0};
  } else {
    c = double2{aa / r, 0.0};
    sn = zscale(double2{a.x / aa, a.y / aa}, hnext / r);
  }
  hv[j] = zadd(zmul(c, a), zscale(sn, hnext));
  for (int i = 0; i <= j; ++i) st->h[i * M + j] = hv[i];
  const double2 gj = st->g[j];
  st->g[j + 1] = zmul(zneg(zconj(sn)), gj);
  st->g[j] = zmul(c, gj);
  st->cs[j] = c;
  st->sn[j] = sn;
  st->used = j + 1;
  if (zabs(st->g[j + 1]) <= tol_abs || hnext == 0.0) st->done = 1;
}

__global__ __launch_bounds__(kRedThreads) void k_rich_begin(GmState* st, const double2* __restrict__ bsq, int count,
                                                            double tol) {
  __shared__ double red[kRedThreads / 64];
  const double tot = rich_sum(bsq, count, red);
  if (threadIdx.x != 0) return;
  st->done = 0;
  st->used = 0;
  st->beta = __builtin_huge_val();
  st->tol_abs = tol * sqrt(tot);
}

__global__ __launch_bounds__(kRedThreads) void k_rich_check(const double2* __restrict__ partial, int count,
                                                            GmState* st, int max_steps) {
  if (st->done) return;  // uniform: every thread reads the same flag before any synchronisation
  __shared__ double red[kRedThreads / 64];
  const double tot = rich_sum(partial, count, red);
  if (threadIdx.x == 0) rich_update_state(st, tot, max_steps);
}

}  // namespace bwpm
