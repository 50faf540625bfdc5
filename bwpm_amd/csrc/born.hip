// Non-templated Born kernels (one definition for the whole library).
#include "born.hpp"

namespace bwpm {

// out[b] = Σ_i partial[b·kRedBlocks + i] in a fixed tree order (deterministic).
__global__ __launch_bounds__(kRedThreads) void k_dot_final(const double2* __restrict__ partial,
                                                           double2* __restrict__ out) {
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

}  // namespace bwpm
