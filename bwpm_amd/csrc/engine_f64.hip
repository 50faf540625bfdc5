// FP64 instantiation of the engine (parity mode against the FP64 reference).
#include "engine_impl.hpp"

namespace bwpm {

template class Engine<double, false>;

std::unique_ptr<EngineBase> make_engine_f64(Spec spec) {
  return std::make_unique<Engine<double, false>>(std::move(spec));
}

void run_continuous_lane_f64(LaneFeed& feed, int slots, int chunk, hipStream_t stream) {
  Engine<double, false>::continuous(feed, slots, chunk, stream);
}

}  // namespace bwpm

namespace bwpm {

// Flux Σ_k Re(k_z)·(|Ê_x|² + |Ê_y|² + |Ê_z|²), Ê_z = −d·k⊥·Ê⊥/k_z, per grid-stride share.
__global__ __launch_bounds__(kRedThreads) void k_flux_partial(const px2<double>* __restrict__ sp, const double* kx,
                                                              const double* ky, int nx, long long npix, zd index,
                                                              double k0, double d, double2* __restrict__ partial) {
  __shared__ double acc[kRedThreads];
  double s = 0;
  for (long long q = blockIdx.x * static_cast<long long>(kRedThreads) + threadIdx.x; q < npix;
       q += static_cast<long long>(kRedBlocks) * kRedThreads) {
    const double a = kx[q % nx], b = ky[q / nx];
    const zd kz = kz_of(index, k0, a * a + b * b);
    const px2<double> e = sp[q];
    const zd num{-(a * e.a.x + b * e.b.x) * d, -(a * e.a.y + b * e.b.y) * d};
    const zd ez = zdiv(num, kz);
    s += kz.x * (e.a.x * e.a.x + e.a.y * e.a.y + e.b.x * e.b.x + e.b.y * e.b.y + ez.x * ez.x + ez.y * ez.y);
  }
  acc[threadIdx.x] = s;
  __syncthreads();
  for (int w = kRedThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) acc[threadIdx.x] += acc[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = double2{acc[0], 0.0};
}

double power_flux_device(const c128* field, long long ny, long long nx, double dy, double dx, double k0, c128 index,
                         int direction, int device) {
  HIP_CHECK(hipSetDevice(device));
  require(ny >= 1 && nx >= 1, "field must have shape (ny, nx, 2)");
  const long long npix = ny * nx;
  hipStream_t s;
  HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  double result = 0;
  try {
    std::vector<double> kx(nx), ky(ny);
    for (long long i = 0; i < nx; ++i) kx[i] = fftfreq(i, nx, dx);
    for (long long j = 0; j < ny; ++j) ky[j] = fftfreq(j, ny, dy);
    DevBuf dkx = upload_vec(kx, s), dky = upload_vec(ky, s);
    DevBuf plane(npix * sizeof(px2<double>)), partial(kRedBlocks * sizeof(double2)), out(sizeof(double2));
    HIP_CHECK(hipMemcpyAsync(plane.as<void>(), field, npix * sizeof(px2<double>), hipMemcpyHostToDevice, s));
    Twiddles<double> tx, ty;
    tx.build(nx, s);
    ty.build(ny, s);
    PlaneSet<double> ps{};
    ps.in[0] = plane.as<px2<double>>();
    ps.out[0] = plane.as<px2<double>>();
    ps.scale = 1.0;
    const size_t lrow = static_cast<size_t>(nx) * 2 * sizeof(cx<double>);
    const size_t lcol = static_cast<size_t>(ny) * 2 * sizeof(cx<double>);
    require(lrow <= 160 * 1024 && lcol <= 160 * 1024, "grid too large for the flux transform");
    HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fft2_rows<double, false, false>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fft2_cols<double, false, false>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    hipLaunchKernelGGL((k_fft2_rows<double, false, false>), dim3(static_cast<unsigned>(ny), 1), dim3(kThreads), lrow, s,
                       ps, static_cast<int>(ny), static_cast<int>(nx), 1, 2, tx.plan);
    HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL((k_fft2_cols<double, false, false>), dim3(static_cast<unsigned>(nx), 1), dim3(kThreads), lcol, s,
                       ps, static_cast<int>(ny), static_cast<int>(nx), 1, 2, ty.plan);
    HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_flux_partial, dim3(kRedBlocks), dim3(kRedThreads), 0, s, plane.as<px2<double>>(),
                       dkx.as<double>(), dky.as<double>(), static_cast<int>(nx), npix, zd{index.real(), index.imag()}, k0,
                       static_cast<double>(direction), partial.as<double2>());
    HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(kRedThreads), 0, s, partial.as<double2>(), out.as<double2>());
    HIP_CHECK(hipGetLastError());
    double2 h{};
    HIP_CHECK(hipMemcpyAsync(&h, out.as<void>(), sizeof(double2), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    result = h.x * dx * dy / static_cast<double>(npix);
  } catch (...) {
    (void)hipStreamDestroy(s);
    throw;
  }
  HIP_CHECK(hipStreamDestroy(s));
  return result;
}

}  // namespace bwpm
