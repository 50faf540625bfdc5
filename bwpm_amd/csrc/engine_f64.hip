// FP64 instantiation of the engine (parity mode against the FP64 reference).
#include "engine_impl.hpp"

#include <mutex>

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

namespace {

// The flux of a plane on the device (overwritten by its spectrum), given its k tables, FFT plans and
// reduction buffers, on stream s: the part both flux entry points share.
double flux_of_plane(px2<double>* plane, const double* dkx, const double* dky, const FftPlan& px, const FftPlan& py,
                     double2* partial, double2* out, long long ny, long long nx, double dy, double dx, double k0,
                     c128 index, int direction, hipStream_t s) {
  const long long npix = ny * nx;
  PlaneSet<double> ps{};
  ps.in[0] = plane;
  ps.out[0] = plane;
  ps.scale = 1.0;
  const size_t lrow = static_cast<size_t>(nx) * 2 * sizeof(cx<double>);
  const size_t lcol = static_cast<size_t>(ny) * 2 * sizeof(cx<double>);
  require(lrow <= 160 * 1024 && lcol <= 160 * 1024, "grid too large for the flux transform");
  HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fft2_rows<double, false, false>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
  HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fft2_cols<double, false, false>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
  hipLaunchKernelGGL((k_fft2_rows<double, false, false>), dim3(static_cast<unsigned>(ny), 1), dim3(kThreads), lrow, s,
                     ps, static_cast<int>(ny), static_cast<int>(nx), 1, 2, px);
  HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL((k_fft2_cols<double, false, false>), dim3(static_cast<unsigned>(nx), 1), dim3(kThreads), lcol, s,
                     ps, static_cast<int>(ny), static_cast<int>(nx), 1, 2, py);
  HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_flux_partial, dim3(kRedBlocks), dim3(kRedThreads), 0, s, plane, dkx, dky, static_cast<int>(nx),
                     npix, zd{index.real(), index.imag()}, k0, static_cast<double>(direction), partial);
  HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(kRedThreads), 0, s, partial, out, nullptr);
  HIP_CHECK(hipGetLastError());
  double2 h{};
  HIP_CHECK(hipMemcpyAsync(&h, out, sizeof(double2), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  return h.x * dx * dy / static_cast<double>(npix);
}

}  // namespace

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
    result = flux_of_plane(plane.as<px2<double>>(), dkx.as<double>(), dky.as<double>(), tx.plan, ty.plan,
                           partial.as<double2>(), out.as<double2>(), ny, nx, dy, dx, k0, index, direction, s);
  } catch (...) {
    (void)hipStreamDestroy(s);
    throw;
  }
  HIP_CHECK(hipStreamDestroy(s));
  return result;
}

// ---- the flux without allocation ------------------------------------------------------------------
//
// bwpm_power_flux_noalloc works in `scratch`, a device array of the library: the plane and its
// spectrum, the k tables, the twiddle tables, the reduction partials and the result. Its size
// bounds the grid (kNoallocMaxLine points per line, kNoallocMaxPixels pixels). The array is the
// process's on each device; host code addresses it only through the device address that
// hipGetSymbolAddress returns for the device (taken once per device, with it current).

namespace {

constexpr long long kNoallocMaxLine = 4096;
constexpr long long kNoallocMaxPixels = 512LL * 512LL;
constexpr int kNoallocMaxDevices = 64;
constexpr size_t kOffPlane = 0;
constexpr size_t kOffKx = kOffPlane + kNoallocMaxPixels * sizeof(px2<double>);
constexpr size_t kOffKy = kOffKx + kNoallocMaxLine * sizeof(double);
constexpr size_t kOffTwx = kOffKy + kNoallocMaxLine * sizeof(double);
constexpr size_t kOffTwy = kOffTwx + kNoallocMaxLine * sizeof(cx<double>);
constexpr size_t kOffPartial = kOffTwy + kNoallocMaxLine * sizeof(cx<double>);
constexpr size_t kOffOut = kOffPartial + kRedBlocks * sizeof(double2);
constexpr size_t kScratchBytes = kOffOut + 256;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

double power_flux_noalloc(const c128* field, long long ny, long long nx, double dy, double dx, double k0, c128 index,
                          int direction, int device) {
  require(ny >= 1 && nx >= 1, "field must have shape (ny, nx, 2)");
  require(ny <= kNoallocMaxLine && nx <= kNoallocMaxLine && ny * nx <= kNoallocMaxPixels,
          "grid " + std::to_string(ny) + " x " + std::to_string(nx) + " exceeds the no-allocation flux limit of " +
              std::to_string(kNoallocMaxLine) + " points per line and " + std::to_string(kNoallocMaxPixels) +
              " pixels");
  require(device >= 0 && device < kNoallocMaxDevices, "device index out of range");
  // calls share the scratch array and the host staging below: one at a time in the process
  static std::mutex mu;
  static unsigned char* base[kNoallocMaxDevices] = {};
  static double hkx[kNoallocMaxLine], hky[kNoallocMaxLine];
  static cx<double> htwx[kNoallocMaxLine], htwy[kNoallocMaxLine];
  std::lock_guard<std::mutex> lock(mu);
  HIP_CHECK(hipSetDevice(device));
  if (!base[device]) {
    void* p = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    base[device] = static_cast<unsigned char*>(p);
  }
  unsigned char* b = base[device];
  for (long long i = 0; i < nx; ++i) hkx[i] = fftfreq(i, nx, dx);
  for (long long j = 0; j < ny; ++j) hky[j] = fftfreq(j, ny, dy);
  for (long long t = 0; t < nx; ++t) twiddle_value(t, nx, htwx[t].x, htwx[t].y);
  for (long long t = 0; t < ny; ++t) twiddle_value(t, ny, htwy[t].x, htwy[t].y);
  const hipStream_t s = nullptr;  // the device's null stream: no stream is created
  HIP_CHECK(hipMemcpyAsync(b + kOffKx, hkx, nx * sizeof(double), hipMemcpyHostToDevice, s));
  HIP_CHECK(hipMemcpyAsync(b + kOffKy, hky, ny * sizeof(double), hipMemcpyHostToDevice, s));
  HIP_CHECK(hipMemcpyAsync(b + kOffTwx, htwx, nx * sizeof(cx<double>), hipMemcpyHostToDevice, s));
  HIP_CHECK(hipMemcpyAsync(b + kOffTwy, htwy, ny * sizeof(cx<double>), hipMemcpyHostToDevice, s));
  HIP_CHECK(hipMemcpyAsync(b + kOffPlane, field, ny * nx * sizeof(px2<double>), hipMemcpyHostToDevice, s));
  // the host staging is read by the copies: they complete before it can be written again
  HIP_CHECK(hipStreamSynchronize(s));
  FftPlan px{}, py{};
  if (nx > 1) plan_radices(nx, px);
  if (ny > 1) plan_radices(ny, py);
  px.n = static_cast<int>(nx);
  py.n = static_cast<int>(ny);
  // FP64 plans read the hi table only (the lo residual is the FP32 paths')
  px.tw_hi = px.tw_lo = b + kOffTwx;
  py.tw_hi = py.tw_lo = b + kOffTwy;
  return flux_of_plane(reinterpret_cast<px2<double>*>(b + kOffPlane), reinterpret_cast<const double*>(b + kOffKx),
                       reinterpret_cast<const double*>(b + kOffKy), px, py, reinterpret_cast<double2*>(b + kOffPartial),
                       reinterpret_cast<double2*>(b + kOffOut), ny, nx, dy, dx, k0, index, direction, s);
}

std::string fft_plan_report() { return plan_report(); }

}  // namespace bwpm
