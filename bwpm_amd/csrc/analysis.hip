// Analysis reductions over field planes on the device: flux through a plane, the flux-weighted
// Fresnel reflectance, the on-axis intensity over propagation distances (axial focus) and the
// Strehl profile of an angular spectrum. They follow src/bwpm/pilot_imc3d.py:19-131,
// experiments/b3_scaling.py:88-97 and src/bwpm/pilot_imc.py:184-197.
//
// A plane (ny, nx, 2) is transformed once (FP64, in a work buffer of the handle); every per-mode
// quantity is then a grid-stride pass with a fixed partition into partial sums that one
// workgroup adds in a fixed order, so results do not depend on scheduling. The distance scans
// are |Σ_m A_m·exp(i·k_z,m·d)|² for every offset d: workgroups own 256 offsets each and a range
// of modes staged through LDS; their partial sums are added per offset in a fixed order.
// Which modes enter a scan (the propagating modes, optionally within an aperture) and their k_z
// are decided on the host in the same IEEE operations as the reference's NumPy expressions.
#include <algorithm>
#include <cmath>
#include <map>
#include <tuple>
#include <vector>

#include "analysis.hpp"
#include "device_mem.hpp"
#include "fft2d.hpp"
#include "interface_math.hpp"

namespace bwpm {

using detail::DevBuf;
using detail::Twiddles;
using detail::upload_vec;

namespace {

constexpr int kAThreads = 256;
constexpr int kABlocks = 256;     // partial sums per plane reduction
constexpr int kTile = 128;        // modes per LDS tile of the scans
constexpr double kGrazing = 1e-9;  // cos θ below this: not propagating (pilot_imc3d.py:17)

struct ScanMode {
  double kz;
  double2 a[3];  // E_x, E_y, E_z amplitudes (only a[0] for scalar scans)
};

// cos θ of mode (kx, ky) in a medium of wavenumber k, |E_s|² and |E_p|² of the transverse
// spectrum e (pilot_imc3d.py:26-52).
__device__ __forceinline__ void split_mode(const px2<double>& e, double kx, double ky, double k, double& cos_t,
                                           double& s2, double& p2) {
#pragma clang fp contract(off)
  const double cos_sq = 1.0 - (kx * kx + ky * ky) / (k * k);
  double c = sqrt(fmax(cos_sq, 0.0));
  c = c > kGrazing ? c : 0.0;
  const double kp = hypot(kx, ky);
  const bool on = kp > 0;
  const double safe = on ? kp : 1.0;
  const double sx = on ? -ky / safe : 0.0, sy = on ? kx / safe : 1.0;
  const double px = on ? kx / safe : 1.0, py = on ? ky / safe : 0.0;
  const double es_r = e.a.x * sx + e.b.x * sy, es_i = e.a.y * sx + e.b.y * sy;
  const double div = c > kGrazing ? c : 1.0;
  const double ep_r = (e.a.x * px + e.b.x * py) / div, ep_i = (e.a.y * px + e.b.y * py) / div;
  cos_t = c;
  s2 = es_r * es_r + es_i * es_i;
  p2 = ep_r * ep_r + ep_i * ep_i;
}

// Power reflectances (R_s, R_p) of n1 → n2 at incidence cosine cos_i; total internal reflection
// gives 1 (pilot_imc3d.py:66-79).
__device__ __forceinline__ void fresnel_rr(double n1, double n2, double cos_i, double& rs2, double& rp2) {
#pragma clang fp contract(off)
  const double r = n1 / n2;
  const double sin_t2 = (r * r) * (1.0 - cos_i * cos_i);
  if (sin_t2 >= 1.0) {
    rs2 = 1.0;
    rp2 = 1.0;
    return;
  }
  const double ct = sqrt(1.0 - sin_t2);
  const double rs = (n1 * cos_i - n2 * ct) / (n1 * cos_i + n2 * ct);
  const double rp = (n2 * cos_i - n1 * ct) / (n2 * cos_i + n1 * ct);
  rs2 = rs * rs;
  rp2 = rp * rp;
}

template <int N>
__device__ __forceinline__ void block_sum(double (&v)[N], double (*sh)[kAThreads]) {
#pragma unroll
  for (int i = 0; i < N; ++i) sh[i][threadIdx.x] = v[i];
  __syncthreads();
  for (int w = kAThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
#pragma unroll
      for (int i = 0; i < N; ++i) sh[i][threadIdx.x] += sh[i][threadIdx.x + w];
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] = sh[i][0];
}

// Partial sums over plane blockIdx.y of the spectra: [0] Σ cos θ (|E_s|² + |E_p|²),
// [1] Σ cos θ (|E_s|² R_s + |E_p|² R_p) (only with `fresnel`).
__global__ __launch_bounds__(kAThreads) void k_flux_sums(const px2<double>* __restrict__ spec, long long npix, int nx,
                                                         const double* __restrict__ kx, const double* __restrict__ ky,
                                                         double k, int fresnel, double n1, double n2,
                                                         double2* __restrict__ partial) {
#pragma clang fp contract(off)
  __shared__ double sh[2][kAThreads];
  const px2<double>* sp = spec + static_cast<long long>(blockIdx.y) * npix;
  double acc[2] = {0.0, 0.0};
  for (long long q = static_cast<long long>(blockIdx.x) * kAThreads + threadIdx.x; q < npix;
       q += static_cast<long long>(kABlocks) * kAThreads) {
    double c, s2, p2;
    split_mode(sp[q], kx[q % nx], ky[q / nx], k, c, s2, p2);
    acc[0] += c * (s2 + p2);
    if (fresnel) {
      double rs2, rp2;
      fresnel_rr(n1, n2, c, rs2, rp2);
      acc[1] += c * (s2 * rs2 + p2 * rp2);
    }
  }
  block_sum(acc, sh);
  if (threadIdx.x == 0) partial[static_cast<long long>(blockIdx.y) * kABlocks + blockIdx.x] = double2{acc[0], acc[1]};
}

// out[plane] = Σ of the kABlocks partials of that plane, added in a fixed order.
__global__ __launch_bounds__(kAThreads) void k_flux_final(const double2* __restrict__ partial, double2* __restrict__ out) {
  __shared__ double sh[2][kAThreads];
  double acc[2] = {0.0, 0.0};
  for (int i = threadIdx.x; i < kABlocks; i += kAThreads) {
    const double2 v = partial[static_cast<long long>(blockIdx.x) * kABlocks + i];
    acc[0] += v.x;
    acc[1] += v.y;
  }
  block_sum(acc, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = double2{acc[0], acc[1]};
}

// Scan modes of the axial focus (pilot_imc3d.py:115-122): the spectrum referred to the grid
// centre, a·exp(i(k_x·(nx//2)·dx + k_y·(ny//2)·dy)), and E_z = −(k_x E_x + k_y E_y)/k_z.
__global__ __launch_bounds__(kAThreads) void k_axial_modes(const px2<double>* __restrict__ spec, const int* __restrict__ q,
                                                           const double* __restrict__ kz, long long m, int nx,
                                                           const double* __restrict__ kx, const double* __restrict__ ky,
                                                           double hx, double dx, double hy, double dy,
                                                           ScanMode* __restrict__ out) {
#pragma clang fp contract(off)
  const long long i = static_cast<long long>(blockIdx.x) * kAThreads + threadIdx.x;
  if (i >= m) return;
  const int p = q[i];
  const double a = kx[p % nx], b = ky[p / nx];
  const double th = a * hx * dx + b * hy * dy;
  double s, c;
  sincos(th, &s, &c);
  const px2<double> e = spec[p];
  const double ax_r = e.a.x * c - e.a.y * s, ax_i = e.a.x * s + e.a.y * c;
  const double ay_r = e.b.x * c - e.b.y * s, ay_i = e.b.x * s + e.b.y * c;
  const double z = kz[i];
  ScanMode o;
  o.kz = z;
  o.a[0] = double2{ax_r, ax_i};
  o.a[1] = double2{ay_r, ay_i};
  o.a[2] = double2{-(a * ax_r + b * ay_r) / z, -(a * ax_i + b * ay_i) / z};
  out[i] = o;
}

// Scalar scan modes: amplitude spec[q] (one complex per mode), or the x component of a plane.
__global__ __launch_bounds__(kAThreads) void k_scalar_modes(const double2* __restrict__ amp, long long stride,
                                                            const int* __restrict__ q, const double* __restrict__ kz,
                                                            long long m, ScanMode* __restrict__ out) {
  const long long i = static_cast<long long>(blockIdx.x) * kAThreads + threadIdx.x;
  if (i >= m) return;
  ScanMode o;
  o.kz = kz[i];
  o.a[0] = amp[(q ? static_cast<long long>(q[i]) : i) * stride];
  o.a[1] = double2{0.0, 0.0};
  o.a[2] = double2{0.0, 0.0};
  out[i] = o;
}

// Σ |A_m| over the scan modes (Strehl normalisation), kABlocks partials.
__global__ __launch_bounds__(kAThreads) void k_abs_sums(const ScanMode* __restrict__ modes, long long m,
                                                        double2* __restrict__ partial) {
  __shared__ double sh[2][kAThreads];
  double acc[2] = {0.0, 0.0};
  for (long long i = static_cast<long long>(blockIdx.x) * kAThreads + threadIdx.x; i < m;
       i += static_cast<long long>(kABlocks) * kAThreads)
    acc[0] += hypot(modes[i].a[0].x, modes[i].a[0].y);
  block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = double2{acc[0], 0.0};
}

// partial[(split·noff + o)·C + c] = Σ_{m in split} A_{m,c}·exp(i k_z,m d_o); one thread per offset,
// modes staged through LDS in tiles (every thread reads the same mode: LDS broadcast).
template <int C>
__global__ __launch_bounds__(kAThreads) void k_scan(const ScanMode* __restrict__ modes, long long m, long long per_split,
                                                    const double* __restrict__ offsets, int noff,
                                                    double2* __restrict__ partial) {
#pragma clang fp contract(off)
  __shared__ ScanMode tile[kTile];
  const int o = blockIdx.x * kAThreads + threadIdx.x;
  const double d = o < noff ? offsets[o] : 0.0;
  const long long m0 = static_cast<long long>(blockIdx.y) * per_split;
  const long long m1 = std::min(m, m0 + per_split);
  double sr[C], si[C];
#pragma unroll
  for (int c = 0; c < C; ++c) sr[c] = si[c] = 0.0;
  for (long long t0 = m0; t0 < m1; t0 += kTile) {
    const int cnt = static_cast<int>(std::min<long long>(kTile, m1 - t0));
    __syncthreads();
    for (int i = threadIdx.x; i < cnt; i += kAThreads) tile[i] = modes[t0 + i];
    __syncthreads();
    for (int i = 0; i < cnt; ++i) {
      double s, c;
      sincos(tile[i].kz * d, &s, &c);
#pragma unroll
      for (int j = 0; j < C; ++j) {
        const double2 a = tile[i].a[j];
        sr[j] += a.x * c - a.y * s;
        si[j] += a.x * s + a.y * c;
      }
    }
  }
  if (o < noff) {
#pragma unroll
    for (int j = 0; j < C; ++j)
      partial[(static_cast<long long>(blockIdx.y) * noff + o) * C + j] = double2{sr[j], si[j]};
  }
}

// out[o] = Σ_c |Σ_split partial|² / div (splits added in order).
template <int C>
__global__ __launch_bounds__(kAThreads) void k_scan_final(const double2* __restrict__ partial, int splits, int noff,
                                                          double div, double* __restrict__ out) {
#pragma clang fp contract(off)
  const int o = blockIdx.x * kAThreads + threadIdx.x;
  if (o >= noff) return;
  double v = 0.0;
#pragma unroll
  for (int j = 0; j < C; ++j) {
    double r = 0.0, i = 0.0;
    for (int s = 0; s < splits; ++s) {
      const double2 p = partial[(static_cast<long long>(s) * noff + o) * C + j];
      r += p.x;
      i += p.y;
    }
    v += r * r + i * i;
  }
  out[o] = v / div;
}

__global__ void k_widen(const px2<float>* __restrict__ in, px2<double>* __restrict__ out, long long n) {
  const long long i = static_cast<long long>(blockIdx.x) * kAThreads + threadIdx.x;
  if (i >= n) return;
  const px2<float> v = in[i];
  out[i] = px2<double>{cx<double>{v.a.x, v.a.y}, cx<double>{v.b.x, v.b.y}};
}

unsigned blocks_for(long long n) { return static_cast<unsigned>((n + kAThreads - 1) / kAThreads); }

}  // namespace

// ---- the handle --------------------------------------------------------------------------------

struct Analysis::Impl {
  int device = 0;
  long long ny = 0, nx = 0, npix = 0;
  double dy = 0, dx = 0, k0 = 0;
  hipStream_t stream = nullptr;
  std::vector<double> kx, ky;  // host, 2π·numpy.fft.fftfreq
  DevBuf dkx, dky;
  Twiddles<double> tx, ty;
  Fft2D<double, false> fft;
  DevBuf work, partial, sums, modes, scan_part, offs, prof;
  size_t work_planes = 0;
  struct ModeList {
    DevBuf q, kz;
    long long m = 0;
  };
  std::map<std::tuple<int, double, double, double>, ModeList> lists;

  void ensure(DevBuf& b, size_t bytes) {
    if (b.bytes() < bytes) b.alloc(bytes);
  }
  // Spectra of `count` planes (≤ kMaxPlanes) in the work buffer.
  void spectra(const void* planes, int where, long long count) {
    const size_t pb = npix * sizeof(px2<double>);
    ensure(work, count * pb);
    for (long long i = 0; i < count; ++i) {
      px2<double>* dst = work.as<px2<double>>() + i * npix;
      if (where == BWPM_PLANE_HOST_C128) {
        HIP_CHECK(hipMemcpyAsync(dst, static_cast<const px2<double>*>(planes) + i * npix, pb, hipMemcpyHostToDevice,
                                 stream));
      } else if (where == BWPM_PLANE_DEVICE_C128) {
        HIP_CHECK(hipMemcpyAsync(dst, static_cast<const px2<double>*>(planes) + i * npix, pb, hipMemcpyDeviceToDevice,
                                 stream));
      } else {
        hipLaunchKernelGGL(k_widen, dim3(blocks_for(npix)), dim3(kAThreads), 0, stream,
                           static_cast<const px2<float>*>(planes) + i * npix, dst, npix);
        HIP_CHECK(hipGetLastError());
      }
    }
    px2<double>* pl[kMaxPlanes] = {};
    for (long long i = 0; i < count; ++i) pl[i] = work.as<px2<double>>() + i * npix;
    fft.planes(pl, static_cast<int>(count), false, nullptr);
  }
  // Σ cos θ (|E_s|² + |E_p|²) [and the Fresnel-weighted sum] of `count` spectra.
  std::vector<double2> flux(long long count, double k, bool fresnel, double n1, double n2) {
    ensure(partial, count * kABlocks * sizeof(double2));
    ensure(sums, count * sizeof(double2));
    hipLaunchKernelGGL(k_flux_sums, dim3(kABlocks, static_cast<unsigned>(count)), dim3(kAThreads), 0, stream,
                       work.as<px2<double>>(), npix, static_cast<int>(nx), dkx.as<double>(), dky.as<double>(), k,
                       fresnel ? 1 : 0, n1, n2, partial.as<double2>());
    HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_flux_final, dim3(static_cast<unsigned>(count)), dim3(kAThreads), 0, stream,
                       partial.as<double2>(), sums.as<double2>());
    HIP_CHECK(hipGetLastError());
    std::vector<double2> h(count);
    HIP_CHECK(hipMemcpyAsync(h.data(), sums.as<void>(), count * sizeof(double2), hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    return h;
  }
  // Modes of a scan and their k_z. kind 0: axial focus, propagating with cos θ > 1e-9 and, with
  // na ≥ 0, |k⊥| ≤ na·k0 (pilot_imc3d.py:43-52,117-120); kind 1: Strehl support
  // sqrt(k_x² + k_y²) ≤ na·k0 with k_z = sqrt(k² − k_x² − k_y²) (b3_scaling.py:89-92).
  const ModeList& mode_list(int kind, double a, double na) {
    auto key = std::make_tuple(kind, a, na, k0);
    auto it = lists.find(key);
    if (it != lists.end()) return it->second;
    std::vector<int> q;
    std::vector<double> kz;
    for (long long j = 0; j < ny; ++j)
      for (long long i = 0; i < nx; ++i) {
        const double x = kx[i], y = ky[j];
        if (kind == 0) {
          const double k = a * k0;
          const double cos_sq = 1.0 - (x * x + y * y) / (k * k);
          double c = std::sqrt(std::max(cos_sq, 0.0));
          c = c > kGrazing ? c : 0.0;
          if (!(c > 0)) continue;
          if (na >= 0 && !(std::hypot(x, y) <= na * k0)) continue;
          q.push_back(static_cast<int>(j * nx + i));
          kz.push_back(k * c);
        } else {
          if (!(std::sqrt(x * x + y * y) <= na * k0)) continue;
          q.push_back(static_cast<int>(j * nx + i));
          kz.push_back(std::sqrt(a * a - x * x - y * y));
        }
      }
    ModeList l;
    l.m = static_cast<long long>(q.size());
    l.q = upload_vec(q, stream);
    l.kz = upload_vec(kz, stream);
    return lists.emplace(key, std::move(l)).first->second;
  }
  // Scan of `m` prepared modes over the offsets; out[o] = Σ_c |Σ_m A_m,c e^{i k_z d_o}|² / div.
  template <int C>
  void scan(long long m, const double* offsets, int noff, double div, double* out) {
    const unsigned oblocks = static_cast<unsigned>((noff + kAThreads - 1) / kAThreads);
    // about 1024 workgroups in all, each with at least one tile of modes
    long long splits = std::max<long long>(1, std::min<long long>((m + kTile - 1) / kTile, 1024 / oblocks));
    const long long per_split = ((m + splits - 1) / splits + kTile - 1) / kTile * kTile;
    splits = std::max<long long>(1, (m + per_split - 1) / per_split);
    ensure(offs, noff * sizeof(double));
    ensure(scan_part, static_cast<size_t>(splits) * noff * C * sizeof(double2));
    ensure(prof, noff * sizeof(double));
    HIP_CHECK(hipMemcpyAsync(offs.as<void>(), offsets, noff * sizeof(double), hipMemcpyHostToDevice, stream));
    if (m > 0) {
      hipLaunchKernelGGL((k_scan<C>), dim3(oblocks, static_cast<unsigned>(splits)), dim3(kAThreads), 0, stream,
                         modes.as<ScanMode>(), m, per_split, offs.as<double>(), noff, scan_part.as<double2>());
      HIP_CHECK(hipGetLastError());
    } else {
      splits = 0;
    }
    hipLaunchKernelGGL((k_scan_final<C>), dim3(oblocks), dim3(kAThreads), 0, stream, scan_part.as<double2>(),
                       static_cast<int>(splits), noff, div, prof.as<double>());
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(out, prof.as<void>(), noff * sizeof(double), hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
  }
  double abs_sum(long long m) {
    ensure(partial, kABlocks * sizeof(double2));
    ensure(sums, sizeof(double2));
    hipLaunchKernelGGL(k_abs_sums, dim3(kABlocks), dim3(kAThreads), 0, stream, modes.as<ScanMode>(), m,
                       partial.as<double2>());
    HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_flux_final, dim3(1), dim3(kAThreads), 0, stream, partial.as<double2>(), sums.as<double2>());
    HIP_CHECK(hipGetLastError());
    double2 h{};
    HIP_CHECK(hipMemcpyAsync(&h, sums.as<void>(), sizeof(double2), hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    return h.x;
  }
};

Analysis::Analysis(int device, long long ny, long long nx, double dy, double dx, double k0) : p_(new Impl) {
  require(ny >= 1 && nx >= 1, "the grid must have ny, nx >= 1");
  require(dy > 0 && dx > 0 && k0 > 0 && std::isfinite(dy) && std::isfinite(dx) && std::isfinite(k0),
          "dy, dx and k0 must be positive");
  require(ny * nx <= (1LL << 31) - 1, "grid too large");
  Impl& s = *p_;
  s.device = device;
  s.ny = ny;
  s.nx = nx;
  s.npix = ny * nx;
  s.dy = dy;
  s.dx = dx;
  s.k0 = k0;
  HIP_CHECK(hipSetDevice(device));
  HIP_CHECK(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
  try {
    s.kx.resize(nx);
    s.ky.resize(ny);
    for (long long i = 0; i < nx; ++i) s.kx[i] = fftfreq(i, nx, dx);
    for (long long j = 0; j < ny; ++j) s.ky[j] = fftfreq(j, ny, dy);
    s.dkx = upload_vec(s.kx, s.stream);
    s.dky = upload_vec(s.ky, s.stream);
    s.tx.build(nx, s.stream);
    s.ty.build(ny, s.stream);
    s.fft.init(ny, nx, s.tx.plan, s.ty.plan, s.stream);
    HIP_CHECK(hipStreamSynchronize(s.stream));
  } catch (...) {
    (void)hipStreamDestroy(s.stream);
    throw;
  }
}

Analysis::~Analysis() {
  (void)hipSetDevice(p_->device);
  (void)hipStreamSynchronize(p_->stream);
  const hipStream_t st = p_->stream;
  p_.reset();
  (void)hipStreamDestroy(st);
}

namespace {
void check_where(int where) {
  require(where == BWPM_PLANE_HOST_C128 || where == BWPM_PLANE_DEVICE_C128 || where == BWPM_PLANE_DEVICE_C64,
          "where must be BWPM_PLANE_HOST_C128, BWPM_PLANE_DEVICE_C128 or BWPM_PLANE_DEVICE_C64");
}
void check_offsets(const double* offsets, long long noff) {
  require(offsets != nullptr && noff >= 1 && noff <= (1LL << 24), "offsets must hold 1 … 2^24 values");
}
}  // namespace

std::vector<double> Analysis::plane_power(const void* planes, int where, long long count, double index) {
  check_where(where);
  require(planes != nullptr && count >= 1, "planes missing");
  require(std::isfinite(index) && index > 0, "index must be positive");
  Impl& s = *p_;
  HIP_CHECK(hipSetDevice(s.device));
  std::vector<double> out;
  const size_t esz = where == BWPM_PLANE_DEVICE_C64 ? sizeof(px2<float>) : sizeof(px2<double>);
  for (long long i0 = 0; i0 < count; i0 += kMaxPlanes) {
    const long long c = std::min<long long>(kMaxPlanes, count - i0);
    s.spectra(static_cast<const char*>(planes) + i0 * s.npix * esz, where, c);
    for (const double2& v : s.flux(c, index * s.k0, false, 0, 0)) out.push_back(index * v.x);
  }
  return out;
}

double Analysis::reference_reflectance(const void* plane, int where, double n1, double n2) {
  check_where(where);
  require(plane != nullptr, "plane missing");
  require(std::isfinite(n1) && std::isfinite(n2) && n1 > 0 && n2 > 0, "n1 and n2 must be positive");
  Impl& s = *p_;
  HIP_CHECK(hipSetDevice(s.device));
  s.spectra(plane, where, 1);
  const double2 v = s.flux(1, n1 * s.k0, true, n1, n2)[0];
  if (v.x == 0.0) throw Error(kEinval, "the field carries no propagating flux");
  return v.y / v.x;
}

void Analysis::axial_profile(const void* plane, int where, double index, const double* offsets, long long noff,
                             double na, double* intensity) {
  check_where(where);
  check_offsets(offsets, noff);
  require(plane != nullptr && intensity != nullptr, "null argument");
  require(std::isfinite(index) && index > 0, "index must be positive");
  require(std::isfinite(na), "na must be finite (< 0: no aperture)");
  Impl& s = *p_;
  HIP_CHECK(hipSetDevice(s.device));
  s.spectra(plane, where, 1);
  const Impl::ModeList& l = s.mode_list(0, index, na);
  s.ensure(s.modes, std::max<long long>(l.m, 1) * sizeof(ScanMode));
  if (l.m > 0) {
    hipLaunchKernelGGL(k_axial_modes, dim3(blocks_for(l.m)), dim3(kAThreads), 0, s.stream, s.work.as<px2<double>>(),
                       l.q.as<int>(), l.kz.as<double>(), l.m, static_cast<int>(s.nx), s.dkx.as<double>(),
                       s.dky.as<double>(), static_cast<double>(s.nx / 2), s.dx, static_cast<double>(s.ny / 2), s.dy,
                       s.modes.as<ScanMode>());
    HIP_CHECK(hipGetLastError());
  }
  const double n2 = static_cast<double>(s.npix) * static_cast<double>(s.npix);
  s.scan<3>(l.m, offsets, static_cast<int>(noff), n2, intensity);
}

void Analysis::strehl_profile(const void* spectrum, int where, double k, double na, const double* offsets,
                              long long noff, double* values) {
  require(where == BWPM_PLANE_HOST_C128 || where == BWPM_PLANE_DEVICE_C128,
          "the spectrum must be complex128 (host or device)");
  check_offsets(offsets, noff);
  require(spectrum != nullptr && values != nullptr, "null argument");
  require(std::isfinite(k) && k > 0 && std::isfinite(na) && na > 0, "k and na must be positive");
  Impl& s = *p_;
  HIP_CHECK(hipSetDevice(s.device));
  const Impl::ModeList& l = s.mode_list(1, k, na);
  require(l.m > 0, "no mode lies within the aperture");
  const size_t bytes = s.npix * sizeof(double2);
  s.ensure(s.work, bytes);
  HIP_CHECK(hipMemcpyAsync(s.work.as<void>(), spectrum, bytes,
                           where == BWPM_PLANE_HOST_C128 ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, s.stream));
  s.ensure(s.modes, l.m * sizeof(ScanMode));
  hipLaunchKernelGGL(k_scalar_modes, dim3(blocks_for(l.m)), dim3(kAThreads), 0, s.stream, s.work.as<double2>(), 1LL,
                     l.q.as<int>(), l.kz.as<double>(), l.m, s.modes.as<ScanMode>());
  HIP_CHECK(hipGetLastError());
  const double a = s.abs_sum(l.m);
  s.scan<1>(l.m, offsets, static_cast<int>(noff), a * a, values);
}

void Analysis::strehl_modes(const double* ky, const void* spectrum, long long m, double k, const double* offsets,
                            long long noff, double* values) {
  check_offsets(offsets, noff);
  require(ky != nullptr && spectrum != nullptr && values != nullptr && m >= 1, "null argument");
  require(std::isfinite(k) && k > 0, "k must be positive");
  Impl& s = *p_;
  HIP_CHECK(hipSetDevice(s.device));
  // k_z = sqrt(max(k² − k_y², 0)) (pilot_imc.py:194)
  std::vector<double> kz(m);
  for (long long i = 0; i < m; ++i) kz[i] = std::sqrt(std::max(k * k - ky[i] * ky[i], 0.0));
  DevBuf dkz = upload_vec(kz, s.stream);
  DevBuf amp(m * sizeof(double2));
  HIP_CHECK(hipMemcpyAsync(amp.as<void>(), spectrum, m * sizeof(double2), hipMemcpyHostToDevice, s.stream));
  s.ensure(s.modes, m * sizeof(ScanMode));
  hipLaunchKernelGGL(k_scalar_modes, dim3(blocks_for(m)), dim3(kAThreads), 0, s.stream, amp.as<double2>(), 1LL,
                     static_cast<const int*>(nullptr), dkz.as<double>(), m, s.modes.as<ScanMode>());
  HIP_CHECK(hipGetLastError());
  const double a = s.abs_sum(m);
  s.scan<1>(m, offsets, static_cast<int>(noff), a * a, values);
}

}  // namespace bwpm
