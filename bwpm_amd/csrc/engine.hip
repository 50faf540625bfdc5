// Non-templated parts of the engine: precision dispatch, batched lanes, the exported FP64
// interface functions and the HBM copy probe.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <chrono>
#include <map>
#include <thread>
#include <tuple>

#include "engine.hpp"
#include "interface_math.hpp"

namespace bwpm {

namespace {

class Buf {
 public:
  explicit Buf(size_t bytes) {
    const hipError_t e = hipMalloc(&p_, bytes);
    if (e != hipSuccess) throw Error(kEnomem, std::string("device allocation failed: ") + hipGetErrorString(e));
  }
  ~Buf() { (void)hipFree(p_); }
  Buf(const Buf&) = delete;
  Buf& operator=(const Buf&) = delete;
  template <typename U>
  U* as() const { return static_cast<U*>(p_); }

 private:
  void* p_ = nullptr;
};

}  // namespace

// Exported interface functions on arbitrary k arrays: full row-major 2×2 (which = 0: T, 1: R).
__global__ void k_interface_eval(const double* kx, const double* ky, long long n, int which, zd n1,
                                 zd n2, double k0, zd* out) {
  const long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const Mat2 m = which == 0 ? transmission(kx[i], ky[i], n1, n2, k0) : reflection(kx[i], ky[i], n1, n2, k0);
  out[4 * i] = m.m00;
  out[4 * i + 1] = m.m01;
  out[4 * i + 2] = m.m01;
  out[4 * i + 3] = m.m11;
}

__global__ void k_fresnel_eval(const double* kx, const double* ky, long long n, zd n1, zd n2, double k0, zd* out) {
  const long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const Fresnel f = fresnel(kx[i], ky[i], n1, n2, k0);
  out[6 * i] = f.r_s;
  out[6 * i + 1] = f.r_p;
  out[6 * i + 2] = f.t_s;
  out[6 * i + 3] = f.t_p;
  out[6 * i + 4] = f.kz_in;
  out[6 * i + 5] = f.kz_out;
}

__global__ void k_kz_eval(zd n, double k0, const double* kp2, long long cnt, zd* out) {
  const long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
  if (i < cnt) out[i] = kz_of(n, k0, kp2[i]);
}


std::unique_ptr<EngineBase> make_engine(Spec spec) {
  switch (spec.precision) {
    case 64: return make_engine_f64(std::move(spec));
    case 48:
      return make_engine_mixed(std::move(spec));
    case 40:
      return make_engine_bfp(std::move(spec));
    case 32: return make_engine_f32(std::move(spec), true);
    case 31: return make_engine_f32(std::move(spec), false);
    default: throw Error(kEinval, "precision must be 64, 48, 40, 32 or 31");
  }
}

namespace {

void check_geometry(const Spec& s) {
  const LabelGeometry& g = s.geom;
  require(g.background >= 0 && static_cast<size_t>(g.background) < s.indices.size(), "labels must index into `indices`");
  for (const GeomPrim& p : g.prims) {
    require(p.kind >= 0 && p.kind <= 2, "geometry primitive kind must be 0 (plane), 1 (layer) or 2 (cylinder)");
    require(p.label >= 0 && static_cast<size_t>(p.label) < s.indices.size(), "labels must index into `indices`");
  }
}

// d + a·x + b·y of a plane primitive and the distance test of a cylinder: the pixel-centre rule's
// expressions, shared by both ways of making labels so that they agree bit for bit.
inline double plane_rhs(const GeomPrim& p, double x, double y) { return p.d + p.a * x + p.b * y; }
inline bool in_disc(const GeomPrim& p, double x, double y) {
  const double u = x - p.cx, v = y - p.cy;
  return std::sqrt(u * u + v * v) <= p.r;
}
inline double slab_z(const Spec& s, long long j) { return s.geom.z0 + (static_cast<double>(j) + 0.5) * s.dz; }

void pixel_coords(const Spec& s, std::vector<double>& xs, std::vector<double>& ys) {
  xs.resize(s.nx);
  ys.resize(s.ny);
  for (long long m = 0; m < s.nx; ++m) xs[m] = static_cast<double>(m - s.nx / 2) * s.dx;
  for (long long n = 0; n < s.ny; ++n) ys[n] = static_cast<double>(n - s.ny / 2) * s.dy;
}

// Runs f(j) for j < n on up to 16 host threads.
template <typename F>
void parallel_for(long long n, F f) {
  const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  std::atomic<long long> next{0};
  for (unsigned t = 0; t < nt; ++t)
    th.emplace_back([&] {
      for (long long j = next++; j < n; j = next++) f(j);
    });
  for (auto& t : th) t.join();
}

}  // namespace

void geometry_labels(const Spec& s, int32_t* out) {
  check_geometry(s);
  const LabelGeometry& g = s.geom;
  const long long ny = s.ny, nx = s.nx, npix = ny * nx;
  std::vector<double> xs, ys;
  pixel_coords(s, xs, ys);
  parallel_for(s.nz, [&](long long j) {
    const double z = slab_z(s, j);
    int32_t* o = out + j * npix;
    std::fill(o, o + npix, g.background);
    for (const GeomPrim& p : g.prims) {
      if (p.kind != 0 && !(p.zlo < z && z <= p.zhi)) continue;
      if (p.kind == 1) {
        std::fill(o, o + npix, p.label);
        continue;
      }
      for (long long n = 0; n < ny; ++n)
        for (long long m = 0; m < nx; ++m) {
          const bool in = p.kind == 0 ? z > plane_rhs(p, xs[m], ys[n]) : in_disc(p, xs[m], ys[n]);
          if (in) o[n * nx + m] = p.label;
        }
    }
  });
}

void geometry_planes(const Spec& s, std::vector<std::vector<int32_t>>& planes, std::vector<int>& plane_of) {
  check_geometry(s);
  const LabelGeometry& g = s.geom;
  const long long ny = s.ny, nx = s.nx, npix = ny * nx, nz = s.nz;
  std::vector<double> xs, ys;
  pixel_coords(s, xs, ys);
  // per primitive, once: the planes' right-hand sides and their range, the cylinders' discs
  struct Pre {
    std::vector<double> rhs;
    std::vector<char> disc;
    double lo = 0, hi = 0;
    bool any = false, all = false;
  };
  std::vector<Pre> pre(g.prims.size());
  for (size_t k = 0; k < g.prims.size(); ++k) {
    const GeomPrim& p = g.prims[k];
    Pre& q = pre[k];
    if (p.kind == 0) {
      q.rhs.resize(npix);
      for (long long n = 0; n < ny; ++n)
        for (long long m = 0; m < nx; ++m) q.rhs[n * nx + m] = plane_rhs(p, xs[m], ys[n]);
      const auto mm = std::minmax_element(q.rhs.begin(), q.rhs.end());
      q.lo = *mm.first;
      q.hi = *mm.second;
    } else if (p.kind == 2) {
      q.disc.resize(npix);
      for (long long n = 0; n < ny; ++n)
        for (long long m = 0; m < nx; ++m) q.disc[n * nx + m] = in_disc(p, xs[m], ys[n]) ? 1 : 0;
      q.any = std::any_of(q.disc.begin(), q.disc.end(), [](char c) { return c != 0; });
      q.all = std::all_of(q.disc.begin(), q.disc.end(), [](char c) { return c != 0; });
    }
  }
  // pass 1: the slabs whose primitives are all uniform over the grid, and their label
  std::vector<int32_t> uni_label(nz, -1);
  for (long long j = 0; j < nz; ++j) {
    const double z = slab_z(s, j);
    bool uni = true;
    int32_t lab = g.background;
    for (size_t k = 0; k < g.prims.size(); ++k) {
      const GeomPrim& p = g.prims[k];
      const Pre& q = pre[k];
      int state;  // the pixels a primitive covers in this slab: 0 none, 1 all, 2 some
      if (p.kind == 0)
        state = z > q.hi ? 1 : (z > q.lo ? 2 : 0);  // z > rhs at all, some or no pixels
      else if (!(p.zlo < z && z <= p.zhi))
        state = 0;
      else
        state = p.kind == 1 ? 1 : (q.all ? 1 : (q.any ? 2 : 0));
      if (state == 1) {  // overwrites every pixel: uniform again
        uni = true;
        lab = p.label;
      } else if (state == 2) {
        uni = false;
      }
    }
    if (uni) uni_label[j] = lab;
  }
  // pass 2: the other slabs' planes, in parallel
  std::vector<long long> mixed;
  for (long long j = 0; j < nz; ++j)
    if (uni_label[j] < 0) mixed.push_back(j);
  std::vector<std::vector<int32_t>> own(mixed.size());
  parallel_for(static_cast<long long>(mixed.size()), [&](long long i) {
    const long long j = mixed[i];
    const double z = slab_z(s, j);
    std::vector<int32_t>& o = own[i];
    o.assign(npix, g.background);
    for (size_t k = 0; k < g.prims.size(); ++k) {
      const GeomPrim& p = g.prims[k];
      if (p.kind != 0 && !(p.zlo < z && z <= p.zhi)) continue;
      if (p.kind == 1) {
        std::fill(o.begin(), o.end(), p.label);
        continue;
      }
      const Pre& q = pre[k];
      for (long long t = 0; t < npix; ++t)
        if (p.kind == 0 ? z > q.rhs[t] : q.disc[t] != 0) o[t] = p.label;
    }
  });
  // assemble: one plane per uniform label; a mixed plane unless it equals the previous slab's
  planes.clear();
  plane_of.assign(nz, -1);
  std::map<int32_t, int> uni_plane;
  size_t next_mixed = 0;
  for (long long j = 0; j < nz; ++j) {
    if (uni_label[j] >= 0) {
      auto it = uni_plane.find(uni_label[j]);
      if (it == uni_plane.end()) {
        planes.emplace_back(npix, uni_label[j]);
        it = uni_plane.emplace(uni_label[j], static_cast<int>(planes.size()) - 1).first;
      }
      plane_of[j] = it->second;
      continue;
    }
    std::vector<int32_t>& o = own[next_mixed++];
    if (j > 0 && uni_label[j - 1] < 0 && planes[plane_of[j - 1]] == o) {
      plane_of[j] = plane_of[j - 1];
    } else {
      planes.push_back(std::move(o));
      plane_of[j] = static_cast<int>(planes.size()) - 1;
    }
  }
}

LaneGraphCounters& lane_graph_counters() {
  static LaneGraphCounters c;
  return c;
}

void run_continuous_lane(int precision, LaneFeed& feed, int slots, int chunk, hipStream_t stream) {
  if (precision == 64)
    run_continuous_lane_f64(feed, slots, chunk, stream);
  else if (precision == 48)
    run_continuous_lane_mixed(feed, slots, chunk, stream);
  else if (precision == 40)
    run_continuous_lane_bfp(feed, slots, chunk, stream);
  else
    run_continuous_lane_f32(feed, slots, chunk, stream, precision == 32);
}

size_t estimate_device_bytes(const Spec& s, int iterations) {
  const bool f64 = s.precision == 64 || s.precision == 48 || s.precision == 40;
  const size_t elem = f64 ? 16 : 8;                                           // complex
  const size_t selem = s.precision == 48 || s.precision == 40 ? 8 : elem;     // streaming planes
  const size_t npix = static_cast<size_t>(s.ny) * static_cast<size_t>(s.nx);
  const size_t plane = 2 * npix * elem;            // (E_x, E_y) plane
  const size_t splane = 2 * npix * selem;
  const size_t nit = static_cast<size_t>(std::max(iterations, 1));
  std::vector<c128> media;
  for (const c128& v : s.indices)
    if (std::find(media.begin(), media.end(), v) == media.end()) media.push_back(v);
  if (std::find(media.begin(), media.end(), s.index_before) == media.end()) media.push_back(s.index_before);
  const size_t m = media.size();
  // staging (complex128), input, e, per-iteration outputs; X and up to four product planes
  size_t bytes = 2 * npix * 16 + plane * (2 + 2 * nit) + splane * 5;
  // phase tables per medium and 2×2 tables per ordered medium pair and kind (hi and lo parts)
  bytes += m * 2 * npix * elem + (m > 1 ? m * (m - 1) * 4 * 6 * npix * elem : 0);
  // sparse interface values (two sets) and lists, bounded by two planes
  bytes += 2 * plane + 2 * npix * 8;
  // Born: potential and emitted planes per slab, per-host tables (the per-k struct of 13 complex
  // values, 12 field planes and 3 operator planes), Krylov basis and 13 work planes
  if (!s.fluct.empty() || (s.synth.on && s.synth.sigma > 0))
    bytes += static_cast<size_t>(s.nz) * (npix * elem + 2 * plane) + m * npix * 28 * elem + (31 + 13) * plane;
  return bytes;
}

double run_batch(std::vector<EngineBase*>& engines, int iterations, int max_lane, bool time_kernels,
                 KernelTimes* kt) {
  require(max_lane >= 1, "lane size must be >= 1");
  // group into lanes: equal precision, grid and device; solves with δn run one per lane
  std::map<std::tuple<int, long long, long long, int>, std::vector<EngineBase*>> groups;
  std::vector<std::vector<EngineBase*>> lanes;
  for (EngineBase* e : engines) {
    if (e->has_born()) {
      lanes.push_back({e});
      continue;
    }
    auto& g = groups[std::make_tuple(e->precision(), e->ny(), e->nx(), e->device())];
    g.push_back(e);
    if (static_cast<int>(g.size()) == max_lane) {
      lanes.push_back(g);
      g.clear();
    }
  }
  for (auto& kv : groups)
    if (!kv.second.empty()) lanes.push_back(kv.second);
  // one host thread and one stream per lane (the stream of the lane's first engine)
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<std::thread> threads;
  std::vector<std::string> errors(lanes.size());
  std::vector<int> codes(lanes.size(), kOk);
  std::vector<KernelTimes> times(lanes.size());
  for (size_t i = 0; i < lanes.size(); ++i)
    threads.emplace_back([&, i] {
      try {
        auto& lane = lanes[i];
        HIP_CHECK(hipSetDevice(lane[0]->device()));
        std::vector<std::vector<long long>> keeps(lane.size());
        lane[0]->run_lane(lane, iterations, keeps, lane[0]->stream(), time_kernels, &times[i]);
      } catch (const Error& e) {
        errors[i] = e.what();
        codes[i] = e.code;
      } catch (const std::exception& e) {
        errors[i] = e.what();
        codes[i] = kEinternal;
      }
    });
  for (auto& t : threads) t.join();
  for (size_t i = 0; i < lanes.size(); ++i)
    if (codes[i] != kOk) throw Error(codes[i], errors[i]);
  if (kt) {
    *kt = KernelTimes{};
    for (const KernelTimes& t : times) {
      kt->col_ms += t.col_ms;
      kt->row_ms += t.row_ms;
      kt->col_launches += t.col_launches;
      kt->row_launches += t.row_launches;
    }
  }
  const auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double, std::milli>(t1 - t0).count();
}

void interface_matrices(int which, const double* kx, const double* ky, long long n, c128 n_in, c128 n_out,
                        double k0, c128* out, int device) {
  HIP_CHECK(hipSetDevice(device));
  if (n <= 0) return;
  hipStream_t s;
  HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  try {
    Buf dkx(n * sizeof(double)), dky(n * sizeof(double)), dout(4 * n * sizeof(zd));
    HIP_CHECK(hipMemcpyAsync(dkx.as<void>(), kx, n * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(dky.as<void>(), ky, n * sizeof(double), hipMemcpyHostToDevice, s));
    const unsigned grid = static_cast<unsigned>((n + 255) / 256);
    hipLaunchKernelGGL(k_interface_eval, dim3(grid), dim3(256), 0, s, dkx.as<double>(), dky.as<double>(), n, which,
                       zd{n_in.real(), n_in.imag()}, zd{n_out.real(), n_out.imag()}, k0, dout.as<zd>());
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(out, dout.as<void>(), 4 * n * sizeof(zd), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipStreamDestroy(s);
    throw;
  }
  HIP_CHECK(hipStreamDestroy(s));
}

__global__ void k_copy_probe(const float4* __restrict__ src, float4* __restrict__ dst, long long n) {
  for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x)
    dst[i] = src[i];
}

double copy_probe(long long bytes, int reps, int device) {
  HIP_CHECK(hipSetDevice(device));
  require(bytes >= 16 && reps >= 1, "copy probe needs bytes >= 16 and reps >= 1");
  const long long n = bytes / 16;
  Buf a(n * 16), b(n * 16);
  hipStream_t s;
  HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  // one float4 per thread (a grid-stride loop with fewer workgroups reads HBM more slowly on this
  // device: 6.2 against about 4.5 TB/s at 1 GiB)
  const unsigned grid = static_cast<unsigned>(std::min<long long>((n + 255) / 256, 0x7fffffffLL));
  hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, s, a.as<float4>(), b.as<float4>(), n);
  HIP_CHECK(hipEventRecord(e0, s));
  for (int r = 0; r < reps; ++r)
    hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, s, a.as<float4>(), b.as<float4>(), n);
  HIP_CHECK(hipEventRecord(e1, s));
  HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0;
  HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(s);
  return ms / reps;
}

void fresnel_coefficients(const double* kx, const double* ky, long long n, c128 n_in, c128 n_out, double k0,
                          c128* out, int device) {
  HIP_CHECK(hipSetDevice(device));
  if (n <= 0) return;
  hipStream_t s;
  HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  try {
    Buf dkx(n * sizeof(double)), dky(n * sizeof(double)), dout(6 * n * sizeof(zd));
    HIP_CHECK(hipMemcpyAsync(dkx.as<void>(), kx, n * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(dky.as<void>(), ky, n * sizeof(double), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_fresnel_eval, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s, dkx.as<double>(),
                       dky.as<double>(), n, zd{n_in.real(), n_in.imag()}, zd{n_out.real(), n_out.imag()}, k0,
                       dout.as<zd>());
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(out, dout.as<void>(), 6 * n * sizeof(zd), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipStreamDestroy(s);
    throw;
  }
  HIP_CHECK(hipStreamDestroy(s));
}

void longitudinal_wavenumbers(c128 index, double k0, const double* kp2, long long n, c128* out, int device) {
  HIP_CHECK(hipSetDevice(device));
  if (n <= 0) return;
  hipStream_t s;
  HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  try {
    Buf dk(n * sizeof(double)), dout(n * sizeof(zd));
    HIP_CHECK(hipMemcpyAsync(dk.as<void>(), kp2, n * sizeof(double), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_kz_eval, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s,
                       zd{index.real(), index.imag()}, k0, dk.as<double>(), n, dout.as<zd>());
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(out, dout.as<void>(), n * sizeof(zd), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipStreamDestroy(s);
    throw;
  }
  HIP_CHECK(hipStreamDestroy(s));
}

}  // namespace bwpm
