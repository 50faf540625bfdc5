// extern "C" entry points of libbwpm_hip.so (declared in include/bwpm_hip.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "../../include/bwpm_hip.h"
#include "analysis.hpp"
#include "engine.hpp"
#include "serve.hpp"

using bwpm::c128;

struct bwpm_solver {
  std::unique_ptr<bwpm::EngineBase> engine;
  int device = 0;
};

struct bwpm_result {
  bwpm::RunOutput out;
  long long npix = 0;
};

struct bwpm_queue {
  std::unique_ptr<bwpm::Queue> q;
};

struct bwpm_analysis {
  std::unique_ptr<bwpm::Analysis> a;
};

namespace {

thread_local std::string g_error;

template <typename F>
int guarded(F&& f) {
  try {
    f();
    g_error.clear();
    return BWPM_OK;
  } catch (const bwpm::Error& e) {
    g_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_error = "host allocation failed";
    return BWPM_ENOMEM;
  } catch (const std::exception& e) {
    g_error = e.what();
    return BWPM_EINTERNAL;
  } catch (...) {
    g_error = "unknown error";
    return BWPM_EINTERNAL;
  }
}

void copy_field(const std::vector<c128>& v, double* out) {
  std::memcpy(out, v.data(), v.size() * sizeof(c128));
}

int interface_call(int which, const double* kx, const double* ky, int64_t n, double a, double b, double c,
                   double d, double k0, double* out) {
  return guarded([&] {
    bwpm::require(n >= 0, "n must be >= 0");
    bwpm::require(n == 0 || (kx && ky && out), "null array");
    int dev = 0;
    (void)hipGetDevice(&dev);
    bwpm::interface_matrices(which, kx, ky, n, c128(a, b), c128(c, d), k0, reinterpret_cast<c128*>(out), dev);
  });
}


bwpm::Spec make_spec(int64_t ny, int64_t nx, double dy, double dx, double dz, double k0, const double* indices,
                   int64_t m, const int32_t* labels, int64_t nz, double ib_re, double ib_im, const double* mask,
                   const double* fluct, const int32_t* labels_before, int device, int precision,
                   int uniform_label = -1) {
  bwpm::require(ny >= 1 && nx >= 1 && nz >= 1, "labels must have shape (nz, ny, nx)");
  bwpm::require(m >= 0 && (m == 0 || indices), "indices missing");
  bwpm::require(labels != nullptr || uniform_label >= 0, "labels missing");
  bwpm::require(precision == BWPM_FP64 || precision == BWPM_MIXED || precision == BWPM_BFP ||
                    precision == BWPM_FP32 || precision == BWPM_FP32_PLAIN,
                "precision must be BWPM_FP64, BWPM_MIXED, BWPM_BFP, BWPM_FP32 or BWPM_FP32_PLAIN");
  bwpm::Spec s;
  s.ny = ny;
  s.nx = nx;
  s.nz = nz;
  s.dy = dy;
  s.dx = dx;
  s.dz = dz;
  s.k0 = k0;
  s.indices.assign(reinterpret_cast<const c128*>(indices), reinterpret_cast<const c128*>(indices) + m);
  if (labels)
    s.labels.assign(labels, labels + nz * ny * nx);
  else
    s.uniform_label = uniform_label;
  s.index_before = c128(ib_re, ib_im);
  if (mask) s.mask.assign(mask, mask + ny * nx);
  if (fluct) {
    const c128* f = reinterpret_cast<const c128*>(fluct);
    s.fluct.assign(f, f + nz * ny * nx);
  }
  if (labels_before) s.labels_before.assign(labels_before, labels_before + ny * nx);
  s.precision = precision;
  s.device = device;
  return s;
}

// Spec of a solve whose labels come from a geometry (kGeomStride doubles per primitive: kind,
// label, a, b, d, zlo, zhi, cx, cy, r) and, with sigma > 0, whose δn is generated on the device.
constexpr int kGeomStride = 10;
bwpm::Spec geometry_spec(int64_t ny, int64_t nx, double dy, double dx, double dz, double k0, const double* indices,
                         int64_t m, int64_t nz, double z0, int background, const double* prims, int64_t nprims,
                         double ib_re, double ib_im, const double* mask, const int32_t* labels_before, double sigma,
                         double corr, uint64_t seed, int device, int precision) {
  bwpm::require(nprims >= 0 && (nprims == 0 || prims), "geometry primitives missing");
  bwpm::require(std::isfinite(z0), "z0 must be finite");
  bwpm::Spec s = make_spec(ny, nx, dy, dx, dz, k0, indices, m, nullptr, nz, ib_re, ib_im, mask, nullptr, labels_before,
                           device, precision, 0);
  s.uniform_label = -1;
  s.geom.on = true;
  s.geom.z0 = z0;
  s.geom.background = background;
  for (int64_t i = 0; i < nprims; ++i) {
    const double* v = prims + kGeomStride * i;
    // bounds may be infinite (an unbounded layer); nothing may be NaN
    for (int k = 0; k < kGeomStride; ++k) bwpm::require(!std::isnan(v[k]), "geometry values must not be NaN");
    for (int k : {0, 1, 2, 3, 4, 7, 8, 9})
      bwpm::require(std::isfinite(v[k]), "geometry kinds, labels, planes and cylinders must be finite");
    bwpm::GeomPrim p;
    p.kind = static_cast<int>(v[0]);
    p.label = static_cast<int>(v[1]);
    bwpm::require(static_cast<double>(p.kind) == v[0] && static_cast<double>(p.label) == v[1],
                  "geometry kind and label must be integers");
    p.a = v[2];
    p.b = v[3];
    p.d = v[4];
    p.zlo = v[5];
    p.zhi = v[6];
    p.cx = v[7];
    p.cy = v[8];
    p.r = v[9];
    s.geom.prims.push_back(p);
  }
  if (sigma > 0) {
    s.synth.on = true;
    s.synth.sigma = sigma;
    s.synth.corr = corr;
    s.synth.seed = seed;
  }
  return s;
}

}  // namespace

extern "C" {

const char* bwpm_last_error(void) { return g_error.c_str(); }

int bwpm_geometry_labels(int64_t ny, int64_t nx, double dy, double dx, double dz, int64_t nz, double z0,
                         int background, const double* prims, int64_t nprims, int64_t m, int32_t* out) {
  return guarded([&] {
    bwpm::require(out != nullptr && ny >= 1 && nx >= 1 && nz >= 1 && m >= 1, "null output or empty grid");
    std::vector<double> idx(2 * static_cast<size_t>(m), 1.0);  // labels are checked against m indices
    bwpm::Spec s = geometry_spec(ny, nx, dy, dx, dz, 1.0, idx.data(), m, nz, z0, background, prims, nprims, 1.0, 0.0,
                                 nullptr, nullptr, 0.0, 0.0, 0, 0, BWPM_FP64);
    bwpm::geometry_labels(s, out);
  });
}

int bwpm_solver_create_geometry(int64_t ny, int64_t nx, double dy, double dx, double dz, double k0,
                                const double* indices, int64_t m, int64_t nz, double z0, int background,
                                const double* prims, int64_t nprims, double ib_re, double ib_im, const double* mask,
                                const int32_t* labels_before, double sigma, double corr, uint64_t seed, int device,
                                int precision, bwpm_solver** out) {
  return guarded([&] {
    bwpm::require(out != nullptr, "null output handle");
    *out = nullptr;
    bwpm::Spec s = geometry_spec(ny, nx, dy, dx, dz, k0, indices, m, nz, z0, background, prims, nprims, ib_re, ib_im,
                                 mask, labels_before, sigma, corr, seed, device, precision);
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
      throw bwpm::Error(bwpm::kEdevice, "no GPU device available");
    bwpm::require(device >= 0 && device < count, "device index out of range");
    auto h = std::make_unique<bwpm_solver>();
    h->engine = bwpm::make_engine(std::move(s));
    h->device = device;
    *out = h.release();
  });
}

int bwpm_device_count(int* count) {
  return guarded([&] {
    bwpm::require(count != nullptr, "null pointer");
    int n = 0;
    const hipError_t e = hipGetDeviceCount(&n);
    *count = e == hipSuccess ? n : 0;
  });
}

int bwpm_transmission_matrix(const double* kx, const double* ky, int64_t n, double n_in_re, double n_in_im,
                             double n_out_re, double n_out_im, double k0, double* out) {
  return interface_call(0, kx, ky, n, n_in_re, n_in_im, n_out_re, n_out_im, k0, out);
}

int bwpm_reflection_matrix(const double* kx, const double* ky, int64_t n, double n_in_re, double n_in_im,
                           double n_out_re, double n_out_im, double k0, double* out) {
  return interface_call(1, kx, ky, n, n_in_re, n_in_im, n_out_re, n_out_im, k0, out);
}

int bwpm_longitudinal_wavenumber(double n_re, double n_im, double k0, const double* kp2, int64_t n, double* out) {
  return guarded([&] {
    bwpm::require(n >= 0, "n must be >= 0");
    bwpm::require(n == 0 || (kp2 && out), "null array");
    int dev = 0;
    (void)hipGetDevice(&dev);
    bwpm::longitudinal_wavenumbers(c128(n_re, n_im), k0, kp2, n, reinterpret_cast<c128*>(out), dev);
  });
}

int bwpm_fresnel_coefficients(const double* kx, const double* ky, int64_t n, double n_in_re, double n_in_im,
                              double n_out_re, double n_out_im, double k0, double* out) {
  return guarded([&] {
    bwpm::require(n >= 0, "n must be >= 0");
    bwpm::require(n == 0 || (kx && ky && out), "null array");
    int dev = 0;
    (void)hipGetDevice(&dev);
    bwpm::fresnel_coefficients(kx, ky, n, c128(n_in_re, n_in_im), c128(n_out_re, n_out_im), k0,
                               reinterpret_cast<c128*>(out), dev);
  });
}

int bwpm_solver_create(int64_t ny, int64_t nx, double dy, double dx, double dz, double k0, const double* indices,
                       int64_t m, const int32_t* labels, int64_t nz, double ib_re, double ib_im, const double* mask,
                       const double* fluct, int device, int precision, bwpm_solver** out) {
  return bwpm_solver_create_ex(ny, nx, dy, dx, dz, k0, indices, m, labels, nz, ib_re, ib_im, mask, fluct, nullptr,
                               device, precision, out);
}

int bwpm_solver_create_ex(int64_t ny, int64_t nx, double dy, double dx, double dz, double k0, const double* indices,
                          int64_t m, const int32_t* labels, int64_t nz, double ib_re, double ib_im, const double* mask,
                          const double* fluct, const int32_t* labels_before, int device, int precision,
                          bwpm_solver** out) {
  return guarded([&] {
    bwpm::require(out != nullptr, "null output handle");
    *out = nullptr;
    bwpm::Spec s = make_spec(ny, nx, dy, dx, dz, k0, indices, m, labels, nz, ib_re, ib_im, mask, fluct,
                             labels_before, device, precision);
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
      throw bwpm::Error(bwpm::kEdevice, "no GPU device available");
    bwpm::require(device >= 0 && device < count, "device index out of range");
    auto h = std::make_unique<bwpm_solver>();
    h->engine = bwpm::make_engine(std::move(s));
    h->device = device;
    *out = h.release();
  });
}

int bwpm_solver_create_synth(int64_t ny, int64_t nx, double dy, double dx, double dz, double k0,
                             const double* indices, int64_t m, const int32_t* labels, int uniform_label, int64_t nz,
                             double ib_re, double ib_im, const double* mask, const int32_t* labels_before,
                             double sigma, double corr, uint64_t seed, int device, int precision, bwpm_solver** out) {
  return guarded([&] {
    bwpm::require(out != nullptr, "null output handle");
    *out = nullptr;
    bwpm::Spec s = make_spec(ny, nx, dy, dx, dz, k0, indices, m, labels, nz, ib_re, ib_im, mask, nullptr,
                             labels_before, device, precision, labels ? -1 : uniform_label);
    s.synth.on = true;
    s.synth.sigma = sigma;
    s.synth.corr = corr;
    s.synth.seed = seed;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
      throw bwpm::Error(bwpm::kEdevice, "no GPU device available");
    bwpm::require(device >= 0 && device < count, "device index out of range");
    auto h = std::make_unique<bwpm_solver>();
    h->engine = bwpm::make_engine(std::move(s));
    h->device = device;
    *out = h.release();
  });
}

int bwpm_solver_fluctuation(const bwpm_solver* s, double* out) {
  return guarded([&] {
    bwpm::require(s && out, "null argument");
    const std::vector<c128> v = s->engine->fluctuation();
    std::memcpy(out, v.data(), v.size() * sizeof(c128));
  });
}

int bwpm_solver_destroy(bwpm_solver* s) {
  return guarded([&] { delete s; });
}

int bwpm_solver_run(bwpm_solver* s, const double* field, int iterations, const int64_t* keep, int64_t nkeep,
                    bwpm_result** out) {
  return guarded([&] {
    bwpm::require(s && field && out, "null argument");
    bwpm::require(iterations >= 0, "iterations must be >= 0");
    *out = nullptr;
    std::vector<long long> k;
    for (int64_t i = 0; i < nkeep; ++i) k.push_back(keep[i]);
    auto r = std::make_unique<bwpm_result>();
    s->engine->run(reinterpret_cast<const c128*>(field), iterations, k, r->out);
    r->npix = static_cast<long long>(r->out.transmitted.size() / 2);
    *out = r.release();
  });
}

int bwpm_result_field(const bwpm_result* r, int which, double* out) {
  return guarded([&] {
    bwpm::require(r && out, "null argument");
    copy_field(which == 0 ? r->out.transmitted : r->out.reflected, out);
  });
}

int bwpm_result_iterations(const bwpm_result* r, int64_t* n) {
  return guarded([&] {
    bwpm::require(r && n, "null argument");
    *n = static_cast<int64_t>(r->out.transmitted_iter.size());
  });
}

int bwpm_result_iter_field(const bwpm_result* r, int which, int64_t it, double* out) {
  return guarded([&] {
    bwpm::require(r && out, "null argument");
    const auto& v = which == 0 ? r->out.transmitted_iter : r->out.reflected_iter;
    bwpm::require(it >= 0 && it < static_cast<int64_t>(v.size()), "iteration out of range");
    copy_field(v[it], out);
  });
}

int bwpm_result_plane(const bwpm_result* r, int which, int64_t p, double* out, int* present) {
  return guarded([&] {
    bwpm::require(r && out && present, "null argument");
    const auto& m = which == 0 ? r->out.forward : r->out.backward;
    auto it = m.find(p);
    *present = it != m.end() ? 1 : 0;
    if (it != m.end()) copy_field(it->second, out);
  });
}

int bwpm_result_stored_values(const bwpm_result* r, int64_t* v) {
  return guarded([&] {
    bwpm::require(r && v, "null argument");
    *v = r->out.stored_values;
  });
}

int bwpm_result_stats(const bwpm_result* r, double* gpu_ms, int64_t* slab_steps, int64_t* gmres_iterations) {
  return guarded([&] {
    bwpm::require(r != nullptr, "null argument");
    if (gpu_ms) *gpu_ms = r->out.gpu_ms;
    if (slab_steps) *slab_steps = r->out.slab_steps;
    if (gmres_iterations) *gmres_iterations = r->out.gmres_iterations;
  });
}

int bwpm_result_free(bwpm_result* r) {
  return guarded([&] { delete r; });
}

int bwpm_solver_upload(bwpm_solver* s, const double* field) {
  return guarded([&] {
    bwpm::require(s && field, "null argument");
    s->engine->upload(reinterpret_cast<const c128*>(field));
  });
}

int bwpm_solver_run_resident(bwpm_solver* s, int iterations, int time_kernels, double* gpu_ms, double* col_ms,
                             double* row_ms, int64_t* col_launches, int64_t* row_launches) {
  return guarded([&] {
    bwpm::require(s != nullptr, "null argument");
    bwpm::require(iterations >= 0, "iterations must be >= 0");
    bwpm::KernelTimes kt;
    const double ms = s->engine->run_resident(iterations, time_kernels != 0, &kt);
    if (gpu_ms) *gpu_ms = ms;
    if (col_ms) *col_ms = kt.col_ms;
    if (row_ms) *row_ms = kt.row_ms;
    if (col_launches) *col_launches = kt.col_launches;
    if (row_launches) *row_launches = kt.row_launches;
  });
}

int bwpm_solver_device_bytes(const bwpm_solver* s, int64_t* bytes) {
  return guarded([&] {
    bwpm::require(s && bytes, "null argument");
    *bytes = static_cast<int64_t>(s->engine->device_bytes());
  });
}

int bwpm_solver_fetch(bwpm_solver* s, int iterations, bwpm_result** out) {
  return guarded([&] {
    bwpm::require(s && out, "null argument");
    bwpm::require(iterations >= 0, "iterations must be >= 0");
    *out = nullptr;
    auto r = std::make_unique<bwpm_result>();
    s->engine->download(r->out, iterations, {});
    r->npix = static_cast<long long>(r->out.transmitted.size() / 2);
    *out = r.release();
  });
}

int bwpm_batch_run(bwpm_solver* const* solvers, int64_t n, int iterations, int max_lane, int time_kernels,
                   double* wall_ms, double* col_ms, double* row_ms, int64_t* col_launches, int64_t* row_launches) {
  return guarded([&] {
    bwpm::require(n >= 0 && (n == 0 || solvers), "null solver array");
    bwpm::require(iterations >= 0, "iterations must be >= 0");
    std::vector<bwpm::EngineBase*> es;
    for (int64_t i = 0; i < n; ++i) {
      bwpm::require(solvers[i] != nullptr, "null solver in batch");
      es.push_back(solvers[i]->engine.get());
    }
    bwpm::KernelTimes kt;
    const double ms = bwpm::run_batch(es, iterations, max_lane, time_kernels != 0, &kt);
    if (wall_ms) *wall_ms = ms;
    if (col_ms) *col_ms = kt.col_ms;
    if (row_ms) *row_ms = kt.row_ms;
    if (col_launches) *col_launches = kt.col_launches;
    if (row_launches) *row_launches = kt.row_launches;
  });
}

int bwpm_copy_probe(int64_t bytes, int reps, int device, double* ms) {
  return guarded([&] {
    bwpm::require(ms != nullptr, "null argument");
    *ms = bwpm::copy_probe(bytes, reps, device);
  });
}

int bwpm_queue_create(int device, double hbm_budget_bytes, int slots, int chunk, bwpm_queue** out) {
  return guarded([&] {
    bwpm::require(out != nullptr, "null output handle");
    *out = nullptr;
    auto h = std::make_unique<bwpm_queue>();
    h->q = std::make_unique<bwpm::Queue>(device, hbm_budget_bytes, slots, chunk);
    *out = h.release();
  });
}

int bwpm_queue_destroy(bwpm_queue* q) {
  return guarded([&] { delete q; });
}

int bwpm_queue_submit(bwpm_queue* q, int64_t ny, int64_t nx, double dy, double dx, double dz, double k0,
                      const double* indices, int64_t m, const int32_t* labels, int64_t nz, double ib_re, double ib_im,
                      const double* mask, const double* fluct, const int32_t* labels_before, int precision,
                      const double* field, int iterations, int64_t* ticket) {
  return guarded([&] {
    bwpm::require(q && field && ticket, "null argument");
    bwpm::Spec s = make_spec(ny, nx, dy, dx, dz, k0, indices, m, labels, nz, ib_re, ib_im, mask, fluct,
                             labels_before, 0, precision);
    const c128* f = reinterpret_cast<const c128*>(field);
    *ticket = q->q->submit(std::move(s), std::vector<c128>(f, f + 2 * ny * nx), iterations);
  });
}

int bwpm_queue_submit_synth(bwpm_queue* q, int64_t ny, int64_t nx, double dy, double dx, double dz, double k0,
                            const double* indices, int64_t m, const int32_t* labels, int uniform_label, int64_t nz,
                            double ib_re, double ib_im, const double* mask, const int32_t* labels_before,
                            double sigma, double corr, uint64_t seed, int precision, const double* field,
                            int iterations, int64_t* ticket) {
  return guarded([&] {
    bwpm::require(q && field && ticket, "null argument");
    bwpm::Spec s = make_spec(ny, nx, dy, dx, dz, k0, indices, m, labels, nz, ib_re, ib_im, mask, nullptr,
                             labels_before, 0, precision, labels ? -1 : uniform_label);
    s.synth.on = true;
    s.synth.sigma = sigma;
    s.synth.corr = corr;
    s.synth.seed = seed;
    const c128* f = reinterpret_cast<const c128*>(field);
    *ticket = q->q->submit(std::move(s), std::vector<c128>(f, f + 2 * ny * nx), iterations);
  });
}

int bwpm_queue_submit_geometry(bwpm_queue* q, int64_t ny, int64_t nx, double dy, double dx, double dz, double k0,
                               const double* indices, int64_t m, int64_t nz, double z0, int background,
                               const double* prims, int64_t nprims, double ib_re, double ib_im, const double* mask,
                               const int32_t* labels_before, double sigma, double corr, uint64_t seed, int precision,
                               const double* field, int iterations, int64_t* ticket) {
  return guarded([&] {
    bwpm::require(q && field && ticket, "null argument");
    bwpm::Spec s = geometry_spec(ny, nx, dy, dx, dz, k0, indices, m, nz, z0, background, prims, nprims, ib_re, ib_im,
                                 mask, labels_before, sigma, corr, seed, 0, precision);
    const c128* f = reinterpret_cast<const c128*>(field);
    *ticket = q->q->submit(std::move(s), std::vector<c128>(f, f + 2 * ny * nx), iterations);
  });
}

int bwpm_queue_state(bwpm_queue* q, int64_t ticket, int* state) {
  return guarded([&] {
    bwpm::require(q && state, "null argument");
    *state = q->q->state(ticket);
  });
}

int bwpm_queue_wait(bwpm_queue* q, int64_t ticket, double timeout_s, int* state) {
  return guarded([&] {
    bwpm::require(q && state, "null argument");
    *state = q->q->wait(ticket, timeout_s);
  });
}

int bwpm_queue_result(bwpm_queue* q, int64_t ticket, bwpm_result** out) {
  return guarded([&] {
    bwpm::require(q && out, "null argument");
    *out = nullptr;
    auto r = std::make_unique<bwpm_result>();
    r->out = q->q->take(ticket);
    r->npix = static_cast<long long>(r->out.transmitted.size() / 2);
    *out = r.release();
  });
}

int bwpm_queue_stats(bwpm_queue* q, double* out) {
  return guarded([&] {
    bwpm::require(q && out, "null argument");
    const bwpm::QueueStats st = q->q->stats();
    const double v[BWPM_QUEUE_NSTATS] = {st.p50_ms,
                                         st.p99_ms,
                                         st.mean_ms,
                                         st.updates_per_s,
                                         static_cast<double>(st.completed),
                                         static_cast<double>(st.failed),
                                         static_cast<double>(st.pending),
                                         static_cast<double>(st.running),
                                         st.used_bytes,
                                         st.budget_bytes,
                                         st.peak_bytes};
    std::memcpy(out, v, sizeof(v));
  });
}

int bwpm_analysis_create(int device, int64_t ny, int64_t nx, double dy, double dx, double k0, bwpm_analysis** out) {
  return guarded([&] {
    bwpm::require(out != nullptr, "null argument");
    *out = nullptr;
    auto h = std::make_unique<bwpm_analysis>();
    h->a = std::make_unique<bwpm::Analysis>(device, ny, nx, dy, dx, k0);
    *out = h.release();
  });
}

int bwpm_analysis_destroy(bwpm_analysis* a) {
  return guarded([&] { delete a; });
}

int bwpm_plane_power(bwpm_analysis* a, const void* planes, int where, int64_t count, double index, double* power) {
  return guarded([&] {
    bwpm::require(a && power, "null argument");
    const std::vector<double> v = a->a->plane_power(planes, where, count, index);
    std::memcpy(power, v.data(), v.size() * sizeof(double));
  });
}

int bwpm_reference_reflectance(bwpm_analysis* a, const void* plane, int where, double n1, double n2,
                               double* reflectance) {
  return guarded([&] {
    bwpm::require(a && reflectance, "null argument");
    *reflectance = a->a->reference_reflectance(plane, where, n1, n2);
  });
}

int bwpm_axial_profile(bwpm_analysis* a, const void* plane, int where, double index, const double* offsets,
                       int64_t noff, double na, double* intensity) {
  return guarded([&] {
    bwpm::require(a != nullptr, "null argument");
    a->a->axial_profile(plane, where, index, offsets, noff, na, intensity);
  });
}

int bwpm_strehl_profile(bwpm_analysis* a, const void* spectrum, int where, double k, double na,
                        const double* offsets, int64_t noff, double* values) {
  return guarded([&] {
    bwpm::require(a != nullptr, "null argument");
    a->a->strehl_profile(spectrum, where, k, na, offsets, noff, values);
  });
}

int bwpm_strehl_modes(bwpm_analysis* a, const double* ky, const double* spectrum, int64_t m, double k,
                      const double* offsets, int64_t noff, double* values) {
  return guarded([&] {
    bwpm::require(a != nullptr, "null argument");
    a->a->strehl_modes(ky, spectrum, m, k, offsets, noff, values);
  });
}

int bwpm_lane_graph_stats(int64_t* built, int64_t* replays, int64_t* fallbacks) {
  return guarded([&] {
    bwpm::require(built && replays && fallbacks, "null argument");
    const bwpm::LaneGraphCounters& c = bwpm::lane_graph_counters();
    *built = c.built.load();
    *replays = c.replays.load();
    *fallbacks = c.fallbacks.load();
  });
}

int bwpm_fft_plans(char* buf, int64_t cap, int64_t* len) {
  return guarded([&] {
    bwpm::require(len != nullptr && cap >= 0 && (buf != nullptr || cap == 0), "null argument");
    const std::string r = bwpm::fft_plan_report();
    *len = static_cast<int64_t>(r.size());
    if (cap > 0) {
      const size_t n = std::min(static_cast<size_t>(cap - 1), r.size());
      std::memcpy(buf, r.data(), n);
      buf[n] = '\0';
    }
  });
}

int bwpm_solver_plane(const bwpm_solver* s, int which, int64_t plane, const void** ptr, int* precision) {
  return guarded([&] {
    bwpm::require(s && ptr && precision, "null argument");
    HIP_CHECK(hipSetDevice(s->engine->device()));
    HIP_CHECK(hipStreamSynchronize(s->engine->stream()));
    *ptr = s->engine->device_plane(which, plane);
    const int p = s->engine->precision();
    *precision = p == BWPM_FP64 || p == BWPM_MIXED || p == BWPM_BFP ? BWPM_FP64 : BWPM_FP32;
  });
}

}  // extern "C"

namespace bwpm {
double power_flux_device(const c128* field, long long ny, long long nx, double dy, double dx, double k0, c128 index,
                         int direction, int device);
double power_flux_noalloc(const c128* field, long long ny, long long nx, double dy, double dx, double k0, c128 index,
                          int direction, int device);
}

extern "C" int bwpm_power_flux(const double* field, int64_t ny, int64_t nx, double dy, double dx, double k0,
                               double n_re, double n_im, int direction, int device, double* out) {
  return guarded([&] {
    bwpm::require(field && out, "null argument");
    bwpm::require(direction == 1 || direction == -1, "direction must be +1 or -1");
    *out = bwpm::power_flux_device(reinterpret_cast<const c128*>(field), ny, nx, dy, dx, k0, c128(n_re, n_im),
                                   direction, device);
  });
}

extern "C" int bwpm_power_flux_noalloc(const double* field, int64_t ny, int64_t nx, double dy, double dx, double k0,
                                       double n_re, double n_im, int direction, int device, double* out) {
  return guarded([&] {
    bwpm::require(field && out, "null argument");
    bwpm::require(direction == 1 || direction == -1, "direction must be +1 or -1");
    *out = bwpm::power_flux_noalloc(reinterpret_cast<const c128*>(field), ny, nx, dy, dx, k0, c128(n_re, n_im),
                                    direction, device);
  });
}
