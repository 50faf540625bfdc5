// Host side of one solve: set-up of the per-boundary label-pair structure, operator tables and
// device buffers, and the device-resident bidirectional driver.
#pragma once

#include <complex>
#include <map>
#include <memory>
#include <vector>
#include <atomic>

#include "common.hpp"

namespace bwpm {

using c128 = std::complex<double>;

// Gaussian-correlated real index fluctuation generated on the device: seeded normal noise
// filtered by exp(−|q|²·corr²/4) (periodic in x, y and z) and scaled to standard deviation sigma.
struct FluctSynth {
  bool on = false;
  double sigma = 0, corr = 0;  // corr in mm
  unsigned long long seed = 0;
};

// Label volume from a geometry by the pixel-centre rule (src/bwpm/geometry.py:24-35): slab j,
// pixel (n, m) takes the label of the last primitive containing its centre
// (x, y, z) = ((m − nx//2)·dx, (n − ny//2)·dy, z0 + (j + ½)·dz), else the background.
struct GeomPrim {
  int kind = 0;  // 0: above the plane, z > d + a·x + b·y; 1: layer, zlo < z ≤ zhi;
                 // 2: cylinder, sqrt((x − cx)² + (y − cy)²) ≤ r and zlo < z ≤ zhi
  int label = 0;
  double a = 0, b = 0, d = 0, zlo = 0, zhi = 0, cx = 0, cy = 0, r = 0;
};
struct LabelGeometry {
  bool on = false;
  int background = 0;
  double z0 = 0;
  std::vector<GeomPrim> prims;
};
// Writes the (nz, ny, nx) labels of s.geom into out (several host threads).
void geometry_labels(const struct Spec& s, int32_t* out);
// The label planes of s.geom without the (nz, ny, nx) volume (the same pixel-centre rule as
// geometry_labels): slab j's labels are planes[plane_of[j]]. A slab where every primitive is
// uniform over the grid (layers, planes that do not cross the slab, cylinders outside their z
// range) shares one uniform plane per label; the other slabs get planes of their own, and
// neighbouring equal planes are stored once.
void geometry_planes(const struct Spec& s, std::vector<std::vector<int32_t>>& planes, std::vector<int>& plane_of);

struct Spec {
  long long ny = 0, nx = 0, nz = 0;
  double dy = 0, dx = 0, dz = 0, k0 = 0;
  std::vector<c128> indices;
  std::vector<int32_t> labels;  // (nz, ny, nx), or empty with uniform_label >= 0
  int uniform_label = -1;       // >= 0: every pixel of every slab has this label
  c128 index_before{1.0, 0.0};
  std::vector<double> mask;     // (ny, nx) or empty
  std::vector<c128> fluct;      // (nz, ny, nx) or empty
  std::vector<int32_t> labels_before;  // labels of plane 0 (ny, nx); empty: index_before everywhere
  FluctSynth synth;             // δn generated on the device (fluct must be empty)
  LabelGeometry geom;           // labels generated from a geometry (labels must be empty)
  int precision = 32;           // 64: FP64; 48: complex64 streaming planes, FP64 arithmetic;
                                // 40: block-floating-point streaming planes, FP64 arithmetic;
                                // 32: FP32 compensated; 31: FP32 plain
  int device = 0;
};

struct RunOutput {
  std::vector<c128> transmitted, reflected;           // (ny, nx, 2)
  std::vector<std::vector<c128>> transmitted_iter, reflected_iter;
  std::map<long long, std::vector<c128>> forward, backward;
  long long stored_values = 0;
  // diagnostics of the last run
  double gpu_ms = 0;            // device time of the whole run (HIP events)
  long long slab_steps = 0;     // slab steps executed
  long long gmres_iterations = 0;
};

// Launch statistics of the dominant kernels, measured with HIP events on the solve's stream.
struct KernelTimes {
  double col_ms = 0, row_ms = 0;
  long long col_launches = 0, row_launches = 0;
};

class EngineBase {
 public:
  virtual ~EngineBase() = default;
  // Full solve from a host field (ny, nx, 2) complex128.
  virtual void run(const c128* field, int iterations, const std::vector<long long>& keep,
                   RunOutput& out) = 0;
  // Device-resident variant for timing: the field is already on the device (set by upload);
  // results stay on the device. Returns the device time in ms.
  virtual void upload(const c128* field) = 0;
  virtual double run_resident(int iterations, bool time_kernels, KernelTimes* kt) = 0;
  virtual void download(RunOutput& out, int iterations, const std::vector<long long>& keep) = 0;
  virtual hipStream_t stream() const = 0;
  virtual size_t device_bytes() const = 0;
  virtual int precision() const = 0;
  virtual long long ny() const = 0;
  virtual long long nx() const = 0;
  virtual int device() const = 0;
  virtual bool has_born() const = 0;
  virtual long long nz() const = 0;
  // δn of every slab (nz, ny, nx) as held on the device (zeros without δn).
  virtual std::vector<c128> fluctuation() const = 0;
  // Device plane (ny, nx) of px2<T> from the last run: 0 transmitted, 1 reflected (final
  // iteration), 2 forward / 3 backward field at kept plane p. Valid until the next run.
  virtual const void* device_plane(int which, long long p) const = 0;
  // Run the step programs of `members` (all of this engine's type, grid and device; this engine
  // among them or not) in lockstep on `stream`: one column and one row launch per step serve every
  // member. Fields must have been uploaded. Returns the device time in ms.
  virtual double run_lane(std::vector<EngineBase*>& members, int iterations,
                          const std::vector<std::vector<long long>>& keeps, hipStream_t stream,
                          bool time_kernels, KernelTimes* kt) = 0;
};

// The engine takes the spec over (the label volume is moved, not copied).
std::unique_ptr<EngineBase> make_engine(Spec spec);
std::unique_ptr<EngineBase> make_engine_f32(Spec spec, bool compensated);
std::unique_ptr<EngineBase> make_engine_f64(Spec spec);
std::unique_ptr<EngineBase> make_engine_mixed(Spec spec);
std::unique_ptr<EngineBase> make_engine_bfp(Spec spec);

// Batched solve of many engines: members are grouped into lanes of equal (precision, grid,
// device, no δn); each lane runs in lockstep on its own stream, lanes run concurrently.
// Results are left on the device (download per engine). Returns the wall time in ms.
double run_batch(std::vector<EngineBase*>& engines, int iterations, int max_lane, bool time_kernels = false,
                 KernelTimes* kt = nullptr);

// ---- continuous lanes (serving) ---------------------------------------------------------------

// A solve admitted into a lane: its engine (field uploaded), its iterations and the caller's tag.
struct LaneMember {
  EngineBase* engine = nullptr;
  int iterations = 0;
  long long tag = 0;
};

// What a continuous lane asks of its owner.
class LaneFeed {
 public:
  virtual ~LaneFeed() = default;
  // A prepared solve for a free slot, or a member with engine == nullptr; must not block.
  virtual LaneMember next() = 0;
  // The member's last slab step is on the lane stream before `done` (the callee owns the event).
  virtual void finished(const LaneMember& m, hipEvent_t done) = 0;
  // Called while the lane is empty: blocks until work may be there; false ends the lane.
  virtual bool wait() = 0;
};

// Runs a lane of up to `slots` solves of one grid and precision on `stream`: members join free
// slots between chunks of `chunk` slab steps and leave when their program ends. Every member's
// result is bit-identical to its solve alone (each slot's work depends on its own descriptors
// only). Solves with index fluctuations are not accepted (they run one per lane). Chunks whose
// steps are all plain launches replay one captured graph per shape for the lane's lifetime, so a
// long-lived lane of small grids is not bound by launch overhead.
void run_continuous_lane(int precision, LaneFeed& feed, int slots, int chunk, hipStream_t stream);

// Process-wide counts of the continuous lanes' graphs: built, replayed, and lanes that fell back
// to direct launches because a graph could not be built.
struct LaneGraphCounters {
  std::atomic<long long> built{0}, replays{0}, fallbacks{0};
};
LaneGraphCounters& lane_graph_counters();
// The radix plans the plan tuner chose in this process, one line per (precision, direction, line
// length): "<precision> <col|row> <n> <plan> <µs of plan 0> <µs of plan 1>".
std::string fft_plan_report();
void run_continuous_lane_f32(LaneFeed& feed, int slots, int chunk, hipStream_t stream, bool compensated);
void run_continuous_lane_f64(LaneFeed& feed, int slots, int chunk, hipStream_t stream);
void run_continuous_lane_mixed(LaneFeed& feed, int slots, int chunk, hipStream_t stream);
void run_continuous_lane_bfp(LaneFeed& feed, int slots, int chunk, hipStream_t stream);

// Upper bound of the device bytes of one solve of `spec` with `iterations` (admission by HBM
// footprint before the engine exists).
size_t estimate_device_bytes(const Spec& spec, int iterations);

// FP64 evaluation of the exported interface functions on the device.
void interface_matrices(int which, const double* kx, const double* ky, long long n, c128 n_in,
                        c128 n_out, double k0, c128* out, int device);
// Fresnel coefficients r_s, r_p, t_s, t_p, k_z,in, k_z,out per point (6 complex each).
void fresnel_coefficients(const double* kx, const double* ky, long long n, c128 n_in, c128 n_out, double k0,
                          c128* out, int device);
void longitudinal_wavenumbers(c128 index, double k0, const double* kp2, long long n, c128* out,
                              int device);

// Average duration (ms) of a plain device-to-device copy kernel over `bytes` (HBM calibration of
// the traffic counters and of the achievable bandwidth).
double copy_probe(long long bytes, int reps, int device);

}  // namespace bwpm
