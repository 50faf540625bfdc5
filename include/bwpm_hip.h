/* C ABI of the MI355X BWPM engine (libbwpm_hip.so).
 *
 * Drop-in boundary for the reference's compiled module `bwpm._core`
 * (cpp/python/bindings.cpp:122-180): the interface matrices, the longitudinal wavenumber and the
 * solver object with its `run`. Plain C types only; complex numbers are pairs of doubles
 * (re, im) in C order, fields are (ny, nx, 2) complex128, i.e. 4·ny·nx doubles.
 *
 * Conventions (as the reference binding):
 * - every input is copied before the call returns; no caller memory is retained;
 * - status codes: BWPM_EINVAL for invalid arguments (the reference's std::invalid_argument →
 *   ValueError), BWPM_ENOCONV, BWPM_EDEVICE, BWPM_ENOMEM, BWPM_EINTERNAL (std::runtime_error →
 *   RuntimeError); bwpm_last_error() returns the message of the last failure of the calling thread;
 * - one solver handle must not run concurrently from two threads (the reference Solver is not
 *   re-entrant either, cpp/src/solver.cpp:73-106); distinct handles may run in parallel.
 */
#ifndef BWPM_HIP_H
#define BWPM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BWPM_OK 0
#define BWPM_EINVAL 1
#define BWPM_ENOCONV 2
#define BWPM_EDEVICE 3
#define BWPM_ENOMEM 4
#define BWPM_EINTERNAL 5

/* precision of the bulk fields */
#define BWPM_FP64 64       /* complex128 throughout (parity mode against the FP64 reference) */
#define BWPM_MIXED 48      /* complex64 storage of the two streaming planes of every slab step,
                              FP64 arithmetic and every other buffer FP64 */
#define BWPM_BFP 40        /* the two streaming planes in a block-floating-point complex format
                              of complex64 size (two 28-bit mantissas sharing an 8-bit exponent),
                              FP64 arithmetic and every other buffer FP64 */
#define BWPM_FP32 32       /* complex64 bulk, FP64 tables, compensated twiddles/operators */
#define BWPM_FP32_PLAIN 31 /* complex64 bulk without compensation (diagnostic) */

typedef struct bwpm_solver bwpm_solver;
typedef struct bwpm_result bwpm_result;

/* Message of the last failed call on this thread ("" if none). */
const char* bwpm_last_error(void);

/* Number of visible GPUs. Replaces `_core.cuda_device_count` (cpp/python/bindings.cpp:157,176). */
int bwpm_device_count(int* count);

/* T of wende2023 eqs. (2), (3) at n points, FP64, row-major 2×2 per point (8n doubles out).
 * Replaces `_core.transmission_matrix` (cpp/python/bindings.cpp:124-130, cpp/src/interface.cpp:30-40). */
int bwpm_transmission_matrix(const double* kx, const double* ky, int64_t n, double n_in_re,
                             double n_in_im, double n_out_re, double n_out_im, double k0,
                             double* out);

/* R of wende2023 eqs. (8), (9); layout as above.
 * Replaces `_core.reflection_matrix` (cpp/python/bindings.cpp:131-137, cpp/src/interface.cpp:42-52). */
int bwpm_reflection_matrix(const double* kx, const double* ky, int64_t n, double n_in_re,
                           double n_in_im, double n_out_re, double n_out_im, double k0,
                           double* out);

/* k_z at n values of k⊥² (2n doubles out), with the grazing rule.
 * Replaces `_core.longitudinal_wavenumber` (cpp/python/bindings.cpp:178-179, cpp/src/interface.cpp:8-18). */
int bwpm_longitudinal_wavenumber(double n_re, double n_im, double k0, const double* kp2, int64_t n,
                                 double* out);

/* Fresnel coefficients (Convention A) at n points: r_s, r_p, t_s, t_p, k_z,in, k_z,out (12n doubles).
 * Replaces `bwpm.interface.fresnel_coefficients` (src/bwpm/interface.py:68-90, cpp/src/interface.cpp:20-28). */
int bwpm_fresnel_coefficients(const double* kx, const double* ky, int64_t n, double n_in_re,
                              double n_in_im, double n_out_re, double n_out_im, double k0,
                              double* out);

/* Poynting flux Σ_k Re(k_z)(|Ê_x|² + |Ê_y|² + |Ê_z|²)·dx·dy/N of a field (ny, nx, 2) in a
 * homogeneous medium, E_z from transversality for waves of `direction` (±1), FP64 on the device
 * with a deterministic reduction. The 2D reference's power_flux is the case nx = 1, dx = 1.
 * Replaces `bwpm.fpwpm.power_flux` and `bwpm.adapter.port_power` (src/bwpm/fpwpm.py:205-216,
 * src/bwpm/adapter.py:267-284), the basis of the energy residual (src/bwpm/fast.py:148-155). */
int bwpm_power_flux(const double* field, int64_t ny, int64_t nx, double dy, double dx, double k0,
                    double n_re, double n_im, int direction, int device, double* out);

/* bwpm_power_flux without device allocation, for callers that cannot allocate device memory (code
 * under graph capture, tight serving loops, harnesses that forbid it); it has no counterpart in the
 * reference. Same arguments and the same bits as bwpm_power_flux. Its device buffers (the plane and
 * its spectrum, k and twiddle tables, reduction partials) are one static array of the library per
 * device, about 8.3 MB, held until the process ends; the grid is limited to ny, nx <= 4096 and
 * ny * nx <= 262144 pixels (larger grids: BWPM_EINVAL). Calls in one process take turns on a lock;
 * the array keeps the last call's contents, and every call writes all it reads. */
int bwpm_power_flux_noalloc(const double* field, int64_t ny, int64_t nx, double dy, double dx, double k0,
                            double n_re, double n_im, int direction, int device, double* out);

/* Solver over a labelled volume. labels: (nz, ny, nx) int32 into `indices` (m complex);
 * mask: (ny, nx) or NULL; fluct: δn (nz, ny, nx) complex or NULL; precision: one of the BWPM_FP64 …
 * BWPM_FP32_PLAIN codes above.
 * Replaces `_core.CudaSolver(...)` / `_core.Solver(...)` (cpp/python/bindings.cpp:139-151,158-171,
 * cpp/src/solver.cpp:27-63; cpp/src/solver_cuda.cu:643-719). */
int bwpm_solver_create(int64_t ny, int64_t nx, double dy, double dx, double dz, double k0,
                       const double* indices, int64_t m, const int32_t* labels, int64_t nz,
                       double index_before_re, double index_before_im, const double* mask,
                       const double* fluct, int device, int precision, bwpm_solver** out);

/* As bwpm_solver_create with the labels of plane 0 given per pixel (labels_before: (ny, nx) int32,
 * -1 = index_before; NULL = -1 everywhere). Serves the single step of `Propagator2D.step`
 * (src/bwpm/fpwpm.py:134-161), whose incoming plane may carry any labels. */
int bwpm_solver_create_ex(int64_t ny, int64_t nx, double dy, double dx, double dz, double k0,
                          const double* indices, int64_t m, const int32_t* labels, int64_t nz,
                          double index_before_re, double index_before_im, const double* mask,
                          const double* fluct, const int32_t* labels_before, int device,
                          int precision, bwpm_solver** out);

/* A solver whose index fluctuation is generated on the device instead of copied from the host
 * (no counterpart in the reference, which takes a (nz, ny, nx) complex array, bindings.cpp:115):
 * real δn, Gaussian-correlated as the reference's test volumes (tests/test_fluctuation_3d.py:34-41):
 * seeded normal noise filtered by exp(−|q|²·corr²/4) (corr in mm), periodic in x, y and z,
 * scaled to standard deviation sigma. labels may be NULL for a volume of one label
 * (uniform_label) in every pixel; every slab must then lie in one host index. Other arguments as
 * bwpm_solver_create_ex. */
int bwpm_solver_create_synth(int64_t ny, int64_t nx, double dy, double dx, double dz, double k0,
                             const double* indices, int64_t m, const int32_t* labels, int uniform_label, int64_t nz,
                             double ib_re, double ib_im, const double* mask, const int32_t* labels_before,
                             double sigma, double corr, uint64_t seed, int device, int precision, bwpm_solver** out);
/* Label volumes from a geometry, by the pixel-centre rule of src/bwpm/geometry.py:24-35 (no
 * counterpart in the reference's C++ core, which takes a (nz, ny, nx) int32 array, bindings.cpp:102).
 * Slab j, pixel (n, m) has its centre at x = (m − nx//2)·dx, y = (n − ny//2)·dy,
 * z = z0 + (j + ½)·dz and takes the label of the last primitive containing it, else `background`.
 * prims: nprims × 10 doubles (kind, label, a, b, d, zlo, zhi, cx, cy, r):
 *   kind 0, above a plane: z > d + a·x + b·y (a tilted or planar interface);
 *   kind 1, a layer: zlo < z <= zhi;
 *   kind 2, a cylinder along z: sqrt((x − cx)² + (y − cy)²) <= r and zlo < z <= zhi.
 * bwpm_geometry_labels writes the (nz, ny, nx) labels (m: the number of indices they select). */
int bwpm_geometry_labels(int64_t ny, int64_t nx, double dy, double dx, double dz, int64_t nz, double z0,
                         int background, const double* prims, int64_t nprims, int64_t m, int32_t* out);
/* A solver whose labels are made from a geometry inside the library (no label array crosses the
 * ABI); with sigma > 0 its δn is generated on the device as in bwpm_solver_create_synth. */
int bwpm_solver_create_geometry(int64_t ny, int64_t nx, double dy, double dx, double dz, double k0,
                                const double* indices, int64_t m, int64_t nz, double z0, int background,
                                const double* prims, int64_t nprims, double ib_re, double ib_im, const double* mask,
                                const int32_t* labels_before, double sigma, double corr, uint64_t seed, int device,
                                int precision, bwpm_solver** out);
/* The δn the solver holds, (nz, ny, nx) complex128 (zeros without δn). */
int bwpm_solver_fluctuation(const bwpm_solver* s, double* out);
int bwpm_solver_destroy(bwpm_solver* s);

/* Bidirectional solve (iterations = 0: forward only) of field (ny, nx, 2), keeping the planes in
 * `keep`. Replaces `.run(field, iterations, keep)` (cpp/python/bindings.cpp:60-86,152-153,172-173;
 * cpp/src/solver.cpp:444-528). */
int bwpm_solver_run(bwpm_solver* s, const double* field, int iterations, const int64_t* keep,
                    int64_t nkeep, bwpm_result** out);

/* Result accessors (the dict of cpp/python/bindings.cpp:71-85). which: 0 transmitted, 1 reflected. */
int bwpm_result_field(const bwpm_result* r, int which, double* out);
int bwpm_result_iterations(const bwpm_result* r, int64_t* n);
int bwpm_result_iter_field(const bwpm_result* r, int which, int64_t it, double* out);
/* Kept plane p; which: 0 forward, 1 backward. *present = 0 if the plane was not kept. */
int bwpm_result_plane(const bwpm_result* r, int which, int64_t p, double* out, int* present);
int bwpm_result_stored_values(const bwpm_result* r, int64_t* v);
/* Device time of the run (HIP events on the solve's stream), slab steps and GMRES iterations. */
int bwpm_result_stats(const bwpm_result* r, double* gpu_ms, int64_t* slab_steps,
                      int64_t* gmres_iterations);
int bwpm_result_free(bwpm_result* r);

/* Device-resident timing path (no counterpart in the reference; for benchmarks and serving):
 * upload a field once, run with the field and all state resident in HBM. With time_kernels != 0
 * every slab-step kernel is bracketed by HIP events (slower; for per-kernel durations). */
int bwpm_solver_upload(bwpm_solver* s, const double* field);
int bwpm_solver_run_resident(bwpm_solver* s, int iterations, int time_kernels, double* gpu_ms,
                             double* col_ms, double* row_ms, int64_t* col_launches,
                             int64_t* row_launches);
int bwpm_solver_device_bytes(const bwpm_solver* s, int64_t* bytes);

/* Results of the last resident or batched run of `s` (no kept planes). */
int bwpm_solver_fetch(bwpm_solver* s, int iterations, bwpm_result** out);

/* Batched solve (no counterpart in the reference; the persistent multi-solve path): every solver
 * must have its field uploaded. Solvers of equal precision, grid and device are grouped into
 * lanes of at most `max_lane` that step in lockstep, one column and one row launch per slab step
 * for the whole lane; lanes run concurrently on their own streams. Results stay on the device
 * until bwpm_solver_fetch. Each result is bit-identical to the solver's run alone. With
 * time_kernels != 0 every slab-step launch is bracketed by HIP events on its lane's stream and the
 * summed durations and launch counts are returned (any output pointer may be NULL). */
int bwpm_batch_run(bwpm_solver* const* solvers, int64_t n, int iterations, int max_lane,
                   int time_kernels, double* wall_ms, double* col_ms, double* row_ms,
                   int64_t* col_launches, int64_t* row_launches);

/* Average duration of a plain copy kernel (k_copy_probe) over `bytes` on `device`: the HBM
 * calibration of the traffic counters and the achievable-bandwidth reference of the benchmark
 * (no counterpart in the reference). */
int bwpm_copy_probe(int64_t bytes, int reps, int device, double* ms);

/* ---- serving: request queue with continuous batching (no counterpart in the reference, which
 * runs one solve per call, src/bwpm/fast.py:165-200) --------------------------------------------
 *
 * Requests are admitted in arrival order while their device footprint fits the HBM budget;
 * admitted solves of one (precision, grid) share a continuous lane in which a request joins a free
 * slot between chunks of slab steps and leaves when its program ends, so requests that arrive
 * while others are on the GPU overlap with them. Each result is bit-identical to the same solve
 * run alone. Solves with index fluctuations run one at a time. All calls are thread-safe. */
typedef struct bwpm_queue bwpm_queue;

#define BWPM_TICKET_PENDING 0  /* waiting for admission */
#define BWPM_TICKET_ADMITTED 1 /* on the device */
#define BWPM_TICKET_DONE 2     /* result ready (bwpm_queue_result) */
#define BWPM_TICKET_FAILED 3   /* bwpm_queue_result returns its error */

/* hbm_budget_bytes <= 0: 90 % of the device memory free now; slots: solves per lane (1 … 1024);
 * chunk: slab steps between admissions into a running lane. */
int bwpm_queue_create(int device, double hbm_budget_bytes, int slots, int chunk, bwpm_queue** out);
/* Finishes the admitted requests, fails the pending ones, frees the queue. */
int bwpm_queue_destroy(bwpm_queue* q);
/* Arguments as bwpm_solver_create_ex (device: the queue's) plus the input field and iterations. */
int bwpm_queue_submit(bwpm_queue* q, int64_t ny, int64_t nx, double dy, double dx, double dz, double k0,
                      const double* indices, int64_t m, const int32_t* labels, int64_t nz, double ib_re,
                      double ib_im, const double* mask, const double* fluct, const int32_t* labels_before,
                      int precision, const double* field, int iterations, int64_t* ticket);
/* As bwpm_queue_submit with δn generated on the device (arguments as bwpm_solver_create_synth). */
int bwpm_queue_submit_synth(bwpm_queue* q, int64_t ny, int64_t nx, double dy, double dx, double dz, double k0,
                            const double* indices, int64_t m, const int32_t* labels, int uniform_label, int64_t nz,
                            double ib_re, double ib_im, const double* mask, const int32_t* labels_before,
                            double sigma, double corr, uint64_t seed, int precision, const double* field,
                            int iterations, int64_t* ticket);
/* As bwpm_queue_submit with the labels made from a geometry (arguments as bwpm_solver_create_geometry). */
int bwpm_queue_submit_geometry(bwpm_queue* q, int64_t ny, int64_t nx, double dy, double dx, double dz, double k0,
                               const double* indices, int64_t m, int64_t nz, double z0, int background,
                               const double* prims, int64_t nprims, double ib_re, double ib_im, const double* mask,
                               const int32_t* labels_before, double sigma, double corr, uint64_t seed, int precision,
                               const double* field, int iterations, int64_t* ticket);
int bwpm_queue_state(bwpm_queue* q, int64_t ticket, int* state);
/* Waits until the ticket is done or failed, or timeout_s passes (timeout_s < 0: no limit). */
int bwpm_queue_wait(bwpm_queue* q, int64_t ticket, double timeout_s, int* state);
/* Result of a done ticket (as bwpm_solver_run; kept planes are not served); the ticket is then
 * forgotten. A failed ticket returns its error code and message. */
int bwpm_queue_result(bwpm_queue* q, int64_t ticket, bwpm_result** out);
/* out[BWPM_QUEUE_NSTATS]: p50, p99 and mean latency (ms, submit → result on the host), completed
 * grid-point·slab updates per second (since the first submit), completed, failed, pending and
 * admitted requests, device bytes in use, the budget and the peak in use. */
#define BWPM_QUEUE_NSTATS 11
int bwpm_queue_stats(bwpm_queue* q, double* out);
/* Process-wide counts of the graphs of continuous lanes: graphs built (node by node, no stream
 * capture), graph replays, and lanes that ran their chunks as direct launches because a graph
 * could not be built (no counterpart in the reference). */
int bwpm_lane_graph_stats(int64_t* built, int64_t* replays, int64_t* fallbacks);
/* The radix plans the plan tuner chose in this process for the register FFT kernels, as text: one
 * line "<precision code> <col|row> <line length> <plan 0|1> <µs of plan 0> <µs of plan 1>" per
 * tuned key. Writes at most cap − 1 characters and a terminating NUL into buf (cap may be 0 to
 * query the length) and the full length into *len. Plans are chosen once per device, precision,
 * line length and direction before the first solve of that key; BWPM_PLAN=0 or 1 fixes them (no
 * counterpart in the reference, whose FFTs are pocketfft/cuFFT plans). */
int bwpm_fft_plans(char* buf, int64_t cap, int64_t* len);

/* ---- analysis reductions over planes (the step after the path) ---------------------------------
 *
 * Replace the NumPy analysis of kept planes in `bwpm.pilot_imc3d` (src/bwpm/pilot_imc3d.py:19-131),
 * the Strehl ratio of experiments/b3_scaling.py:88-97 and `bwpm.pilot_imc.strehl_from_spectrum`
 * (src/bwpm/pilot_imc.py:184-197). The transforms, per-mode weights and sums run in FP64 on the
 * device with deterministic reductions; a plane can stay on the device (a kept plane or output of
 * a solver, bwpm_solver_plane). A handle is bound to one grid and device and is not safe for
 * concurrent calls. */
typedef struct bwpm_analysis bwpm_analysis;

#define BWPM_PLANE_HOST_C128 0   /* host memory, complex128 (ny, nx, 2) */
#define BWPM_PLANE_DEVICE_C128 1 /* device memory, complex128 */
#define BWPM_PLANE_DEVICE_C64 2  /* device memory, complex64 (planes of an FP32 solver) */

/* Grid (ny, nx) with steps dy, dx (mm) at vacuum wavenumber k0 = 2π/λ (1/mm), as `Grid3D`
 * (src/bwpm/volume3d.py:38-67). */
int bwpm_analysis_create(int device, int64_t ny, int64_t nx, double dy, double dx, double k0, bwpm_analysis** out);
int bwpm_analysis_destroy(bwpm_analysis* a);
/* power[count]: flux n·Σ cos θ (|E_s|² + |E_p|²) of each plane in a medium of index n
 * (`plane_power`, src/bwpm/pilot_imc3d.py:55-63). */
int bwpm_plane_power(bwpm_analysis* a, const void* planes, int where, int64_t count, double index, double* power);
/* Flux-weighted Fresnel reflectance of n1 → n2 for the field (`reference_reflectance`,
 * src/bwpm/pilot_imc3d.py:82-97); BWPM_EINVAL if the field carries no propagating flux. */
int bwpm_reference_reflectance(bwpm_analysis* a, const void* plane, int where, double n1, double n2,
                               double* reflectance);
/* intensity[noff]: on-axis |E_x|² + |E_y|² + |E_z|² after each distance offsets[i] (mm), from the
 * propagating modes, with na >= 0 only those with |k⊥| <= na·k0 (na < 0: no aperture). The host
 * picks the first maximum, as `axial_focus` does (src/bwpm/pilot_imc3d.py:100-131). */
int bwpm_axial_profile(bwpm_analysis* a, const void* plane, int where, double index, const double* offsets,
                       int64_t noff, double na, double* intensity);
/* values[noff]: |Σ A e^{i k_z d}|² / (Σ|A|)² over the modes of the spectrum (ny, nx) complex128 with
 * |k⊥| <= na·k0, k_z = sqrt(k² − k⊥²) (experiments/b3_scaling.py:88-97). where: host or device
 * complex128. */
int bwpm_strehl_profile(bwpm_analysis* a, const void* spectrum, int where, double k, double na,
                        const double* offsets, int64_t noff, double* values);
/* values[noff] for m modes with transverse wavenumbers ky[m] and host amplitudes spectrum[m]
 * (complex128), k_z = sqrt(max(k² − k_y², 0)) (`strehl_from_spectrum`, src/bwpm/pilot_imc.py:184-197);
 * the handle supplies the device only. */
int bwpm_strehl_modes(bwpm_analysis* a, const double* ky, const double* spectrum, int64_t m, double k,
                      const double* offsets, int64_t noff, double* values);

/* Device address of a plane of the solver's last run, valid until its next run or destroy:
 * which 0 = transmitted, 1 = reflected (final iteration), 2 = forward / 3 = backward field at a
 * kept plane `plane`. precision receives BWPM_FP64 (complex128 planes) or BWPM_FP32 (complex64).
 * Feeds the analysis calls without a copy to the host (no counterpart in the reference). */
int bwpm_solver_plane(const bwpm_solver* s, int which, int64_t plane, const void** ptr, int* precision);

#ifdef __cplusplus
}
#endif

#endif /* BWPM_HIP_H */
