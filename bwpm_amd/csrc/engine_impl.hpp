// Device-resident bidirectional BWPM engine (reference semantics: Solver::run,
// cpp/src/solver.cpp:444-528): per-solve set-up of label pairs, sparse interface lists and
// operator tables; a static step program per run; a lane runner that executes the programs of
// many solves of one grid in lockstep, one column and one row launch per step for all of them.
#pragma once

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <unordered_map>

#include "born.hpp"
#include "device_mem.hpp"
#include "engine.hpp"
#include "kernels.hpp"
#include "kernels_fast.hpp"
#include "kernels_plain.hpp"

namespace bwpm {

namespace detail {

// ---- per-boundary structure ------------------------------------------------------------------

// Boundary q separates plane q (slab q − 1, or the medium before the volume for q = 0) from
// plane q + 1 (slab q). Forward step q crosses it left to right, backward step q + 1 right to left.
struct Boundary {
  std::vector<std::pair<int32_t, int32_t>> pairs;  // (label left, label right), sorted
  std::vector<int> mleft, mright;                   // medium ids of each pair
  std::vector<int> rpair;                           // pair index of each R product
  long long tsel_off = -1;                          // into the tsel byte array (pairs > 1)
  long long rowuse_off = -1;                        // ny per-row product-use masks (pairs > 1)
  bool mixed_rows = false;                          // over a quarter of the rows select several T pairs
  long long sp_off = 0, sp_count = 0;               // sparse interface entries
  long long rowptr_off = -1;                        // ny + 1 offsets (sp_count > 0)
  int fwd_prod = 0, bwd_prod = 0;                   // offsets into the product descriptor array
  int source = -1;                                  // boundary whose structure this one reuses
};

// Operator table of one (kind, medium pair): hi and lo parts.
struct OpTable {
  DevBuf hi, lo;
};

// Operator tables are shared by every engine of the process with the same device, precision,
// grid, wavelength, slab thickness, kind and media: solves of one structure (different incident
// fields, polarisations or δn realisations) then read one table, which stays in the caches. The
// reference keeps the same cache per solver object (Solver::op, cpp/src/solver.cpp:73-106).
struct TableKey {
  int device = 0, precision = 0, kind = 0;  // kind −1: phase table of a homogeneous pair
  long long ny = 0, nx = 0;
  double dy = 0, dx = 0, dz = 0, k0 = 0, nr_re = 0, nr_im = 0, ns_re = 0, ns_im = 0;
  bool operator<(const TableKey& o) const {
    return std::tie(device, precision, kind, ny, nx, dy, dx, dz, k0, nr_re, nr_im, ns_re, ns_im) <
           std::tie(o.device, o.precision, o.kind, o.ny, o.nx, o.dy, o.dx, o.dz, o.k0, o.nr_re, o.nr_im, o.ns_re,
                    o.ns_im);
  }
};

inline std::mutex& table_mutex() {
  static std::mutex m;
  return m;
}
inline std::map<TableKey, std::weak_ptr<OpTable>>& table_cache() {
  static std::map<TableKey, std::weak_ptr<OpTable>> c;
  return c;
}

// The live table of `key`, or a new one from make() (which returns it complete on the device).
template <typename F>
std::shared_ptr<OpTable> shared_table(const TableKey& key, F&& make) {
  std::lock_guard<std::mutex> lock(table_mutex());
  auto& cache = table_cache();
  for (auto it = cache.begin(); it != cache.end();) it = it->second.expired() ? cache.erase(it) : std::next(it);
  auto it = cache.find(key);
  if (it != cache.end())
    if (auto live = it->second.lock()) return live;
  std::shared_ptr<OpTable> t = make();
  cache[key] = t;
  return t;
}

inline int pow2_floor(int v) {
  int p = 1;
  while (p * 2 <= v) p *= 2;
  return p;
}

// ---- kernel geometry ---------------------------------------------------------------------------

// Line lengths with register kernels: powers of two and 3·2^k, 9·2^k (fft_reg.hpp's plans).
template <typename T>
inline bool fast_size(long long n) {
  switch (n) {
    case 64: case 128: case 192: case 256: case 384: case 512: case 576: case 768: case 1024: case 1152:
    case 1536: case 2048: case 3072: case 4096:
      return true;
    default:
      return false;
  }
}

// Calls f.template operator()<N>() with the compile-time line length N = n (a fast_size).
template <typename F>
inline auto with_fast_size(long long n, F&& f) {
  switch (n) {
    case 64: return f.template operator()<64>();
    case 128: return f.template operator()<128>();
    case 192: return f.template operator()<192>();
    case 256: return f.template operator()<256>();
    case 384: return f.template operator()<384>();
    case 512: return f.template operator()<512>();
    case 576: return f.template operator()<576>();
    case 768: return f.template operator()<768>();
    case 1024: return f.template operator()<1024>();
    case 1152: return f.template operator()<1152>();
    case 1536: return f.template operator()<1536>();
    case 2048: return f.template operator()<2048>();
    case 3072: return f.template operator()<3072>();
    case 4096: return f.template operator()<4096>();
    default: throw Error(kEinternal, "no register kernels for lines of " + std::to_string(n) + " points");
  }
}

// Calls f.template operator()<PL>() with radix plan `variant` of line length n (0: the default
// plan, 1: the alternative; lengths with a single plan take it for either).
template <typename F>
inline auto with_fast_plan(long long n, int variant, F&& f) {
  return with_fast_size(n, [&]<int N>() {
    if constexpr (reg_plan_count<N>() > 1)
      if (variant == 1) return f.template operator()<RegPlanV<N, 1>>();
    return f.template operator()<RegPlanV<N, 0>>();
  });
}
// columns / rows per workgroup of the register kernels (kernels_fast.hpp)
template <typename T, bool COMP>
constexpr int fast_w(int n) {
  return fast_lines(n, true, static_cast<int>(sizeof(cx<T>)), COMP);
}
template <typename T, bool COMP>
constexpr int fast_rw(int n) {
  return fast_lines(n, false, static_cast<int>(sizeof(cx<T>)), COMP);
}

// Launch geometry of one grid size (shared by every solve of a lane).
struct Geometry {
  bool col_fast = false, row_fast = false;
  int col_w = 1;                  // fast: columns per workgroup (fast_w or 1)
  int cols1 = 1, cols2 = 1;       // generic: columns per workgroup for one / several products
  int ls_col1 = 1, ls_col2 = 1;
  int rw = 1;                     // fast: rows per workgroup (fast_rw or 1)
  int rows = 1, ls_row = 1;       // generic
  bool lean_col = false, lean_row = false;  // generic: the plan's radices are 2, 3, 4 and 8 only
  // generic, for launches of few workgroups: lines per workgroup within kNarrowLds of LDS, so that
  // several workgroups share a CU
  int cols2n = 1, ls_col2n = 1, rowsn = 1, ls_rown = 1;
  // register grids: the plain steps take the persistent kernels of kernels_plain.hpp (else the
  // general register kernels); chosen per lane shape by the tuner
  bool plain_col = true, plain_row = true;
  // register grids: the radix plan of the column and of the row transforms (0: default, 1: the
  // alternative), fixed per device, precision and line length by the plan tuner
  int plan_col = 0, plan_row = 0;
};


// ---- lines-per-workgroup autotuning ----------------------------------------------------------

// The register kernels can run with W columns (rows) per workgroup or with one. Which is faster
// depends on the grid, the number of solves per launch and the device, so lanes measure it: the
// first uniform stretches of slab steps of a new (device, precision, grid, slots) key run each
// candidate once, timed with events, and later launches use the fastest. Every candidate computes
// the same values per line, so tuning does not change results.
struct TuneKey {
  int device = 0, precision = 0, slots = 0;
  long long ny = 0, nx = 0;
  bool operator<(const TuneKey& o) const {
    return std::tie(device, precision, slots, ny, nx) < std::tie(o.device, o.precision, o.slots, o.ny, o.nx);
  }
};
// A tuning candidate: columns and rows per workgroup, and whether plain steps take the persistent
// column / row kernels.
struct TuneChoice {
  int cw = 1, rw = 1;
  bool pc = true, pr = true;
  bool operator==(const TuneChoice& o) const { return cw == o.cw && rw == o.rw && pc == o.pc && pr == o.pr; }
};
// Timed stretches per candidate: each candidate is timed this many times, in rounds over all
// candidates, and keeps its fastest time (one sample each let a noisy stretch pick a column
// kernel 12 % slower than the best for the headline lane).
inline constexpr int kTuneSamples = 3;
struct TuneState {
  std::vector<TuneChoice> cand;
  std::vector<double> us;  // fastest measured time per solve-step, < 0: not yet
  std::vector<int> n;      // samples taken
  int best = -1;
};
inline std::mutex& tune_mutex() {
  static std::mutex m;
  return m;
}
inline std::map<TuneKey, TuneState>& tune_cache() {
  static std::map<TuneKey, TuneState> c;
  return c;
}

// Candidate to run next for `key` (given the kernel geometry `g`): the best one when tuning is
// done (trial = −1), else the next untimed candidate (trial = its index).
inline TuneChoice tune_pick(const TuneKey& key, const Geometry& g, int& trial) {
  std::lock_guard<std::mutex> lock(tune_mutex());
  TuneState& st = tune_cache()[key];
  if (st.cand.empty()) {
    for (int cw : {g.col_w, 1})
      for (int rw : {g.rw, 1})
        for (bool pc : {true, false})
          for (bool pr : {true, false}) {
            if (!g.col_fast && (cw != g.col_w || !pc)) continue;
            if (!g.row_fast && (rw != g.rw || !pr)) continue;
            const TuneChoice c{cw, rw, pc, pr};
            if (std::find(st.cand.begin(), st.cand.end(), c) == st.cand.end()) st.cand.push_back(c);
          }
    st.us.assign(st.cand.size(), -1.0);
    st.n.assign(st.cand.size(), 0);
    if (st.cand.size() == 1) st.best = 0;
    // BWPM_LINES="<columns>,<rows>[,<plain columns>,<plain rows>]" fixes the choice when it is a
    // candidate (reproducible profiles: hardware counters slow kernels unevenly, which can sway the
    // timed choice); the plain flags default to 1
    if (const char* env = std::getenv("BWPM_LINES")) {
      int cw = 0, rw = 0, pc = 1, pr = 1;
      if (std::sscanf(env, "%d,%d,%d,%d", &cw, &rw, &pc, &pr) >= 2) {
        const TuneChoice want{cw, rw, pc != 0 && g.col_fast, pr != 0 && g.row_fast};
        const auto it = std::find(st.cand.begin(), st.cand.end(), want);
        if (it != st.cand.end()) st.best = static_cast<int>(it - st.cand.begin());
      }
    }
  }
  trial = -1;
  if (st.best >= 0) return st.cand[st.best];
  // the candidate with the fewest samples (the first of them): rounds over all candidates
  const auto it = std::min_element(st.n.begin(), st.n.end());
  trial = static_cast<int>(it - st.n.begin());
  return st.cand[trial];
}

inline void tune_report(const TuneKey& key, int trial, double us_per_solve_step) {
  std::lock_guard<std::mutex> lock(tune_mutex());
  TuneState& st = tune_cache()[key];
  if (trial < 0 || trial >= static_cast<int>(st.us.size()) || st.best >= 0) return;
  if (st.us[trial] < 0 || us_per_solve_step < st.us[trial]) st.us[trial] = us_per_solve_step;
  ++st.n[trial];
  if (std::all_of(st.n.begin(), st.n.end(), [](int k) { return k >= kTuneSamples; }))
    st.best = static_cast<int>(std::min_element(st.us.begin(), st.us.end()) - st.us.begin());
}

// Workgroups of `kern` (blocks of `threads`) the current device holds at once.
inline int resident_groups(const void* kern, int threads) {
  int dev = 0, cus = 0, per = 0;
  HIP_CHECK(hipGetDevice(&dev));
  HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, threads, 0));
  return std::max(1, cus * std::max(per, 1));
}

// The plain kernels (kernels_plain.hpp) run only where they need no scratch memory: at 4096-point
// lines (1024-thread workgroups, 128 registers per thread) the FP64 column kernel spilled 792 bytes
// per thread and gave run-dependent, wrong columns in a 2D solve, while the general kernel of the
// same shape was exact. Those shapes keep the general kernels.
inline bool spill_free(const void* kern) {
  hipFuncAttributes fa{};
  HIP_CHECK(hipFuncGetAttributes(&fa, kern));
  return fa.localSizeBytes == 0;
}

// Grid of a persistent launch over `tiles` tiles with room for `cap` resident workgroups: as few
// workgroups as give every one the same number of tiles (± 1) in ceil(tiles / cap) rounds. With
// `align` (the XCD count, for kernels that map tiles with xcd_tile) and `tiles` a multiple of it,
// the grid is rounded up to a multiple of `align`, so that the tiles b, b + P, b + 2P, … of
// workgroup b keep b's residue: workgroup b runs on XCD b mod align, and every tile the mapping
// gives to one XCD is then processed there (with a grid of 683, neighbouring column groups ran on
// different XCDs and each XCD fetched the 128-byte lines they share: twice the column reads).
inline int persistent_grid(long long tiles, int cap, int align = 1) {
  const long long rounds = (tiles + cap - 1) / cap;
  long long g = std::max<long long>(1, (tiles + rounds - 1) / rounds);
  if (align > 1 && tiles % align == 0 && cap >= align) {
    g = std::min<long long>((g + align - 1) / align * align, std::max<long long>(align, cap / align * align));
    g = std::min<long long>(g, tiles);
  }
  return static_cast<int>(g);
}

// A row step every solve of the launch takes in the plain row kernel (kernels_plain.hpp): one
// transmitted product, no reflections, sparse sources, Born increment or Born plane.
template <typename T, typename ST>
inline bool row_is_plain(const RowDesc<T, ST>& r) {
  return r.active == 0 || (r.nt == 1 && r.nr == 0 && !(r.rowptr && r.sp_add) && r.dense == nullptr && r.inc == nullptr);
}

}  // namespace detail

using namespace detail;

// Where a launch goes: onto a stream, or appended as a kernel node to a graph (after the previous
// node, so the nodes of a chunk run in order). Building graphs node by node instead of capturing
// a stream keeps them immune to what other threads of the process do meanwhile.
struct Emit {
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphNode_t last = nullptr;

  template <typename... P, typename... A>
  void operator()(void (*kern)(P...), dim3 grid, dim3 block, size_t lds, A... args) {
    launch_args<P...>(kern, grid, block, lds, static_cast<P>(args)...);
  }

 private:
  template <typename... P>
  void launch_args(void (*kern)(P...), dim3 grid, dim3 block, size_t lds, P... args) {
    if (!graph) {
      hipLaunchKernelGGL(kern, grid, block, lds, stream, args...);
      return;
    }
    void* params[] = {static_cast<void*>(&args)...};
    hipKernelNodeParams np{};
    np.func = reinterpret_cast<void*>(kern);
    np.gridDim = grid;
    np.blockDim = block;
    np.sharedMemBytes = static_cast<unsigned>(lds);
    np.kernelParams = params;
    hipGraphNode_t node = nullptr;
    HIP_CHECK(hipGraphAddKernelNode(&node, graph, last ? &last : nullptr, last ? 1 : 0, &np));
    last = node;
  }
};

// Launches of one slab step over a lane of S solves of one grid: the kernel geometry and the
// twiddle tables (shared, so a lane runner may outlive the solve it took them from).
template <typename T, bool COMP, typename ST = T>
struct Launcher {
  long long ny = 0, nx = 0, npix = 0;
  Geometry geo;
  std::shared_ptr<Twiddles<T>> twx, twy;
  // solves of a launch that do work (0: all S); the generic kernels size their workgroups by it,
  // since the workgroups of idle slots return at once
  int busy = 0;

  // bit 0: the generic column kernel takes the narrow geometry; bit 1: the generic row kernel
  int narrow_mask(int S) const {
    const long long n = busy > 0 ? busy : S;
    return (n * ((nx + geo.cols2 - 1) / geo.cols2) < kNarrowBelow ? 1 : 0) |
           (n * ((ny + geo.rows - 1) / geo.rows) < kNarrowBelow ? 2 : 0);
  }

  // multi: some solve of the launch has several products at this step (register kernels keep
  // their spectrum in LDS); the result bits do not depend on it.
  void col(const ColDesc<T, ST>* d, int S, hipStream_t s, bool multi) const {
    Emit e;
    e.stream = s;
    col(d, S, e, multi);
  }
  void row(const RowDesc<T, ST>* d, int S, hipStream_t s, bool multi) const {
    Emit e;
    e.stream = s;
    row(d, S, e, multi);
  }

  void col(const ColDesc<T, ST>* d, int S, Emit& s, bool multi) const {
    if (geo.col_fast && COMP == (sizeof(T) == 4)) {
      with_fast_plan(ny, geo.plan_col, [&]<typename PL>() { col_fast_n<PL>(d, S, s, multi); });
      return;
    }
    // generic: one LDS buffer of ny × ls per block (products reload the spectrum from X)
    const bool narrow = (narrow_mask(S) & 1) != 0;
    ColGeom g{static_cast<int>(ny), static_cast<int>(nx), narrow ? geo.cols2n : geo.cols2,
              narrow ? geo.ls_col2n : geo.ls_col2, npix, twy->plan};
    const size_t lds = static_cast<size_t>(ny) * g.ls * sizeof(cx<T>);
    const dim3 grid(static_cast<unsigned>((nx + g.cols - 1) / g.cols), S);
    if (geo.lean_col)
      s(&k_col_step<T, COMP, ST, true>, grid, dim3(kThreads), lds, d, g);
    else
      s(&k_col_step<T, COMP, ST>, grid, dim3(kThreads), lds, d, g);
  }

  void row(const RowDesc<T, ST>* d, int S, Emit& s, bool multi) const {
    if (geo.row_fast && COMP == (sizeof(T) == 4)) {
      with_fast_plan(nx, geo.plan_row, [&]<typename PL>() { row_fast_n<PL>(d, S, s, multi); });
      return;
    }
    const bool narrow = (narrow_mask(S) & 2) != 0;
    RowGeom g{static_cast<int>(ny), static_cast<int>(nx), narrow ? geo.rowsn : geo.rows, narrow ? geo.ls_rown : geo.ls_row,
              npix, twx->plan};
    const size_t lds = 2 * static_cast<size_t>(nx) * g.ls * sizeof(cx<T>);
    const dim3 grid(static_cast<unsigned>((ny + g.rows - 1) / g.rows), S);
    if (geo.lean_row)
      s(&k_row_step<T, COMP, ST, true>, grid, dim3(kThreads), lds, d, g);
    else
      s(&k_row_step<T, COMP, ST>, grid, dim3(kThreads), lds, d, g);
  }

 private:
  // The LDS-stash form is instantiated where its extra LDS fits a workgroup (GROUP: component
  // lines per bank group, col_group_lines(W) or kRowGroupLines).
  template <int N, int W, int GROUP>
  static constexpr bool stash_fits() {
    constexpr int cs = static_cast<int>(sizeof(cx<T>));
    return fast_lds_bytes(N, 2 * W, GROUP, cs, COMP) + 2LL * W * lds_stride(N, GROUP, cs) * cs <= 163840;
  }
  template <typename PL, int W>
  void col_launch(const ColDesc<T, ST>* d, int S, Emit& s, bool multi, const cx<T>* th, const cx<T>* tl) const {
    const dim3 grid(static_cast<unsigned>(nx / W), S);
    if constexpr (stash_fits<PL::N, W, col_group_lines(W)>())
      if (multi) {
        s(&k_col_fast<PL, W, T, COMP, true, ST>, grid, dim3(PL::N / 4 * W), 0, d, static_cast<int>(nx), npix, th, tl);
        return;
      }
    s(&k_col_fast<PL, W, T, COMP, false, ST>, grid, dim3(PL::N / 4 * W), 0, d, static_cast<int>(nx), npix, th, tl);
  }
  template <typename PL, int RW>
  void row_launch(const RowDesc<T, ST>* d, int S, Emit& s, bool multi, const cx<T>* th, const cx<T>* tl) const {
    const dim3 grid(static_cast<unsigned>(ny / RW), S);
    if constexpr (stash_fits<PL::N, RW, kRowGroupLines>())
      if (multi) {
        s(&k_row_fast<PL, RW, T, COMP, true, ST>, grid, dim3(PL::N / 4 * RW), 0, d, npix, th, tl);
        return;
      }
    s(&k_row_fast<PL, RW, T, COMP, false, ST>, grid, dim3(PL::N / 4 * RW), 0, d, npix, th, tl);
  }
 public:
  // Plain steps (every solve one column product; one transmitted row product and no reflections,
  // sparse sources or Born terms) on the persistent kernels of kernels_plain.hpp; false if the grid
  // has no register kernels for this direction.
  bool col_plain(const ColDesc<T, ST>* d, int S, Emit& s) const {
    if (!geo.col_fast || !geo.plain_col) return false;
    return with_fast_plan(ny, geo.plan_col, [&]<typename PL>() { return col_plain_n<PL>(d, S, s); });
  }
  bool row_plain(const RowDesc<T, ST>* d, int S, Emit& s) const {
    if (!geo.row_fast || !geo.plain_row) return false;
    return with_fast_plan(nx, geo.plan_row, [&]<typename PL>() { return row_plain_n<PL>(d, S, s); });
  }

 private:
  template <typename PL, int W>
  bool col_plain_launch(const ColDesc<T, ST>* d, int S, Emit& s) const {
    auto kern = &k_col_plain<PL, W, T, COMP, ST>;
    static const bool fits = spill_free(reinterpret_cast<const void*>(kern));
    if (!fits) return false;
    static const int cap = resident_groups(reinterpret_cast<const void*>(kern), PL::N / 4 * W);
    const int groups = static_cast<int>(nx / W);
    s(kern, dim3(static_cast<unsigned>(persistent_grid(static_cast<long long>(S) * groups, cap, kXcds))),
      dim3(PL::N / 4 * W), 0, d, S, groups, static_cast<int>(nx), twy->hi.template as<cx<T>>(),
      twy->lo.template as<cx<T>>());
    return true;
  }
  template <typename PL, int RW>
  bool row_plain_launch(const RowDesc<T, ST>* d, int S, Emit& s) const {
    auto kern = &k_row_plain<PL, RW, T, COMP, ST>;
    static const bool fits = spill_free(reinterpret_cast<const void*>(kern));
    if (!fits) return false;
    static const int cap = resident_groups(reinterpret_cast<const void*>(kern), PL::N / 4 * RW);
    const int groups = static_cast<int>(ny / RW);
    s(kern, dim3(static_cast<unsigned>(persistent_grid(static_cast<long long>(S) * groups, cap))), dim3(PL::N / 4 * RW), 0,
      d, S, groups, twx->hi.template as<cx<T>>(), twx->lo.template as<cx<T>>());
    return true;
  }
  template <typename PL>
  bool col_plain_n(const ColDesc<T, ST>* d, int S, Emit& s) const {
    constexpr int W = fast_w<T, COMP>(PL::N);
    return geo.col_w == W ? col_plain_launch<PL, W>(d, S, s) : col_plain_launch<PL, 1>(d, S, s);
  }
  template <typename PL>
  bool row_plain_n(const RowDesc<T, ST>* d, int S, Emit& s) const {
    constexpr int RW = fast_rw<T, COMP>(PL::N);
    return geo.rw == RW ? row_plain_launch<PL, RW>(d, S, s) : row_plain_launch<PL, 1>(d, S, s);
  }

  template <typename PL>
  void col_fast_n(const ColDesc<T, ST>* d, int S, Emit& s, bool multi) const {
    constexpr int W = fast_w<T, COMP>(PL::N);
    const cx<T>* th = twy->hi.template as<cx<T>>();
    const cx<T>* tl = twy->lo.template as<cx<T>>();
    if (geo.col_w == W)
      col_launch<PL, W>(d, S, s, multi, th, tl);
    else
      col_launch<PL, 1>(d, S, s, multi, th, tl);
  }
  template <typename PL>
  void row_fast_n(const RowDesc<T, ST>* d, int S, Emit& s, bool multi) const {
    constexpr int RW = fast_rw<T, COMP>(PL::N);
    const cx<T>* th = twx->hi.template as<cx<T>>();
    const cx<T>* tl = twx->lo.template as<cx<T>>();
    if (geo.rw == RW)
      row_launch<PL, RW>(d, S, s, multi, th, tl);
    else
      row_launch<PL, 1>(d, S, s, multi, th, tl);
  }
};

// ---- radix-plan tuning ---------------------------------------------------------------------------

// The radix plan of the register transforms (fft_reg.hpp's RegPlans: a default and an alternative
// per line length) is chosen once per device, precision, line length and direction, before any
// solve of that key runs. Every solve of a process then uses one plan, so results stay
// bit-identical between a solve alone, in a lane and in the queue. The tuner times the persistent
// plain kernel of each plan (or the general kernel where the plain one needs scratch memory) on
// zeroed scratch planes: a lane of S solves of min(n, 512) lines of n points (at most 4 M pixels),
// one warm launch and the best of three timed launches per plan, alternating.
// BWPM_PLAN=0 (or 1) fixes the default (or the alternative) for every length.
struct PlanKey {
  int device = 0, precision = 0;
  long long n = 0;
  bool col = true;
  bool operator<(const PlanKey& o) const {
    return std::tie(device, precision, n, col) < std::tie(o.device, o.precision, o.n, o.col);
  }
};
struct PlanChoice {
  int variant = 0;
  double us[2] = {-1.0, -1.0};  // timed launch of each plan (µs), −1: not timed
};
inline std::mutex& plan_mutex() {
  static std::mutex m;
  return m;
}
inline std::map<PlanKey, PlanChoice>& plan_cache() {
  static std::map<PlanKey, PlanChoice> c;
  return c;
}

// Every tuned choice as text: one line "<precision> <col|row> <n> <variant> <µs plan 0> <µs plan 1>".
inline std::string plan_report() {
  std::lock_guard<std::mutex> lock(plan_mutex());
  std::string out;
  for (const auto& [k, c] : plan_cache()) {
    char line[160];
    std::snprintf(line, sizeof(line), "%d %s %lld %d %.3f %.3f\n", k.precision, k.col ? "col" : "row", k.n,
                  c.variant, c.us[0], c.us[1]);
    out += line;
  }
  return out;
}

template <typename T, bool COMP, typename ST>
PlanChoice plan_time(long long n, bool col, const Geometry& g0, hipStream_t stream);

template <typename T, bool COMP, typename ST>
int plan_pick(int device, int precision, long long n, bool col, const Geometry& g0, hipStream_t stream) {
  if (!with_fast_size(n, [&]<int N>() { return reg_plan_count<N>() > 1; })) return 0;
  if (const char* env = std::getenv("BWPM_PLAN")) {
    if (env[0] == '0' || env[0] == '1') return env[0] - '0';
  }
  std::lock_guard<std::mutex> lock(plan_mutex());
  const PlanKey key{device, precision, n, col};
  auto it = plan_cache().find(key);
  if (it != plan_cache().end()) return it->second.variant;

  // a device too full for the scratch planes keeps the default plan (and does not retry)
  try {
    plan_cache()[key] = plan_time<T, COMP, ST>(n, col, g0, stream);
  } catch (const Error& e) {
    if (e.code != kEnomem) throw;
    (void)hipGetLastError();
    plan_cache()[key] = PlanChoice{};
  }
  return plan_cache()[key].variant;
}

// Times both radix plans of line length n (plan_pick) on zeroed scratch planes.
template <typename T, bool COMP, typename ST>
PlanChoice plan_time(long long n, bool col, const Geometry& g0, hipStream_t stream) {
  const long long lines = std::min<long long>(n, 512);
  const int S = static_cast<int>(std::max<long long>(1, std::min<long long>(16, (4LL << 20) / (lines * n))));
  Launcher<T, COMP, ST> L;
  L.ny = col ? n : lines;
  L.nx = col ? lines : n;
  L.npix = L.ny * L.nx;
  L.geo = g0;
  if (col) {
    L.geo.col_fast = true;
    L.geo.plain_col = true;
    const int w = fast_w<T, COMP>(static_cast<int>(n));
    L.geo.col_w = (L.nx % w == 0) ? w : 1;
  } else {
    L.geo.row_fast = true;
    L.geo.plain_row = true;
    const int r = fast_rw<T, COMP>(static_cast<int>(n));
    L.geo.rw = (L.ny % r == 0) ? r : 1;
  }
  auto tw = std::make_shared<Twiddles<T>>();
  tw->build(n, stream);
  L.twx = tw;
  L.twy = tw;
  const size_t plane = static_cast<size_t>(L.npix) * sizeof(px2<ST>);
  DevBuf X, P, tab, lo, dprod, dcol, drow;
  X.alloc(plane * S);
  P.alloc(plane * S);
  tab.alloc(static_cast<size_t>(L.npix) * sizeof(cx<T>));
  lo.alloc(static_cast<size_t>(L.npix) * sizeof(cx<T>));
  HIP_CHECK(hipMemsetAsync(X.as<void>(), 0, plane * S, stream));
  HIP_CHECK(hipMemsetAsync(P.as<void>(), 0, plane * S, stream));
  HIP_CHECK(hipMemsetAsync(tab.as<void>(), 0, tab.bytes(), stream));
  HIP_CHECK(hipMemsetAsync(lo.as<void>(), 0, lo.bytes(), stream));
  const std::vector<Prod<T>> prod{Prod<T>{0, tab.as<const cx<T>>(), lo.as<const cx<T>>()}};
  dprod = upload_vec(prod, stream);
  std::vector<ColDesc<T, ST>> cd(S);
  std::vector<RowDesc<T, ST>> rd(S);
  for (int i = 0; i < S; ++i) {
    px2<ST>* xi = X.as<px2<ST>>() + static_cast<long long>(i) * L.npix;
    px2<ST>* pi = P.as<px2<ST>>() + static_cast<long long>(i) * L.npix;
    cd[i] = ColDesc<T, ST>{xi, pi, dprod.as<const Prod<T>>(), 1, 1, nullptr};
    rd[i] = RowDesc<T, ST>{};
    rd[i].active = 1;
    rd[i].nt = 1;
    rd[i].prod = pi;
    rd[i].x_out = xi;
  }
  dcol = upload_vec(cd, stream);
  drow = upload_vec(rd, stream);
  auto run = [&](int v) {
    Emit e;
    e.stream = stream;
    if (col) {
      L.geo.plan_col = v;
      if (!L.col_plain(dcol.as<const ColDesc<T, ST>>(), S, e)) L.col(dcol.as<const ColDesc<T, ST>>(), S, stream, false);
    } else {
      L.geo.plan_row = v;
      if (!L.row_plain(drow.as<const RowDesc<T, ST>>(), S, e)) L.row(drow.as<const RowDesc<T, ST>>(), S, stream, false);
    }
    HIP_CHECK(hipGetLastError());
  };
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  PlanChoice c;
  run(0);
  run(1);
  for (int rep = 0; rep < 3; ++rep)
    for (int v = 0; v < 2; ++v) {
      HIP_CHECK(hipEventRecord(e0, stream));
      run(v);
      HIP_CHECK(hipEventRecord(e1, stream));
      HIP_CHECK(hipEventSynchronize(e1));
      float ms = 0.f;
      HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
      const double us = 1e3 * ms;
      if (c.us[v] < 0 || us < c.us[v]) c.us[v] = us;
    }
  HIP_CHECK(hipStreamSynchronize(stream));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  c.variant = c.us[1] < c.us[0] ? 1 : 0;
  return c;
}

// A lane step: one column launch and one row launch over every solve of the lane, or a Born
// sub-step of a single-solve lane.
enum OpKind : int { kOpPair = 0, kOpBornFwd = 1, kOpBornBwd = 2 };

template <typename T, typename ST = T>
struct Op {
  int kind = kOpPair;
  ColDesc<T, ST> col{};
  RowDesc<T, ST> row{};
  // the row kernel merges several T products of a row in LDS (its interface variant): many rows
  // mix pairs, or storage is narrower than the arithmetic (a merge through a product plane would
  // round); otherwise rows that use one product take it directly in the plain variant
  bool row_stash = false;
  bool col_phase = false;  // the column work is one phase product (the plain column kernel's case)
  long long slab = 0;
  bool bidir = false;
};

// ST: storage of the two streaming planes (X and the product planes); T: the arithmetic.
template <typename T, bool COMP, typename ST = T>
class Engine final : public EngineBase {
 public:
  explicit Engine(Spec s) : sp_(std::move(s)) {
    HIP_CHECK(hipSetDevice(sp_.device));
    HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    try {
      setup();
    } catch (...) {
      (void)hipStreamDestroy(stream_);
      throw;
    }
  }
  ~Engine() override {
    (void)hipSetDevice(sp_.device);
    (void)hipStreamSynchronize(stream_);
    born_.reset();
    (void)hipStreamDestroy(stream_);
  }

  hipStream_t stream() const override { return stream_; }
  size_t device_bytes() const override { return bytes_; }
  int precision() const override { return sp_.precision; }
  long long ny() const override { return sp_.ny; }
  long long nx() const override { return sp_.nx; }
  int device() const override { return sp_.device; }
  bool has_born() const override { return born_ != nullptr; }
  long long nz() const override { return sp_.nz; }
  std::vector<c128> fluctuation() const override {
    HIP_CHECK(hipSetDevice(sp_.device));
    HIP_CHECK(hipStreamSynchronize(stream_));
    if (!born_) return std::vector<c128>(static_cast<size_t>(sp_.nz * npix_), c128(0.0, 0.0));
    return born_->fluctuation();
  }

  const void* device_plane(int which, long long p) const override {
    require(ran_, "the solver has not run yet");
    const int nit = std::max(last_iterations_, 1);
    if (which == 0) return iter_t_.as<px2<T>>() + (nit - 1) * npix_;
    if (which == 1) {
      require(last_iterations_ > 0, "a run without iterations has no reflected field");
      return iter_r_.as<px2<T>>() + (nit - 1) * npix_;
    }
    require(which == 2 || which == 3, "which must be 0 (transmitted), 1 (reflected), 2 (forward) or 3 (backward)");
    const int slot = keep_slot(p);
    require(slot >= 0 && p >= 0 && p <= sp_.nz, "plane " + std::to_string(p) + " was not kept in the last run");
    require(which == 2 || last_iterations_ > 0, "a run without iterations has no backward field");
    return (which == 2 ? keepf_ : keepb_).as<px2<T>>() + slot * npix_;
  }

  void run(const c128* field, int iterations, const std::vector<long long>& keep, RunOutput& out) override {
    HIP_CHECK(hipSetDevice(sp_.device));
    upload(field);
    std::vector<EngineBase*> one{this};
    std::vector<std::vector<long long>> keeps{keep};
    out.gpu_ms = run_lane(one, iterations, keeps, stream_, false, nullptr);
    download(out, iterations, keep);
  }

  void upload(const c128* field) override {
    HIP_CHECK(hipSetDevice(sp_.device));
    HIP_CHECK(hipMemcpyAsync(stage_.as<void>(), field, npix_ * 2 * sizeof(c128), hipMemcpyHostToDevice, stream_));
    launch_1d(k_from_c128<T>, npix_, stage_.as<zd>(), field_in_.as<px2<T>>(), npix_);
    HIP_CHECK(hipStreamSynchronize(stream_));
  }

  double run_resident(int iterations, bool time_kernels, KernelTimes* kt) override {
    HIP_CHECK(hipSetDevice(sp_.device));
    std::vector<EngineBase*> one{this};
    std::vector<std::vector<long long>> keeps{{}};
    return run_lane(one, iterations, keeps, stream_, time_kernels, kt);
  }

  void download(RunOutput& out, int iterations, const std::vector<long long>& keep) override {
    HIP_CHECK(hipSetDevice(sp_.device));
    const int nit = std::max(iterations, 1);
    auto fetch = [&](const px2<T>* src) {
      std::vector<c128> v(npix_ * 2);
      launch_1d(k_to_c128<T>, npix_, src, stage_.as<zd>(), npix_);
      HIP_CHECK(hipMemcpyAsync(v.data(), stage_.as<void>(), npix_ * 2 * sizeof(c128), hipMemcpyDeviceToHost, stream_));
      HIP_CHECK(hipStreamSynchronize(stream_));
      return v;
    };
    out.transmitted = fetch(iter_t_.as<px2<T>>() + (nit - 1) * npix_);
    out.transmitted_iter.clear();
    out.reflected_iter.clear();
    if (iterations > 0) {
      out.reflected = fetch(iter_r_.as<px2<T>>() + (nit - 1) * npix_);
      for (int it = 0; it < nit; ++it) {
        out.transmitted_iter.push_back(fetch(iter_t_.as<px2<T>>() + it * npix_));
        out.reflected_iter.push_back(fetch(iter_r_.as<px2<T>>() + it * npix_));
      }
    } else {
      out.reflected.assign(npix_ * 2, c128(0.0, 0.0));
    }
    out.forward.clear();
    out.backward.clear();
    for (long long p : keep) {
      const int slot = keep_slot(p);
      if (slot < 0 || p < 0 || p > sp_.nz) continue;  // planes 0 … nz exist (solver.cpp:474,493)
      out.forward[p] = fetch(keepf_.as<px2<T>>() + slot * npix_);
      if (iterations > 0) out.backward[p] = fetch(keepb_.as<px2<T>>() + slot * npix_);
    }
    out.stored_values = iterations > 0 ? 2 * total_sparse_ : 0;
    out.slab_steps = last_steps_;
    out.gmres_iterations = born_ ? born_->iterations() : 0;
  }

  // Static step program of one run (reference order, solver.cpp:469-525):
  //  forward:  e = field; for p: [FwdR → src_bwd[p]] FwdT (+ mask·Born increment, R1 → born_bwd[p+1]);
  //            e += src_fwd[p+1] + born_fwd[p+1]
  //  backward: e = born_bwd[nz]; for p = nz…1: [R3 → born_fwd[p]; e += R2] [BwdR → src_fwd[p]] BwdT;
  //            e += src_bwd[p−1] + born_bwd[p−1]
  std::vector<Op<T, ST>> program(int iterations, const std::vector<long long>& keep) {
    if (born_) born_->reset_iterations();  // the result reports this run's GMRES iterations
    const long long nz = sp_.nz;
    const int nit = std::max(iterations, 1);
    const bool bidir = iterations > 0;
    keep_ = keep;
    last_iterations_ = iterations;
    ran_ = true;
    const size_t pb = npix_ * sizeof(px2<T>);
    if (iter_t_.bytes() < nit * pb) iter_t_.alloc(nit * pb);
    if (bidir && iter_r_.bytes() < nit * pb) iter_r_.alloc(nit * pb);
    if (!keep_.empty()) {
      if (keepf_.bytes() < keep_.size() * pb) keepf_.alloc(keep_.size() * pb);
      if (bidir && keepb_.bytes() < keep_.size() * pb) keepb_.alloc(keep_.size() * pb);
    }
    px2<ST>* X = X_.as<px2<ST>>();
    px2<T>* E = e_.as<px2<T>>();
    std::vector<Op<T, ST>> ops;
    long long steps = 0;
    auto born_at = [&](long long slab) { return born_ && slab >= 0 && slab < nz && has_host_[slab]; };
    for (int it = 0; it < nit; ++it) {
      const bool last = it == nit - 1;
      auto kept = [&](DevBuf& buf, long long p) -> px2<T>* {
        if (!last) return nullptr;
        const int slot = keep_slot(p);
        return slot < 0 ? nullptr : buf.as<px2<T>>() + slot * npix_;
      };
      px2<T>* trans = iter_t_.as<px2<T>>() + it * npix_;
      {  // forward start: e = field (+ src_fwd[0], never present)
        Op<T, ST> o;
        o.row.active = 1;
        o.row.e_in = field_in_.as<px2<T>>();
        o.row.e_out2 = kept(keepf_, 0);
        o.row.e_out = born_at(0) ? E : nullptr;
        o.row.x_out = X;
        ops.push_back(o);
      }
      for (long long p = 0; p < nz; ++p) {
        const Boundary& B = bnd_[p];
        const bool with_r = bidir && !B.rpair.empty();
        if (born_at(p)) {
          Op<T, ST> b;
          b.kind = kOpBornFwd;
          b.slab = p;
          b.bidir = bidir;
          ops.push_back(b);
        }
        Op<T, ST> o;
        o.col = col_desc(B, true, with_r);
        o.col_phase = o.col.nprod == 1 && B.mleft[0] == B.mright[0];
        o.row = row_desc(B, p, with_r);
        o.row_stash = B.mixed_rows || sizeof(ST) != sizeof(T);
        o.row.inc = born_at(p) ? born_->increment() : nullptr;
        if (B.sp_count > 0) {
          o.row.sp_add = bidir ? sp_vals(fwd_vals_, p) : nullptr;
          o.row.sp_store = with_r ? sp_vals(bwd_vals_, p) : nullptr;
        }
        o.row.dense = (born_ && it > 0) ? born_->fwd_plane(p + 1) : nullptr;
        const bool final_plane = p + 1 == nz;
        o.row.e_out = final_plane ? trans : (born_at(p + 1) ? E : nullptr);
        o.row.e_out2 = kept(keepf_, p + 1);
        o.row.x_out = final_plane ? nullptr : X;
        ops.push_back(o);
        ++steps;
      }
      if (!bidir) break;
      px2<T>* refl = iter_r_.as<px2<T>>() + it * npix_;
      {  // backward start: e = 0 (+ src_bwd[nz], never present) + born_bwd[nz]
        Op<T, ST> o;
        o.row.active = 1;
        o.row.dense = born_ ? born_->bwd_plane(nz) : nullptr;
        o.row.e_out2 = kept(keepb_, nz);
        o.row.e_out = born_at(nz - 1) ? E : nullptr;
        o.row.x_out = X;
        ops.push_back(o);
      }
      for (long long p = nz; p >= 1; --p) {
        const Boundary& B = bnd_[p - 1];
        if (born_at(p - 1)) {
          Op<T, ST> b;  // R3 from e at plane p into born_fwd[p], then e += R2 (no mask) and the row FFT
          b.kind = kOpBornBwd;
          b.slab = p - 1;
          b.bidir = true;
          ops.push_back(b);
          Op<T, ST> o;
          o.row.active = 1;
          o.row.e_in = E;
          o.row.inc = born_->increment();
          o.row.x_out = X;
          ops.push_back(o);
        }
        const bool with_r = !B.rpair.empty();
        Op<T, ST> o;
        o.col = col_desc(B, false, with_r);
        o.col_phase = o.col.nprod == 1 && B.mleft[0] == B.mright[0];
        o.row = row_desc(B, p - 1, with_r);
        o.row_stash = B.mixed_rows || sizeof(ST) != sizeof(T);
        if (B.sp_count > 0) {
          o.row.sp_add = sp_vals(bwd_vals_, p - 1);
          o.row.sp_store = with_r ? sp_vals(fwd_vals_, p - 1) : nullptr;
        }
        o.row.dense = born_ ? born_->bwd_plane(p - 1) : nullptr;
        const bool final_plane = p - 1 == 0;
        o.row.e_out = final_plane ? refl : (born_at(p - 2) ? E : nullptr);
        o.row.e_out2 = kept(keepb_, p - 1);
        o.row.x_out = final_plane ? nullptr : X;
        ops.push_back(o);
        ++steps;
      }
    }
    // Rows of one point (2D solves, nx = 1): the row transforms are the identity, so a step whose
    // row work only copies its one product into X (no mask, sources, Born terms or outputs) has the
    // column kernel write straight into X, in place, and no row launch.
    if (sp_.nx == 1)
      for (Op<T, ST>& o : ops)
        if (o.kind == kOpPair && row_is_copy(o, X)) {
          o.col.out = X;
          o.row = RowDesc<T, ST>{};
        }
    last_steps_ = steps;
    return ops;
  }

  static bool row_is_copy(const Op<T, ST>& o, const px2<ST>* X) {
    const RowDesc<T, ST>& r = o.row;
    return o.col.in != nullptr && o.col.nprod == 1 && o.col.nt == 1 && o.col.rowuse == nullptr && r.active != 0 &&
           r.nt == 1 && r.nr == 0 && r.prod == o.col.out && r.tsel == nullptr && r.rowuse == nullptr &&
           r.mask == nullptr && r.mask_y == nullptr && r.e_in == nullptr && r.inc == nullptr && r.rowptr == nullptr &&
           r.sp_add == nullptr && r.sp_store == nullptr && r.dense == nullptr && r.e_out == nullptr &&
           r.e_out2 == nullptr && r.x_out == X;
  }

  // Run the programs of every member (same grid and precision) in lockstep on `stream`.
  double run_lane(std::vector<EngineBase*>& members, int iterations, const std::vector<std::vector<long long>>& keeps,
                  hipStream_t stream, bool time_kernels, KernelTimes* kt) override {
    const int S = static_cast<int>(members.size());
    std::vector<Engine*> es;
    for (EngineBase* b : members) {
      auto* e = dynamic_cast<Engine*>(b);
      require(e && e->sp_.ny == sp_.ny && e->sp_.nx == sp_.nx && e->sp_.device == sp_.device,
              "a batch lane needs solves of equal grid, precision and device");
      es.push_back(e);
    }
    std::vector<std::vector<Op<T, ST>>> progs(S);
    size_t G = 0;
    for (int i = 0; i < S; ++i) {
      progs[i] = es[i]->program(iterations, keeps[i]);
      G = std::max(G, progs[i].size());
      if (S > 1)
        for (const Op<T, ST>& o : progs[i])
          require(o.kind == kOpPair, "solves with index fluctuations run one per lane");
    }
    // every member's buffers (tables, kept planes) must be ready before the lane starts
    for (Engine* e : es)
      if (e->stream_ != stream) HIP_CHECK(hipStreamSynchronize(e->stream_));
    std::vector<ColDesc<T, ST>> hcol(G * S);
    std::vector<RowDesc<T, ST>> hrow(G * S);
    std::vector<char> has_col(G, 0), has_row(G, 0), multi_col(G, 0), multi_row(G, 0), plain_row(G, 1), plain_col(G, 1);
    for (size_t g = 0; g < G; ++g)
      for (int i = 0; i < S; ++i) {
        if (g >= progs[i].size() || progs[i][g].kind != kOpPair) continue;
        const Op<T, ST>& o = progs[i][g];
        hcol[g * S + i] = o.col;
        hrow[g * S + i] = o.row;
        has_col[g] |= o.col.in != nullptr;
        has_row[g] |= o.row.active != 0;
        multi_col[g] |= o.col.in != nullptr && o.col.nprod > 1;
        multi_row[g] |= o.row.active != 0 && o.row.nt > 1 && o.row_stash;
        plain_row[g] &= row_is_plain(o.row);
        plain_col[g] &= o.col.in == nullptr || o.col_phase;
      }
    if (lane_col_.bytes() < hcol.size() * sizeof(ColDesc<T, ST>)) lane_col_.alloc(hcol.size() * sizeof(ColDesc<T, ST>));
    if (lane_row_.bytes() < hrow.size() * sizeof(RowDesc<T, ST>)) lane_row_.alloc(hrow.size() * sizeof(RowDesc<T, ST>));
    HIP_CHECK(hipMemcpyAsync(lane_col_.as<void>(), hcol.data(), hcol.size() * sizeof(ColDesc<T, ST>),
                             hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipMemcpyAsync(lane_row_.as<void>(), hrow.data(), hrow.size() * sizeof(RowDesc<T, ST>),
                             hipMemcpyHostToDevice, stream));
    if (total_sparse_ > 0 || S > 1)
      for (Engine* e : es)
        if (e->total_sparse_ > 0)
          launch_1d_on(stream, k_fill_zero<T>, e->total_sparse_, e->fwd_vals_.as<px2<T>>(), e->total_sparse_);

    KernelTimes times;
    hipEvent_t t0, t1, k0 = nullptr, k1 = nullptr;
    HIP_CHECK(hipEventCreate(&t0));
    HIP_CHECK(hipEventCreate(&t1));
    if (time_kernels) {
      HIP_CHECK(hipEventCreate(&k0));
      HIP_CHECK(hipEventCreate(&k1));
    }
    auto timed = [&](bool col, auto&& launch) {
      if (!time_kernels) {
        launch();
        HIP_CHECK(hipGetLastError());
        return;
      }
      HIP_CHECK(hipEventRecord(k0, stream));
      launch();
      HIP_CHECK(hipGetLastError());
      HIP_CHECK(hipEventRecord(k1, stream));
      HIP_CHECK(hipEventSynchronize(k1));
      float ms = 0;
      HIP_CHECK(hipEventElapsedTime(&ms, k0, k1));
      (col ? times.col_ms : times.row_ms) += ms;
      ++(col ? times.col_launches : times.row_launches);
    };
    HIP_CHECK(hipEventRecord(t0, stream));
    // lines per workgroup: the tuned choice for this lane shape; untuned shapes time each
    // candidate on a uniform stretch of kTuneSteps slab steps first (not while timing kernels)
    constexpr size_t kTuneSteps = 32;
    const TuneKey tk{sp_.device, precision_code(), S, sp_.ny, sp_.nx};
    Launcher<T, COMP, ST> Lc = launch_;
    for (size_t g = 0; g < G;) {
      int trial = -1;
      const TuneChoice choice = tune_pick(tk, launch_.geo, trial);
      Lc.geo.col_w = choice.cw;
      Lc.geo.rw = choice.rw;
      Lc.geo.plain_col = choice.pc;
      Lc.geo.plain_row = choice.pr;
      bool uniform = !time_kernels && trial >= 0 && g + kTuneSteps <= G;
      for (size_t q = g; uniform && q < g + kTuneSteps; ++q)
        uniform = (S > 1 || progs[0][q].kind == kOpPair) && has_col[q] && has_row[q] && !multi_col[q] &&
                  !multi_row[q] && plain_col[q] && plain_row[q];
      const size_t end = uniform ? g + kTuneSteps : std::min(G, g + kTuneSteps);
      hipEvent_t a = nullptr, b = nullptr;
      if (uniform) {
        HIP_CHECK(hipEventCreate(&a));
        HIP_CHECK(hipEventCreate(&b));
        HIP_CHECK(hipEventRecord(a, stream));
      }
      for (; g < end; ++g) {
        if (S == 1 && progs[0][g].kind != kOpPair) {
          const Op<T, ST>& o = progs[0][g];
          if (o.kind == kOpBornFwd)
            born_->forward(e_.template as<px2<T>>(), o.slab, o.bidir);
          else
            born_->backward(e_.template as<px2<T>>(), o.slab);
          continue;
        }
        const ColDesc<T, ST>* dc = lane_col_.as<ColDesc<T, ST>>() + g * S;
        const RowDesc<T, ST>* dr = lane_row_.as<RowDesc<T, ST>>() + g * S;
        Emit em;
        em.stream = stream;
        if (has_col[g])
          timed(true, [&] {
            if (!plain_col[g] || !Lc.col_plain(dc, S, em)) Lc.col(dc, S, stream, multi_col[g] != 0);
          });
        if (has_row[g])
          timed(false, [&] {
            if (!plain_row[g] || !Lc.row_plain(dr, S, em)) Lc.row(dr, S, stream, multi_row[g] != 0);
          });
      }
      if (uniform) {
        HIP_CHECK(hipEventRecord(b, stream));
        HIP_CHECK(hipEventSynchronize(b));
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, a, b));
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        tune_report(tk, trial, 1e3 * ms / (static_cast<double>(kTuneSteps) * S));
      }
    }
    HIP_CHECK(hipEventRecord(t1, stream));
    HIP_CHECK(hipEventSynchronize(t1));
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, t0, t1));
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (time_kernels) {
      (void)hipEventDestroy(k0);
      (void)hipEventDestroy(k1);
    }
    if (kt) *kt = times;
    return ms;
  }

  // Continuous lane (see run_continuous_lane): members join between chunks of slab steps.
  static void continuous(LaneFeed& feed, int S, int chunk, hipStream_t stream) {
    require(S >= 1 && chunk >= 1, "lane slots and chunk must be >= 1");
    std::vector<Engine*> slot(S, nullptr);
    std::vector<LaneMember> mem(S);
    std::vector<std::vector<Op<T, ST>>> prog(S);
    std::vector<size_t> pc(S, 0);
    Launcher<T, COMP, ST> L;
    bool have_launcher = false;
    const size_t n = static_cast<size_t>(chunk) * S;
    DevBuf dcol(n * sizeof(ColDesc<T, ST>)), drow(n * sizeof(RowDesc<T, ST>));
    // descriptor staging in pinned memory, two sets: a set is rewritten only after the chunk that
    // copied it has completed (at most two chunks are in flight)
    PinnedBuf hcol[2] = {PinnedBuf(n * sizeof(ColDesc<T, ST>)), PinnedBuf(n * sizeof(ColDesc<T, ST>))};
    PinnedBuf hrow[2] = {PinnedBuf(n * sizeof(RowDesc<T, ST>)), PinnedBuf(n * sizeof(RowDesc<T, ST>))};
    std::vector<char> has_col(chunk), has_row(chunk), multi_col(chunk), multi_row(chunk), plain_row(chunk),
        plain_col(chunk);
    std::map<std::tuple<size_t, int, int, int, int>, hipGraphExec_t> graphs;
    bool graphs_ok = true;
    auto drop_graphs = [&] {
      for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
      graphs.clear();
    };
    hipEvent_t prev = nullptr;
    auto settle = [&] {
      if (!prev) return;
      HIP_CHECK(hipEventSynchronize(prev));
      (void)hipEventDestroy(prev);
      prev = nullptr;
    };
    int set = 0;
    try {
      for (;;) {
        bool any = false;
        for (int i = 0; i < S; ++i) {
          if (!slot[i]) {
            const LaneMember m = feed.next();
            if (m.engine) {
              auto* e = dynamic_cast<Engine*>(m.engine);
              require(e != nullptr, "a lane needs solves of equal grid, precision and device");
              if (have_launcher)
                require(e->sp_.ny == L.ny && e->sp_.nx == L.nx, "a lane needs solves of equal grid");
              prog[i] = e->program(m.iterations, {});
              for (const Op<T, ST>& o : prog[i])
                require(o.kind == kOpPair, "solves with index fluctuations run one per lane");
              HIP_CHECK(hipStreamSynchronize(e->stream_));
              if (e->total_sparse_ > 0)
                launch_1d_on(stream, k_fill_zero<T>, e->total_sparse_, e->fwd_vals_.template as<px2<T>>(),
                             e->total_sparse_);
              if (!have_launcher) {
                L = e->launch_;
                have_launcher = true;
              }
              slot[i] = e;
              mem[i] = m;
              pc[i] = 0;
            }
          }
          any |= slot[i] != nullptr;
        }
        if (!any) {
          settle();
          if (!feed.wait()) break;
          continue;
        }
        size_t G = 0;
        for (int i = 0; i < S; ++i)
          if (slot[i]) G = std::max(G, std::min<size_t>(chunk, prog[i].size() - pc[i]));
        auto* hc = hcol[set].template as<ColDesc<T, ST>>();
        auto* hr = hrow[set].template as<RowDesc<T, ST>>();
        std::fill(hc, hc + G * S, ColDesc<T, ST>{});
        std::fill(hr, hr + G * S, RowDesc<T, ST>{});
        std::fill(has_col.begin(), has_col.end(), 0);
        std::fill(has_row.begin(), has_row.end(), 0);
        std::fill(multi_col.begin(), multi_col.end(), 0);
        std::fill(multi_row.begin(), multi_row.end(), 0);
        std::fill(plain_row.begin(), plain_row.end(), 1);
        std::fill(plain_col.begin(), plain_col.end(), 1);
        for (size_t g = 0; g < G; ++g)
          for (int i = 0; i < S; ++i) {
            if (!slot[i] || pc[i] + g >= prog[i].size()) continue;
            const Op<T, ST>& o = prog[i][pc[i] + g];
            hc[g * S + i] = o.col;
            hr[g * S + i] = o.row;
            has_col[g] |= o.col.in != nullptr;
            has_row[g] |= o.row.active != 0;
            multi_col[g] |= o.col.in != nullptr && o.col.nprod > 1;
            multi_row[g] |= o.row.active != 0 && o.row.nt > 1 && o.row_stash;
            plain_row[g] &= row_is_plain(o.row);
            plain_col[g] &= o.col.in == nullptr || o.col_phase;
          }
        HIP_CHECK(hipMemcpyAsync(dcol.as<void>(), hc, G * S * sizeof(ColDesc<T, ST>), hipMemcpyHostToDevice, stream));
        HIP_CHECK(hipMemcpyAsync(drow.as<void>(), hr, G * S * sizeof(RowDesc<T, ST>), hipMemcpyHostToDevice, stream));
        // uniform chunk: every step is one plain column launch and either one plain row launch
        // or, in every step, none (rows of one point whose row work is a copy)
        bool uniform = true;
        const bool rows_all = G > 0 && has_row[0];
        for (size_t g = 0; g < G; ++g)
          uniform = uniform && has_col[g] && has_row[g] == rows_all && !multi_col[g] && plain_col[g] &&
                    (!rows_all || (!multi_row[g] && plain_row[g]));
        int active = 0;
        for (int i = 0; i < S; ++i) active += slot[i] != nullptr;
        const TuneKey tk{slot_device(slot), precision_code(), S, L.ny, L.nx};
        int trial = -1;
        const TuneChoice choice = tune_pick(tk, L.geo, trial);
        Launcher<T, COMP, ST> Lc = L;
        Lc.geo.col_w = choice.cw;
        Lc.geo.rw = choice.rw;
        Lc.geo.plain_col = choice.pc;
        Lc.geo.plain_row = choice.pr;
        Lc.busy = active;
        auto launch_chunk = [&](Emit& em) {
          const bool eager = em.graph == nullptr;
          for (size_t g = 0; g < G; ++g) {
            if (has_col[g]) {
              const ColDesc<T, ST>* dc = dcol.as<ColDesc<T, ST>>() + g * S;
              if (!plain_col[g] || !Lc.col_plain(dc, S, em)) Lc.col(dc, S, em, multi_col[g] != 0);
              if (eager) HIP_CHECK(hipGetLastError());
            }
            if (has_row[g]) {
              const RowDesc<T, ST>* dr = drow.as<RowDesc<T, ST>>() + g * S;
              if (!plain_row[g] || !Lc.row_plain(dr, S, em)) Lc.row(dr, S, em, multi_row[g] != 0);
              if (eager) HIP_CHECK(hipGetLastError());
            }
          }
        };
        Emit eager;
        eager.stream = stream;
        if (uniform && rows_all && trial >= 0) {
          // a timed trial of a tuning candidate
          hipEvent_t a, b;
          HIP_CHECK(hipEventCreate(&a));
          HIP_CHECK(hipEventCreate(&b));
          HIP_CHECK(hipEventRecord(a, stream));
          launch_chunk(eager);
          HIP_CHECK(hipEventRecord(b, stream));
          HIP_CHECK(hipEventSynchronize(b));
          float ms = 0;
          HIP_CHECK(hipEventElapsedTime(&ms, a, b));
          (void)hipEventDestroy(a);
          (void)hipEventDestroy(b);
          tune_report(tk, trial, 1e3 * ms / (static_cast<double>(G) * active));
        } else if (uniform && graphs_ok) {
          // uniform chunks replay one captured graph per shape (descriptors live in dcol/drow)
          const auto gk = std::make_tuple(G, Lc.geo.col_w, Lc.geo.rw, Lc.narrow_mask(S),
                                          Lc.geo.plain_col * 2 + Lc.geo.plain_row + (rows_all ? 0 : 4));
          auto it = graphs.find(gk);
          if (it == graphs.end()) {
            // the graph is built node by node (no stream capture, so other threads of the
            // process may allocate, copy and synchronise meanwhile)
            Emit build;
            hipGraphExec_t exec = nullptr;
            bool ok = hipGraphCreate(&build.graph, 0) == hipSuccess;
            if (ok) {
              try {
                launch_chunk(build);
              } catch (const Error&) {
                ok = false;
              }
              ok = ok && hipGraphInstantiate(&exec, build.graph, nullptr, nullptr, 0) == hipSuccess;
              (void)hipGraphDestroy(build.graph);
            }
            if (!ok) {
              // no graph for this lane: clear the error, run this chunk directly from now on
              (void)hipGetLastError();
              if (exec) (void)hipGraphExecDestroy(exec);
              graphs_ok = false;
              ++lane_graph_counters().fallbacks;
              launch_chunk(eager);
            } else {
              it = graphs.emplace(gk, exec).first;
              ++lane_graph_counters().built;
            }
          }
          if (it != graphs.end()) {
            HIP_CHECK(hipGraphLaunch(it->second, stream));
            ++lane_graph_counters().replays;
          }
        } else {
          launch_chunk(eager);
        }
        for (int i = 0; i < S; ++i) {
          if (!slot[i]) continue;
          if (pc[i] + G >= prog[i].size()) {
            hipEvent_t done;
            HIP_CHECK(hipEventCreateWithFlags(&done, hipEventDisableTiming));
            HIP_CHECK(hipEventRecord(done, stream));
            feed.finished(mem[i], done);
            slot[i] = nullptr;
          } else {
            pc[i] += G;
          }
        }
        hipEvent_t cur;
        HIP_CHECK(hipEventCreateWithFlags(&cur, hipEventDisableTiming));
        HIP_CHECK(hipEventRecord(cur, stream));
        settle();
        prev = cur;
        set ^= 1;
      }
    } catch (...) {
      if (prev) {
        (void)hipEventSynchronize(prev);
        (void)hipEventDestroy(prev);
      }
      (void)hipStreamSynchronize(stream);
      drop_graphs();
      throw;
    }
    (void)hipStreamSynchronize(stream);
    drop_graphs();
  }

  static int precision_code() {
    if (std::is_same<ST, bfp>::value) return 40;
    return sizeof(ST) < sizeof(T) ? 48 : COMP ? 32 : (sizeof(T) == 8 ? 64 : 31);
  }
  static int slot_device(const std::vector<Engine*>& slot) {
    for (const Engine* e : slot)
      if (e) return e->sp_.device;
    return 0;
  }

 private:
  template <typename K, typename... A>
  void launch_1d(K kern, long long n, A... args) {
    launch_1d_on(stream_, kern, n, args...);
  }
  template <typename K, typename... A>
  static void launch_1d_on(hipStream_t s, K kern, long long n, A... args) {
    if (n <= 0) return;
    const int bs = 256;
    const long long grid = (n + bs - 1) / bs;
    hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(grid)), dim3(bs), 0, s, args...);
    HIP_CHECK(hipGetLastError());
  }

  ColDesc<T, ST> col_desc(const Boundary& B, bool forward, bool with_r) {
    ColDesc<T, ST> c{};
    c.in = X_.as<px2<ST>>();
    c.out = prod_.as<px2<ST>>();
    c.prods = prods_.as<Prod<T>>() + (forward ? B.fwd_prod : B.bwd_prod);
    c.nprod = static_cast<int>(B.pairs.size()) + (with_r ? static_cast<int>(B.rpair.size()) : 0);
    c.nt = static_cast<int>(B.pairs.size());
    c.rowuse = B.rowuse_off >= 0 ? rowuse_.as<uint32_t>() + B.rowuse_off : nullptr;
    return c;
  }

  RowDesc<T, ST> row_desc(const Boundary& B, long long q, bool with_r) {
    RowDesc<T, ST> r{};
    r.active = 1;
    r.nt = static_cast<int>(B.pairs.size());
    r.nr = with_r ? static_cast<int>(B.rpair.size()) : 0;
    r.prod = prod_.as<px2<ST>>();
    r.tsel = B.tsel_off >= 0 ? tsel_.as<uint8_t>() + B.tsel_off : nullptr;
    r.rowuse = B.rowuse_off >= 0 ? rowuse_.as<uint32_t>() + B.rowuse_off : nullptr;
    r.mask = mask_.as<T>();
    r.mask_y = mask_y_.as<double>();
    r.mask_x = mask_x_.as<double>();
    if (B.sp_count > 0) {
      const long long list = list_of(q);
      r.rowptr = rowptr_.as<int>() + B.rowptr_off;
      r.sp_pix = sp_pix_.as<int>() + list;
      r.sp_rid = sp_rid_.as<uint8_t>() + list;
    }
    return r;
  }

  px2<T>* sp_vals(DevBuf& b, long long q) { return b.as<px2<T>>() + bnd_[q].sp_off; }

  int keep_slot(long long p) const {
    for (size_t i = 0; i < keep_.size(); ++i)
      if (keep_[i] == p) return static_cast<int>(i);
    return -1;
  }

  int medium_of_label(int32_t l) const { return l < 0 ? before_id_ : label_medium_[l]; }
  const int32_t* plane_labels(long long p) const {
    if (p == 0) return before_.data();
    if (!slab_plane_.empty()) return slab_plane_[p - 1];
    return sp_.uniform_label >= 0 ? uniform_plane_.data() : sp_.labels.data() + (p - 1) * npix_;
  }

  void validate() {
    require(sp_.ny >= 1 && sp_.nx >= 1, "grid must have ny, nx >= 1");
    require(sp_.nz >= 1, "labels must have shape (nz, ny, nx) with nz >= 1");
    // the slab kernels form byte offsets into a plane as unsigned 32-bit values (two complex values
    // per pixel)
    require(sp_.ny * sp_.nx <= (1LL << 32) / (2 * static_cast<long long>(sizeof(cx<T>))),
            "grid of ny·nx = " + std::to_string(sp_.ny * sp_.nx) + " pixels exceeds the plane size limit");
    if (sp_.geom.on) {
      require(sp_.labels.empty() && sp_.uniform_label < 0, "a geometry replaces the label array");
      // the label planes of the slabs, without a (nz, ny, nx) volume (set-up only)
      std::vector<int> plane_of;
      geometry_planes(sp_, geo_planes_, plane_of);
      for (int i : plane_of) slab_plane_.push_back(geo_planes_[i].data());
    }
    const Spec& s = sp_;
    if (s.uniform_label >= 0)
      require(s.labels.empty() && static_cast<size_t>(s.uniform_label) < s.indices.size(),
              "a uniform label must index into `indices` (and no label array be given)");
    else if (slab_plane_.empty())
      require(static_cast<long long>(s.labels.size()) == s.nz * s.ny * s.nx, "labels must have shape (nz, ny, nx)");
    require(!s.synth.on || (s.fluct.empty() && s.synth.sigma >= 0 && s.synth.corr >= 0),
            "a synthesised fluctuation needs sigma >= 0, corr >= 0 and no fluctuation array");
    require(s.mask.empty() || static_cast<long long>(s.mask.size()) == s.ny * s.nx, "mask must have shape (ny, nx)");
    int32_t lo = 0, hi = 0;
    if (!s.labels.empty()) {
      const auto mm = std::minmax_element(s.labels.begin(), s.labels.end());
      lo = *mm.first;
      hi = *mm.second;
    }
    require(lo >= 0 && static_cast<size_t>(hi) < s.indices.size(), "labels must index into `indices`");
    require(s.fluct.empty() || static_cast<long long>(s.fluct.size()) == s.nz * s.ny * s.nx,
            "fluctuation must have the shape of labels (nz, ny, nx)");
    require(s.labels_before.empty() || static_cast<long long>(s.labels_before.size()) == s.ny * s.nx,
            "labels_before must have shape (ny, nx)");
    if (!s.labels_before.empty()) {
      const auto mm = std::minmax_element(s.labels_before.begin(), s.labels_before.end());
      require(*mm.first >= -1 && *mm.second < static_cast<long long>(s.indices.size()),
              "labels must index into `indices`");
    }
  }

  void setup() {
    validate();
    const Spec& s = sp_;
    npix_ = s.ny * s.nx;
    std::vector<c128> media;  // distinct complex index values
    auto id_of = [&](c128 v) {
      for (size_t i = 0; i < media.size(); ++i)
        if (media[i] == v) return static_cast<int>(i);
      media.push_back(v);
      return static_cast<int>(media.size() - 1);
    };
    before_id_ = id_of(s.index_before);
    for (const c128& v : s.indices) label_medium_.push_back(id_of(v));
    media_ = media;
    if (s.labels_before.empty())
      before_.assign(npix_, -1);
    else
      before_ = s.labels_before;
    if (s.uniform_label >= 0) uniform_plane_.assign(npix_, s.uniform_label);

    // Born hosts per slab (validation as the reference constructor, solver.cpp:47-62)
    has_host_.assign(s.nz, 0);
    host_.assign(s.nz, c128(0.0, 0.0));
    if (s.synth.on && s.synth.sigma > 0) {
      // a generated δn fills every slab: each slab must lie in one host index (a label plane
      // shared with the previous slab was checked already)
      const int32_t* checked = nullptr;
      for (long long j = 0; j < s.nz; ++j) {
        const int32_t* lab = plane_labels(j + 1);
        const c128 idx = s.indices[lab[0]];
        if (lab != checked)
          for (long long q = 1; q < npix_; ++q)
            if (s.indices[lab[q]] != idx)
              throw Error(kEinval, "slab " + std::to_string(j) + ": fluctuation must lie in one host index");
        checked = lab;
        has_host_[j] = 1;
        host_[j] = idx;
      }
    } else if (!s.fluct.empty()) {
      for (long long j = 0; j < s.nz; ++j)
        for (long long q = 0; q < npix_; ++q) {
          const c128 d = s.fluct[j * npix_ + q];
          if (d == c128(0.0, 0.0)) continue;
          const c128 idx = s.indices[plane_labels(j + 1)[q]];
          if (has_host_[j] && host_[j] != idx)
            throw Error(kEinval, "slab " + std::to_string(j) + ": fluctuation must lie in one host index");
          has_host_[j] = 1;
          host_[j] = idx;
        }
    }

    // k tables (host, rounded like 2π·numpy.fft.fftfreq)
    std::vector<double> kx(s.nx), ky(s.ny);
    for (long long i = 0; i < s.nx; ++i) kx[i] = fftfreq(i, s.nx, s.dx);
    for (long long j = 0; j < s.ny; ++j) ky[j] = fftfreq(j, s.ny, s.dy);
    kx_ = upload_vec(kx, stream_);
    ky_ = upload_vec(ky, stream_);
    twx_ = std::make_shared<Twiddles<T>>();
    twy_ = std::make_shared<Twiddles<T>>();
    twx_->build(s.nx, stream_);
    twy_->build(s.ny, stream_);

    build_boundaries();
    build_products();
    choose_geometry();
    launch_.ny = s.ny;
    launch_.nx = s.nx;
    launch_.npix = npix_;
    launch_.geo = geo_;
    launch_.twx = twx_;
    launch_.twy = twy_;

    const size_t pb = npix_ * sizeof(px2<T>);
    stage_.alloc(npix_ * 2 * sizeof(c128));
    field_in_.alloc(pb);
    X_.alloc(npix_ * sizeof(px2<ST>));
    prod_.alloc(npix_ * sizeof(px2<ST>) * std::max(1, max_prod_));
    e_.alloc(pb);
    const size_t sp_bytes = std::max<long long>(1, total_sparse_) * sizeof(px2<T>);
    fwd_vals_.alloc(sp_bytes);
    bwd_vals_.alloc(sp_bytes);
    if (!s.mask.empty()) {
      // a mask that is a row profile times a column profile (the reference's absorbing_mask_2d) is
      // applied from the two profiles, if their FP64 product gives every value of the plane bit for
      // bit; the profiles are the plane's centre column and centre row
      std::vector<double> my(s.ny), mx(s.nx);
      for (long long y = 0; y < s.ny; ++y) my[y] = s.mask[y * s.nx + s.nx / 2];
      for (long long x = 0; x < s.nx; ++x) mx[x] = s.mask[(s.ny / 2) * s.nx + x];
      bool separable = true;
      for (long long y = 0; separable && y < s.ny; ++y)
        for (long long x = 0; x < s.nx; ++x)
          if (!(my[y] * mx[x] == s.mask[y * s.nx + x])) {
            separable = false;
            break;
          }
      if (separable) {
        mask_y_ = upload_vec(my, stream_);
        mask_x_ = upload_vec(mx, stream_);
      } else {
        std::vector<T> m(npix_);
        for (long long q = 0; q < npix_; ++q) m[q] = static_cast<T>(s.mask[q]);
        mask_ = upload_vec(m, stream_);
      }
    }
    if (std::any_of(has_host_.begin(), has_host_.end(), [](char c) { return c != 0; }))
      born_ = std::make_unique<Born<T, COMP>>(s, has_host_, host_, kx_.as<double>(), ky_.as<double>(), twx_->plan,
                                              twy_->plan, stream_,
                                              born_tolerance(sizeof(T) == 8, !std::is_same<ST, T>::value));
    HIP_CHECK(hipStreamSynchronize(stream_));
    bytes_ = stage_.bytes() + field_in_.bytes() + X_.bytes() + prod_.bytes() + e_.bytes() + fwd_vals_.bytes() +
             bwd_vals_.bytes() + mask_.bytes() + mask_y_.bytes() + mask_x_.bytes() + ops_bytes_ + tsel_.bytes() + rowuse_.bytes() + sp_pix_.bytes() +
             sp_rid_.bytes() + rowptr_.bytes() + (born_ ? born_->device_bytes() : 0);
    // the label volume and δn now live in the boundary structures and device tables
    sp_.labels = std::vector<int32_t>();
    sp_.fluct = std::vector<c128>();
  }

  void choose_geometry() {
    const Spec& s = sp_;
    Geometry& g = geo_;
    const bool fast_prec = COMP == (sizeof(T) == 4);  // register kernels: FP32 compensated and FP64
    g.col_fast = fast_prec && fast_size<T>(s.ny);
    g.row_fast = fast_prec && fast_size<T>(s.nx);
    if (g.col_fast) {
      const int w = fast_w<T, COMP>(static_cast<int>(s.ny));
      g.col_w = (s.nx % w == 0) ? w : 1;
    }
    if (g.row_fast) {
      const int r = fast_rw<T, COMP>(static_cast<int>(s.nx));
      g.rw = (s.ny % r == 0) ? r : 1;
    }
    if (g.col_fast) g.plan_col = plan_pick<T, COMP, ST>(s.device, precision_code(), s.ny, true, g, stream_);
    if (g.row_fast) g.plan_row = plan_pick<T, COMP, ST>(s.device, precision_code(), s.nx, false, g, stream_);
    const int elems = Budget<T>::elems;
    auto ls_of = [](int lines) { return lines <= 2 ? lines : (lines | 1); };
    if (!g.col_fast) {
      int w = 16;
      while (w > 1 && (s.ny * ls_of(2 * w) > elems || s.ny * ls_of(2 * w) * (long long)sizeof(cx<T>) > kLdsMax))
        w /= 2;
      w = std::min<long long>(w, pow2_floor(static_cast<int>(std::max<long long>(1, s.nx))));
      require(s.ny * ls_of(2 * w) <= elems && s.ny * ls_of(2 * w) * (long long)sizeof(cx<T>) <= kLdsMax,
              "grid size ny = " + std::to_string(s.ny) + " exceeds the column kernel's LDS budget");
      g.cols2 = w;
      g.ls_col2 = ls_of(2 * w);
      while (w > 1 && s.ny * ls_of(2 * w) * (long long)sizeof(cx<T>) > kNarrowLds) w /= 2;
      g.cols2n = w;
      g.ls_col2n = ls_of(2 * w);
      g.lean_col = lean_plan(twy_->plan);
      HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_col_step<T, COMP, ST>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, kLdsMax));
      HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_col_step<T, COMP, ST, true>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, kLdsMax));
    }
    if (!g.row_fast) {
      int rw = static_cast<int>(std::min<long long>(s.ny, 256));
      while (rw > 1 && (s.nx * ls_of(2 * rw) > elems || 2 * s.nx * ls_of(2 * rw) * (long long)sizeof(cx<T>) > kLdsMax))
        rw /= 2;
      require(s.nx * ls_of(2 * rw) <= elems && 2 * s.nx * ls_of(2 * rw) * (long long)sizeof(cx<T>) <= kLdsMax,
              "grid size nx = " + std::to_string(s.nx) + " exceeds the row kernel's LDS budget");
      g.rows = rw;
      g.ls_row = ls_of(2 * rw);
      while (rw > 1 && 2 * s.nx * ls_of(2 * rw) * (long long)sizeof(cx<T>) > kNarrowLds) rw /= 2;
      g.rowsn = rw;
      g.ls_rown = ls_of(2 * rw);
      g.lean_row = lean_plan(twx_->plan);
      HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_row_step<T, COMP, ST>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, kLdsMax));
      HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_row_step<T, COMP, ST, true>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, kLdsMax));
    }
  }

  // Label pairs, tsel maps and sparse interface lists per boundary.
  void build_boundaries() {
    const Spec& s = sp_;
    const long long nz = s.nz;
    bnd_.resize(nz);
    std::vector<uint8_t> tsel;
    std::vector<uint32_t> rowuse;
    std::vector<int32_t> sp_pix, rowptr;
    std::vector<uint8_t> sp_rid;
    total_sparse_ = 0;
    // same[p]: label plane p equals plane p − 1 (one comparison per plane; the same plane object
    // needs none)
    std::vector<char> same(nz + 1, 0);
    for (long long p = 1; p <= nz; ++p)
      same[p] = plane_labels(p) == plane_labels(p - 1) ||
                std::memcmp(plane_labels(p), plane_labels(p - 1), npix_ * sizeof(int32_t)) == 0;
    for (long long q = 0; q < nz; ++q) {
      const int32_t* a = plane_labels(q);
      const int32_t* b = plane_labels(q + 1);
      Boundary& B = bnd_[q];
      // reuse the previous boundary's structure when both label planes repeat
      if (q >= 1 && same[q] && same[q + 1]) {
        const int src = bnd_[q - 1].source >= 0 ? bnd_[q - 1].source : static_cast<int>(q - 1);
        B = bnd_[src];
        B.source = src;
        // sparse values are per boundary: give this boundary its own value slots
        B.sp_off = total_sparse_;
        total_sparse_ += B.sp_count;
        continue;
      }
      // distinct pairs (sorted like the reference's std::set); neighbouring pixels mostly share
      // their pair, so the map is consulted only where the pair changes
      std::map<std::pair<int32_t, int32_t>, int> pid;
      {
        int32_t la = a[0], lb = b[0];
        pid.emplace(std::make_pair(la, lb), 0);
        for (long long i = 1; i < npix_; ++i)
          if (a[i] != la || b[i] != lb) {
            la = a[i];
            lb = b[i];
            pid.emplace(std::make_pair(la, lb), 0);
          }
      }
      require(pid.size() <= 255, "boundary " + std::to_string(q) + " has more than 255 label pairs");
      int k = 0;
      for (auto& kv : pid) {
        kv.second = k++;
        B.pairs.push_back(kv.first);
        B.mleft.push_back(medium_of_label(kv.first.first));
        B.mright.push_back(medium_of_label(kv.first.second));
      }
      std::vector<int> rid_of_pair(B.pairs.size(), -1);
      for (size_t i = 0; i < B.pairs.size(); ++i)
        if (B.mleft[i] != B.mright[i]) {
          rid_of_pair[i] = static_cast<int>(B.rpair.size());
          B.rpair.push_back(static_cast<int>(i));
        }
      if (B.pairs.size() > 1) {
        B.tsel_off = static_cast<long long>(tsel.size());
        tsel.resize(tsel.size() + npix_);
        int32_t la = a[0], lb = b[0];
        uint8_t id = static_cast<uint8_t>(pid.at({la, lb}));
        for (long long i = 0; i < npix_; ++i) {
          if (a[i] != la || b[i] != lb) {
            la = a[i];
            lb = b[i];
            id = static_cast<uint8_t>(pid.at({la, lb}));
          }
          tsel[B.tsel_off + i] = id;
        }
        // pairs each row selects, and whether it holds interface pixels (row_use_bit, kRowUseR)
        B.rowuse_off = static_cast<long long>(rowuse.size());
        rowuse.resize(rowuse.size() + s.ny, 0u);
        long long mixed = 0;
        for (long long y = 0; y < s.ny; ++y) {
          uint32_t u = 0;
          for (long long x = 0; x < s.nx; ++x) {
            const int pi = tsel[B.tsel_off + y * s.nx + x];
            u |= 1u << row_use_bit(pi);
            if (rid_of_pair[pi] >= 0) u |= 1u << kRowUseR;
          }
          rowuse[B.rowuse_off + y] = u;
          if (!single_t(static_cast<int>(u))) ++mixed;
        }
        B.mixed_rows = 4 * mixed > s.ny;
      }
      if (!B.rpair.empty()) {
        B.sp_off = total_sparse_;
        B.rowptr_off = static_cast<long long>(rowptr.size());
        long long cnt = 0;
        for (long long y = 0; y < s.ny; ++y) {
          rowptr.push_back(static_cast<int32_t>(cnt));
          for (long long x = 0; x < s.nx; ++x) {
            const long long i = y * s.nx + x;
            const int pi = B.pairs.size() > 1 ? tsel[B.tsel_off + i] : 0;
            if (rid_of_pair[pi] >= 0) {
              sp_pix.push_back(static_cast<int32_t>(i));
              sp_rid.push_back(static_cast<uint8_t>(rid_of_pair[pi]));
              ++cnt;
            }
          }
        }
        rowptr.push_back(static_cast<int32_t>(cnt));
        B.sp_count = cnt;
        // the list entries of this structure start at this offset of sp_pix / sp_rid
        list_off_[q] = static_cast<long long>(sp_pix.size()) - cnt;
        total_sparse_ += cnt;
      }
    }
    check_boundary_tables(tsel, rowuse, sp_pix, sp_rid, rowptr);
    tsel_ = upload_vec(tsel, stream_);
    rowuse_ = upload_vec(rowuse, stream_);
    sp_pix_ = upload_vec(sp_pix, stream_);
    sp_rid_ = upload_vec(sp_rid, stream_);
    rowptr_ = upload_vec(rowptr, stream_);
  }

  // Every index the slab kernels read through the interface tables lies inside its table: the
  // pair map (tsel, npix entries below the pair count), the per-row use masks (ny entries), the
  // sparse row offsets (ny + 1, from 0 to the entry count, non-decreasing), the sparse pixels (each
  // inside its own row) and their R products (below the R count), and the sparse value slots.
  // Checked on the host for every boundary before the tables go to the device; a violation is an
  // internal error of the set-up, reported instead of launching kernels on it.
  void check_boundary_tables(const std::vector<uint8_t>& tsel, const std::vector<uint32_t>& rowuse,
                             const std::vector<int32_t>& sp_pix, const std::vector<uint8_t>& sp_rid,
                             const std::vector<int32_t>& rowptr) const {
    const long long ny = sp_.ny, nx = sp_.nx;
    auto fail = [](long long q, const std::string& what) {
      throw Error(kEinternal, "interface tables of boundary " + std::to_string(q) + ": " + what);
    };
    if (sp_rid.size() != sp_pix.size()) fail(-1, "sparse pixel and product lists differ in length");
    for (long long q = 0; q < sp_.nz; ++q) {
      const Boundary& B = bnd_[q];
      const long long np = static_cast<long long>(B.pairs.size());
      const long long nr = static_cast<long long>(B.rpair.size());
      if (np < 1 || np > 255) fail(q, "pair count " + std::to_string(np));
      for (int r : B.rpair)
        if (r < 0 || r >= np) fail(q, "R product of pair " + std::to_string(r));
      if (np > 1) {
        if (B.tsel_off < 0 || B.tsel_off + npix_ > static_cast<long long>(tsel.size())) fail(q, "pair map extent");
        if (B.rowuse_off < 0 || B.rowuse_off + ny > static_cast<long long>(rowuse.size())) fail(q, "row masks extent");
        if (B.source < 0)
          for (long long i = 0; i < npix_; ++i)
            if (tsel[B.tsel_off + i] >= np) fail(q, "pair map value at pixel " + std::to_string(i));
      } else if (B.tsel_off >= 0 || B.rowuse_off >= 0) {
        fail(q, "pair map on a boundary of one pair");
      }
      if (nr == 0) {
        if (B.sp_count != 0 || B.rowptr_off >= 0) fail(q, "sparse entries without R products");
        continue;
      }
      if (B.sp_off < 0 || B.sp_off + B.sp_count > total_sparse_) fail(q, "sparse value slots");
      if (B.rowptr_off < 0 || B.rowptr_off + ny + 1 > static_cast<long long>(rowptr.size())) fail(q, "row offsets extent");
      const int32_t* rp = rowptr.data() + B.rowptr_off;
      if (rp[0] != 0 || rp[ny] != B.sp_count) fail(q, "row offsets do not span the entries");
      const long long list = list_of(q);
      if (list < 0 || list + B.sp_count > static_cast<long long>(sp_pix.size())) fail(q, "sparse list extent");
      if (B.source >= 0) continue;  // the tables of a reused structure were checked at their source
      for (long long y = 0; y < ny; ++y) {
        if (rp[y + 1] < rp[y]) fail(q, "row offsets decrease at row " + std::to_string(y));
        for (long long e = rp[y]; e < rp[y + 1]; ++e) {
          const long long pix = sp_pix[list + e];
          if (pix < y * nx || pix >= (y + 1) * nx) fail(q, "sparse pixel outside its row " + std::to_string(y));
          if (sp_rid[list + e] >= nr) fail(q, "sparse entry of R product " + std::to_string(sp_rid[list + e]));
        }
      }
    }
  }

  // Operator tables and per-step product descriptor lists.
  void build_products() {
    const Spec& s = sp_;
    std::vector<Prod<T>> prods;
    max_prod_ = 1;
    for (long long q = 0; q < s.nz; ++q) {
      Boundary& B = bnd_[q];
      // forward step q: T kinds (FwdT) for every pair, then R kinds (FwdR)
      B.fwd_prod = static_cast<int>(prods.size());
      for (size_t i = 0; i < B.pairs.size(); ++i) prods.push_back(t_prod(kFwdT, B.mleft[i], B.mright[i]));
      for (int pi : B.rpair) prods.push_back(matrix_prod(kFwdR, B.mleft[pi], B.mright[pi]));
      // backward step q + 1: r = medium on the right (plane q + 1), s = medium on the left
      B.bwd_prod = static_cast<int>(prods.size());
      for (size_t i = 0; i < B.pairs.size(); ++i) prods.push_back(t_prod(kBwdT, B.mright[i], B.mleft[i]));
      for (int pi : B.rpair) prods.push_back(matrix_prod(kBwdR, B.mright[pi], B.mleft[pi]));
      max_prod_ = std::max<int>(max_prod_, static_cast<int>(B.pairs.size() + B.rpair.size()));
    }
    prods_ = upload_vec(prods, stream_);
  }

  Prod<T> t_prod(int kind, int mr, int ms) {
    if (mr == ms) return phase_prod(mr);
    return matrix_prod(kind, mr, ms);
  }

  TableKey table_key(int kind, int mr, int ms) const {
    TableKey k;
    k.device = sp_.device;
    k.precision = COMP ? 32 : (sizeof(T) == 8 ? 64 : 31);
    k.kind = kind;
    k.ny = sp_.ny;
    k.nx = sp_.nx;
    k.dy = sp_.dy;
    k.dx = sp_.dx;
    k.dz = sp_.dz;
    k.k0 = sp_.k0;
    k.nr_re = media_[mr].real();
    k.nr_im = media_[mr].imag();
    k.ns_re = media_[ms].real();
    k.ns_im = media_[ms].imag();
    return k;
  }

  Prod<T> phase_prod(int m) {
    auto it = phase_ops_.find(m);
    if (it == phase_ops_.end()) {
      auto t = shared_table(table_key(-1, m, m), [&] {
        auto nt = std::make_shared<OpTable>();
        nt->hi.alloc(npix_ * sizeof(cx<T>));
        nt->lo.alloc(npix_ * sizeof(cx<T>));
        const c128 n = media_[m];
        launch_1d(k_build_phase<T>, npix_, kx_.as<double>(), ky_.as<double>(), static_cast<int>(sp_.ny),
                  static_cast<int>(sp_.nx), zd{n.real(), n.imag()}, sp_.k0, sp_.dz,
                  1.0 / static_cast<double>(npix_), nt->hi.as<cx<T>>(), nt->lo.as<cx<T>>());
        HIP_CHECK(hipStreamSynchronize(stream_));
        return nt;
      });
      ops_bytes_ += 2 * npix_ * sizeof(cx<T>);
      it = phase_ops_.emplace(m, std::move(t)).first;
    }
    return Prod<T>{0, it->second->hi.as<cx<T>>(), it->second->lo.as<cx<T>>()};
  }

  Prod<T> matrix_prod(int kind, int mr, int ms) {
    const long long key = (static_cast<long long>(kind) << 40) | (static_cast<long long>(mr) << 20) | ms;
    auto it = matrix_ops_.find(key);
    if (it == matrix_ops_.end()) {
      auto t = shared_table(table_key(kind, mr, ms), [&] {
        auto nt = std::make_shared<OpTable>();
        nt->hi.alloc(3 * npix_ * sizeof(cx<T>));
        nt->lo.alloc(3 * npix_ * sizeof(cx<T>));
        const c128 nr = media_[mr], ns = media_[ms];
        launch_1d(k_build_matrix<T>, npix_, kx_.as<double>(), ky_.as<double>(), static_cast<int>(sp_.ny),
                  static_cast<int>(sp_.nx), kind, zd{nr.real(), nr.imag()}, zd{ns.real(), ns.imag()}, sp_.k0,
                  sp_.dz, 1.0 / static_cast<double>(npix_), nt->hi.as<cx<T>>(), nt->lo.as<cx<T>>());
        HIP_CHECK(hipStreamSynchronize(stream_));
        return nt;
      });
      ops_bytes_ += 6 * npix_ * sizeof(cx<T>);
      it = matrix_ops_.emplace(key, std::move(t)).first;
    }
    return Prod<T>{1, it->second->hi.as<cx<T>>(), it->second->lo.as<cx<T>>()};
  }

  long long list_of(long long q) const {
    const Boundary& B = bnd_[q];
    return list_off_.at(B.source >= 0 ? B.source : q);
  }

  static constexpr long long kLdsMax = 160 * 1024;

  Spec sp_;
  hipStream_t stream_{};
  long long npix_ = 0;
  std::vector<c128> media_;
  std::vector<int> label_medium_;
  int before_id_ = 0;
  std::vector<int32_t> before_;
  std::vector<int32_t> uniform_plane_;  // the label plane of a uniform volume
  // a geometry's distinct label planes and each slab's plane (set-up only; no volume is made)
  std::vector<std::vector<int32_t>> geo_planes_;
  std::vector<const int32_t*> slab_plane_;
  std::vector<char> has_host_;
  std::vector<c128> host_;
  DevBuf kx_, ky_;
  std::shared_ptr<Twiddles<T>> twx_, twy_;
  Launcher<T, COMP, ST> launch_;
  Geometry geo_;
  std::vector<Boundary> bnd_;
  std::unordered_map<long long, long long> list_off_;
  long long total_sparse_ = 0;
  DevBuf tsel_, rowuse_, sp_pix_, sp_rid_, rowptr_;
  std::map<int, std::shared_ptr<OpTable>> phase_ops_;
  std::map<long long, std::shared_ptr<OpTable>> matrix_ops_;
  size_t ops_bytes_ = 0;
  DevBuf prods_;
  int max_prod_ = 1;
  DevBuf stage_, field_in_, X_, prod_, e_, fwd_vals_, bwd_vals_, mask_, mask_y_, mask_x_;
  DevBuf iter_t_, iter_r_, keepf_, keepb_;
  DevBuf lane_col_, lane_row_;
  std::vector<long long> keep_;
  std::unique_ptr<Born<T, COMP>> born_;
  long long last_steps_ = 0;
  int last_iterations_ = 0;
  bool ran_ = false;
  size_t bytes_ = 0;
};

}  // namespace bwpm
