// Host side of the gfx950 pressure solver: level planning, device hierarchy,
// the multigrid cycle program, the PCG driver and the C ABI declared in
// include/hundun_flow.h.  All device work of one plan is issued on the plan's
// HIP stream; the host reads back only the scalars a Krylov decision needs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>
#include <new>
#include <vector>

#include "hf_assembly.hpp"
#include "hf_kernels.hpp"
#include "hf_krylov.hpp"
#include "hf_line.hpp"
#include "hundun_flow.h"

namespace hf {
namespace {

constexpr int kMaxLevels = 32;

// Kernel classes reported by hf_plan_profile.
enum ProfileClass : int {
  P_RB_FINEST = 0,    // red-black colour pass on level 0
  P_RB_COARSE = 1,    // red-black colour pass on grid levels >= 1
  P_DEFECT = 2,       // level stencil / defect
  P_RESTRICT = 3,
  P_PROLONG = 4,
  P_TAIL = 5,         // single-workgroup coarse levels
  P_OPERATOR = 6,     // Krylov A p + dot
  P_UPDATE_XR = 7,    // x += a p, r -= a q, |r|^2
  P_TRUE_RESIDUAL = 8,
  P_MG_ENTRY_EXIT = 9,
  P_PROJECT = 10,
  P_UPDATE_P = 11,
  P_FINALIZE = 12,
  P_OTHER = 13,
  P_SWEEP_FINEST = 14,  // fused two-colour sweep on level 0
  P_SWEEP_COARSE = 15,  // fused two-colour sweep on grid levels >= 1
  P_LINE = 16,          // line relaxation (any level)
  P_RB_FIRST = 17,      // first colour pass, or fused sweep, from x = 0
  P_SWEEP_MEAN = 18,    // fused sweep that also sums the level-0 mean
  kProfileClasses = 19
};
constexpr uint32_t kDetailPlan = 9101U;
constexpr uint32_t kDetailAlloc = 9102U;
constexpr uint32_t kDetailHip = 9103U;
constexpr uint32_t kDetailNumeric = 9104U;
constexpr uint32_t kDetailCoefficients = 9105U;
constexpr uint32_t kDetailBreakdown = 9106U;
constexpr uint32_t kDetailNonFinite = 9107U;
constexpr uint32_t kDetailLine = 9108U;  // line relaxation pivot vanished
constexpr uint32_t kDetailLineOff = 9109U;  // line levels not enabled
constexpr uint32_t kDetailScratch = 9110U;  // beyond the static scratch array
constexpr uint32_t kDetailAssembly = 9111U;  // pressure system inputs refused

hf_status ok() { return {HF_OK, 0U}; }
hf_status fail(int code, uint32_t detail) { return {code, detail}; }

int64_t blocks_for(int64_t work) {
  return std::max<int64_t>(1, (work + kBlock - 1) / kBlock);
}

// The numeric hierarchy of one level.  A plan holds the published set in
// its Level and, once a second refresh arrives, a spare set that the refresh
// builds into and that is swapped in only when the whole hierarchy passed
// its checks (the active/inactive hierarchy storage of src/solver_mg.cpp:
// 4700-4770).
struct Coeffs {
  double volume = 0.0;  // sum of fluid volumes (constant projection)
  // FP64 hierarchy (level 0 doubles as the Krylov operator)
  double* D64 = nullptr;
  double* fx64 = nullptr;
  double* fy64 = nullptr;
  double* fz64 = nullptr;
  double* vol = nullptr;
  uint8_t* act = nullptr;
  // storage-precision copies (alias the FP64 arrays when prec == FP64)
  void* DT = nullptr;
  void* fxT = nullptr;
  void* fyT = nullptr;
  void* fzT = nullptr;
};

struct Level : Coeffs {
  int n[3];
  int64_t cells = 0, nfx = 0, nfy = 0, nfz = 0;
  int coarsen_mask = 0;
  int line_mask = 0;
  int prec = HF_PRECISION_FP64;
  int fused = 1;
  bool tail = false;
  Coeffs spare;  // inactive hierarchy (allocated at the first refresh)
  // workspace slots in storage precision; x and b are the current buffers of
  // a ping-pong pair on levels smoothed by the fused sweep kernel
  void* x = nullptr;
  void* b = nullptr;
  void* r = nullptr;
  void* x2 = nullptr;
  void* b2 = nullptr;
  // line relaxation scratch: Thomas c' and d' (compute precision) and the
  // periodic ghost planes of the widest line axis
  void* lcw = nullptr;
  void* lcd = nullptr;
  void* lghost = nullptr;
  bool cheb = false;   // Chebyshev-Jacobi smoothing (x2 holds x_next)
  bool sweep = false;  // fused two-colour sweep kernel
  int sweep_zc = 0;    // planes per block of the sweep kernel
  int sweep_gx = 0, sweep_gy = 0, sweep_blocks = 0;
  // launch geometry: one tile per block (cells / colour cells) and chunked
  Tiles tc{};
  Tiles tcol{};
  Tiles tr{};
  Tiles tred{};  // kRedPer tiles per block: reduction-only kernels
  Tiles tz{};    // 2-D tiles x z-chunks of t.per planes: z-marching stencils
  Tiles tzp{};   // tz over pairs of cells of a row (k_prolong_z2)
};

constexpr int kFoldSpan = 2048;   // partials per block of the first fold
constexpr int kFoldStride = 4096;  // capacity per accumulator of the fold

int ceil_log2(int v) {
  int l = 0;
  while ((1 << l) < v) ++l;
  return l;
}

// Tile geometry over an nx x ny x nz iteration space.
Tiles make_tiles(int nx, int ny, int nz, bool chunked) {
  Tiles t{};
  t.nx = nx;
  t.ny = ny;
  t.nz = nz;
  t.txl = std::min(6, ceil_log2(std::max(nx, 1)));
  const int ty = kBlock >> t.txl;
  t.gx = (nx + (1 << t.txl) - 1) >> t.txl;
  t.gy = (ny + ty - 1) / ty;
  t.count = t.gx * t.gy * nz;
  // one tile per block: streaming in dispatch order keeps HBM pages hot;
  // reductions fold their per-block partials in a second pass instead
  (void)chunked;
  t.per = 1;
  t.blocks = t.count;
  return t;
}

Chunk make_chunk(int n) {
  Chunk c{};
  c.n = n;
  c.chunk = kBlock;  // one element per thread
  return c;
}

int chunk_blocks(const Chunk& c) {
  return std::max(1, (c.n + c.chunk - 1) / c.chunk);
}

// One step of the host-built cycle program (src/solver_mg.cpp:4192-4338).
struct CycleOp {
  int kind;  // T_SMOOTH, T_RESTRICT, T_PROLONG
  int level;
  int sweeps;
  int reverse;
  int retain;
  int from_zero;  // smoothing starts from x = 0 (first visit of a level)
};

// Executable item: one grid-level op, or a contiguous run of ops on tail
// levels issued as one k_tail launch.
struct ExecItem {
  bool tail = false;
  CycleOp op{};
  int tail_begin = 0;  // index into tail_ops
  int tail_count = 0;
  // a retaining smoothing whose defect and the following restriction run
  // as one kernel (k_defect_restrict); that restriction item is skipped
  bool defect_restrict = false;
  bool skip = false;
};

size_t elem_size(int prec) {
  return prec == HF_PRECISION_FP64 ? 8 : (prec == HF_PRECISION_FP32 ? 4 : 2);
}

template <class F>
void dispatch(int prec, F&& f) {
  if (prec == HF_PRECISION_FP64)
    f(double{});
  else if (prec == HF_PRECISION_FP32)
    f(float{});
  else
    f(bf16{});
}

template <class T>
Lv<T> view(const Level& L) {
  Lv<T> v;
  v.nx = L.n[0];
  v.ny = L.n[1];
  v.nz = L.n[2];
  v.cells = static_cast<int>(L.cells);
  v.D = static_cast<const T*>(L.DT);
  v.fx = static_cast<const T*>(L.fxT);
  v.fy = static_cast<const T*>(L.fyT);
  v.fz = static_cast<const T*>(L.fzT);
  v.act = L.act;
  v.vol = L.vol;
  v.x = static_cast<T*>(L.x);
  v.b = static_cast<T*>(L.b);
  v.r = static_cast<T*>(L.r);
  return v;
}

Lv<double> op_view(const Level& L) {
  Lv<double> v;
  v.nx = L.n[0];
  v.ny = L.n[1];
  v.nz = L.n[2];
  v.cells = static_cast<int>(L.cells);
  v.D = L.D64;
  v.fx = L.fx64;
  v.fy = L.fy64;
  v.fz = L.fz64;
  v.act = L.act;
  v.vol = L.vol;
  v.x = nullptr;
  v.b = nullptr;
  v.r = nullptr;
  return v;
}

}  // namespace
}  // namespace hf

using namespace hf;

namespace {
// Brackets the enclosed launch with a pair of events when profiling is on.
struct ProfScope {
  hf_plan* p;
  explicit ProfScope(hf_plan* plan, int cls);
  ~ProfScope();
};
#define HF_PROF(plan, cls) ProfScope hf_prof_scope_(plan, cls)
}  // namespace

struct hf_plan {
  hf_grid_desc grid{};
  int device = 0;  // HIP device current at creation (owner of the stream)
  hf_policy policy{};
  hipStream_t stream = nullptr;
  int n[3] = {0, 0, 0};
  int64_t cells = 0;
  Bc bc{};
  int periodic_mask = 0;
  std::vector<double> faces[3];
  std::vector<double> widths[3];
  double* d_faces[3] = {nullptr, nullptr, nullptr};
  double* d_widths[3] = {nullptr, nullptr, nullptr};
  std::vector<Level> levels;
  int tail_first = 0;  // first level handled by the tail kernel
  bool row_pairs = true;   // pair kernels on even rows (HF_ROW_PAIRS)
  bool cheb_first = true;  // b-and-D-only first Chebyshev stage (HF_CHEB_FIRST)
  bool krylov_pairs = true;  // FGMRES level-0 kernels over row pairs (HF_KRYLOV_PAIRS)
  std::vector<CycleOp> program;
  std::vector<ExecItem> exec;
  std::vector<TailOp> tail_ops_host;
  TailOp* d_tail_ops = nullptr;
  TailCheb tail_cheb{};  // stage factors of the tail's Chebyshev levels
  TailLevel* d_tail_levels = nullptr;
  size_t tail_lds_bytes = 0;
  bool coefficients_ready = false;
  bool has_activity = false;
  // the published hierarchy applies z = D^-1 r (diagonal shortcut)
  bool shortcut = false;
  unsigned long long* d_ratio = nullptr;  // shortcut selection maximum
  // FGMRES basis and reduction buffers (allocated by the first solve)
  int fg_capacity = 0;
  double* fg_v = nullptr;        // restart + 1 basis vectors
  double* fg_z = nullptr;        // restart preconditioned vectors
  double* fg_partials = nullptr;  // (restart + 2) accumulators
  double* fg_folded = nullptr;
  double* fg_scalars = nullptr;
  double* fg_host = nullptr;      // pinned mirror of fg_scalars
  // raw activity masks (copies) for the adjoint
  uint8_t* fxa = nullptr;
  uint8_t* fya = nullptr;
  uint8_t* fza = nullptr;
  Chunk chunk0{};  // 1-D chunking of level-0 vectors
  // reductions
  double* partials = nullptr;
  int64_t partial_stride = 0;
  double* folded = nullptr;
  double* scalars = nullptr;
  double* host_scalars = nullptr;  // pinned
  int* d_flag = nullptr;
  // Krylov workspace (5 vectors, solve_pcg slots 0..4) + adjoint scratch
  double* kx = nullptr;
  double* kr = nullptr;
  double* kp = nullptr;
  double* kq = nullptr;
  double* adj_rhs = nullptr;
  double* adj_lambda = nullptr;
  uint64_t device_bytes = 0;
  uint64_t reductions = 0;
  // kernel-class timing with HIP events on the plan stream
  bool profiling = false;
  std::vector<hipEvent_t> events;
  std::vector<int> event_class;  // class of the launch starting at 2*i
  size_t event_pairs = 0;
  double class_ms[kProfileClasses] = {};
  uint64_t class_launches[kProfileClasses] = {};
  std::vector<void*> allocations;
  std::vector<size_t> allocation_bytes;
};

namespace {

ProfScope::ProfScope(hf_plan* plan, int cls) : p(plan) {
  if (!p->profiling) return;
  const size_t need = 2 * (p->event_pairs + 1);
  while (p->events.size() < need) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) {
      p->profiling = false;
      return;
    }
    p->events.push_back(e);
  }
  p->event_class.resize(p->event_pairs + 1);
  p->event_class[p->event_pairs] = cls;
  (void)hipEventRecord(p->events[2 * p->event_pairs], p->stream);
}

ProfScope::~ProfScope() {
  if (!p->profiling) return;
  (void)hipEventRecord(p->events[2 * p->event_pairs + 1], p->stream);
  ++p->event_pairs;
}

// Accumulates the recorded launches (the stream must be idle).
void profile_collect(hf_plan* p) {
  if (!p->profiling) return;
  for (size_t i = 0; i < p->event_pairs; ++i) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, p->events[2 * i], p->events[2 * i + 1]) ==
        hipSuccess) {
      p->class_ms[p->event_class[i]] += ms;
      p->class_launches[p->event_class[i]] += 1;
    }
  }
  p->event_pairs = 0;
}

hf_status hip_check(hipError_t e) {
  return e == hipSuccess ? ok() : fail(HF_INVALID_PLAN, kDetailHip);
}

template <class T>
hf_status dalloc(hf_plan* p, T** out, size_t count) {
  void* ptr = nullptr;
  const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  const hipError_t e = hipMalloc(&ptr, bytes);
  if (e != hipSuccess) {
    // the status carries no HIP code; name it once on stderr
    std::fprintf(stderr, "hundun_flow: hipMalloc(%zu bytes) failed: %s\n",
                 bytes, hipGetErrorString(e));
    *out = nullptr;
    return fail(HF_ALLOCATION_FAILURE, kDetailAlloc);
  }
  p->allocations.push_back(ptr);
  p->allocation_bytes.push_back(bytes);
  p->device_bytes += bytes;
  *out = static_cast<T*>(ptr);
  return ok();
}

// Frees one allocation of the plan (null is a no-op).
void dfree(hf_plan* p, void* ptr) {
  if (ptr == nullptr) return;
  for (size_t i = 0; i < p->allocations.size(); ++i) {
    if (p->allocations[i] != ptr) continue;
    (void)hipFree(ptr);
    p->device_bytes -= p->allocation_bytes[i];
    p->allocations.erase(p->allocations.begin() + static_cast<long>(i));
    p->allocation_bytes.erase(p->allocation_bytes.begin() +
                              static_cast<long>(i));
    return;
  }
}

hf_status dalloc_raw(hf_plan* p, void** out, size_t bytes) {
  char* c = nullptr;
  const hf_status s = dalloc(p, &c, bytes);
  *out = c;
  return s;
}

#define HF_TRY(expr)                         \
  do {                                       \
    const hf_status hf_try_s = (expr);       \
    if (hf_try_s.code != HF_OK) return hf_try_s; \
  } while (0)

#define HF_LAUNCH(kernel, grid, block, shared, stream, ...)             \
  hipLaunchKernelGGL(kernel, dim3(static_cast<uint32_t>(grid)),         \
                     dim3(block), shared, stream, __VA_ARGS__)

// Device arrays of one numeric hierarchy set of level L.
hf_status alloc_coeffs(hf_plan* p, const Level& L, Coeffs& k) {
  HF_TRY(dalloc(p, &k.D64, L.cells));
  HF_TRY(dalloc(p, &k.fx64, L.nfx));
  HF_TRY(dalloc(p, &k.fy64, L.nfy));
  HF_TRY(dalloc(p, &k.fz64, L.nfz));
  HF_TRY(dalloc(p, &k.vol, L.cells));
  HF_TRY(dalloc(p, &k.act, L.cells));
  if (L.prec == HF_PRECISION_FP64) {
    k.DT = k.D64;
    k.fxT = k.fx64;
    k.fyT = k.fy64;
    k.fzT = k.fz64;
    return ok();
  }
  const size_t es = elem_size(L.prec);
  HF_TRY(dalloc_raw(p, &k.DT, es * L.cells));
  HF_TRY(dalloc_raw(p, &k.fxT, es * L.nfx));
  HF_TRY(dalloc_raw(p, &k.fyT, es * L.nfy));
  HF_TRY(dalloc_raw(p, &k.fzT, es * L.nfz));
  return ok();
}



// Level planner, src/solver_mg_workspace.cpp:113-143, :181-199, :317-453.
void plan_levels(hf_plan* p) {
  double mn[3], mx[3], rep[3];
  for (int a = 0; a < 3; ++a) {
    mn[a] = INFINITY;
    mx[a] = 0.0;
    rep[a] = 0.0;
    for (double w : p->widths[a]) {
      mn[a] = std::min(mn[a], w);
      mx[a] = std::max(mx[a], w);
      rep[a] += w;
    }
    rep[a] /= static_cast<double>(p->widths[a].size());
  }
  const double largest = std::max({rep[0], rep[1], rep[2]});
  int line_mask = 0;
  for (int a = 0; a < 3; ++a) {
    const bool internal = mn[a] > 0.0 &&
                          mx[a] / mn[a] >= p->policy.anisotropy_threshold;
    const bool cross = rep[a] > 0.0 &&
                       largest / rep[a] >= p->policy.anisotropy_threshold;
    if (internal || cross) line_mask |= 1 << a;
  }
  int preferred;
  if ((line_mask & (line_mask - 1)) != 0)
    preferred = 7;
  else
    preferred = (7 & ~line_mask) == 0 ? 7 : (7 & ~line_mask);
  int g[3] = {p->n[0], p->n[1], p->n[2]};
  p->levels.clear();
  for (int l = 0; l < p->policy.maximum_levels; ++l) {
    Level L;
    std::memcpy(L.n, g, sizeof g);
    L.line_mask = line_mask;
    int next[3] = {g[0], g[1], g[2]};
    int mask = 0;
    for (int a = 0; a < 3; ++a) {
      if ((preferred & (1 << a)) == 0) continue;
      if (g[a] <= p->policy.minimum_coarse_extent) continue;
      next[a] = (g[a] + 1) / 2;
      mask |= 1 << a;
    }
    L.coarsen_mask = mask;
    L.cells = (int64_t)g[0] * g[1] * g[2];
    L.nfx = (int64_t)(g[0] + 1) * g[1] * g[2];
    L.nfy = (int64_t)g[0] * (g[1] + 1) * g[2];
    L.nfz = (int64_t)g[0] * g[1] * (g[2] + 1);
    L.fused = 1;
    for (int a = 0; a < 3; ++a)
      if (p->bc.m[2 * a] == 2 && (g[a] & 1)) L.fused = 0;
    p->levels.push_back(L);
    if (mask == 0 || (int)p->levels.size() == p->policy.maximum_levels) break;
    std::memcpy(g, next, sizeof g);
  }
}

// The fused two-colour sweep (one pass over the level instead of two) runs on
// the finest level; HF_FUSED_SWEEP=0 selects two colour passes there too.
// Coarser levels of fewer than sweep_min_cells() cells have too few tiles per
// plane to hide its plane-step barriers and keep the two-pass route.
bool sweep_enabled() {
  const char* env = std::getenv("HF_FUSED_SWEEP");
  return env == nullptr || std::atoi(env) != 0;
}

// Coarse levels run the fused sweep from this many cells (measured: at
// 512x256x256 the 256x128x128 level gains, the 128x64x64 one does not);
// HF_SWEEP_MIN_CELLS overrides it.
int64_t sweep_min_cells() {
  const char* env = std::getenv("HF_SWEEP_MIN_CELLS");
  if (env != nullptr) {
    const long long v = std::atoll(env);
    if (v >= 0) return v;
  }
  return int64_t{1} << 21;
}

// Planes per block of the fused sweep on every level that takes it (0: the
// planner's choice); HF_SWEEP_ZC sets it.
int sweep_zc_override() {
  const char* env = std::getenv("HF_SWEEP_ZC");
  if (env != nullptr) {
    const int v = std::atoi(env);
    if (v > 0) return v;
  }
  return 0;
}

// The retained defect of a fused-sweep level and its restriction run as one
// kernel where the layout allows (defect_restrict_ok); HF_DEFECT_RESTRICT=0
// keeps the two launches.
bool defect_restrict_enabled() {
  const char* env = std::getenv("HF_DEFECT_RESTRICT");
  return env == nullptr || std::atoi(env) != 0;
}

// HF_ROW_PAIRS=0: the per-cell prolongation and Chebyshev stage on every
// level (the pair kernels are bit for bit the same; the switch lets the
// parity tests compare the two).
bool row_pairs_enabled() {
  const char* env = std::getenv("HF_ROW_PAIRS");
  return env == nullptr || std::atoi(env) != 0;
}

// HF_KRYLOV_PAIRS=0: the FGMRES operator, residual and dot kernels one
// cell per lane on every grid (the pair forms add each lane's two products
// in cell order, so their inner products round differently; the switch
// lets the tests compare the two).
bool krylov_pairs_enabled() {
  const char* env = std::getenv("HF_KRYLOV_PAIRS");
  return env == nullptr || std::atoi(env) != 0;
}

// HF_CHEB_FIRST=0: a from-zero Chebyshev visit starts with the general
// stage over an explicit zero x (stored by the entry and the restriction).
bool cheb_first_enabled() {
  const char* env = std::getenv("HF_CHEB_FIRST");
  return env == nullptr || std::atoi(env) != 0;
}

// Line-relaxation plans are created only with HF_LINE_RELAXATION=1: their
// device path matches the oracle in a CPU build of these kernels but has not
// yet run cleanly on a device, so it stays off by default.
bool line_relaxation_enabled() {
  const char* env = std::getenv("HF_LINE_RELAXATION");
  return env != nullptr && std::atoi(env) != 0;
}


int64_t tail_threshold() {
  const char* env = std::getenv("HF_TAIL_CELLS");
  if (env != nullptr) {
    const long long v = std::atoll(env);
    if (v >= 0) return v;
  }
  return 8192;
}

// Build the cycle program exactly as mg_cycle recurses
// (src/solver_mg.cpp:4192-4338).
void emit_cycle(hf_plan* p, int level, int counter, int from_zero) {
  const int last = static_cast<int>(p->levels.size()) - 1;
  if (level == last) {
    p->program.push_back(
        {T_SMOOTH, level, p->policy.coarse_sweeps, 0, 0, from_zero});
    return;
  }
  p->program.push_back(
      {T_SMOOTH, level, p->policy.pre_sweeps, 0, 1, from_zero});
  p->program.push_back({T_RESTRICT, level, 0, 0, 0, 0});
  emit_cycle(p, level + 1, counter, 1);
  if (counter > 1) emit_cycle(p, level + 1, counter - 1, 0);
  p->program.push_back({T_PROLONG, level, 0, 0, 0, 0});
  p->program.push_back({T_SMOOTH, level, p->policy.post_sweeps, 1, 0, 0});
}

// An op belongs to the tail when every level it touches is a tail level.
bool op_on_tail(const hf_plan* p, const CycleOp& op) {
  const int lo = op.level;
  return lo >= p->tail_first;
}

// k_defect_restrict applies on level l: a fused-sweep or Chebyshev level
// (its retained defect comes from one stencil pass after the smoothing),
// coarsened along every axis with even extents, full z-marching tiles of
// even width and height, even z-chunks.
bool defect_restrict_ok(const hf_plan* p, int l) {
  if (!defect_restrict_enabled() || l + 1 >= static_cast<int>(p->levels.size()))
    return false;
  const Level& L = p->levels[l];
  const Level& Cn = p->levels[l + 1];
  const Tiles& t = L.tz;
  const int tw = 1 << t.txl, th = kBlock >> t.txl;
  return (L.sweep || L.cheb) && L.coarsen_mask == 7 && L.n[0] % 2 == 0 &&
         L.n[1] % 2 == 0 && L.n[2] % 2 == 0 && Cn.n[0] * 2 == L.n[0] &&
         Cn.n[1] * 2 == L.n[1] && Cn.n[2] * 2 == L.n[2] && tw % 2 == 0 &&
         th % 2 == 0 && L.n[0] % tw == 0 && L.n[1] % th == 0 &&
         t.per % 2 == 0;
}

void build_exec(hf_plan* p) {
  p->program.clear();
  p->exec.clear();
  p->tail_ops_host.clear();
  emit_cycle(p, 0, p->policy.f_cycle ? 2 : 1, 1);
  size_t i = 0;
  while (i < p->program.size()) {
    if (!op_on_tail(p, p->program[i])) {
      ExecItem e;
      e.op = p->program[i];
      p->exec.push_back(e);
      ++i;
      continue;
    }
    ExecItem e;
    e.tail = true;
    e.tail_begin = static_cast<int>(p->tail_ops_host.size());
    while (i < p->program.size() && op_on_tail(p, p->program[i])) {
      const CycleOp& op = p->program[i];
      TailOp t{};
      t.kind = static_cast<uint8_t>(op.kind);
      t.level = static_cast<uint8_t>(op.level);
      t.reverse = static_cast<uint8_t>(op.reverse);
      t.retain = static_cast<uint8_t>(op.retain);
      t.sweeps = static_cast<uint16_t>(op.sweeps);
      p->tail_ops_host.push_back(t);
      ++i;
    }
    e.tail_count = static_cast<int>(p->tail_ops_host.size()) - e.tail_begin;
    p->exec.push_back(e);
  }
  for (size_t n = 0; n + 1 < p->exec.size(); ++n) {
    ExecItem& a = p->exec[n];
    ExecItem& b = p->exec[n + 1];
    if (a.tail || b.tail || a.op.kind != T_SMOOTH || !a.op.retain ||
        b.op.kind != T_RESTRICT || b.op.level != a.op.level ||
        !defect_restrict_ok(p, a.op.level))
      continue;
    a.defect_restrict = true;
    b.skip = true;
  }
}

// ---- reductions -----------------------------------------------------------
// acc0: index of the first accumulator read (partials of accumulator a sit
// at p->partials + a * partial_stride).
void finalize(hf_plan* p, int64_t blocks, int nacc, int op, int slot,
              double volume = 0.0, int unit_linear = 1, int acc0 = 0) {
  HF_PROF(p, P_FINALIZE);
  double* first = p->partials + static_cast<int64_t>(acc0) * p->partial_stride;
  const double* src = first;
  int64_t count = blocks;
  int64_t stride = p->partial_stride;
  if (blocks > kFoldSpan) {
    // first pass: fixed spans of partials summed in order (deterministic)
    const int64_t spans = (blocks + kFoldSpan - 1) / kFoldSpan;
    hipLaunchKernelGGL(k_fold_partials, dim3(static_cast<uint32_t>(spans),
                                             static_cast<uint32_t>(nacc)),
                       dim3(kBlock), 0, p->stream, first, blocks,
                       p->partial_stride, p->folded, kFoldStride);
    src = p->folded;
    count = spans;
    stride = kFoldStride;
  }
  HF_LAUNCH(k_finalize, 1, 1024, 0, p->stream, src, count, stride, nacc,
            p->scalars, op, slot, volume, unit_linear);
  ++p->reductions;
}

// Copies the scalar block to the host.  A line relaxation failure since the
// last read is reported (and cleared) as a numerical failure, the deferred
// status of src/solver_mg.cpp:3522-3524.
hf_status read_scalars(hf_plan* p) {
  if (hipMemcpyAsync(p->host_scalars, p->scalars, S_COUNT * sizeof(double),
                     hipMemcpyDeviceToHost, p->stream) != hipSuccess)
    return fail(HF_INVALID_PLAN, kDetailHip);
  if (hipStreamSynchronize(p->stream) != hipSuccess)
    return fail(HF_INVALID_PLAN, kDetailHip);
  if (p->host_scalars[S_LINE_FAIL] != 0.0) {
    p->host_scalars[S_LINE_FAIL] = 0.0;
    if (hipMemcpy(p->scalars + S_LINE_FAIL, p->host_scalars + S_LINE_FAIL,
                  sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      return fail(HF_INVALID_PLAN, kDetailHip);
    return fail(HF_NUMERICAL_FAILURE, kDetailLine);
  }
  return ok();
}

// The scaled 2-norm of a cells-long device vector (global_norm,
// src/solver_krylov.cpp:650-711), taken where a plain sum of squares is not
// finite: *finite is false when an element is not finite (the reference's
// non-finite failure), else *norm is max|v| sqrt(sum (v / max|v|)^2).
hf_status stable_norm(hf_plan* p, const double* v, double* norm,
                      bool* finite) {
  const Chunk ch = p->chunk0;
  const int nb = chunk_blocks(ch);
  HF_LAUNCH(k_absmax, nb, kBlock, 0, p->stream, ch, v, p->partials);
  HF_LAUNCH(k_max_fold, 1, 1024, 0, p->stream,
            static_cast<const double*>(p->partials), static_cast<int64_t>(nb),
            p->scalars, static_cast<int>(S_NSCALE));
  HF_TRY(read_scalars(p));
  const double scale = p->host_scalars[S_NSCALE];
  *norm = 0.0;
  *finite = std::isfinite(scale);
  if (!*finite || scale == 0.0) return ok();
  HF_LAUNCH(k_scaled_squares, nb, kBlock, 0, p->stream, ch, v,
            static_cast<const double*>(p->scalars + S_NSCALE), p->partials);
  finalize(p, nb, 1, F_STORE, S_NSUM);
  HF_TRY(read_scalars(p));
  *norm = scale * std::sqrt(p->host_scalars[S_NSUM]);
  *finite = std::isfinite(*norm);
  return ok();
}

// ---- multigrid ------------------------------------------------------------
Widths widths(const hf_plan* p) {
  return Widths{p->d_widths[0], p->d_widths[1], p->d_widths[2]};
}

// r = b - A x (defect) or r = A x on level l, z-marching.
template <class T>
void level_apply(hf_plan* p, const Level& L, const Lv<T>& v, int defect) {
  HF_LAUNCH((k_stencil_z<T, kZApply>), L.tz.blocks, kBlock, 0, p->stream,
            L.tz, v, p->bc, static_cast<const T*>(v.x),
            static_cast<const T*>(v.b), v.r, static_cast<double*>(nullptr),
            defect);
}

// Mean of a level field for the constant projection: scalars[S_MEAN0 + l].
template <class T>
void level_mean(hf_plan* p, int l, const T* field) {
  Level& L = p->levels[l];
  {
    HF_PROF(p, P_PROJECT);
    if (l == 0)
      HF_LAUNCH(k_vol_dot_finest<T>, L.tred.blocks, kBlock, 0, p->stream,
                L.tred, field, L.act, widths(p), p->partials);
    else
      HF_LAUNCH(k_vol_dot<T>, L.tred.blocks, kBlock, 0, p->stream, L.tred,
                field, static_cast<const double*>(L.vol), p->partials);
  }
  finalize(p, L.tred.blocks, 1, F_MEAN, S_MEAN0 + l, L.volume);
}

template <class T>
void project_apply(hf_plan* p, int l, T* field) {
  Level& L = p->levels[l];
  HF_PROF(p, P_PROJECT);
  HF_LAUNCH(k_project_apply<T>, L.tc.blocks, kBlock, 0, p->stream, L.tc,
            field, static_cast<const uint8_t*>(L.act),
            static_cast<const double*>(p->scalars + S_MEAN0 + l));
}

// mean_partials: the last sweep also writes the per-block partial sums of the
// level's constant-projection mean of its output (k_rb_sweep with kMean).
template <class T>
void sweep_smooth(hf_plan* p, int l, int sweeps, int reverse, int retain,
                  int from_zero, double* mean_partials = nullptr) {
  Level& L = p->levels[l];
  const int rb_class = l == 0 ? P_SWEEP_FINEST : P_SWEEP_COARSE;
  const double* mean = (from_zero && p->policy.null_space_constant)
                           ? p->scalars + S_MEAN0 + l
                           : nullptr;
  for (int sweep = 0; sweep < sweeps; ++sweep) {
    const bool zero = from_zero && sweep == 0;
    SweepArgs<T> a;
    a.L = view<T>(L);
    a.bc = p->bc;
    a.xout = static_cast<T*>(L.x2);
    a.bout = (zero && mean != nullptr) ? static_cast<T*>(L.b2) : nullptr;
    a.mean = zero ? mean : nullptr;
    a.first = reverse ? 1 : 0;
    a.from_zero = zero ? 1 : 0;
    a.zc = L.sweep_zc;
    a.gx = L.sweep_gx;
    a.gy = L.sweep_gy;
    const bool with_mean = mean_partials != nullptr && sweep == sweeps - 1;
    a.partials = with_mean ? mean_partials : nullptr;
    a.wx = p->d_widths[0];
    a.wy = p->d_widths[1];
    a.wz = p->d_widths[2];
    {
      HF_PROF(p, zero ? P_RB_FIRST : with_mean ? P_SWEEP_MEAN : rb_class);
      if (zero)
        HF_LAUNCH((k_rb_sweep<T, true>), L.sweep_blocks, kSwThreads, 0,
                  p->stream, a);
      else if (with_mean)
        HF_LAUNCH((k_rb_sweep<T, false, true>), L.sweep_blocks, kSwThreads, 0,
                  p->stream, a);
      else
        HF_LAUNCH((k_rb_sweep<T, false>), L.sweep_blocks, kSwThreads, 0,
                  p->stream, a);
    }
    std::swap(L.x, L.x2);
    if (a.bout != nullptr) std::swap(L.b, L.b2);
  }
  if (retain) {
    HF_PROF(p, P_DEFECT);
    level_apply<T>(p, L, view<T>(L), 1);
  }
}

template <class T>
void point_smooth(hf_plan* p, int l, int sweeps, int reverse, int retain,
                  int from_zero, double* mean_partials);

// Line geometry of axis ax on level L: lines (a, b), E cells each.
template <class T>
LineArgs<T> line_args(hf_plan* p, Level& L, int ax) {
  using C = typename Compute<T>::type;
  LineArgs<T> g{};
  g.L = view<T>(L);
  g.modes = 0;
  for (int f = 0; f < 6; ++f) g.modes |= (p->bc.m[f] & 3) << (2 * f);
  g.ax = ax;
  g.A = ax == 0 ? L.n[1] : L.n[0];
  g.B = ax == 2 ? L.n[1] : L.n[2];
  g.E = L.n[ax];
  g.cw = static_cast<C*>(L.lcw);
  g.cd = static_cast<C*>(L.lcd);
  const int a_axis = ax == 0 ? 1 : 0;
  const int b_axis = ax == 2 ? 1 : 2;
  T* ghost = static_cast<T*>(L.lghost);
  g.ghost_a = (p->periodic_mask >> a_axis) & 1 ? ghost : nullptr;
  g.ghost_b = (p->periodic_mask >> b_axis) & 1
                  ? ghost + (g.ghost_a != nullptr ? 2 * g.B * g.E : 0)
                  : nullptr;
  g.flag = p->scalars + S_LINE_FAIL;
  return g;
}

// smooth() on a line level (src/solver_mg.cpp:3540-3590): each line axis in
// x, y, z order (reverse: z, y, x), all sweeps of one axis before the next.
template <class T>
void line_smooth(hf_plan* p, int l, int sweeps, int reverse) {
  Level& L = p->levels[l];
  for (int i = 0; i < 3; ++i) {
    const int ax = reverse ? 2 - i : i;
    if ((L.line_mask & (1 << ax)) == 0) continue;
    if (L.n[ax] > kLineMaximumExtent) {
      point_smooth<T>(p, l, sweeps, reverse, 0, 0, nullptr);
      continue;
    }
    const LineArgs<T> g = line_args<T>(p, L, ax);
    HF_PROF(p, P_LINE);
    if (p->policy.line_order == 0) {
      HF_LAUNCH(k_line_lex<T>, 1, 1024, 0, p->stream, g, sweeps, reverse);
      continue;
    }
    const int ghosts = (g.ghost_a != nullptr ? 2 * g.B * g.E : 0) +
                       (g.ghost_b != nullptr ? 2 * g.A * g.E : 0);
    const int lines = (g.A + 1) / 2 * g.B;
    for (int sweep = 0; sweep < sweeps; ++sweep) {
      for (int pass = 0; pass < 2; ++pass) {
        const int colour = ((reverse ? 1 : 0) + pass) & 1;
        if (ghosts > 0)
          HF_LAUNCH(k_line_ghost<T>, blocks_for(ghosts), kBlock, 0, p->stream,
                    g);
        HF_LAUNCH(k_line_zebra<T>, blocks_for(lines), kBlock, 0, p->stream, g,
                  colour);
      }
    }
  }
}

// Stage factors of Chebyshev-Jacobi smoothing over the spectrum
// [2 fraction, 2] of D^-1 A (src/solver_mg.cpp:2916-3090): stage s moves x
// by d = pf d + rf (b - A x) / D, from pf = 0, rf = 1 / theta at s = 0 and
// the three-term recurrence of rho after it.  One sequence serves the grid
// levels and the tail's table, so both apply the same values.
struct ChebStage {
  double theta, delta, sigma, previous_rho;
  int stage = 0;
  explicit ChebStage(double fraction) {
    const double upper = 2.0;
    const double lower = fraction * upper;
    theta = (upper + lower) * 0.5;
    delta = (upper - lower) * 0.5;
    sigma = theta / delta;
    previous_rho = 1.0 / sigma;
  }
  void next(double* rf, double* pf) {
    const double rho = stage == 0 ? 0.0 : 1.0 / (2.0 * sigma - previous_rho);
    *pf = stage == 0 ? 0.0 : rho * previous_rho;
    *rf = stage == 0 ? 1.0 / theta : 2.0 * rho / delta;
    if (stage != 0) previous_rho = rho;
    ++stage;
  }
};

// Chebyshev-Jacobi smoothing of level l (src/solver_mg.cpp:2916-3090):
// `stages` Jacobi stages with the three-term recurrence over the spectrum
// [2 fraction, 2] of D^-1 A, from the x the level holds (zero on a first
// visit: the entry or the restriction stored it), then the retained defect.
template <class T>
void cheb_smooth(hf_plan* p, int l, int stages, int retain, int from_zero) {
  using C = typename Compute<T>::type;
  Level& L = p->levels[l];
  ChebStage stage_factors(p->policy.chebyshev_lower_fraction);
  const int rb_class = l == 0 ? P_RB_FINEST : P_RB_COARSE;
  for (int stage = 0; stage < stages; ++stage) {
    double rf = 0.0, pf = 0.0;
    stage_factors.next(&rf, &pf);
    // the last stage's direction is read by no later stage
    T* dout = stage + 1 < stages ? static_cast<T*>(L.r) : nullptr;
    if (stage == 0 && from_zero && p->cheb_first) {
      // x = 0 (stored by the entry or the restriction): b and D suffice
      HF_PROF(p, rb_class);
      const int64_t groups = (L.cells + 3) / 4;
      HF_LAUNCH(k_cheb_first<T>, (groups + kBlock - 1) / kBlock, kBlock, 0,
                p->stream, static_cast<int>(L.cells), view<T>(L),
                static_cast<T*>(L.x2), static_cast<C>(rf), dout);
    } else if (p->row_pairs && L.n[0] % 2 == 0) {
      HF_PROF(p, rb_class);
      HF_LAUNCH(k_cheb_stage2<T>, L.tzp.blocks, kBlock, 0, p->stream, L.tzp,
                view<T>(L), p->bc, static_cast<const T*>(L.x),
                static_cast<T*>(L.x2), static_cast<C>(rf), static_cast<C>(pf),
                stage == 0 ? 1 : 0, dout);
    } else {
      HF_PROF(p, rb_class);
      HF_LAUNCH(k_cheb_stage<T>, L.tz.blocks, kBlock, 0, p->stream, L.tz,
                view<T>(L), p->bc, static_cast<const T*>(L.x),
                static_cast<T*>(L.x2), static_cast<C>(rf), static_cast<C>(pf),
                stage == 0 ? 1 : 0, dout);
    }
    std::swap(L.x, L.x2);
  }
  if (retain) {
    HF_PROF(p, P_DEFECT);
    level_apply<T>(p, L, view<T>(L), 1);
  }
}

template <class T>
void grid_smooth(hf_plan* p, int l, int sweeps, int reverse, int retain,
                 int from_zero, double* mean_partials = nullptr) {
  Level& L = p->levels[l];
  if (L.cheb) {
    cheb_smooth<T>(p, l, sweeps, retain, from_zero);
    return;
  }
  if (L.line_mask != 0) {
    // x starts from an explicit zero on line levels; the retained defect is
    // recomputed after line smoothing (compute_residual, :4256-4262)
    line_smooth<T>(p, l, sweeps, reverse);
    if (retain) {
      HF_PROF(p, P_DEFECT);
      level_apply<T>(p, L, view<T>(L), 1);
    }
    return;
  }
  point_smooth<T>(p, l, sweeps, reverse, retain, from_zero, mean_partials);
}

template <class T>
void point_smooth(hf_plan* p, int l, int sweeps, int reverse, int retain,
                  int from_zero, double* mean_partials) {
  Level& L = p->levels[l];
  if (L.sweep) {
    sweep_smooth<T>(p, l, sweeps, reverse, retain, from_zero, mean_partials);
    return;
  }
  const Lv<T> v = view<T>(L);
  const int rb_class = l == 0 ? P_RB_FINEST : P_RB_COARSE;
  // a pending constant projection of b is folded into the first pass
  const double* mean = (from_zero && p->policy.null_space_constant)
                           ? p->scalars + S_MEAN0 + l
                           : nullptr;
  if (!L.fused && !from_zero) {
    HF_PROF(p, P_DEFECT);
    level_apply<T>(p, L, v, 0);
  }
  for (int sweep = 0; sweep < sweeps; ++sweep) {
    const int first = reverse ? 1 : 0;
    for (int pass = 0; pass < 2; ++pass) {
      const int color = (first + pass) & 1;
      const bool zero_pass = from_zero && sweep == 0 && pass == 0;
      {
        HF_PROF(p, zero_pass ? P_RB_FIRST : rb_class);
        if (zero_pass)
          HF_LAUNCH(k_rb_first<T>, L.tc.blocks, kBlock, 0, p->stream, L.tc, v,
                    mean);
        else if (L.fused)
          HF_LAUNCH(k_rb<T>, L.tcol.blocks, kBlock, 0, p->stream, L.tcol, v,
                    p->bc, color);
        else
          HF_LAUNCH(k_rb_separated<T>, L.tcol.blocks, kBlock, 0, p->stream,
                    L.tcol, v, color);
      }
      const bool final_update = sweep + 1 == sweeps && pass == 1;
      if (!final_update || retain) {
        if (!L.fused || (final_update && retain)) {
          HF_PROF(p, P_DEFECT);
          level_apply<T>(p, L, v, (retain && final_update) ? 1 : 0);
        }
      }
    }
  }
}

void exec_grid_op(hf_plan* p, const CycleOp& op,
                  double* mean_partials = nullptr,
                  bool defect_restrict = false) {
  Level& L = p->levels[op.level];
  if (op.kind == T_SMOOTH) {
    // a one-stage Chebyshev smoothing from zero before the fused defect and
    // restriction: that kernel forms the stage's x from b and D itself.
    // Each cell then divides five times instead of once: measured, this
    // pays on FP64 levels and on launch-bound small levels, while a large
    // FP32 level runs faster as two passes (at 512x256x256: about 350 us
    // fused against 79 + 210 us).
    const bool first_stage =
        defect_restrict && L.cheb && op.from_zero && op.sweeps == 1 &&
        p->cheb_first &&
        (L.prec == HF_PRECISION_FP64 || L.cells < (int64_t{1} << 20));
    if (!first_stage) {
      dispatch(L.prec, [&](auto t) {
        using T = decltype(t);
        grid_smooth<T>(p, op.level, op.sweeps, op.reverse,
                       defect_restrict ? 0 : op.retain, op.from_zero,
                       mean_partials);
      });
    }
    if (!defect_restrict) return;
    // the retained defect and the restriction of the next op, in one pass
    Level& Cd = p->levels[op.level + 1];
    const bool proj = p->policy.null_space_constant != 0;
    const bool into_tail = Cd.tail || Cd.line_mask != 0 || Cd.cheb;
    dispatch(L.prec, [&](auto tf) {
      using Tf = decltype(tf);
      dispatch(Cd.prec, [&](auto tc) {
        using Tc = decltype(tc);
        if (first_stage) {
          using C = typename Compute<Tf>::type;
          double rf0 = 0.0, pf0 = 0.0;
          ChebStage(p->policy.chebyshev_lower_fraction).next(&rf0, &pf0);
          HF_PROF(p, P_DEFECT);
          HF_LAUNCH((k_defect_restrict<Tf, Tc, true>), L.tz.blocks, kBlock, 0,
                    p->stream, L.tz, view<Tf>(L), p->bc, view<Tc>(Cd),
                    into_tail ? 1 : 0,
                    proj ? p->partials : static_cast<double*>(nullptr),
                    static_cast<C>(rf0), static_cast<Tf*>(L.x2));
          std::swap(L.x, L.x2);
        } else {
          HF_PROF(p, P_DEFECT);
          HF_LAUNCH((k_defect_restrict<Tf, Tc>), L.tz.blocks, kBlock, 0,
                    p->stream, L.tz, view<Tf>(L), p->bc, view<Tc>(Cd),
                    into_tail ? 1 : 0,
                    proj ? p->partials : static_cast<double*>(nullptr));
        }
        if (proj) {
          finalize(p, L.tz.blocks, 1, F_MEAN, S_MEAN0 + op.level + 1,
                   Cd.volume);
          if (into_tail)
            project_apply<Tc>(p, op.level + 1, static_cast<Tc*>(Cd.b));
        }
      });
    });
    return;
  }
  Level& Cn = p->levels[op.level + 1];
  if (op.kind == T_RESTRICT) {
    const bool proj = p->policy.null_space_constant != 0;
    // a point-smoothed grid level starts from zero and applies the
    // projection in its first pass; the tail and line levels need x = 0 and
    // a projected b up front
    const bool into_tail = Cn.tail || Cn.line_mask != 0 || Cn.cheb;
    dispatch(L.prec, [&](auto tf) {
      using Tf = decltype(tf);
      dispatch(Cn.prec, [&](auto tc) {
        using Tc = decltype(tc);
        {
          HF_PROF(p, P_RESTRICT);
          HF_LAUNCH((k_restrict<Tf, Tc>), Cn.tr.blocks, kBlock, 0, p->stream,
                    Cn.tr, view<Tf>(L), view<Tc>(Cn), L.coarsen_mask,
                    into_tail ? 1 : 0,
                    proj ? p->partials : static_cast<double*>(nullptr));
        }
        if (proj) {
          finalize(p, Cn.tr.blocks, 1, F_MEAN, S_MEAN0 + op.level + 1,
                   Cn.volume);
          if (into_tail) project_apply<Tc>(p, op.level + 1, static_cast<Tc*>(Cn.b));
        }
      });
    });
    return;
  }
  dispatch(Cn.prec, [&](auto tc) {
    using Tc = decltype(tc);
    dispatch(L.prec, [&](auto tf) {
      using Tf = decltype(tf);
      HF_PROF(p, P_PROLONG);
      if (p->row_pairs && L.n[0] % 2 == 0)
        HF_LAUNCH((k_prolong_z2<Tc, Tf>), L.tzp.blocks, kBlock, 0, p->stream,
                  L.tzp, view<Tc>(Cn), view<Tf>(L), L.coarsen_mask,
                  p->periodic_mask);
      else
        HF_LAUNCH((k_prolong_z<Tc, Tf>), L.tz.blocks, kBlock, 0, p->stream,
                  L.tz, view<Tc>(Cn), view<Tf>(L), L.coarsen_mask,
                  p->periodic_mask);
    });
  });
}

// Whether the cycle ends with a fused level-0 sweep that can sum the mean of
// its output: the last op smooths level 0 and its last sweep is not the
// from-zero one.
bool mean_in_sweep(const hf_plan* p) {
  if (p->exec.empty() || p->exec.back().tail) return false;
  const CycleOp& op = p->exec.back().op;
  const Level& L0 = p->levels[0];
  return op.kind == T_SMOOTH && op.level == 0 && L0.sweep && op.sweeps >= 1 &&
         !(op.sweeps == 1 && op.from_zero);
}

// mean_partials: passed to the cycle's final op (see mean_in_sweep).
void run_cycle(hf_plan* p, double* mean_partials = nullptr) {
  for (size_t n = 0; n < p->exec.size(); ++n) {
    const ExecItem& e = p->exec[n];
    if (e.skip) continue;
    if (!e.tail) {
      exec_grid_op(p, e.op, n + 1 == p->exec.size() ? mean_partials : nullptr,
                   e.defect_restrict);
      continue;
    }
    HF_PROF(p, P_TAIL);
    HF_LAUNCH(k_tail, 1, 1024, p->tail_lds_bytes, p->stream, p->d_tail_levels,
              p->tail_first,
              p->d_tail_ops + e.tail_begin, e.tail_count, p->bc,
              p->periodic_mask, static_cast<int>(p->policy.null_space_constant),
              p->tail_cheb);
  }
}

// z = M r on device arrays; optional fused Krylov dots (r.z, z.q).
// entered: b of level 0 (and, with the null space, the partial sums of its
// mean in accumulator 1) were stored by the PCG update (k_update_xr_entry);
// only when level 0 is a grid level (not the tail, not line relaxed).
// The operands from which the exit forms z (level-0 candidate, projection
// mean, scale), for a consumer that forms z itself (k_update_p_exit).
struct ExitOperands {
  const void* x = nullptr;
  const double* mean = nullptr;
  const double* alpha = nullptr;
};

// z == nullptr: the exit stores no z (it still forms the dots), and `exit`
// receives the operands z is formed from.
hf_status mg_apply(hf_plan* p, const double* r, double* z, const double* q,
                   bool dots, bool entered = false,
                   ExitOperands* exit = nullptr, int64_t entry_blocks = 0) {
  Level& L0 = p->levels[0];
  const bool proj = p->policy.null_space_constant != 0;
  // whole cycle inside the tail kernel, or a line level 0: x = 0 and the
  // projected b are stored before the cycle
  const bool tail0 = L0.tail || L0.line_mask != 0 || L0.cheb;
  // (a grid Chebyshev level 0 starts with a from-zero stage, which reads
  // b and D only: its x needs no zero)
  const bool zero_x =
      L0.tail || L0.line_mask != 0 || (L0.cheb && !p->cheb_first);
  dispatch(L0.prec, [&](auto t) {
    using T = decltype(t);
    T* b = static_cast<T*>(L0.b);
    T* x = static_cast<T*>(L0.x);
    if (!entered) {
      HF_PROF(p, P_MG_ENTRY_EXIT);
      HF_LAUNCH(k_mg_entry<T>, L0.tred.blocks, kBlock, 0, p->stream, L0.tred, r,
                static_cast<const uint8_t*>(L0.act), widths(p), b,
                zero_x ? x : static_cast<T*>(nullptr),
                proj ? p->partials : static_cast<double*>(nullptr));
    }
    if (proj) {
      // an entry made by the caller leaves its partials in accumulator 1
      // (entry_blocks of them when not in the reduction layout)
      finalize(p, entered && entry_blocks > 0 ? entry_blocks : L0.tred.blocks,
               1, F_MEAN, S_MEAN0, L0.volume, 1, entered ? 1 : 0);
      if (tail0) project_apply<T>(p, 0, b);
    }
  });
  // the cycle's last fused sweep also sums the level-0 mean of its output
  const bool sweep_mean = proj && !p->shortcut && mean_in_sweep(p);
  if (p->shortcut) {
    // z = D^-1 r (src/solver_mg.cpp:4926-4936); the shortcut is only
    // selected without a null space
    const Chunk ch = make_chunk(static_cast<int>(L0.cells));
    dispatch(L0.prec, [&](auto t) {
      using T = decltype(t);
      HF_PROF(p, P_OTHER);
      HF_LAUNCH(k_diag_solve<T>, chunk_blocks(ch), kBlock, 0, p->stream, ch,
                view<T>(L0));
    });
  } else {
    run_cycle(p, sweep_mean ? p->partials : nullptr);
  }
  const double* alpha = nullptr;
  const double* mean = nullptr;
  dispatch(L0.prec, [&](auto t) {
    using T = decltype(t);
    T* x = static_cast<T*>(L0.x);
    if (proj) {
      if (sweep_mean)
        finalize(p, L0.sweep_blocks, 1, F_MEAN, S_MEAN0, L0.volume);
      else
        level_mean<T>(p, 0, x);
      mean = p->scalars + S_MEAN0;
    }
    if (!p->policy.unit_linear) {
      // residual-minimising scaling needs the projected candidate stored
      if (proj) {
        project_apply<T>(p, 0, x);
        mean = nullptr;
      }
      {
        HF_PROF(p, P_DEFECT);
        level_apply<T>(p, L0, view<T>(L0), 1);
      }
      {
        HF_PROF(p, P_MG_ENTRY_EXIT);
        HF_LAUNCH(k_mg_proj3<T>, L0.tred.blocks, kBlock, 0, p->stream, L0.tred, r,
                  static_cast<const uint8_t*>(L0.act),
                  static_cast<const T*>(L0.r), p->partials,
                  static_cast<int>(p->partial_stride));
      }
      finalize(p, L0.tred.blocks, 3, F_MGALPHA, 0, 0.0, 0);
      alpha = p->scalars + S_MGALPHA;
    }
    HF_PROF(p, P_MG_ENTRY_EXIT);
    HF_LAUNCH(k_mg_exit<T>, L0.tred.blocks, kBlock, 0, p->stream, L0.tred,
              static_cast<const T*>(x), static_cast<const uint8_t*>(L0.act),
              mean, alpha, z, dots ? r : static_cast<const double*>(nullptr),
              q, dots ? p->partials : static_cast<double*>(nullptr),
              static_cast<int>(p->partial_stride));
    if (exit != nullptr) {
      exit->x = x;
      exit->mean = mean;
      exit->alpha = alpha;
    }
  });
  return ok();
}

hf_status upload_tail(hf_plan* p) {
  const int L = static_cast<int>(p->levels.size());
  const int count = L - p->tail_first;
  if (count <= 0) return ok();
  std::vector<TailLevel> host(count);
  int64_t lds = 0;
  for (int l = p->tail_first; l < L; ++l) {
    TailLevel& t = host[l - p->tail_first];
    t.L = view<double>(p->levels[l]);
    t.coarsen_mask = p->levels[l].coarsen_mask;
    t.fused = p->levels[l].fused;
    t.cheb = p->levels[l].cheb ? 1 : 0;
    t.volume = p->levels[l].volume;
    t.lds = static_cast<int>(lds);
    lds += 3 * p->levels[l].cells;
  }
  if (hipMemcpyAsync(p->d_tail_levels, host.data(),
                     count * sizeof(TailLevel), hipMemcpyHostToDevice,
                     p->stream) != hipSuccess)
    return fail(HF_INVALID_PLAN, kDetailHip);
  if (!p->tail_ops_host.empty() &&
      hipMemcpyAsync(p->d_tail_ops, p->tail_ops_host.data(),
                     p->tail_ops_host.size() * sizeof(TailOp),
                     hipMemcpyHostToDevice, p->stream) != hipSuccess)
    return fail(HF_INVALID_PLAN, kDetailHip);
  return hip_check(hipStreamSynchronize(p->stream));
}

AxisGeom axis_geom(const hf_plan* p) {
  AxisGeom g;
  for (int a = 0; a < 3; ++a) {
    g.faces[a] = p->d_faces[a];
    g.n[a] = p->n[a];
  }
  for (int f = 0; f < 6; ++f) g.kind[f] = static_cast<int8_t>(p->grid.boundary[f]);
  return g;
}

Lv<double> krylov_view(const hf_plan* p) { return op_view(p->levels[0]); }

}  // namespace

// Reduction scratch of hf_pcg_solve_static: the per-block partials, the
// folded partials and the FP64 scalar block of one solve, in one array that
// each GPU holds for the life of the process.  In host code `scratch` is a
// stand-in for the symbol, not a device address: kernels get the address
// that hipGetSymbolAddress returned for the device (static_scratch below).
constexpr size_t kStaticScratchBytes = size_t{8} << 20;
static __device__ __align__(256) double
    scratch[kStaticScratchBytes / sizeof(double)];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch[kMaxDevices] = {};

// Device address of `scratch` on `device`, looked up once per device with
// that device current (each GPU holds its own copy at its own address).
hf_status static_scratch(int device, double** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices)
    return fail(HF_INVALID_PLAN, kDetailScratch);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch[device] == nullptr) {
    int current = 0;
    if (hipGetDevice(&current) != hipSuccess)
      return fail(HF_INVALID_PLAN, kDetailHip);
    if (current != device && hipSetDevice(device) != hipSuccess)
      return fail(HF_INVALID_PLAN, kDetailHip);
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (current != device) (void)hipSetDevice(current);
    if (e != hipSuccess || ptr == nullptr)
      return fail(HF_INVALID_PLAN, kDetailHip);
    g_scratch[device] = static_cast<double*>(ptr);
  }
  *out = g_scratch[device];
  return ok();
}

// Doubles of reduction scratch one solve of plan p uses.
int64_t reduction_doubles(const hf_plan* p) {
  return 3 * p->partial_stride + 3 * int64_t{kFoldStride} + S_COUNT;
}

}  // namespace

// ---- FGMRES (src/solver_krylov.cpp:1557-2650) ------------------------------
namespace {

// Slot of fg_scalars that a non-finite basis normalisation sets.
constexpr int kFgFlag = kFgMaxRestart + 2;
constexpr int kFgSlots = kFgMaxRestart + 4;

// Basis, preconditioned vectors and reduction buffers for restart R, kept
// by the plan; a longer restart replaces them.
hf_status fg_reserve(hf_plan* p, int restart) {
  if (p->fg_capacity >= restart) return ok();
  dfree(p, p->fg_v);
  dfree(p, p->fg_z);
  dfree(p, p->fg_partials);
  dfree(p, p->fg_folded);
  p->fg_v = p->fg_z = p->fg_partials = p->fg_folded = nullptr;
  p->fg_capacity = 0;
  const int64_t N = p->cells;
  const int64_t blocks = p->levels[0].tz.blocks;
  HF_TRY(dalloc(p, &p->fg_v, static_cast<size_t>((restart + 1) * N)));
  HF_TRY(dalloc(p, &p->fg_z, static_cast<size_t>(restart * N)));
  HF_TRY(dalloc(p, &p->fg_partials,
                static_cast<size_t>((restart + 2) * blocks)));
  HF_TRY(dalloc(p, &p->fg_folded,
                static_cast<size_t>((restart + 2) * int64_t{kFoldStride})));
  if (p->fg_scalars == nullptr) HF_TRY(dalloc(p, &p->fg_scalars, kFgSlots));
  if (p->fg_host == nullptr &&
      hipHostMalloc(reinterpret_cast<void**>(&p->fg_host),
                    kFgSlots * sizeof(double), hipHostMallocDefault) !=
          hipSuccess) {
    p->fg_host = nullptr;
    return fail(HF_ALLOCATION_FAILURE, kDetailAlloc);
  }
  p->fg_capacity = restart;
  return ok();
}

// single_reduction_fgmres_norm (src/solver_krylov.cpp:565-604): the new
// vector's norm from bar = w.w less the squared projections; 0 ok, 1 the
// recurrence went negative (unsafe), -1 non-finite.
int fg_norm_recurrence(const double* h, size_t count, double bar, double* t,
                       double* next) {
  if (!std::isfinite(bar)) return -1;
  double v = bar;
  for (size_t row = 0; row < count; ++row) {
    const double projection = h[row];
    if (!std::isfinite(projection)) return -1;
    const double square = projection * projection;
    if (!std::isfinite(square)) return -1;
    v = v - square;
    if (!std::isfinite(v)) return -1;
  }
  *t = v;
  if (v < 0.0) {
    *next = 0.0;
    return 1;
  }
  *next = std::sqrt(v);
  return std::isfinite(*next) ? 0 : -1;
}

// Back substitution of the rotated (upper triangular) Hessenberg columns,
// column-major with `rows` entries per column; 0, or the termination.
int fg_back_substitute(const std::vector<double>& h, size_t rows,
                       const std::vector<double>& g, size_t count,
                       std::vector<double>& y) {
  for (size_t r = count; r > 0; --r) {
    const size_t row = r - 1;
    double value = g[row];
    for (size_t upper = row + 1; upper < count; ++upper)
      value -= h[upper * rows + row] * y[upper];
    const double divisor = h[row * rows + row];
    if (divisor == 0.0 || !std::isfinite(divisor)) return HF_TERM_BREAKDOWN;
    y[row] = value / divisor;
    if (!std::isfinite(y[row])) return HF_TERM_NON_FINITE;
  }
  return 0;
}

// k_fg_stencil<kFgArnoldi, nd> and k_fg_dots<nd> (their row-pair forms
// when `pairs`) for a dot count nd in [1, kFgDots] known at run time.
template <int N = 1, class... Args>
void launch_fg_arnoldi(bool pairs, int nd, int64_t blocks, hipStream_t stream,
                       Args... args) {
  if constexpr (N <= kFgDots) {
    if (nd == N) {
      if (pairs)
        HF_LAUNCH((k_fg_stencil2<kFgArnoldi, N>), blocks, kBlock, 0, stream,
                  args...);
      else
        HF_LAUNCH((k_fg_stencil<kFgArnoldi, N>), blocks, kBlock, 0, stream,
                  args...);
      return;
    }
    launch_fg_arnoldi<N + 1>(pairs, nd, blocks, stream, args...);
  }
}

template <int N = 1, class... Args>
void launch_fg_dots(bool pairs, int nd, int64_t blocks, hipStream_t stream,
                    Args... args) {
  if constexpr (N <= kFgDots) {
    if (nd == N) {
      if (pairs)
        HF_LAUNCH(k_fg_dots2<N>, blocks, kBlock, 0, stream, args...);
      else
        HF_LAUNCH(k_fg_dots<N>, blocks, kBlock, 0, stream, args...);
      return;
    }
    launch_fg_dots<N + 1>(pairs, nd, blocks, stream, args...);
  }
}

hf_status fgmres_impl(hf_plan* p, const double* rhs, double* solution,
                      const hf_fgmres_control* ctl, hf_solve_result* res) {
  const size_t R = ctl->restart;
  const size_t rows = R + 1;
  const int64_t N = p->cells;
  const Lv<double> A = krylov_view(p);
  const Level& L0 = p->levels[0];
  // the level-0 stencils and dots run over pairs of cells of a row when
  // rows have even length (one partial per block of the layout used)
  const bool pairs = p->krylov_pairs && L0.n[0] % 2 == 0;
  const Tiles& tz = pairs ? L0.tzp : L0.tz;
  const Chunk ch = p->chunk0;
  const int nbv = chunk_blocks(ch);
  // the basis combinations take two elements per thread when the vector
  // length is even (every basis row then starts 16-byte aligned)
  const int comb_pairs = N % 2 == 0 ? 1 : 0;
  Chunk ch2 = ch;
  ch2.chunk = comb_pairs ? 2 * kBlock : kBlock;
  const int nbv2 = chunk_blocks(ch2);
  const int vrows = ctl->volume_rows ? 1 : 0;
  const Widths wd = widths(p);
  const int64_t pst = tz.blocks;
  double* x = p->kx;
  double* ax = p->kq;
  double* scaled = p->kr;
  auto V = [&](size_t r) { return p->fg_v + static_cast<int64_t>(r) * N; };
  auto Z = [&](size_t r) { return p->fg_z + static_cast<int64_t>(r) * N; };
  std::vector<double> h(rows * R, 0.0), cs(R, 0.0), sn(R, 0.0), g(rows, 0.0),
      y(R, 0.0);
  FgScales scales{};
  uint64_t restarts = 0;
  const double eps = std::numeric_limits<double>::epsilon();
  // a basis normalisation is checked for finiteness at the next read; a
  // failing one reports the counts of the moment it was made
  bool flag_pending = false;
  uint64_t ops_at = 0, pcs_at = 0;
  uint32_t it_at = 0;
  auto finish = [&](hf_status st, int term) {
    res->status = st;
    res->termination = term;
    res->norm_breakdown_restarts = restarts;
    return st;
  };
  auto fail_term = [&](int code, int term) {
    return finish(fail(code, term == HF_TERM_BREAKDOWN ? kDetailBreakdown
                                                       : kDetailNonFinite),
                  term);
  };
  // sums of `nacc` accumulators of fg_partials into fg_host; `bad` reports a
  // pending non-finite normalisation
  auto reduce = [&](int nacc, bool* bad) -> hf_status {
    const double* src = p->fg_partials;
    int64_t count = pst;
    int64_t stride = pst;
    if (pst > kFoldSpan) {
      const int64_t spans = (pst + kFoldSpan - 1) / kFoldSpan;
      hipLaunchKernelGGL(k_fold_partials,
                         dim3(static_cast<uint32_t>(spans),
                              static_cast<uint32_t>(nacc)),
                         dim3(kBlock), 0, p->stream, p->fg_partials, pst, pst,
                         p->fg_folded, static_cast<int64_t>(kFoldStride));
      src = p->fg_folded;
      count = spans;
      stride = kFoldStride;
    }
    {
      HF_PROF(p, P_FINALIZE);
      HF_LAUNCH(k_finalize_many, 1, 1024, 0, p->stream, src, count, stride,
                nacc, p->fg_scalars);
    }
    ++p->reductions;
    if (hipMemcpyAsync(p->fg_host, p->fg_scalars, kFgSlots * sizeof(double),
                       hipMemcpyDeviceToHost, p->stream) != hipSuccess ||
        hipStreamSynchronize(p->stream) != hipSuccess)
      return fail(HF_INVALID_PLAN, kDetailHip);
    *bad = false;
    if (flag_pending) {
      flag_pending = false;
      if (p->fg_host[kFgFlag] != 0.0) {
        res->operator_applies = ops_at;
        res->preconditioner_applies = pcs_at;
        res->iterations = it_at;
        *bad = true;
      }
    }
    return ok();
  };
  // out = b - A' vec and its norm (FG_TRUE_RESIDUAL); term != 0 on failure
  auto true_residual = [&](const double* vec, double* out, double* norm,
                           int* term) -> hf_status {
    {
      HF_PROF(p, P_TRUE_RESIDUAL);
      if (pairs)
        HF_LAUNCH((k_fg_stencil2<kFgResidual, 0>), tz.blocks, kBlock, 0,
                  p->stream, tz, A, p->bc, vec, rhs, out,
                  static_cast<const double*>(nullptr), int64_t{0}, 0, 0.0,
                  vrows, wd, p->fg_partials, pst);
      else
        HF_LAUNCH((k_fg_stencil<kFgResidual, 0>), tz.blocks, kBlock, 0,
                  p->stream, tz, A, p->bc, vec, rhs, out,
                  static_cast<const double*>(nullptr), int64_t{0}, 0, 0.0,
                  vrows, wd, p->fg_partials, pst);
    }
    bool bad = false;
    HF_TRY(reduce(1, &bad));
    if (bad) {
      *term = HF_TERM_NON_FINITE;
      return ok();
    }
    res->operator_applies++;
    *norm = std::sqrt(p->fg_host[0]);
    if (!std::isfinite(*norm)) {
      bool finite = false;
      HF_TRY(stable_norm(p, out, norm, &finite));
      if (!finite) *norm = std::numeric_limits<double>::quiet_NaN();
    }
    *term = std::isfinite(*norm) ? 0 : HF_TERM_OPERATOR_FAILURE;
    return ok();
  };
  // dst = src + sum s[row] basis_row (scale: then times post, checked)
  // A normalised basis vector is the next preconditioner input; its
  // combine also stores the multigrid's level-0 entry when that entry is a
  // plain copy: a grid level 0 whose first smoothing reads no x (point
  // levels, Chebyshev with the b-and-D-only first stage), no constant
  // projection, no volume-scaled rows.
  const bool entry_in_combine =
      !vrows && p->policy.null_space_constant == 0 && !L0.tail &&
      L0.line_mask == 0 && (!L0.cheb || p->cheb_first);
  bool entered_next = false;
  auto combine = [&](double* dst, const double* src, const double* basis,
                     const double* s, size_t count, double post, int scale) {
    FgScales k{};
    for (size_t r = 0; r < count; ++r) k.s[r] = s[r];
    HF_PROF(p, P_OTHER);
    if (scale && entry_in_combine) {
      dispatch(L0.prec, [&](auto t) {
        using T = decltype(t);
        HF_LAUNCH(k_fg_combine_entry<T>, nbv2, kBlock, 0, p->stream, ch2, dst,
                  src, basis, N, k, static_cast<int>(count), post,
                  p->fg_scalars + kFgFlag,
                  static_cast<const uint8_t*>(L0.act), static_cast<T*>(L0.b),
                  comb_pairs);
      });
      entered_next = true;
    } else {
      HF_LAUNCH(k_fg_combine, nbv2, kBlock, 0, p->stream, ch2, dst, src, basis,
                N, k, static_cast<int>(count), post, scale,
                p->fg_scalars + kFgFlag, comb_pairs);
    }
    if (scale) {
      flag_pending = true;
      ops_at = res->operator_applies;
      pcs_at = res->preconditioner_applies;
      it_at = res->iterations;
    }
  };
  // dots basis_row . wv for rows [0, count) into accumulators [acc0, ...)
  auto basis_dots = [&](const double* wv, size_t row0, size_t count,
                        int acc0) {
    for (size_t r = row0; r < row0 + count; r += kFgDots) {
      const int nd = static_cast<int>(std::min<size_t>(kFgDots, row0 + count - r));
      launch_fg_dots(pairs, nd, tz.blocks, p->stream, tz, wv,
                     static_cast<const double*>(p->fg_v), N,
                     static_cast<int>(r), acc0 + static_cast<int>(r - row0),
                     p->fg_partials, pst);
    }
  };
  auto norm_of = [&](const double* wv, double* norm, bool* bad) -> hf_status {
    if (pairs)
      HF_LAUNCH(k_fg_dots2<1>, tz.blocks, kBlock, 0, p->stream, tz, wv, wv,
                int64_t{0}, 0, 0, p->fg_partials, pst);
    else
      HF_LAUNCH(k_fg_dots<1>, tz.blocks, kBlock, 0, p->stream, tz, wv, wv,
                int64_t{0}, 0, 0, p->fg_partials, pst);
    HF_TRY(reduce(1, bad));
    *norm = std::sqrt(p->fg_host[0]);
    return ok();
  };
  auto publish = [&](const double* from) {
    HF_LAUNCH(k_copy, nbv, kBlock, 0, p->stream, ch, solution, from);
  };

  // ||b||
  HF_LAUNCH(k_dot, nbv, kBlock, 0, p->stream, ch, rhs, rhs, p->partials);
  finalize(p, nbv, 1, F_STORE, S_SUM);
  HF_TRY(read_scalars(p));
  double rhs_norm = std::sqrt(p->host_scalars[S_SUM]);
  if (!std::isfinite(rhs_norm)) {
    bool finite = false;
    HF_TRY(stable_norm(p, rhs, &rhs_norm, &finite));
    if (!finite) rhs_norm = std::numeric_limits<double>::quiet_NaN();
  }
  res->rhs_norm = rhs_norm;
  if (!std::isfinite(rhs_norm))
    return fail_term(HF_NUMERICAL_FAILURE, HF_TERM_NON_FINITE);
  const double tol =
      std::max(ctl->absolute_tolerance, ctl->relative_tolerance * rhs_norm);
  res->true_residual_limit = tol;
  if (rhs_norm == 0.0) {
    HF_LAUNCH(k_fill, nbv, kBlock, 0, p->stream, ch, solution, 0.0);
    return finish(ok(), HF_TERM_ZERO_RHS);
  }
  {
    const Chunk sc = make_chunk(kFgSlots);
    HF_LAUNCH(k_fill, chunk_blocks(sc), kBlock, 0, p->stream, sc,
              p->fg_scalars, 0.0);
  }
  HF_LAUNCH(k_copy, nbv, kBlock, 0, p->stream, ch, x, solution);
  int term = 0;
  HF_TRY(true_residual(x, V(0), &res->initial_true_residual, &term));
  if (term) return fail_term(HF_NUMERICAL_FAILURE, term);
  res->final_true_residual = res->initial_true_residual;
  res->recursive_residual = res->initial_true_residual;
  if (!std::isfinite(tol))
    return fail_term(HF_NUMERICAL_FAILURE, HF_TERM_NON_FINITE);
  if (res->initial_true_residual <= tol) {
    publish(x);
    return finish(ok(), HF_TERM_CONVERGED);
  }
  while (res->iterations < ctl->maximum_iterations) {
    const double beta = res->final_true_residual;
    combine(V(0), V(0), nullptr, nullptr, 0, 1.0 / beta, 1);
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    bool restarted = false;
    for (size_t col = 0;
         col < R && res->iterations < ctl->maximum_iterations; ++col) {
      // z_col = M v_col (M applied to V v_col with volume rows)
      const double* src = V(col);
      if (vrows) {
        HF_PROF(p, P_OTHER);
        HF_LAUNCH(k_scale_volume, L0.tc.blocks, kBlock, 0, p->stream, L0.tc,
                  scaled, static_cast<const double*>(V(col)), wd);
        src = scaled;
      }
      HF_TRY(mg_apply(p, src, Z(col), nullptr, false, entered_next));
      entered_next = false;
      res->preconditioner_applies++;
      if (L0.line_mask != 0) {
        const hf_status ls = read_scalars(p);
        if (ls.code != HF_OK)
          return finish(ls, ls.detail == kDetailLine
                                ? HF_TERM_PRECONDITIONER_FAILURE
                                : HF_TERM_INVALID_PLAN);
      }
      // w = A' z_col - shift v_col, its projections on v_0..v_col and w.w
      const size_t dots = col + 1;
      {
        HF_PROF(p, P_OPERATOR);
        launch_fg_arnoldi(pairs,
                          static_cast<int>(std::min<size_t>(dots, kFgDots)),
                          tz.blocks, p->stream, tz, A, p->bc,
                          static_cast<const double*>(Z(col)),
                          static_cast<const double*>(V(col)), ax,
                          static_cast<const double*>(p->fg_v), N,
                          static_cast<int>(dots), 1.0, vrows, wd,
                          p->fg_partials, pst);
        if (dots > static_cast<size_t>(kFgDots))
          basis_dots(ax, kFgDots, dots - kFgDots, kFgDots);
      }
      bool bad = false;
      HF_TRY(reduce(static_cast<int>(dots) + 1, &bad));
      if (bad) return fail_term(HF_NUMERICAL_FAILURE, HF_TERM_NON_FINITE);
      res->operator_applies++;
      double* hc = h.data() + col * rows;
      for (size_t row = 0; row <= dots; ++row) hc[row] = p->fg_host[row];
      for (size_t row = 0; row <= dots; ++row)
        if (!std::isfinite(hc[row]))
          return fail_term(HF_NUMERICAL_FAILURE, HF_TERM_NON_FINITE);
      const double bar = hc[dots];
      for (size_t row = 0; row < dots; ++row) scales.s[row] = -hc[row];
      double t = 0.0, next_norm = 0.0;
      const int nst = fg_norm_recurrence(hc, dots, bar, &t, &next_norm);
      if (nst < 0) return fail_term(HF_NUMERICAL_FAILURE, HF_TERM_NON_FINITE);
      const bool unsafe = nst == 1;
      if (unsafe && col != 0) {
        // discard the column, publish the shorter cycle and restart
        // (src/solver_krylov.cpp:2232-2340)
        res->iterations++;
        const int bt = fg_back_substitute(h, rows, g, col, y);
        if (bt) return fail_term(HF_NUMERICAL_FAILURE, bt);
        combine(x, x, p->fg_z, y.data(), col, 1.0, 0);
        HF_TRY(true_residual(x, V(0), &res->final_true_residual, &term));
        if (term) return fail_term(HF_NUMERICAL_FAILURE, term);
        res->recursive_residual = res->final_true_residual;
        if (res->final_true_residual <= tol) {
          publish(x);
          return finish(ok(), HF_TERM_CONVERGED);
        }
        if (res->iterations == ctl->maximum_iterations)
          return finish(fail(HF_REJECTED_STEP, kDetailBreakdown),
                        HF_TERM_MAXIMUM_ITERATIONS);
        ++restarts;
        std::fill(h.begin(), h.end(), 0.0);
        std::fill(cs.begin(), cs.end(), 0.0);
        std::fill(sn.begin(), sn.end(), 0.0);
        std::fill(g.begin(), g.end(), 0.0);
        std::fill(y.begin(), y.end(), 0.0);
        restarted = true;
        break;
      }
      bool explicit_pass = false;
      if (unsafe || next_norm == 0.0) {
        // second orthogonalisation pass (src/solver_krylov.cpp:2342-2400)
        combine(ax, ax, p->fg_v, scales.s, dots, 1.0, 0);
        basis_dots(ax, 0, dots, 0);
        HF_TRY(reduce(static_cast<int>(dots), &bad));
        if (bad) return fail_term(HF_NUMERICAL_FAILURE, HF_TERM_NON_FINITE);
        for (size_t row = 0; row < dots; ++row) {
          const double yr = p->fg_host[row];
          hc[row] += yr;
          if (!std::isfinite(hc[row]))
            return fail_term(HF_NUMERICAL_FAILURE, HF_TERM_NON_FINITE);
          scales.s[row] = -yr;
        }
        combine(ax, ax, p->fg_v, scales.s, dots, 1.0, 0);
        HF_TRY(norm_of(ax, &next_norm, &bad));
        if (bad || !std::isfinite(next_norm))
          return fail_term(HF_NUMERICAL_FAILURE, HF_TERM_NON_FINITE);
        explicit_pass = true;
      }
      hc[col] += 1.0;  // the Arnoldi shift
      if (!std::isfinite(hc[col]))
        return fail_term(HF_NUMERICAL_FAILURE, HF_TERM_NON_FINITE);
      hc[col + 1] = next_norm;
      for (size_t row = 0; row < col; ++row) {
        const double upper = hc[row], lower = hc[row + 1];
        const double rotated = cs[row] * upper + sn[row] * lower;
        hc[row + 1] = -sn[row] * upper + cs[row] * lower;
        hc[row] = rotated;
      }
      const double rn = std::hypot(hc[col], hc[col + 1]);
      if (!(rn > 0.0) || !std::isfinite(rn))
        return fail_term(HF_NUMERICAL_FAILURE, HF_TERM_BREAKDOWN);
      cs[col] = hc[col] / rn;
      sn[col] = hc[col + 1] / rn;
      hc[col] = rn;
      hc[col + 1] = 0.0;
      g[col + 1] = -sn[col] * g[col];
      g[col] = cs[col] * g[col];
      const double recursive = std::fabs(g[col + 1]);
      res->recursive_residual = recursive;
      res->iterations++;
      const bool happy = next_norm == 0.0;
      const bool cycle_end = happy || col + 1 == R ||
                             res->iterations == ctl->maximum_iterations;
      const bool verify = happy || recursive <= tol ||
                          res->iterations % ctl->true_residual_interval == 0 ||
                          cycle_end;
      if (!cycle_end)
        combine(V(col + 1), ax, p->fg_v, scales.s, explicit_pass ? 0 : dots,
                1.0 / next_norm, 1);
      if (!verify) continue;
      const int bt = fg_back_substitute(h, rows, g, dots, y);
      if (bt) return fail_term(HF_NUMERICAL_FAILURE, bt);
      if (!cycle_end) {
        // audit a trial solution in the unused last basis slot
        double* trial = V(R);
        combine(trial, x, p->fg_z, y.data(), dots, 1.0, 0);
        double audited = 0.0;
        HF_TRY(true_residual(trial, ax, &audited, &term));
        if (term) return fail_term(HF_NUMERICAL_FAILURE, term);
        res->final_true_residual = audited;
        if (audited <= tol) {
          publish(trial);
          return finish(ok(), HF_TERM_CONVERGED);
        }
        continue;
      }
      combine(x, x, p->fg_z, y.data(), dots, 1.0, 0);
      HF_TRY(true_residual(x, V(0), &res->final_true_residual, &term));
      if (term) return fail_term(HF_NUMERICAL_FAILURE, term);
      if (res->final_true_residual <= tol) {
        publish(x);
        return finish(ok(), HF_TERM_CONVERGED);
      }
      if (res->iterations == ctl->maximum_iterations)
        return finish(fail(HF_REJECTED_STEP, kDetailBreakdown),
                      HF_TERM_MAXIMUM_ITERATIONS);
      if (happy) {
        const double guard = 64.0 * eps * std::max(beta, tol);
        if (!(res->final_true_residual + guard < beta))
          return fail_term(HF_NUMERICAL_FAILURE, HF_TERM_BREAKDOWN);
        ++restarts;
      }
      restarted = true;
      break;
    }
    if (!restarted) break;
  }
  return finish(fail(HF_REJECTED_STEP, kDetailBreakdown),
                HF_TERM_MAXIMUM_ITERATIONS);
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

hf_status hf_policy_default(hf_policy* policy) {
  if (policy == nullptr) return fail(HF_INVALID_PLAN, kDetailPlan);
  policy->anisotropy_threshold = 4.0;
  policy->pre_sweeps = 2;
  policy->post_sweeps = 2;
  policy->maximum_levels = 32;
  policy->minimum_coarse_extent = 3;
  policy->coarse_sweeps = 24;
  policy->f_cycle = 0;
  policy->null_space_constant = 1;
  policy->unit_linear = 1;
  policy->precision = HF_PRECISION_FP64;
  policy->coarse_face_scaling = 0;
  policy->line_order = 0;
  policy->point_smoother = 0;
  policy->chebyshev_lower_fraction = 0.3;
  policy->diagonal_shortcut_ratio = 0.0;
  return ok();
}

hf_status hf_plan_create(const hf_grid_desc* grid, const hf_policy* policy,
                         void* stream, hf_plan** out) {
  if (out == nullptr) return fail(HF_INVALID_PLAN, kDetailPlan);
  *out = nullptr;
  if (grid == nullptr || policy == nullptr || grid->nx <= 0 || grid->ny <= 0 ||
      grid->nz <= 0 || grid->faces_x == nullptr || grid->faces_y == nullptr ||
      grid->faces_z == nullptr)
    return fail(HF_INVALID_PLAN, kDetailPlan);
  for (int a = 0; a < 3; ++a) {
    const int lo = grid->boundary[2 * a], hi = grid->boundary[2 * a + 1];
    if (lo < 0 || lo > 2 || hi < 0 || hi > 2) return fail(HF_INVALID_PLAN, kDetailPlan);
    if ((lo == HF_BC_PERIODIC) != (hi == HF_BC_PERIODIC))
      return fail(HF_INVALID_PLAN, kDetailPlan);
  }
  if (policy->pre_sweeps == 0 || policy->post_sweeps == 0 ||
      policy->coarse_sweeps == 0 || policy->maximum_levels < 2 ||
      policy->maximum_levels > kMaxLevels || policy->minimum_coarse_extent < 2 ||
      !(policy->anisotropy_threshold > 1.0) || policy->precision > 2 ||
      policy->coarse_face_scaling > 1 || policy->line_order > 1 ||
      policy->point_smoother > 1 ||
      !(policy->chebyshev_lower_fraction > 0.0) ||
      !(policy->chebyshev_lower_fraction < 1.0) ||
      !(policy->diagonal_shortcut_ratio >= 0.0) ||
      !(policy->diagonal_shortcut_ratio < 1.0))
    return fail(HF_INVALID_PLAN, kDetailPlan);
  {
    // every kernel indexes cells and faces with 32-bit integers
    const int64_t nx = grid->nx, ny = grid->ny, nz = grid->nz;
    const int64_t limit = std::numeric_limits<int32_t>::max();
    if (nx * ny * nz > limit || (nx + 1) * ny * nz > limit ||
        nx * (ny + 1) * nz > limit || nx * ny * (nz + 1) > limit)
      return fail(HF_INVALID_PLAN, kDetailPlan);
  }
  hf_plan* p = new (std::nothrow) hf_plan();
  if (p == nullptr) return fail(HF_ALLOCATION_FAILURE, kDetailAlloc);
  p->grid = *grid;
  p->policy = *policy;
  p->stream = static_cast<hipStream_t>(stream);
  if (hipGetDevice(&p->device) != hipSuccess) {
    delete p;
    return fail(HF_INVALID_PLAN, kDetailHip);
  }
  p->n[0] = grid->nx;
  p->n[1] = grid->ny;
  p->n[2] = grid->nz;
  p->cells = (int64_t)grid->nx * grid->ny * grid->nz;
  const double* f[3] = {grid->faces_x, grid->faces_y, grid->faces_z};
  for (int a = 0; a < 3; ++a) {
    p->faces[a].assign(f[a], f[a] + p->n[a] + 1);
    p->widths[a].resize(p->n[a]);
    for (int i = 0; i < p->n[a]; ++i) {
      const double w = p->faces[a][i + 1] - p->faces[a][i];
      if (!(w > 0.0) || !std::isfinite(w)) {
        delete p;
        return fail(HF_INVALID_PLAN, kDetailPlan);
      }
      p->widths[a][i] = w;
    }
  }
  for (int face = 0; face < 6; ++face)
    p->bc.m[face] = static_cast<int8_t>(grid->boundary[face] == HF_BC_DIRICHLET
                                            ? 0
                                            : (grid->boundary[face] == HF_BC_NEUMANN ? 1 : 2));
  p->periodic_mask = (grid->boundary[0] == HF_BC_PERIODIC ? 1 : 0) |
                     (grid->boundary[2] == HF_BC_PERIODIC ? 2 : 0) |
                     (grid->boundary[4] == HF_BC_PERIODIC ? 4 : 0);
  plan_levels(p);
  if (p->levels.size() < 2) {
    delete p;
    return fail(HF_INVALID_PLAN, kDetailPlan);
  }
  const bool line_plan = p->levels[0].line_mask != 0;
  if (line_plan && !line_relaxation_enabled()) {
    delete p;
    return fail(HF_INVALID_PLAN, kDetailLineOff);
  }
  // the tail is the longest run of coarsest levels whose x, b, r fit the
  // tail kernel's LDS and each stay under the cell threshold
  const int64_t threshold = tail_threshold();
  p->row_pairs = row_pairs_enabled();
  p->cheb_first = cheb_first_enabled();
  p->krylov_pairs = krylov_pairs_enabled();
  p->tail_first = static_cast<int>(p->levels.size());
  int64_t lds_doubles = 0;
  // line levels (every level of a line plan) are smoothed by the line
  // kernels, never in the tail; Chebyshev levels join it while their stage
  // counts fit its table of stage factors
  const bool cheb_plan = policy->point_smoother == 1;
  const bool cheb_fits =
      std::max(policy->pre_sweeps, policy->post_sweeps) <= kTailChebStages;
  for (int l = static_cast<int>(p->levels.size()) - 1;
       l >= 0 && !line_plan && (!cheb_plan || cheb_fits); --l) {
    const int64_t need = lds_doubles + 3 * p->levels[l].cells;
    if (p->levels[l].cells <= threshold &&
        need * static_cast<int64_t>(sizeof(double)) <= kTailLdsBytes) {
      p->tail_first = l;
      lds_doubles = need;
    } else {
      break;
    }
  }
  {
    ChebStage stage_factors(policy->chebyshev_lower_fraction);
    for (int s = 0; s < kTailChebStages; ++s)
      stage_factors.next(&p->tail_cheb.rf[s], &p->tail_cheb.pf[s]);
  }
  // the coarsest level is always FP64 and, on point plans, handled by the
  // tail kernel
  if (!line_plan && p->tail_first > static_cast<int>(p->levels.size()) - 1)
    p->tail_first = static_cast<int>(p->levels.size()) - 1;
  int64_t tail_doubles = 0;
  for (size_t l = 0; l < p->levels.size(); ++l) {
    Level& L = p->levels[l];
    L.tail = static_cast<int>(l) >= p->tail_first;
    if (L.tail) tail_doubles += 3 * L.cells;
    const bool coarsest = l + 1 == p->levels.size();
    L.prec = (L.tail || coarsest) ? HF_PRECISION_FP64 : policy->precision;
    // Chebyshev-Jacobi on every point level but the coarsest
    // (src/solver_mg.cpp:1048-1062)
    L.cheb = cheb_plan && !coarsest && L.line_mask == 0;
    L.tc = make_tiles(L.n[0], L.n[1], L.n[2], false);
    L.tcol = make_tiles((L.n[0] + 1) / 2, L.n[1], L.n[2], false);
    L.tr = make_tiles(L.n[0], L.n[1], L.n[2], true);
    L.tred = L.tr;
    L.tred.per = kRedPer;
    L.tred.blocks = (L.tr.count + kRedPer - 1) / kRedPer;
    {
      // z-chunks of a few planes (measured fastest: many blocks in flight,
      // each reusing its planes)
      L.tz = L.tr;
      const int gxy = L.tz.gx * L.tz.gy;
      const int64_t want = static_cast<int64_t>(L.n[2]) * gxy / 32768;
      // (an even number of planes, so defect and restriction can fuse)
      int per = static_cast<int>(std::max<int64_t>(2, std::min<int64_t>(64, want)));
      L.tz.per = per + (per & 1);
      L.tz.blocks = gxy * ((L.n[2] + L.tz.per - 1) / L.tz.per);
      // pairs of cells of a row (even x extents), the same z-chunks
      L.tzp = make_tiles(std::max(1, L.n[0] / 2), L.n[1], L.n[2], false);
      L.tzp.per = L.tz.per;
      L.tzp.blocks = L.tzp.gx * L.tzp.gy *
                     ((L.n[2] + L.tzp.per - 1) / L.tzp.per);
    }
    // fused sweeps on grid levels with the fused colour route and rows long
    // enough for the tiles to stay mostly busy
    // (its buffer loads take 32-bit byte offsets: every array under 4 GiB)
    const int64_t widest =
        std::max({static_cast<int64_t>(L.n[0] + 1) * L.n[1] * L.n[2],
                  static_cast<int64_t>(L.n[0]) * (L.n[1] + 1) * L.n[2],
                  static_cast<int64_t>(L.n[0]) * L.n[1] * (L.n[2] + 1)});
    // (each lane owns a pair of cells of a row: nx even)
    L.sweep = (l == 0 || L.cells >= sweep_min_cells()) && !L.tail &&
              L.line_mask == 0 && L.fused && !L.cheb &&
              L.n[0] >= 32 && L.n[0] % 2 == 0 && L.n[1] >= 8 &&
              sweep_enabled() &&
              widest * static_cast<int64_t>(elem_size(L.prec)) <
                  (int64_t{1} << 32);
    if (L.sweep) {
      L.sweep_gx = (L.n[0] + kSwTX - 1) / kSwTX;
      L.sweep_gy = (L.n[1] + kSwTY - 1) / kSwTY;
      const int per_plane = L.sweep_gx * L.sweep_gy;
      // about half a million threads in flight, split into z-chunks
      const int want = std::max(1, (1 << 19) / (kSwThreads * per_plane));
      L.sweep_zc = std::min(32, std::max(4, (L.n[2] + want - 1) / want));
      if (sweep_zc_override() > 0)
        L.sweep_zc = std::min(L.n[2], sweep_zc_override());
      L.sweep_blocks = per_plane * ((L.n[2] + L.sweep_zc - 1) / L.sweep_zc);
    }
  }
  p->chunk0 = make_chunk(static_cast<int>(p->cells));
  p->tail_lds_bytes = static_cast<size_t>(tail_doubles) * sizeof(double);
  if (p->tail_lds_bytes > static_cast<size_t>(kTailLdsBytes)) {
    // the coarsest level alone exceeds the tail kernel's LDS
    delete p;
    return fail(HF_INVALID_PLAN, kDetailPlan);
  }
  build_exec(p);

  hf_status s = ok();
  auto guard = [&](hf_status st) {
    if (st.code != HF_OK && s.code == HF_OK) s = st;
  };
  for (int a = 0; a < 3; ++a) {
    guard(dalloc(p, &p->d_faces[a], p->n[a] + 1));
    guard(dalloc(p, &p->d_widths[a], p->n[a]));
  }
  for (Level& L : p->levels) {
    guard(alloc_coeffs(p, L, L));
    const size_t es = elem_size(L.prec);
    guard(dalloc_raw(p, &L.x, es * L.cells));
    guard(dalloc_raw(p, &L.b, es * L.cells));
    guard(dalloc_raw(p, &L.r, es * L.cells));
    if (L.sweep) {
      guard(dalloc_raw(p, &L.x2, es * L.cells));
      guard(dalloc_raw(p, &L.b2, es * L.cells));
    }
    if (L.cheb) guard(dalloc_raw(p, &L.x2, es * L.cells));
    if (L.line_mask != 0) {
      const size_t cs = L.prec == HF_PRECISION_FP64 ? 8 : 4;
      guard(dalloc_raw(p, &L.lcw, cs * L.cells));
      guard(dalloc_raw(p, &L.lcd, cs * L.cells));
      int64_t ghosts = 1;
      for (int ax = 0; ax < 3; ++ax) {
        if ((L.line_mask & (1 << ax)) == 0) continue;
        const int64_t A = ax == 0 ? L.n[1] : L.n[0];
        const int64_t B = ax == 2 ? L.n[1] : L.n[2];
        ghosts = std::max<int64_t>(ghosts, 2 * (A + B) * L.n[ax]);
      }
      guard(dalloc_raw(p, &L.lghost, es * ghosts));
    }
  }
  // one partial per tile (level 0 has the most) or per 256-element block
  int64_t max_blocks = blocks_for(p->cells);
  for (const Level& L : p->levels)
    max_blocks = std::max<int64_t>(max_blocks, L.tc.count);
  p->partial_stride = max_blocks + 8;
  guard(dalloc(p, &p->partials, 3 * p->partial_stride));
  guard(dalloc(p, &p->folded, 3 * kFoldStride));
  guard(dalloc(p, &p->scalars, S_COUNT));
  guard(dalloc(p, &p->d_flag, 4));
  guard(dalloc(p, &p->d_ratio, 1));
  guard(dalloc(p, &p->kx, p->cells));
  guard(dalloc(p, &p->kr, p->cells));
  guard(dalloc(p, &p->kp, p->cells));
  guard(dalloc(p, &p->kq, p->cells));
  const int tail_levels = static_cast<int>(p->levels.size()) - p->tail_first;
  guard(dalloc(p, &p->d_tail_levels, std::max(tail_levels, 1)));
  guard(dalloc(p, &p->d_tail_ops, std::max<size_t>(p->tail_ops_host.size(), 1)));
  if (s.code == HF_OK &&
      hipHostMalloc(reinterpret_cast<void**>(&p->host_scalars),
                    S_COUNT * sizeof(double), hipHostMallocDefault) != hipSuccess) {
    std::fprintf(stderr, "hundun_flow: hipHostMalloc(%zu bytes) failed\n",
                 S_COUNT * sizeof(double));
    s = fail(HF_ALLOCATION_FAILURE, kDetailAlloc);
  }
  if (s.code == HF_OK) {
    for (int a = 0; a < 3; ++a) {
      if (hipMemcpyAsync(p->d_faces[a], p->faces[a].data(),
                         (p->n[a] + 1) * sizeof(double), hipMemcpyHostToDevice,
                         p->stream) != hipSuccess ||
          hipMemcpyAsync(p->d_widths[a], p->widths[a].data(),
                         p->n[a] * sizeof(double), hipMemcpyHostToDevice,
                         p->stream) != hipSuccess)
        s = fail(HF_INVALID_PLAN, kDetailHip);
    }
    if (s.code == HF_OK)
      s = hip_check(hipMemsetAsync(p->scalars, 0, S_COUNT * sizeof(double),
                                   p->stream));
    if (s.code == HF_OK) s = hip_check(hipStreamSynchronize(p->stream));
  }
  if (s.code != HF_OK) {
    hf_plan_destroy(p);
    return s;
  }
  *out = p;
  return ok();
}

hf_status hf_plan_destroy(hf_plan* p) {
  if (p == nullptr) return ok();
  if (p->stream != nullptr) (void)hipStreamSynchronize(p->stream);
  else (void)hipDeviceSynchronize();
  for (void* ptr : p->allocations) (void)hipFree(ptr);
  for (hipEvent_t e : p->events) (void)hipEventDestroy(e);
  if (p->host_scalars != nullptr) (void)hipHostFree(p->host_scalars);
  delete p;
  return ok();
}

hf_status hf_assemble_coefficients(hf_plan* p, const double* mobility,
                                   double* diag, double* bx, double* by,
                                   double* bz) {
  if (p == nullptr || mobility == nullptr || diag == nullptr || bx == nullptr ||
      by == nullptr || bz == nullptr)
    return fail(HF_INVALID_PLAN, kDetailPlan);
  const AxisGeom g = axis_geom(p);
  const Level& L = p->levels[0];
  double* out[3] = {bx, by, bz};
  const int64_t counts[3] = {L.nfx, L.nfy, L.nfz};
  for (int a = 0; a < 3; ++a)
    HF_LAUNCH(k_assemble_faces, blocks_for(counts[a]), kBlock, 0, p->stream, g,
              a, mobility, out[a], counts[a]);
  HF_LAUNCH(k_face_sum_diag, blocks_for(L.cells), kBlock, 0, p->stream,
            p->n[0], p->n[1], p->n[2], bx, by, bz, diag);
  return hip_check(hipGetLastError());
}

hf_status hf_assemble_pressure_system(
    hf_plan* p, const double* bx, const double* by, const double* bz,
    const double* phi_x, const double* phi_y, const double* phi_z,
    const double* rho, const double* rho_n, const double* rho_nm1,
    const double* drho_dp, const double* source, double a0, double a1,
    double a2, double* diag, double* rhs) {
  if (p == nullptr || bx == nullptr || by == nullptr || bz == nullptr ||
      phi_x == nullptr || phi_y == nullptr || phi_z == nullptr ||
      rho == nullptr || rho_n == nullptr || drho_dp == nullptr ||
      diag == nullptr || rhs == nullptr)
    return fail(HF_INVALID_PLAN, kDetailPlan);
  if (!(a0 > 0.0) || !std::isfinite(a0) || !std::isfinite(a1) ||
      !std::isfinite(a2))
    return fail(HF_INVALID_PLAN, kDetailPlan);
  const Level& L = p->levels[0];
  PressureSystem s{};
  s.nx = p->n[0];
  s.ny = p->n[1];
  s.nz = p->n[2];
  s.w = widths(p);
  s.bx = bx;
  s.by = by;
  s.bz = bz;
  s.phix = phi_x;
  s.phiy = phi_y;
  s.phiz = phi_z;
  s.rho = rho;
  s.rho_n = rho_n;
  s.rho_nm1 = rho_nm1;
  s.drho_dp = drho_dp;
  s.source = source;
  s.a0 = a0;
  s.a1 = a1;
  s.a2 = a2;
  s.diag = diag;
  s.rhs = rhs;
  // the preflight checks everything before anything is written
  HF_TRY(hip_check(hipMemsetAsync(p->d_flag, 0, sizeof(int), p->stream)));
  const double* beta[3] = {bx, by, bz};
  const double* phi[3] = {phi_x, phi_y, phi_z};
  const int64_t counts[3] = {L.nfx, L.nfy, L.nfz};
  for (int a = 0; a < 3; ++a)
    HF_LAUNCH(k_check_faces, blocks_for(counts[a]), kBlock, 0, p->stream,
              beta[a], phi[a], counts[a], p->d_flag);
  HF_LAUNCH(k_pressure_system, blocks_for(L.cells), kBlock, 0, p->stream, s, 0,
            p->d_flag);
  int flag = 0;
  HF_TRY(hip_check(hipMemcpyAsync(&flag, p->d_flag, sizeof(int),
                                  hipMemcpyDeviceToHost, p->stream)));
  HF_TRY(hip_check(hipStreamSynchronize(p->stream)));
  if (flag != 0) return fail(HF_NUMERICAL_FAILURE, kDetailAssembly);
  HF_LAUNCH(k_pressure_system, blocks_for(L.cells), kBlock, 0, p->stream, s, 1,
            p->d_flag);
  return hip_check(hipGetLastError());
}

hf_status hf_cold_coefficients(hf_plan* p, const double* proj_x,
                               const double* proj_y, const double* proj_z,
                               const double* lim_x, const double* lim_y,
                               const double* lim_z, const uint8_t* fluid,
                               const double* rho, const double* pressure,
                               double pressure_reference, double dt,
                               int32_t fixed_eos, const int32_t* kinds,
                               double* diag, double* cx, double* cy,
                               double* cz) {
  if (p == nullptr || proj_x == nullptr || proj_y == nullptr ||
      proj_z == nullptr || lim_x == nullptr || lim_y == nullptr ||
      lim_z == nullptr || kinds == nullptr || diag == nullptr ||
      cx == nullptr || cy == nullptr || cz == nullptr ||
      (!fixed_eos && (rho == nullptr || pressure == nullptr)))
    return fail(HF_INVALID_PLAN, kDetailPlan);
  if (!(dt > 0.0) || !std::isfinite(dt) || !std::isfinite(pressure_reference))
    return fail(HF_INVALID_PLAN, kDetailPlan);
  ColdSystem s{};
  for (int f = 0; f < 6; ++f) {
    if (kinds[f] < 0 || kinds[f] > 2) return fail(HF_INVALID_PLAN, kDetailPlan);
    s.kind[f] = static_cast<int8_t>(kinds[f]);
  }
  for (int a = 0; a < 3; ++a)
    if ((kinds[2 * a] == HF_COLD_PERIODIC) !=
        (kinds[2 * a + 1] == HF_COLD_PERIODIC))
      return fail(HF_INVALID_PLAN, kDetailPlan);
  const Level& L = p->levels[0];
  for (int a = 0; a < 3; ++a) s.n[a] = p->n[a];
  s.w = widths(p);
  s.proj[0] = proj_x;
  s.proj[1] = proj_y;
  s.proj[2] = proj_z;
  s.lim[0] = lim_x;
  s.lim[1] = lim_y;
  s.lim[2] = lim_z;
  s.fluid = fluid;
  s.rho = rho;
  s.pressure = pressure;
  s.pressure_reference = pressure_reference;
  s.dt = dt;
  s.fixed_eos = fixed_eos != 0 ? 1 : 0;
  s.diag = diag;
  s.coef[0] = cx;
  s.coef[1] = cy;
  s.coef[2] = cz;
  const int64_t counts[3] = {L.nfx, L.nfy, L.nfz};
  for (int a = 0; a < 3; ++a)
    HF_LAUNCH(k_cold_faces, blocks_for(counts[a]), kBlock, 0, p->stream, s, a,
              counts[a]);
  HF_LAUNCH(k_cold_diag, blocks_for(L.cells), kBlock, 0, p->stream, s);
  return hip_check(hipGetLastError());
}

hf_status hf_plan_set_coefficients(hf_plan* p, const double* diag,
                                   const double* bx, const double* by,
                                   const double* bz, const uint8_t* ca,
                                   const uint8_t* fxa, const uint8_t* fya,
                                   const uint8_t* fza) {
  if (p == nullptr || diag == nullptr || bx == nullptr || by == nullptr ||
      bz == nullptr)
    return fail(HF_INVALID_PLAN, kDetailPlan);
  const bool any = ca || fxa || fya || fza;
  const bool all = ca && fxa && fya && fza;
  if (any && !all) return fail(HF_INVALID_PLAN, kDetailPlan);
  // The first publication builds into the level arrays themselves; a
  // refresh of a published hierarchy builds into the spare set and swaps it
  // in only after every check passed, so a rejected refresh leaves the
  // previous hierarchy solving exactly as before (src/solver_mg.cpp:
  // 4700-4770).
  const bool into_spare = p->coefficients_ready;
  std::vector<Coeffs*> target(p->levels.size());
  for (size_t l = 0; l < p->levels.size(); ++l) {
    Level& L = p->levels[l];
    if (into_spare && L.spare.D64 == nullptr)
      HF_TRY(alloc_coeffs(p, L, L.spare));
    target[l] = into_spare ? &L.spare : static_cast<Coeffs*>(&L);
  }
  Level& L0 = p->levels[0];
  Coeffs& K0 = *target[0];
  // level 0: masking
  HF_LAUNCH(k_mask_faces, blocks_for(L0.nfx), kBlock, 0, p->stream, bx, fxa, K0.fx64, L0.nfx);
  HF_LAUNCH(k_mask_faces, blocks_for(L0.nfy), kBlock, 0, p->stream, by, fya, K0.fy64, L0.nfy);
  HF_LAUNCH(k_mask_faces, blocks_for(L0.nfz), kBlock, 0, p->stream, bz, fza, K0.fz64, L0.nfz);
  Finest f{};
  f.nx = p->n[0];
  f.ny = p->n[1];
  f.nz = p->n[2];
  f.D = diag;
  f.bx = bx;
  f.by = by;
  f.bz = bz;
  f.ca = ca;
  f.fxa = fxa;
  f.fya = fya;
  f.fza = fza;
  f.wx = p->d_widths[0];
  f.wy = p->d_widths[1];
  f.wz = p->d_widths[2];
  f.oD = K0.D64;
  f.ofx = K0.fx64;
  f.ofy = K0.fy64;
  f.ofz = K0.fz64;
  f.vol = K0.vol;
  f.act = K0.act;
  HF_LAUNCH(k_mask_cells, blocks_for(L0.cells), kBlock, 0, p->stream, f);
  HF_TRY(hip_check(hipMemsetAsync(p->d_flag, 0, sizeof(int), p->stream)));
  // diagonal shortcut (src/solver_mg.cpp:4545-4575): the finest level is
  // validated first, then its largest row ratio decides; a shortcut plan
  // never builds or reads the coarse hierarchy
  bool shortcut = false;
  if (p->policy.diagonal_shortcut_ratio != 0.0 &&
      !p->policy.null_space_constant) {
    HF_LAUNCH(k_validate, blocks_for(L0.cells), kBlock, 0, p->stream, L0.n[0],
              L0.n[1], L0.n[2], K0.D64, K0.fx64, K0.fy64, K0.fz64, K0.vol,
              p->d_flag);
    HF_TRY(hip_check(hipMemsetAsync(p->d_ratio, 0, sizeof(unsigned long long),
                                    p->stream)));
    HF_LAUNCH(k_row_ratio_max,
              std::min<int64_t>(blocks_for(L0.cells), kRatioBlocks), kBlock, 0,
              p->stream, L0.n[0], L0.n[1], L0.n[2], K0.D64, K0.fx64, K0.fy64,
              K0.fz64, p->d_ratio);
    int flag0 = 0;
    unsigned long long bits = 0;
    HF_TRY(hip_check(hipMemcpyAsync(&flag0, p->d_flag, sizeof(int),
                                    hipMemcpyDeviceToHost, p->stream)));
    HF_TRY(hip_check(hipMemcpyAsync(&bits, p->d_ratio, sizeof bits,
                                    hipMemcpyDeviceToHost, p->stream)));
    HF_TRY(hip_check(hipStreamSynchronize(p->stream)));
    if (flag0 != 0) return fail(HF_NUMERICAL_FAILURE, kDetailCoefficients);
    double ratio = 0.0;
    std::memcpy(&ratio, &bits, sizeof ratio);
    shortcut = ratio <= p->policy.diagonal_shortcut_ratio;
  }
  const size_t built = shortcut ? 1 : p->levels.size();
  // coarse levels
  for (size_t l = 1; l < built; ++l) {
    const Level& F = p->levels[l - 1];
    const Level& C = p->levels[l];
    const Coeffs& KF = *target[l - 1];
    const Coeffs& KC = *target[l];
    Agg a;
    for (int ax = 0; ax < 3; ++ax) {
      a.fn[ax] = F.n[ax];
      a.cn[ax] = C.n[ax];
    }
    a.mask = F.coarsen_mask;
    Fine64 fine{KF.D64, KF.fx64, KF.fy64, KF.fz64, KF.vol};
    Coarse64 coarse{KC.D64, KC.fx64, KC.fy64, KC.fz64, KC.vol, KC.act};
    HF_LAUNCH(k_coarse_cells, blocks_for(C.cells), kBlock, 0, p->stream, a, fine, coarse);
    // distance-corrected aggregation halves faces whose normal was coarsened
    const bool dc = p->policy.coarse_face_scaling != 0;
    const double sx = dc && (F.coarsen_mask & 1) ? 0.5 : 1.0;
    const double sy = dc && (F.coarsen_mask & 2) ? 0.5 : 1.0;
    const double sz = dc && (F.coarsen_mask & 4) ? 0.5 : 1.0;
    HF_LAUNCH(k_coarse_faces, blocks_for(C.nfx), kBlock, 0, p->stream, a, 0, KF.fx64, KC.fx64, sx);
    HF_LAUNCH(k_coarse_faces, blocks_for(C.nfy), kBlock, 0, p->stream, a, 1, KF.fy64, KC.fy64, sy);
    HF_LAUNCH(k_coarse_faces, blocks_for(C.nfz), kBlock, 0, p->stream, a, 2, KF.fz64, KC.fz64, sz);
    HF_LAUNCH(k_coarse_diag, blocks_for(C.cells), kBlock, 0, p->stream, C.n[0], C.n[1], C.n[2], coarse, p->d_flag);
  }
  for (size_t l = 0; l < built; ++l) {
    const Level& L = p->levels[l];
    const Coeffs& K = *target[l];
    HF_LAUNCH(k_validate, blocks_for(L.cells), kBlock, 0, p->stream, L.n[0], L.n[1], L.n[2], K.D64, K.fx64, K.fy64, K.fz64, K.vol, p->d_flag);
  }
  // storage-precision copies
  for (size_t l = 0; l < built; ++l) {
    const Level& L = p->levels[l];
    const Coeffs& K = *target[l];
    if (L.prec == HF_PRECISION_FP64) continue;
    dispatch(L.prec, [&](auto t) {
      using T = decltype(t);
      HF_LAUNCH(k_convert<T>, blocks_for(L.cells), kBlock, 0, p->stream, K.D64, static_cast<T*>(K.DT), L.cells);
      HF_LAUNCH(k_convert<T>, blocks_for(L.nfx), kBlock, 0, p->stream, K.fx64, static_cast<T*>(K.fxT), L.nfx);
      HF_LAUNCH(k_convert<T>, blocks_for(L.nfy), kBlock, 0, p->stream, K.fy64, static_cast<T*>(K.fyT), L.nfy);
      HF_LAUNCH(k_convert<T>, blocks_for(L.nfz), kBlock, 0, p->stream, K.fz64, static_cast<T*>(K.fzT), L.nfz);
    });
  }
  // projection volumes (sum of fluid volumes per level)
  for (size_t l = 0; l < built; ++l) {
    const Level& L = p->levels[l];
    Coeffs& K = *target[l];
    const Chunk ch = make_chunk(static_cast<int>(L.cells));
    const int nb = chunk_blocks(ch);
    HF_LAUNCH(k_sum, nb, kBlock, 0, p->stream, ch, K.vol,
              static_cast<const uint8_t*>(nullptr), p->partials);
    finalize(p, nb, 1, F_STORE, S_SUM);
    HF_TRY(read_scalars(p));
    K.volume = p->host_scalars[S_SUM];
  }
  int flag = 0;
  HF_TRY(hip_check(hipMemcpyAsync(&flag, p->d_flag, sizeof(int), hipMemcpyDeviceToHost, p->stream)));
  HF_TRY(hip_check(hipStreamSynchronize(p->stream)));
  HF_TRY(hip_check(hipGetLastError()));
  if (flag != 0) return fail(HF_NUMERICAL_FAILURE, kDetailCoefficients);
  if (p->policy.null_space_constant && !(target[0]->volume > 0.0))
    return fail(HF_NUMERICAL_FAILURE, kDetailCoefficients);
  // every check passed: keep the raw activity masks for the adjoint, then
  // publish
  if (all) {
    if (p->fxa == nullptr) {
      HF_TRY(dalloc(p, &p->fxa, L0.nfx));
      HF_TRY(dalloc(p, &p->fya, L0.nfy));
      HF_TRY(dalloc(p, &p->fza, L0.nfz));
    }
    HF_TRY(hip_check(hipMemcpyAsync(p->fxa, fxa, L0.nfx, hipMemcpyDeviceToDevice, p->stream)));
    HF_TRY(hip_check(hipMemcpyAsync(p->fya, fya, L0.nfy, hipMemcpyDeviceToDevice, p->stream)));
    HF_TRY(hip_check(hipMemcpyAsync(p->fza, fza, L0.nfz, hipMemcpyDeviceToDevice, p->stream)));
  }
  p->has_activity = all;
  p->shortcut = shortcut;
  if (into_spare)
    for (Level& L : p->levels) std::swap(static_cast<Coeffs&>(L), L.spare);
  // the tail kernel reads its level descriptors from device memory; a plan
  // whose descriptors could not follow the swap accepts no further solves
  const hf_status uploaded = upload_tail(p);
  p->coefficients_ready = uploaded.code == HF_OK;
  HF_TRY(uploaded);
  return hip_check(hipGetLastError());
}

hf_status hf_operator_apply(hf_plan* p, const double* x, double* y) {
  if (p == nullptr || x == nullptr || y == nullptr || !p->coefficients_ready)
    return fail(HF_INVALID_PLAN, kDetailPlan);
  const Tiles& t = p->levels[0].tz;
  HF_LAUNCH((k_stencil_z<double, kZOpDot>), t.blocks, kBlock, 0, p->stream, t,
            krylov_view(p), p->bc, x, static_cast<const double*>(nullptr), y,
            static_cast<double*>(nullptr), 0);
  return hip_check(hipGetLastError());
}

hf_status hf_mg_apply(hf_plan* p, const double* r, double* z) {
  if (p == nullptr || r == nullptr || z == nullptr || !p->coefficients_ready)
    return fail(HF_INVALID_PLAN, kDetailPlan);
  HF_TRY(mg_apply(p, r, z, nullptr, false));
  HF_TRY(hip_check(hipGetLastError()));
  // line relaxation reports a vanished pivot through the scalar block
  if (p->levels[0].line_mask != 0) return read_scalars(p);
  return ok();
}

static hf_status pcg_impl(hf_plan* p, const double* rhs, double* solution,
                          const hf_solve_control* ctl, hf_solve_result* res) {
  const Lv<double> A = krylov_view(p);
  const Tiles& tr = p->levels[0].tz;  // z-marching Krylov stencils
  // the operator over pairs of cells of a row on rows of even length
  const bool op_pairs = p->krylov_pairs && p->levels[0].n[0] % 2 == 0;
  const Tiles& L0p = p->levels[0].tzp;
  // the PCG update stores the preconditioner's entry when level 0 is a grid
  // level (the tail and line-relaxed entries also zero x; a Chebyshev level 0
  // reads no x in its first stage, and mg_apply projects its b)
  const bool fuse_entry = !p->levels[0].tail &&
                          p->levels[0].line_mask == 0 &&
                          (!p->levels[0].cheb || p->cheb_first);
  const Chunk ch = p->chunk0;
  const int nbv = chunk_blocks(ch);
  double* x = p->kx;
  double* r = p->kr;
  double* pp = p->kp;
  double* q = p->kq;
  const uint64_t red0 = p->reductions;
  auto finish = [&](hf_status st, int term) {
    res->status = st;
    res->termination = term;
    res->reduction_calls = p->reductions - red0;
    return st;
  };
  // r = b - A x with r.r in accumulator 0 and, with_bb, b.b in accumulator
  // 2.  On a grid level 0 the pass also makes the preconditioner's entry
  // from r (accumulator 1), which the restart that follows an unconverged
  // check then skips.
  const int stride = static_cast<int>(p->partial_stride);
  auto true_residual = [&](bool with_bb) -> hf_status {
    HF_PROF(p, P_TRUE_RESIDUAL);
    if (!fuse_entry) {
      HF_LAUNCH((k_stencil_z<double, kZResidual>), tr.blocks, kBlock, 0,
                p->stream, tr, A, p->bc, static_cast<const double*>(x), rhs,
                r, p->partials, with_bb ? 2 * stride : 0);
      return ok();
    }
    Level& L0 = p->levels[0];
    dispatch(L0.prec, [&](auto t) {
      using T = decltype(t);
      HF_LAUNCH(k_residual_entry<T>, tr.blocks, kBlock, 0, p->stream, tr, A,
                p->bc, static_cast<const double*>(x), rhs, r,
                static_cast<const uint8_t*>(L0.act), widths(p),
                static_cast<T*>(L0.b), p->partials, stride,
                p->policy.null_space_constant != 0 ? 1 : 0, with_bb ? 1 : 0);
    });
    return ok();
  };
  // ||b|| and the initial true residual come from one pass and one read; the
  // gates then run in the reference's order (a non-finite ||b||, a zero
  // ||b||, whose residual goes unused, then the residual)
  HF_LAUNCH(k_copy, nbv, kBlock, 0, p->stream, ch, x, solution);
  HF_TRY(true_residual(true));
  finalize(p, tr.blocks, 1, F_STORE, S_SUM, 0.0, 1, 2);
  finalize(p, tr.blocks, 1, F_STORE, S_RNORM2);
  HF_TRY(read_scalars(p));
  double bnorm = std::sqrt(p->host_scalars[S_SUM]);
  if (!std::isfinite(bnorm)) {
    // the squares overflowed, or b holds a non-finite value
    bool finite = false;
    HF_TRY(stable_norm(p, rhs, &bnorm, &finite));
    if (!finite) bnorm = std::numeric_limits<double>::quiet_NaN();
  }
  res->rhs_norm = bnorm;
  if (!std::isfinite(bnorm))
    return finish(fail(HF_NUMERICAL_FAILURE, kDetailNonFinite), HF_TERM_NON_FINITE);
  const double tol = std::max(ctl->absolute_tolerance, ctl->relative_tolerance * bnorm);
  res->true_residual_limit = tol;
  if (bnorm == 0.0) {
    HF_LAUNCH(k_fill, nbv, kBlock, 0, p->stream, ch, solution, 0.0);
    return finish(ok(), HF_TERM_ZERO_RHS);
  }
  res->operator_applies++;
  res->initial_true_residual = std::sqrt(p->host_scalars[S_RNORM2]);
  if (!std::isfinite(res->initial_true_residual)) {
    bool finite = false;
    HF_TRY(stable_norm(p, r, &res->initial_true_residual, &finite));
    if (!finite)
      res->initial_true_residual = std::numeric_limits<double>::quiet_NaN();
  }
  res->final_true_residual = res->initial_true_residual;
  res->recursive_residual = res->initial_true_residual;
  if (!std::isfinite(res->initial_true_residual))
    return finish(fail(HF_NUMERICAL_FAILURE, kDetailNonFinite), HF_TERM_OPERATOR_FAILURE);
  if (res->initial_true_residual <= tol) {
    HF_LAUNCH(k_copy, nbv, kBlock, 0, p->stream, ch, solution, x);
    return finish(ok(), HF_TERM_CONVERGED);
  }
  // p = z = M r: the preconditioner writes z straight into p (nothing else
  // reads this z)
  auto precondition_restart = [&]() -> hf_status {
    HF_TRY(mg_apply(p, r, pp, nullptr, true, fuse_entry, nullptr, tr.blocks));
    res->preconditioner_applies++;
    finalize(p, p->levels[0].tred.blocks, 1, F_RHO, 0);
    return ok();
  };
  HF_TRY(precondition_restart());
  HF_TRY(read_scalars(p));
  {
    const double rho = p->host_scalars[S_RHO];
    if (!(rho > 0.0) || !std::isfinite(rho))
      return finish(fail(HF_NUMERICAL_FAILURE, kDetailBreakdown), HF_TERM_BREAKDOWN);
  }
  // r.z of the latest preconditioner application is tested at the next
  // host read (the reference tests it before the next operator apply,
  // src/solver_krylov.cpp:1520-1544); a failing value reports the breakdown
  // with the counts of the reference, which stopped before the operator
  // apply and the update issued since.
  bool rho_pending = false;
  // x += alpha p of the current iteration left to the direction update (or
  // to the check that needs x first)
  bool x_pending = false;
  while (res->iterations < ctl->maximum_iterations) {
    const uint64_t operator_applies0 = res->operator_applies;
    const uint32_t iterations0 = res->iterations;
    auto rho_failed = [&]() -> bool {
      if (!rho_pending) return false;
      rho_pending = false;
      const double rho = p->host_scalars[S_RHO];
      if (rho > 0.0 && std::isfinite(rho)) return false;
      res->operator_applies = operator_applies0;
      res->iterations = iterations0;
      return true;
    };
    auto fail_rho = [&]() {
      return finish(fail(HF_NUMERICAL_FAILURE, kDetailBreakdown),
                    HF_TERM_BREAKDOWN);
    };
    if (op_pairs) {
      // q = A p over pairs of cells of a row (p.q rounds in another order)
      HF_PROF(p, P_OPERATOR);
      HF_LAUNCH((k_fg_stencil2<kFgOpDot, 0>), L0p.blocks, kBlock, 0,
                p->stream, L0p, A, p->bc, static_cast<const double*>(pp),
                static_cast<const double*>(nullptr), q,
                static_cast<const double*>(nullptr), int64_t{0}, 0, 0.0, 0,
                widths(p), p->partials, int64_t{0});
    } else {
      HF_PROF(p, P_OPERATOR);
      HF_LAUNCH((k_stencil_z<double, kZOpDot>), tr.blocks, kBlock, 0,
                p->stream, tr, A, p->bc, static_cast<const double*>(pp),
                static_cast<const double*>(nullptr), q, p->partials, 0);
    }
    res->operator_applies++;
    finalize(p, op_pairs ? L0p.blocks : tr.blocks, 1, F_ALPHA, 0);
    // an iteration that verifies on schedule replaces r with the true
    // residual, so its update forms x alone and its step checks wait for the
    // true residual's read
    const bool scheduled =
        (res->iterations + 1) % ctl->true_residual_interval == 0 ||
        res->iterations + 1 == ctl->maximum_iterations;
    if (scheduled) {
      HF_PROF(p, P_UPDATE_XR);
      HF_LAUNCH(k_update_x, nbv, kBlock, 0, p->stream, ch, x,
                static_cast<const double*>(pp), p->scalars + S_ALPHA);
    } else if (fuse_entry) {
      // the update also stores the preconditioner's input b (used unless
      // this iteration verifies the true residual, which replaces r); x is
      // updated by the direction update, which reads p anyway
      HF_PROF(p, P_UPDATE_XR);
      Level& L0 = p->levels[0];
      x_pending = true;
      dispatch(L0.prec, [&](auto t) {
        using T = decltype(t);
        HF_LAUNCH(k_update_xr_entry<T>, L0.tred.blocks, kBlock, 0, p->stream,
                  L0.tred, static_cast<double*>(nullptr),
                  static_cast<const double*>(nullptr), r,
                  static_cast<const double*>(q), p->scalars + S_ALPHA,
                  static_cast<const uint8_t*>(L0.act), widths(p),
                  static_cast<T*>(L0.b), p->partials,
                  static_cast<int>(p->partial_stride),
                  p->policy.null_space_constant != 0 ? 1 : 0);
      });
    } else {
      HF_PROF(p, P_UPDATE_XR);
      HF_LAUNCH(k_update_xr, nbv, kBlock, 0, p->stream, ch, x, pp, r, q,
                p->scalars + S_ALPHA, p->partials);
    }
    if (!scheduled)
      finalize(p, fuse_entry ? p->levels[0].tred.blocks : nbv, 1, F_STORE,
               S_RNORM2);
    res->iterations++;
    // p.Ap and alpha (read with the recursive norm, or with the true one)
    auto step_failure = [&]() -> int {
      const double pap = p->host_scalars[S_PAP];
      if (!(pap > 0.0) || !std::isfinite(pap)) return HF_TERM_BREAKDOWN;
      if (!std::isfinite(p->host_scalars[S_ALPHA])) return HF_TERM_NON_FINITE;
      return 0;
    };
    auto fail_step = [&](int term) {
      return finish(fail(HF_NUMERICAL_FAILURE, term == HF_TERM_BREAKDOWN
                                                   ? kDetailBreakdown
                                                   : kDetailNonFinite),
                    term);
    };
    bool verify = scheduled;
    if (!scheduled) {
      HF_TRY(read_scalars(p));
      if (rho_failed()) return fail_rho();
      if (const int term = step_failure()) return fail_step(term);
      double recursive = std::sqrt(p->host_scalars[S_RNORM2]);
      if (!std::isfinite(recursive)) {
        bool finite = false;
        HF_TRY(stable_norm(p, r, &recursive, &finite));
        if (!finite) recursive = std::numeric_limits<double>::quiet_NaN();
      }
      if (!std::isfinite(recursive))
        return finish(fail(HF_NUMERICAL_FAILURE, kDetailNonFinite), HF_TERM_NON_FINITE);
      res->recursive_residual = recursive;
      verify = recursive <= tol;
    }
    if (verify) {
      if (x_pending) {
        HF_PROF(p, P_UPDATE_XR);
        HF_LAUNCH(k_update_x, nbv, kBlock, 0, p->stream, ch, x,
                  static_cast<const double*>(pp), p->scalars + S_ALPHA);
        x_pending = false;
      }
      HF_TRY(true_residual(false));
      finalize(p, tr.blocks, 1, F_STORE, S_RNORM2);
      HF_TRY(read_scalars(p));
      if (rho_failed()) return fail_rho();
      // a step that fails these checks stops before the true residual, as
      // in the reference, so that residual is not counted
      if (scheduled) {
        if (const int term = step_failure()) return fail_step(term);
      }
      res->operator_applies++;
      res->final_true_residual = std::sqrt(p->host_scalars[S_RNORM2]);
      if (!std::isfinite(res->final_true_residual)) {
        bool finite = false;
        HF_TRY(stable_norm(p, r, &res->final_true_residual, &finite));
        if (!finite)
          res->final_true_residual = std::numeric_limits<double>::quiet_NaN();
      }
      res->recursive_residual = res->final_true_residual;
      if (!std::isfinite(res->final_true_residual))
        return finish(fail(HF_NUMERICAL_FAILURE, kDetailNonFinite), HF_TERM_OPERATOR_FAILURE);
      if (res->final_true_residual <= tol) {
        HF_LAUNCH(k_copy, nbv, kBlock, 0, p->stream, ch, solution, x);
        return finish(ok(), HF_TERM_CONVERGED);
      }
      if (res->iterations == ctl->maximum_iterations)
        return finish(fail(HF_REJECTED_STEP, kDetailBreakdown), HF_TERM_MAXIMUM_ITERATIONS);
      HF_TRY(precondition_restart());
      rho_pending = true;
      continue;
    }
    // z is formed again by the direction update instead of being stored
    ExitOperands ex;
    HF_TRY(mg_apply(p, r, nullptr, ctl->flexible ? q : nullptr, true,
                    fuse_entry, &ex));
    res->preconditioner_applies++;
    rho_pending = true;
    finalize(p, p->levels[0].tred.blocks, ctl->flexible ? 2 : 1,
             ctl->flexible ? F_BETA_PR : F_BETA_FR, 0);
    {
      HF_PROF(p, P_UPDATE_P);
      Level& L0 = p->levels[0];
      dispatch(L0.prec, [&](auto t) {
        using T = decltype(t);
        HF_LAUNCH(k_update_p_exit<T>, L0.tred.blocks, kBlock, 0, p->stream,
                  L0.tred, static_cast<const T*>(ex.x),
                  static_cast<const uint8_t*>(L0.act), ex.mean, ex.alpha, pp,
                  static_cast<const double*>(p->scalars + S_BETA),
                  x_pending ? x : static_cast<double*>(nullptr),
                  static_cast<const double*>(p->scalars + S_ALPHA));
      });
      x_pending = false;
    }
  }
  return finish(fail(HF_REJECTED_STEP, kDetailBreakdown), HF_TERM_MAXIMUM_ITERATIONS);
}

hf_status hf_pcg_solve(hf_plan* p, const double* rhs, double* solution,
                       const hf_solve_control* ctl, hf_solve_result* res) {
  if (res != nullptr) std::memset(res, 0, sizeof *res);
  if (p == nullptr || rhs == nullptr || solution == nullptr || ctl == nullptr ||
      res == nullptr || !p->coefficients_ready)
    return fail(HF_INVALID_PLAN, kDetailPlan);
  if (!std::isfinite(ctl->absolute_tolerance) ||
      !std::isfinite(ctl->relative_tolerance) || ctl->absolute_tolerance < 0.0 ||
      ctl->relative_tolerance < 0.0 ||
      (ctl->absolute_tolerance == 0.0 && ctl->relative_tolerance == 0.0) ||
      ctl->maximum_iterations == 0 || ctl->true_residual_interval == 0) {
    res->status = fail(HF_INVALID_PLAN, kDetailPlan);
    res->termination = HF_TERM_INVALID_PLAN;
    return res->status;
  }
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  (void)hipEventRecord(e0, p->stream);
  hf_status st = pcg_impl(p, rhs, solution, ctl, res);
  if (st.code != HF_OK && res->termination == HF_TERM_CONVERGED) {
    // a failure raised outside the Krylov gates (device error, line pivot)
    res->status = st;
    res->termination = st.detail == kDetailLine ? HF_TERM_PRECONDITIONER_FAILURE
                                                : HF_TERM_INVALID_PLAN;
  }
  (void)hipEventRecord(e1, p->stream);
  (void)hipEventSynchronize(e1);
  profile_collect(p);
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  res->solve_milliseconds = ms;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  const hipError_t last = hipGetLastError();
  if (st.code == HF_OK && last != hipSuccess) {
    res->status = fail(HF_INVALID_PLAN, kDetailHip);
    return res->status;
  }
  return st;
}

hf_status hf_pcg_solve_static(hf_plan* p, const double* rhs, double* solution,
                              const hf_solve_control* ctl,
                              hf_solve_result* res) {
  if (res != nullptr) std::memset(res, 0, sizeof *res);
  if (p == nullptr || rhs == nullptr || solution == nullptr || ctl == nullptr ||
      res == nullptr || !p->coefficients_ready)
    return fail(HF_INVALID_PLAN, kDetailPlan);
  auto refuse = [&](hf_status st) {
    res->status = st;
    res->termination = HF_TERM_INVALID_PLAN;
    return st;
  };
  if (!std::isfinite(ctl->absolute_tolerance) ||
      !std::isfinite(ctl->relative_tolerance) || ctl->absolute_tolerance < 0.0 ||
      ctl->relative_tolerance < 0.0 ||
      (ctl->absolute_tolerance == 0.0 && ctl->relative_tolerance == 0.0) ||
      ctl->maximum_iterations == 0 || ctl->true_residual_interval == 0)
    return refuse(fail(HF_INVALID_PLAN, kDetailPlan));
  // a grid whose reductions need more than the static array is refused
  if (reduction_doubles(p) * static_cast<int64_t>(sizeof(double)) >
      static_cast<int64_t>(kStaticScratchBytes)) {
    std::fprintf(stderr,
                 "hundun_flow: hf_pcg_solve_static needs %lld bytes of "
                 "reduction scratch, the static array holds %zu\n",
                 static_cast<long long>(reduction_doubles(p) * 8),
                 kStaticScratchBytes);
    return refuse(fail(HF_INVALID_PLAN, kDetailScratch));
  }
  double* base = nullptr;
  {
    const hf_status st = static_scratch(p->device, &base);
    if (st.code != HF_OK) return refuse(st);
  }
  // the solve's kernels take the static array in place of the plan's
  // reduction buffers; event timing (which creates events) is off
  double* const own[3] = {p->partials, p->folded, p->scalars};
  const bool profiling = p->profiling;
  p->partials = base;
  p->folded = base + 3 * p->partial_stride;
  p->scalars = p->folded + 3 * kFoldStride;
  p->profiling = false;
  // the array keeps the previous call's values: the scalar block (the line
  // failure flag among them) starts from zero
  const Chunk sc = make_chunk(S_COUNT);
  HF_LAUNCH(k_fill, chunk_blocks(sc), kBlock, 0, p->stream, sc, p->scalars,
            0.0);
  const auto t0 = std::chrono::steady_clock::now();
  hf_status st = pcg_impl(p, rhs, solution, ctl, res);
  if (st.code != HF_OK && res->termination == HF_TERM_CONVERGED) {
    res->status = st;
    res->termination = st.detail == kDetailLine ? HF_TERM_PRECONDITIONER_FAILURE
                                                : HF_TERM_INVALID_PLAN;
  }
  const hipError_t synced = hipStreamSynchronize(p->stream);
  const auto t1 = std::chrono::steady_clock::now();
  res->solve_milliseconds =
      std::chrono::duration<double, std::milli>(t1 - t0).count();
  p->partials = own[0];
  p->folded = own[1];
  p->scalars = own[2];
  p->profiling = profiling;
  const hipError_t last = hipGetLastError();
  if (st.code == HF_OK && (synced != hipSuccess || last != hipSuccess)) {
    res->status = fail(HF_INVALID_PLAN, kDetailHip);
    return res->status;
  }
  return st;
}

hf_status hf_fgmres_solve(hf_plan* p, const double* rhs, double* solution,
                          const hf_fgmres_control* ctl,
                          hf_solve_result* res) {
  if (res != nullptr) std::memset(res, 0, sizeof *res);
  if (p == nullptr || rhs == nullptr || solution == nullptr || ctl == nullptr ||
      res == nullptr || !p->coefficients_ready)
    return fail(HF_INVALID_PLAN, kDetailPlan);
  if (!std::isfinite(ctl->absolute_tolerance) ||
      !std::isfinite(ctl->relative_tolerance) || ctl->absolute_tolerance < 0.0 ||
      ctl->relative_tolerance < 0.0 ||
      (ctl->absolute_tolerance == 0.0 && ctl->relative_tolerance == 0.0) ||
      ctl->maximum_iterations == 0 || ctl->true_residual_interval == 0 ||
      ctl->restart == 0 || ctl->restart > static_cast<uint32_t>(kFgMaxRestart)) {
    res->status = fail(HF_INVALID_PLAN, kDetailPlan);
    res->termination = HF_TERM_INVALID_PLAN;
    return res->status;
  }
  {
    const hf_status st = fg_reserve(p, static_cast<int>(ctl->restart));
    if (st.code != HF_OK) {
      res->status = st;
      res->termination = HF_TERM_INVALID_PLAN;
      return st;
    }
  }
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  (void)hipEventRecord(e0, p->stream);
  hf_status st = fgmres_impl(p, rhs, solution, ctl, res);
  if (st.code != HF_OK && res->termination == HF_TERM_CONVERGED) {
    res->status = st;
    res->termination = st.detail == kDetailLine ? HF_TERM_PRECONDITIONER_FAILURE
                                                : HF_TERM_INVALID_PLAN;
  }
  (void)hipEventRecord(e1, p->stream);
  (void)hipEventSynchronize(e1);
  profile_collect(p);
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  res->solve_milliseconds = ms;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  const hipError_t last = hipGetLastError();
  if (st.code == HF_OK && last != hipSuccess) {
    res->status = fail(HF_INVALID_PLAN, kDetailHip);
    return res->status;
  }
  return st;
}

struct hf_vector_engine {
  hipStream_t stream = nullptr;
  int64_t capacity = 0;
  int32_t slots = 0;
  double* vectors = nullptr;
  double* partials = nullptr;
  double* folded = nullptr;
  double* scalars = nullptr;  // [0] the last sum, [1] the non-finite flag
  double* host = nullptr;     // pinned mirror of scalars
};

hf_status hf_vector_engine_destroy(hf_vector_engine* e) {
  if (e == nullptr) return ok();
  (void)hipStreamSynchronize(e->stream);
  for (void* ptr : {static_cast<void*>(e->vectors),
                    static_cast<void*>(e->partials),
                    static_cast<void*>(e->folded),
                    static_cast<void*>(e->scalars)})
    if (ptr != nullptr) (void)hipFree(ptr);
  if (e->host != nullptr) (void)hipHostFree(e->host);
  delete e;
  return ok();
}

hf_status hf_vector_engine_create(int64_t capacity, int32_t slots,
                                  void* stream, hf_vector_engine** out) {
  if (out == nullptr) return fail(HF_INVALID_PLAN, kDetailPlan);
  *out = nullptr;
  // vector kernels index with 32-bit integers
  if (capacity <= 0 || slots < 0 || slots > 256 ||
      capacity > std::numeric_limits<int32_t>::max() - kBlock)
    return fail(HF_INVALID_PLAN, kDetailPlan);
  hf_vector_engine* e = new (std::nothrow) hf_vector_engine();
  if (e == nullptr) return fail(HF_ALLOCATION_FAILURE, kDetailAlloc);
  e->stream = static_cast<hipStream_t>(stream);
  e->capacity = capacity;
  e->slots = slots;
  const int64_t blocks = blocks_for(capacity);
  bool good =
      (slots == 0 ||
       hipMalloc(reinterpret_cast<void**>(&e->vectors),
                 static_cast<size_t>(slots) * capacity * sizeof(double)) ==
           hipSuccess) &&
      hipMalloc(reinterpret_cast<void**>(&e->partials),
                static_cast<size_t>(blocks) * sizeof(double)) == hipSuccess &&
      hipMalloc(reinterpret_cast<void**>(&e->folded),
                kFoldStride * sizeof(double)) == hipSuccess &&
      hipMalloc(reinterpret_cast<void**>(&e->scalars), 2 * sizeof(double)) ==
          hipSuccess &&
      hipHostMalloc(reinterpret_cast<void**>(&e->host), 2 * sizeof(double),
                    hipHostMallocDefault) == hipSuccess;
  if (good)
    good = hipMemsetAsync(e->scalars, 0, 2 * sizeof(double), e->stream) ==
               hipSuccess &&
           hipStreamSynchronize(e->stream) == hipSuccess;
  if (!good) {
    hf_vector_engine_destroy(e);
    return fail(HF_ALLOCATION_FAILURE, kDetailAlloc);
  }
  *out = e;
  return ok();
}

hf_status hf_vector_engine_slot(hf_vector_engine* e, int32_t slot,
                                double** vector) {
  if (e == nullptr || vector == nullptr || slot < 0 || slot >= e->slots)
    return fail(HF_INVALID_PLAN, kDetailPlan);
  *vector = e->vectors + static_cast<int64_t>(slot) * e->capacity;
  return ok();
}

hf_status hf_vector_dot(hf_vector_engine* e, int64_t n, const double* a,
                        const double* b, double* result) {
  if (e == nullptr || a == nullptr || b == nullptr || result == nullptr ||
      n <= 0 || n > e->capacity)
    return fail(HF_INVALID_PLAN, kDetailPlan);
  const Chunk ch = make_chunk(static_cast<int>(n));
  const int64_t blocks = chunk_blocks(ch);
  HF_LAUNCH(k_dot, blocks, kBlock, 0, e->stream, ch, a, b, e->partials);
  const double* src = e->partials;
  int64_t count = blocks;
  if (blocks > kFoldSpan) {
    const int64_t spans = (blocks + kFoldSpan - 1) / kFoldSpan;
    hipLaunchKernelGGL(k_fold_partials, dim3(static_cast<uint32_t>(spans), 1),
                       dim3(kBlock), 0, e->stream, e->partials, blocks,
                       blocks, e->folded, static_cast<int64_t>(kFoldStride));
    src = e->folded;
    count = spans;
  }
  HF_LAUNCH(k_finalize, 1, 1024, 0, e->stream, src, count,
            static_cast<int64_t>(kFoldStride), 1, e->scalars,
            static_cast<int>(F_STORE), 0, 0.0, 1);
  if (hipMemcpyAsync(e->host, e->scalars, 2 * sizeof(double),
                     hipMemcpyDeviceToHost, e->stream) != hipSuccess ||
      hipStreamSynchronize(e->stream) != hipSuccess)
    return fail(HF_INVALID_PLAN, kDetailHip);
  if (e->host[1] != 0.0) {
    // a non-finite vector value since the last sum
    const Chunk one = make_chunk(2);
    HF_LAUNCH(k_fill, 1, kBlock, 0, e->stream, one, e->scalars, 0.0);
    return fail(HF_NUMERICAL_FAILURE, kDetailNonFinite);
  }
  *result = e->host[0];
  return hip_check(hipGetLastError());
}

hf_status hf_vector_axpby(hf_vector_engine* e, int64_t n, double a,
                          const double* left, double b, const double* right,
                          double* dest) {
  if (e == nullptr || left == nullptr || right == nullptr || dest == nullptr ||
      n <= 0 || n > e->capacity)
    return fail(HF_INVALID_PLAN, kDetailPlan);
  const Chunk ch = make_chunk(static_cast<int>(n));
  HF_LAUNCH(k_axpby, chunk_blocks(ch), kBlock, 0, e->stream, ch, a, left, b,
            right, dest, e->scalars + 1);
  return hip_check(hipGetLastError());
}

hf_status hf_vector_copy(hf_vector_engine* e, int64_t n, const double* source,
                         double* destination) {
  if (e == nullptr || source == nullptr || destination == nullptr || n <= 0 ||
      n > e->capacity)
    return fail(HF_INVALID_PLAN, kDetailPlan);
  const Chunk ch = make_chunk(static_cast<int>(n));
  HF_LAUNCH(k_copy_checked, chunk_blocks(ch), kBlock, 0, e->stream, ch,
            destination, source, e->scalars + 1);
  return hip_check(hipGetLastError());
}

hf_status hf_project_mean(hf_plan* p, double* field) {
  if (p == nullptr || field == nullptr || !p->coefficients_ready)
    return fail(HF_INVALID_PLAN, kDetailPlan);
  const Chunk ch = p->chunk0;
  const int nb = chunk_blocks(ch);
  const uint8_t* act = p->levels[0].act;
  HF_LAUNCH(k_count, nb, kBlock, 0, p->stream, ch, act, p->partials);
  finalize(p, nb, 1, F_STORE, S_SUM2);
  HF_TRY(read_scalars(p));
  const double count = p->host_scalars[S_SUM2];
  if (!(count > 0.0)) return fail(HF_NUMERICAL_FAILURE, kDetailNumeric);
  HF_LAUNCH(k_sum, nb, kBlock, 0, p->stream, ch, field, act, p->partials);
  finalize(p, nb, 1, F_MEAN, S_SUM, count);
  HF_LAUNCH(k_subtract_mean, nb, kBlock, 0, p->stream, ch, field,
            p->scalars + S_SUM, act);
  return hip_check(hipGetLastError());
}

hf_status hf_adjoint_gradients(hf_plan* p, const double* mobility,
                               const double* pressure, const double* grad_p,
                               const hf_solve_control* ctl, double* grad_b,
                               double* grad_logrho, hf_solve_result* res) {
  if (res != nullptr) std::memset(res, 0, sizeof *res);
  if (p == nullptr || pressure == nullptr || grad_p == nullptr || ctl == nullptr ||
      res == nullptr || !p->coefficients_ready ||
      (grad_logrho != nullptr && mobility == nullptr))
    return fail(HF_INVALID_PLAN, kDetailPlan);
  const int64_t N = p->cells;
  const int64_t nb = blocks_for(N);
  const Chunk ch = p->chunk0;
  const int nbv = chunk_blocks(ch);
  if (p->adj_rhs == nullptr) {
    HF_TRY(dalloc(p, &p->adj_rhs, N));
    HF_TRY(dalloc(p, &p->adj_lambda, N));
  }
  // adjoint rhs: P grad_p (solids zero, zero fluid mean)
  HF_LAUNCH(k_copy, nbv, kBlock, 0, p->stream, ch, p->adj_rhs, grad_p);
  HF_TRY(hf_project_mean(p, p->adj_rhs));
  HF_TRY(hip_check(hipMemsetAsync(p->adj_lambda, 0, N * sizeof(double), p->stream)));
  const hf_status st = hf_pcg_solve(p, p->adj_rhs, p->adj_lambda, ctl, res);
  if (st.code != HF_OK) return st;
  HF_TRY(hf_project_mean(p, p->adj_lambda));
  if (grad_b != nullptr)
    HF_LAUNCH(k_copy, nbv, kBlock, 0, p->stream, ch, grad_b, p->adj_lambda);
  if (grad_logrho != nullptr) {
    HF_LAUNCH(k_grad_logrho, nb, kBlock, 0, p->stream, axis_geom(p), mobility,
              pressure, p->adj_lambda,
              p->has_activity ? p->fxa : static_cast<uint8_t*>(nullptr),
              p->has_activity ? p->fya : static_cast<uint8_t*>(nullptr),
              p->has_activity ? p->fza : static_cast<uint8_t*>(nullptr),
              grad_logrho);
  }
  return hip_check(hipGetLastError());
}

hf_status hf_plan_level_count(const hf_plan* p, int32_t* levels) {
  if (p == nullptr || levels == nullptr) return fail(HF_INVALID_PLAN, kDetailPlan);
  *levels = static_cast<int32_t>(p->levels.size());
  return ok();
}

hf_status hf_plan_level_info(const hf_plan* p, int32_t level,
                             hf_level_info* info) {
  if (p == nullptr || info == nullptr || level < 0 ||
      level >= static_cast<int32_t>(p->levels.size()))
    return fail(HF_INVALID_PLAN, kDetailPlan);
  const Level& L = p->levels[level];
  info->nx = L.n[0];
  info->ny = L.n[1];
  info->nz = L.n[2];
  info->coarsen_mask = static_cast<uint8_t>(L.coarsen_mask);
  info->line_mask = static_cast<uint8_t>(L.line_mask);
  info->precision = static_cast<uint8_t>(L.prec);
  info->on_tail = L.tail ? 1 : 0;
  return ok();
}

hf_status hf_plan_level_array(const hf_plan* p, int32_t level, int32_t which,
                              double* out) {
  if (p == nullptr || out == nullptr || level < 0 ||
      level >= static_cast<int32_t>(p->levels.size()) || which < 0 || which > 4)
    return fail(HF_INVALID_PLAN, kDetailPlan);
  const Level& L = p->levels[level];
  const double* src[5] = {L.D64, L.fx64, L.fy64, L.fz64, L.vol};
  const int64_t count[5] = {L.cells, L.nfx, L.nfy, L.nfz, L.cells};
  if (hipMemcpyAsync(out, src[which], count[which] * sizeof(double),
                     hipMemcpyDeviceToDevice, p->stream) != hipSuccess)
    return fail(HF_INVALID_PLAN, kDetailHip);
  return hip_check(hipStreamSynchronize(p->stream));
}

hf_status hf_plan_device_bytes(const hf_plan* p, uint64_t* bytes) {
  if (p == nullptr || bytes == nullptr) return fail(HF_INVALID_PLAN, kDetailPlan);
  *bytes = p->device_bytes;
  return ok();
}

hf_status hf_plan_set_profiling(hf_plan* p, int32_t enabled) {
  if (p == nullptr) return fail(HF_INVALID_PLAN, kDetailPlan);
  HF_TRY(hip_check(hipStreamSynchronize(p->stream)));
  profile_collect(p);
  p->profiling = enabled != 0;
  p->event_pairs = 0;
  for (int c = 0; c < kProfileClasses; ++c) {
    p->class_ms[c] = 0.0;
    p->class_launches[c] = 0;
  }
  return ok();
}

hf_status hf_plan_profile(const hf_plan* p, double* ms, uint64_t* launches,
                          int32_t capacity, int32_t* count) {
  if (p == nullptr || count == nullptr) return fail(HF_INVALID_PLAN, kDetailPlan);
  *count = kProfileClasses;
  for (int c = 0; c < kProfileClasses && c < capacity; ++c) {
    if (ms != nullptr) ms[c] = p->class_ms[c];
    if (launches != nullptr) launches[c] = p->class_launches[c];
  }
  return ok();
}

const char* hf_status_message(hf_status s) {
  switch (s.code) {
    case HF_OK: return "ok";
    case HF_INVALID_CASE: return "invalid case";
    case HF_INVALID_PLAN: return "invalid plan";
    case HF_ALLOCATION_FAILURE: return "allocation failure";
    case HF_MPI_FAILURE: return "mpi failure";
    case HF_NUMERICAL_FAILURE: return "numerical failure";
    case HF_REJECTED_STEP: return "rejected step";
    case HF_IO_FAILURE: return "io failure";
    default: return "unknown status";
  }
}

}  // extern "C"
