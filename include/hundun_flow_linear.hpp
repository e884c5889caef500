// hundun_flow_linear.hpp -- C++ drop-in layer over the C ABI (hundun_flow.h).
//
// Mirrors the reference solver_* interface the pressure path is called
// through (versions/v0.4/include/hundun/v04_linear.hpp of the reference):
//   LinearOperator          (:396-405)   -> HipPressureOperator
//   LinearPreconditioner    (:418-452)   -> HipCartesianMgPlan (apply and the
//                                           prepared-batch lifecycle)
//   NativeCartesianMgPlan::compile / update_coefficients (:952-1013)
//   solve_pcg / solve_fgmres (:1070-1082) -> solve_pcg / solve_fgmres below
//   VolumeScaledPreconditioner (src/solver_cold.hpp:378-415) and the
//   per-volume cold rows -> HipVolumeScaledPreconditioner /
//   HipVolumeRowsOperator
// with the same names, argument meaning and Status error convention
// (include/hundun/v04_status.hpp:13-31): nothing throws, every call returns
// Status{code, detail}, failures never publish a solution.
//
// A solve whose operator and preconditioner are this library's pair over one
// plan runs entirely on the device (hf_pcg_solve / hf_fgmres_solve).  Any
// other pair (a caller's own LinearOperator or LinearPreconditioner) runs the
// same Krylov recurrence from the host, calling the virtual applies and the
// device vector engine of the SolverWorkspace for its vectors and sums.
//
// Fields are device-resident: DeviceFieldView carries a device pointer and
// the interior shape (no ghosts; the library fills halos on device).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <utility>
#include <vector>

#include "hundun_flow.h"

namespace hundun::hip {

enum class StatusCode : std::uint16_t {
  ok,
  invalid_case,
  invalid_plan,
  allocation_failure,
  mpi_failure,
  numerical_failure,
  rejected_step,
  io_failure
};

struct Status {
  StatusCode code{StatusCode::ok};
  std::uint32_t detail{};
  constexpr explicit operator bool() const noexcept {
    return code == StatusCode::ok;
  }
};

inline Status from_c(hf_status s) noexcept {
  return {static_cast<StatusCode>(s.code), s.detail};
}

struct Int3 {
  std::int32_t x{}, y{}, z{};
};

template <class T>
struct BasicDeviceFieldView {
  T* base{};
  Int3 interior{};
};
using DeviceFieldView = BasicDeviceFieldView<double>;
using ConstDeviceFieldView = BasicDeviceFieldView<const double>;

enum class LinearOperatorClass : std::uint8_t { spd, nonsymmetric };
enum class LinearPreconditionerClass : std::uint8_t {
  fixed_spd,
  fixed_general,
  flexible
};
enum class LinearTermination : std::uint8_t {
  converged,
  zero_rhs,
  maximum_iterations,
  breakdown,
  non_finite,
  operator_failure,
  preconditioner_failure,
  collective_failure,
  invalid_plan,
  convergence_audit_failure
};
enum class MgBoundaryKind : std::uint8_t { dirichlet, neumann, periodic };
enum class MgNullSpace : std::uint8_t { none, constant };
enum class MgCycleKind : std::uint8_t { v_cycle, f_cycle };
enum class MgCorrectionScaling : std::uint8_t {
  residual_minimizing,
  unit_linear
};
enum class MgPointSmootherKind : std::uint8_t { red_black, chebyshev_jacobi };
enum class MgStoragePrecision : std::uint8_t { fp64, fp32, bf16 };
enum class MgLineOrder : std::uint8_t { lexicographic, zebra };

struct MgHierarchyPolicy {
  double anisotropy_threshold{4.0};
  std::uint8_t pre_sweeps{2U};
  std::uint8_t post_sweeps{2U};
  std::uint8_t maximum_levels{32U};
  std::uint16_t coarse_sweeps{24U};
  std::uint8_t minimum_coarse_extent{3U};
  MgPointSmootherKind point_smoother{MgPointSmootherKind::red_black};
  MgCycleKind cycle{MgCycleKind::v_cycle};
  double chebyshev_lower_spectrum_fraction{0.3};
  // zero selects the ordinary cycle; otherwise D^-1 when every finest row's
  // face sum over its diagonal is at most this value (< 1)
  double diagonal_shortcut_maximum_ratio{0.0};
  // not in the reference; the defaults reproduce the reference hierarchy
  MgStoragePrecision precision{MgStoragePrecision::fp64};
  bool distance_corrected_coarse_faces{false};
  MgLineOrder line_order{MgLineOrder::lexicographic};
};

struct MgBoundarySet {
  MgBoundaryKind x_min{MgBoundaryKind::dirichlet};
  MgBoundaryKind x_max{MgBoundaryKind::dirichlet};
  MgBoundaryKind y_min{MgBoundaryKind::dirichlet};
  MgBoundaryKind y_max{MgBoundaryKind::dirichlet};
  MgBoundaryKind z_min{MgBoundaryKind::dirichlet};
  MgBoundaryKind z_max{MgBoundaryKind::dirichlet};
};

// Explicit face coordinates per axis (runtime axes) on the host.
struct CartesianAxes {
  Int3 cells{};
  const double* faces_x{};
  const double* faces_y{};
  const double* faces_z{};
};

// Device spans of 0/1 activity bytes; all empty = full Cartesian domain.
struct MgDomainActivityView {
  const std::uint8_t* cells{};
  const std::uint8_t* x_faces{};
  const std::uint8_t* y_faces{};
  const std::uint8_t* z_faces{};
};

// Device coefficient arrays: D (cells) and the three face arrays.
struct MgCoefficientViews {
  const double* diagonal{};
  const double* x{};
  const double* y{};
  const double* z{};
};

struct NativeCartesianMgSpec {
  CartesianAxes geometry{};
  MgBoundarySet boundaries{};
  MgNullSpace null_space{MgNullSpace::none};
  MgHierarchyPolicy policy{};
  MgCorrectionScaling correction_scaling{
      MgCorrectionScaling::residual_minimizing};
  MgDomainActivityView activity{};
  void* stream{};  // hipStream_t; null = default stream
};

// Runtime services of a plan (MgRuntimeServices, v04_linear.hpp:344-351):
// on one GPU the halo engines and the reduction engine are the plan's own
// device kernels, so the only service is the stream the plan runs on.
struct MgRuntimeServices {
  void* stream{};  // hipStream_t; null = the spec's stream
};

// MgPlanCounters (v04_linear.hpp:145-156), the fields a single-GPU plan has.
struct MgPlanCounters {
  std::uint64_t symbolic_builds{};
  std::uint64_t numeric_refreshes{};
  std::uint64_t hierarchy_rebuilds{};
  std::uint64_t applications{};
};

struct LinearOperatorCertificate {
  Int3 local_shape{};
  LinearOperatorClass operator_class{LinearOperatorClass::nonsymmetric};
};

// LinearOperatorFailureProvenance (v04_linear.hpp:387-394): the status of the
// most recent failed apply; one GPU is one rank.
struct LinearOperatorFailureProvenance {
  Status status{};
  int lowest_failing_rank{-1};
};

struct LinearPreconditionerCertificate {
  LinearPreconditionerClass preconditioner_class{
      LinearPreconditionerClass::fixed_general};
};

// Prepared-batch lifecycle (v04_linear.hpp:418-452): a solver declares the
// shape and the number of applications it will make; the ticket the
// preconditioner issues binds those applications to its current numeric
// generation, and apply_prepared refuses a stale or foreign ticket.
struct LinearPreconditionerBatchDescriptor {
  Int3 shape{};
  std::uint32_t maximum_applications{};
};

struct LinearPreconditionerBatchTicket {
  std::uintptr_t owner{};
  Int3 shape{};
  std::uint32_t maximum_applications{};
  std::uint64_t generation{};
};

class LinearOperator {
 public:
  virtual LinearOperatorCertificate certificate() const noexcept = 0;
  virtual Status apply(DeviceFieldView x, DeviceFieldView y) const noexcept = 0;
  virtual LinearOperatorFailureProvenance failure_provenance() const noexcept {
    return {};
  }
  virtual ~LinearOperator() = default;
};

class LinearPreconditioner {
 public:
  virtual LinearPreconditionerCertificate certificate() const noexcept = 0;
  virtual Status apply(ConstDeviceFieldView input, DeviceFieldView output,
                       std::uint32_t iteration) noexcept = 0;
  virtual Status prepare_batch(const LinearPreconditionerBatchDescriptor&,
                               LinearPreconditionerBatchTicket&) noexcept {
    return {StatusCode::invalid_plan, 603U};
  }
  virtual Status apply_prepared(ConstDeviceFieldView, DeviceFieldView,
                                std::uint32_t,
                                const LinearPreconditionerBatchTicket&) noexcept {
    return {StatusCode::invalid_plan, 603U};
  }
  virtual ~LinearPreconditioner() = default;

 protected:
  static void issue_batch_ticket(
      LinearPreconditionerBatchTicket& ticket, const LinearPreconditioner* owner,
      const LinearPreconditionerBatchDescriptor& descriptor,
      std::uint64_t generation) noexcept {
    ticket.owner = reinterpret_cast<std::uintptr_t>(owner);
    ticket.shape = descriptor.shape;
    ticket.maximum_applications = descriptor.maximum_applications;
    ticket.generation = generation;
  }
};

struct LinearSolveControl {
  double absolute_tolerance{};
  double relative_tolerance{};
  std::uint32_t maximum_iterations{};
  std::uint32_t true_residual_interval{};
  std::uint32_t restart{};  // FGMRES basis length
  // not in the reference: Polak-Ribiere beta for inexact (mixed-precision) cycles
  bool flexible{};
};

struct LinearSolveInvocation {
  ConstDeviceFieldView rhs{};
  DeviceFieldView solution{};
  LinearSolveControl control{};
};

struct LinearSolveResult {
  Status status{};
  LinearTermination termination{LinearTermination::invalid_plan};
  std::uint32_t iterations{};
  double initial_true_residual{};
  double final_true_residual{};
  double recursive_residual{};
  std::uint64_t reduction_calls{};
  std::uint64_t operator_applies{};
  std::uint64_t preconditioner_applies{};
  std::uint64_t norm_breakdown_restarts{};
  double rhs_norm{};
  double true_residual_limit{};
};

// ResourceCounters: what a solve consumed.
struct ResourceCounters {
  std::uint64_t operator_applies{};
  std::uint64_t preconditioner_applies{};
  std::uint64_t reductions{};
};

// ReductionEngine (src/solver_reduction.cpp): on one GPU every sum is a
// device reduction of the plan or of the workspace's vector engine; the
// engine counts the reductions a solve made and holds the consensus of the
// last status (one rank: the status itself).
class ReductionEngine {
 public:
  Status consensus(Status local) noexcept {
    last_ = local;
    return local;
  }
  int lowest_failing_rank() const noexcept { return last_ ? -1 : 0; }
  std::uint64_t calls() const noexcept { return calls_; }
  void count(std::uint64_t n) noexcept { calls_ += n; }

 private:
  Status last_{};
  std::uint64_t calls_{};
};

// SolverWorkspace: the Krylov vectors of the host-driven solves, held by the
// library's device vector engine (PCG 5 vectors, FGMRES 2 restart + 3).
class SolverWorkspace {
 public:
  SolverWorkspace() noexcept = default;
  ~SolverWorkspace() noexcept { release(); }
  SolverWorkspace(const SolverWorkspace&) = delete;
  SolverWorkspace& operator=(const SolverWorkspace&) = delete;
  SolverWorkspace(SolverWorkspace&& o) noexcept
      : engine_(std::exchange(o.engine_, nullptr)), shape_(o.shape_),
        maximum_restart_(o.maximum_restart_) {}

  static Status bind(Int3 shape, std::uint32_t maximum_restart, void* stream,
                     SolverWorkspace& out) noexcept {
    const std::int64_t n = static_cast<std::int64_t>(shape.x) * shape.y *
                           shape.z;
    if (shape.x <= 0 || shape.y <= 0 || shape.z <= 0 || maximum_restart > 64U)
      return {StatusCode::invalid_plan, 602U};
    const std::int32_t slots =
        std::max<std::int32_t>(5, 2 * static_cast<std::int32_t>(maximum_restart) + 3);
    hf_vector_engine* e = nullptr;
    const Status st = from_c(hf_vector_engine_create(n, slots, stream, &e));
    if (!st) return st;
    out.release();
    out.engine_ = e;
    out.shape_ = shape;
    out.maximum_restart_ = maximum_restart;
    return {};
  }
  hf_vector_engine* engine() const noexcept { return engine_; }
  Int3 shape() const noexcept { return shape_; }
  std::uint32_t maximum_restart() const noexcept { return maximum_restart_; }
  double* vector(std::int32_t slot) const noexcept {
    double* v = nullptr;
    if (engine_ != nullptr) hf_vector_engine_slot(engine_, slot, &v);
    return v;
  }

 private:
  void release() noexcept {
    if (engine_ != nullptr) hf_vector_engine_destroy(engine_);
    engine_ = nullptr;
  }
  hf_vector_engine* engine_{};
  Int3 shape_{};
  std::uint32_t maximum_restart_{};
};

class HipCartesianMgPlan;

// Exact IBM pressure operator over a compiled plan's level-0 coefficients.
class HipPressureOperator final : public LinearOperator {
 public:
  explicit HipPressureOperator(const HipCartesianMgPlan& plan) noexcept;
  LinearOperatorCertificate certificate() const noexcept override;
  Status apply(DeviceFieldView x, DeviceFieldView y) const noexcept override;
  LinearOperatorFailureProvenance failure_provenance() const noexcept override {
    return failure_;
  }
  hf_plan* handle() const noexcept { return plan_; }

 private:
  hf_plan* plan_{};
  Int3 shape_{};
  mutable LinearOperatorFailureProvenance failure_{};
};

// NativeCartesianMgPlan counterpart (move-only, owns the device hierarchy).
class HipCartesianMgPlan final : public LinearPreconditioner {
 public:
  HipCartesianMgPlan() noexcept = default;
  ~HipCartesianMgPlan() noexcept { release(); }
  HipCartesianMgPlan(const HipCartesianMgPlan&) = delete;
  HipCartesianMgPlan& operator=(const HipCartesianMgPlan&) = delete;
  HipCartesianMgPlan(HipCartesianMgPlan&& o) noexcept
      : plan_(std::exchange(o.plan_, nullptr)), shape_(o.shape_),
        spec_(o.spec_), generation_(o.generation_), counters_(o.counters_) {}
  HipCartesianMgPlan& operator=(HipCartesianMgPlan&& o) noexcept {
    if (this != &o) {
      release();
      plan_ = std::exchange(o.plan_, nullptr);
      shape_ = o.shape_;
      spec_ = o.spec_;
      generation_ = o.generation_;
      counters_ = o.counters_;
    }
    return *this;
  }

  // compile (v04_linear.hpp:960-966): plan levels, allocate, build the
  // hierarchy from coefficients.  A failed compile leaves `out` unchanged.
  static Status compile(const NativeCartesianMgSpec& spec,
                        MgRuntimeServices services,
                        MgCoefficientViews coefficients,
                        HipCartesianMgPlan& out,
                        MgPlanCounters* counters = nullptr) noexcept {
    if (spec.geometry.faces_x == nullptr || spec.geometry.faces_y == nullptr ||
        spec.geometry.faces_z == nullptr)
      return {StatusCode::invalid_plan, 9101U};
    hf_grid_desc g{};
    g.nx = spec.geometry.cells.x;
    g.ny = spec.geometry.cells.y;
    g.nz = spec.geometry.cells.z;
    g.faces_x = spec.geometry.faces_x;
    g.faces_y = spec.geometry.faces_y;
    g.faces_z = spec.geometry.faces_z;
    const MgBoundaryKind kinds[6] = {spec.boundaries.x_min, spec.boundaries.x_max,
                                     spec.boundaries.y_min, spec.boundaries.y_max,
                                     spec.boundaries.z_min, spec.boundaries.z_max};
    for (int f = 0; f < 6; ++f) g.boundary[f] = static_cast<int32_t>(kinds[f]);
    hf_policy p{};
    hf_policy_default(&p);
    p.anisotropy_threshold = spec.policy.anisotropy_threshold;
    p.pre_sweeps = spec.policy.pre_sweeps;
    p.post_sweeps = spec.policy.post_sweeps;
    p.maximum_levels = spec.policy.maximum_levels;
    p.coarse_sweeps = spec.policy.coarse_sweeps;
    p.minimum_coarse_extent = spec.policy.minimum_coarse_extent;
    p.f_cycle = spec.policy.cycle == MgCycleKind::f_cycle ? 1 : 0;
    p.null_space_constant = spec.null_space == MgNullSpace::constant ? 1 : 0;
    p.unit_linear =
        spec.correction_scaling == MgCorrectionScaling::unit_linear ? 1 : 0;
    p.precision = static_cast<uint8_t>(spec.policy.precision);
    p.coarse_face_scaling = spec.policy.distance_corrected_coarse_faces ? 1 : 0;
    p.line_order = static_cast<uint8_t>(spec.policy.line_order);
    p.point_smoother = static_cast<uint8_t>(spec.policy.point_smoother);
    p.chebyshev_lower_fraction = spec.policy.chebyshev_lower_spectrum_fraction;
    p.diagonal_shortcut_ratio = spec.policy.diagonal_shortcut_maximum_ratio;
    void* stream = services.stream != nullptr ? services.stream : spec.stream;
    hf_plan* plan = nullptr;
    Status st = from_c(hf_plan_create(&g, &p, stream, &plan));
    if (!st) return st;
    st = from_c(hf_plan_set_coefficients(
        plan, coefficients.diagonal, coefficients.x, coefficients.y,
        coefficients.z, spec.activity.cells, spec.activity.x_faces,
        spec.activity.y_faces, spec.activity.z_faces));
    if (!st) {
      hf_plan_destroy(plan);
      return st;
    }
    out.release();
    out.plan_ = plan;
    out.shape_ = spec.geometry.cells;
    out.spec_ = spec;
    out.spec_.stream = stream;
    out.generation_ += 1;
    out.counters_ = {};
    out.counters_.symbolic_builds = 1;
    out.counters_.numeric_refreshes = 1;
    out.counters_.hierarchy_rebuilds = 1;
    if (counters != nullptr) *counters = out.counters_;
    return {};
  }
  static Status compile(const NativeCartesianMgSpec& spec,
                        MgCoefficientViews coefficients,
                        HipCartesianMgPlan& out) noexcept {
    return compile(spec, MgRuntimeServices{}, coefficients, out, nullptr);
  }

  // Numeric refresh with new coefficients (same activity and geometry):
  // transactional, a rejected refresh keeps the published hierarchy.
  Status update_coefficients(MgCoefficientViews coefficients,
                             MgPlanCounters* counters = nullptr) noexcept {
    if (plan_ == nullptr) return {StatusCode::invalid_plan, 9101U};
    const Status st = from_c(hf_plan_set_coefficients(
        plan_, coefficients.diagonal, coefficients.x, coefficients.y,
        coefficients.z, spec_.activity.cells, spec_.activity.x_faces,
        spec_.activity.y_faces, spec_.activity.z_faces));
    if (st) {
      generation_ += 1;
      counters_.numeric_refreshes += 1;
      counters_.hierarchy_rebuilds += 1;
    }
    if (counters != nullptr) *counters = counters_;
    return st;
  }

  LinearPreconditionerCertificate certificate() const noexcept override {
    // unit-linear FP64 cycles are fixed operators; the residual-minimising
    // scaling and reduced-precision storage make the cycle flexible
    const bool fixed =
        spec_.correction_scaling == MgCorrectionScaling::unit_linear &&
        spec_.policy.precision == MgStoragePrecision::fp64;
    return {fixed ? LinearPreconditionerClass::fixed_general
                  : LinearPreconditionerClass::flexible};
  }

  Status apply(ConstDeviceFieldView input, DeviceFieldView output,
               std::uint32_t) noexcept override {
    if (plan_ == nullptr || !same(input.interior) || !same(output.interior))
      return {StatusCode::invalid_plan, 9101U};
    counters_.applications += 1;
    return from_c(hf_mg_apply(plan_, input.base, output.base));
  }

  Status prepare_batch(const LinearPreconditionerBatchDescriptor& descriptor,
                       LinearPreconditionerBatchTicket& ticket) noexcept override {
    if (plan_ == nullptr || !same(descriptor.shape) ||
        descriptor.maximum_applications == 0U)
      return {StatusCode::invalid_plan, 603U};
    issue_batch_ticket(ticket, this, descriptor, generation_);
    return {};
  }

  Status apply_prepared(ConstDeviceFieldView input, DeviceFieldView output,
                        std::uint32_t iteration,
                        const LinearPreconditionerBatchTicket& ticket) noexcept
      override {
    // a ticket binds this plan and the hierarchy it was issued for
    if (ticket.owner != reinterpret_cast<std::uintptr_t>(this) ||
        ticket.generation != generation_ || !same(ticket.shape) ||
        iteration >= ticket.maximum_applications)
      return {StatusCode::invalid_plan, 603U};
    return apply(input, output, iteration);
  }

  std::size_t level_count() const noexcept {
    int32_t n = 0;
    if (plan_ != nullptr) hf_plan_level_count(plan_, &n);
    return static_cast<std::size_t>(n);
  }
  Status level(std::size_t index, hf_level_info& out) const noexcept {
    if (plan_ == nullptr) return {StatusCode::invalid_plan, 9101U};
    return from_c(hf_plan_level_info(plan_, static_cast<int32_t>(index), &out));
  }
  const MgPlanCounters& counters() const noexcept { return counters_; }
  hf_plan* handle() const noexcept { return plan_; }
  Int3 shape() const noexcept { return shape_; }
  void* stream() const noexcept { return spec_.stream; }

 private:
  bool same(Int3 s) const noexcept {
    return s.x == shape_.x && s.y == shape_.y && s.z == shape_.z;
  }
  void release() noexcept {
    if (plan_ != nullptr) hf_plan_destroy(plan_);
    plan_ = nullptr;
  }
  hf_plan* plan_{};
  Int3 shape_{};
  NativeCartesianMgSpec spec_{};
  std::uint64_t generation_{};
  MgPlanCounters counters_{};
};

inline HipPressureOperator::HipPressureOperator(
    const HipCartesianMgPlan& plan) noexcept
    : plan_(plan.handle()), shape_(plan.shape()) {}

inline LinearOperatorCertificate HipPressureOperator::certificate() const
    noexcept {
  return {shape_, LinearOperatorClass::spd};
}

inline Status HipPressureOperator::apply(DeviceFieldView x,
                                         DeviceFieldView y) const noexcept {
  Status st{StatusCode::invalid_plan, 9101U};
  if (plan_ != nullptr) st = from_c(hf_operator_apply(plan_, x.base, y.base));
  failure_ = st ? LinearOperatorFailureProvenance{}
                : LinearOperatorFailureProvenance{st, 0};
  return st;
}

// The cold pressure pairing (src/solver_cold.hpp:378-415): operator rows per
// unit volume, and the multigrid applied to V r.  Solved on the device by
// solve_fgmres when both wrap the same plan.
class HipVolumeRowsOperator final : public LinearOperator {
 public:
  explicit HipVolumeRowsOperator(const HipCartesianMgPlan& plan) noexcept
      : plan_(plan.handle()), shape_(plan.shape()) {}
  LinearOperatorCertificate certificate() const noexcept override {
    return {shape_, LinearOperatorClass::nonsymmetric};
  }
  // only solve_fgmres applies these rows (on the device)
  Status apply(DeviceFieldView, DeviceFieldView) const noexcept override {
    return {StatusCode::invalid_plan, 603U};
  }
  hf_plan* handle() const noexcept { return plan_; }

 private:
  hf_plan* plan_{};
  Int3 shape_{};
};

class HipVolumeScaledPreconditioner final : public LinearPreconditioner {
 public:
  explicit HipVolumeScaledPreconditioner(HipCartesianMgPlan& plan) noexcept
      : mg_(plan) {}
  LinearPreconditionerCertificate certificate() const noexcept override {
    return mg_.certificate();
  }
  Status apply(ConstDeviceFieldView, DeviceFieldView,
               std::uint32_t) noexcept override {
    return {StatusCode::invalid_plan, 603U};
  }
  hf_plan* handle() const noexcept { return mg_.handle(); }

 private:
  HipCartesianMgPlan& mg_;
};

namespace detail {

inline void fill_result(const hf_solve_result& r, LinearSolveResult& result,
                        ReductionEngine& reductions,
                        ResourceCounters* resources) noexcept {
  result.status = from_c(r.status);
  result.termination = static_cast<LinearTermination>(r.termination);
  result.iterations = r.iterations;
  result.initial_true_residual = r.initial_true_residual;
  result.final_true_residual = r.final_true_residual;
  result.recursive_residual = r.recursive_residual;
  result.reduction_calls = r.reduction_calls;
  result.operator_applies = r.operator_applies;
  result.preconditioner_applies = r.preconditioner_applies;
  result.norm_breakdown_restarts = r.norm_breakdown_restarts;
  result.rhs_norm = r.rhs_norm;
  result.true_residual_limit = r.true_residual_limit;
  reductions.count(r.reduction_calls);
  reductions.consensus(result.status);
  if (resources != nullptr) {
    resources->operator_applies += r.operator_applies;
    resources->preconditioner_applies += r.preconditioner_applies;
    resources->reductions += r.reduction_calls;
  }
}

inline bool same_shape(Int3 a, Int3 b) noexcept {
  return a.x == b.x && a.y == b.y && a.z == b.z;
}

inline LinearSolveResult refuse(Status st, ReductionEngine& reductions) {
  LinearSolveResult result;
  result.status = st;
  result.termination = LinearTermination::invalid_plan;
  reductions.consensus(st);
  return result;
}

// solve_pcg from the host over any operator / preconditioner pair
// (src/solver_krylov.cpp:1299-1555): the device vector engine holds x, r, z,
// p, Ap and makes every sum; the gates are the reference's.
inline LinearSolveResult generic_pcg(const LinearOperator& op,
                                     LinearPreconditioner& pc,
                                     const LinearSolveInvocation& inv,
                                     SolverWorkspace& ws,
                                     ReductionEngine& reductions,
                                     ResourceCounters* resources) noexcept {
  LinearSolveResult res;
  hf_vector_engine* e = ws.engine();
  const Int3 shape = inv.rhs.interior;
  const std::int64_t n = static_cast<std::int64_t>(shape.x) * shape.y * shape.z;
  double* x = ws.vector(0);
  double* r = ws.vector(1);
  double* z = ws.vector(2);
  double* p = ws.vector(3);
  double* ap = ws.vector(4);
  const double* b = inv.rhs.base;
  const LinearSolveControl& ctl = inv.control;
  std::uint64_t sums = 0;
  auto finish = [&](Status st, LinearTermination term) {
    res.status = st;
    res.termination = term;
    res.reduction_calls = sums;
    reductions.count(sums);
    reductions.consensus(st);
    if (resources != nullptr) {
      resources->operator_applies += res.operator_applies;
      resources->preconditioner_applies += res.preconditioner_applies;
      resources->reductions += sums;
    }
    return res;
  };
  auto dot = [&](const double* u, const double* v, double& out) -> Status {
    ++sums;
    return from_c(hf_vector_dot(e, n, u, v, &out));
  };
  const Status numeric{StatusCode::numerical_failure, 605U};
  const Status breakdown{StatusCode::numerical_failure, 604U};
  double bb = 0.0;
  Status st = dot(b, b, bb);
  if (!st) return finish(st, LinearTermination::non_finite);
  const double rhs_norm = std::sqrt(bb);
  res.rhs_norm = rhs_norm;
  if (!std::isfinite(rhs_norm)) return finish(numeric, LinearTermination::non_finite);
  const double tol = std::max(ctl.absolute_tolerance,
                              ctl.relative_tolerance * rhs_norm);
  res.true_residual_limit = tol;
  if (rhs_norm == 0.0) {
    st = from_c(hf_vector_axpby(e, n, 0.0, b, 0.0, b, inv.solution.base));
    return finish(st, LinearTermination::zero_rhs);
  }
  auto true_residual = [&](double& norm) -> Status {
    Status s = op.apply({x, shape}, {ap, shape});
    res.operator_applies++;
    if (!s) return s;
    s = from_c(hf_vector_axpby(e, n, 1.0, b, -1.0, ap, r));
    if (!s) return s;
    double rr = 0.0;
    s = dot(r, r, rr);
    norm = std::sqrt(rr);
    return s;
  };
  st = from_c(hf_vector_copy(e, n, inv.solution.base, x));
  if (!st) return finish(st, LinearTermination::non_finite);
  st = true_residual(res.initial_true_residual);
  if (!st) return finish(st, LinearTermination::operator_failure);
  res.final_true_residual = res.initial_true_residual;
  res.recursive_residual = res.initial_true_residual;
  if (!std::isfinite(res.initial_true_residual))
    return finish(numeric, LinearTermination::operator_failure);
  auto publish = [&]() {
    return from_c(hf_vector_copy(e, n, x, inv.solution.base));
  };
  if (res.initial_true_residual <= tol)
    return finish(publish(), LinearTermination::converged);
  auto precondition = [&](double& rho) -> Status {
    Status s = pc.apply({r, shape}, {z, shape}, res.iterations);
    res.preconditioner_applies++;
    if (!s) return s;
    return dot(r, z, rho);
  };
  double rho = 0.0;
  st = precondition(rho);
  if (!st) return finish(st, LinearTermination::preconditioner_failure);
  if (!(rho > 0.0) || !std::isfinite(rho))
    return finish(breakdown, LinearTermination::breakdown);
  st = from_c(hf_vector_copy(e, n, z, p));
  while (res.iterations < ctl.maximum_iterations) {
    st = op.apply({p, shape}, {ap, shape});
    res.operator_applies++;
    if (!st) return finish(st, LinearTermination::operator_failure);
    double pap = 0.0;
    st = dot(p, ap, pap);
    if (!st) return finish(st, LinearTermination::non_finite);
    if (!(pap > 0.0) || !std::isfinite(pap))
      return finish(breakdown, LinearTermination::breakdown);
    const double alpha = rho / pap;
    if (!std::isfinite(alpha)) return finish(numeric, LinearTermination::non_finite);
    hf_vector_axpby(e, n, 1.0, x, alpha, p, x);
    hf_vector_axpby(e, n, 1.0, r, -alpha, ap, r);
    res.iterations++;
    double rr = 0.0;
    st = dot(r, r, rr);
    if (!st) return finish(st, LinearTermination::non_finite);
    const double recursive = std::sqrt(rr);
    res.recursive_residual = recursive;
    const bool verify = recursive <= tol ||
                        res.iterations % ctl.true_residual_interval == 0 ||
                        res.iterations == ctl.maximum_iterations;
    if (verify) {
      st = true_residual(res.final_true_residual);
      if (!st) return finish(st, LinearTermination::operator_failure);
      if (res.final_true_residual <= tol)
        return finish(publish(), LinearTermination::converged);
      if (res.iterations == ctl.maximum_iterations)
        return finish({StatusCode::rejected_step, 604U},
                      LinearTermination::maximum_iterations);
      st = precondition(rho);
      if (!st) return finish(st, LinearTermination::preconditioner_failure);
      if (!(rho > 0.0) || !std::isfinite(rho))
        return finish(breakdown, LinearTermination::breakdown);
      hf_vector_copy(e, n, z, p);
      continue;
    }
    double rho_next = 0.0;
    st = precondition(rho_next);
    if (!st) return finish(st, LinearTermination::preconditioner_failure);
    if (!(rho_next > 0.0) || !std::isfinite(rho_next))
      return finish(breakdown, LinearTermination::breakdown);
    double beta = rho_next / rho;
    if (ctl.flexible) {
      double zq = 0.0;
      st = dot(z, ap, zq);
      if (!st) return finish(st, LinearTermination::non_finite);
      beta = -alpha * zq / rho;
    }
    hf_vector_axpby(e, n, 1.0, z, beta, p, p);
    rho = rho_next;
  }
  return finish({StatusCode::rejected_step, 604U},
                LinearTermination::maximum_iterations);
}

}  // namespace detail

// solve_pcg (v04_linear.hpp:1070-1075; reference semantics: true residual
// every true_residual_interval iterations and at the recursive gate,
// direction restart after an unconverged check).  This library's pair over
// one plan solves on the device; any other pair runs from the host over the
// workspace's vector engine.
inline LinearSolveResult solve_pcg(const LinearOperator& linear_operator,
                                   LinearPreconditioner& preconditioner,
                                   const LinearSolveInvocation& invocation,
                                   SolverWorkspace& workspace,
                                   ReductionEngine& reductions,
                                   ResourceCounters* resources = nullptr) noexcept {
  const Int3 shape = invocation.rhs.interior;
  if (invocation.rhs.base == nullptr || invocation.solution.base == nullptr ||
      !detail::same_shape(shape, invocation.solution.interior) ||
      !detail::same_shape(shape, linear_operator.certificate().local_shape) ||
      invocation.control.maximum_iterations == 0U ||
      invocation.control.true_residual_interval == 0U)
    return detail::refuse({StatusCode::invalid_plan, 601U}, reductions);
  const auto* op = dynamic_cast<const HipPressureOperator*>(&linear_operator);
  auto* mg = dynamic_cast<HipCartesianMgPlan*>(&preconditioner);
  if (op != nullptr && mg != nullptr && op->handle() == mg->handle() &&
      op->handle() != nullptr) {
    hf_solve_control ctl{};
    ctl.absolute_tolerance = invocation.control.absolute_tolerance;
    ctl.relative_tolerance = invocation.control.relative_tolerance;
    ctl.maximum_iterations = invocation.control.maximum_iterations;
    ctl.true_residual_interval = invocation.control.true_residual_interval;
    ctl.flexible = invocation.control.flexible ? 1 : 0;
    hf_solve_result r{};
    hf_pcg_solve(op->handle(), invocation.rhs.base, invocation.solution.base,
                 &ctl, &r);
    LinearSolveResult result;
    detail::fill_result(r, result, reductions, resources);
    return result;
  }
  if (workspace.engine() == nullptr ||
      !detail::same_shape(workspace.shape(), shape))
    return detail::refuse({StatusCode::invalid_plan, 602U}, reductions);
  return detail::generic_pcg(linear_operator, preconditioner, invocation,
                             workspace, reductions, resources);
}

// solve_fgmres (v04_linear.hpp:1076-1082): restarted single-reduction FGMRES
// on the device for this library's pairs over one plan -- the pressure
// operator with the plan, or the per-volume rows with the volume-scaled plan
// (the cold pairing).  Other pairs are refused with invalid_plan.
inline LinearSolveResult solve_fgmres(const LinearOperator& linear_operator,
                                      LinearPreconditioner& preconditioner,
                                      const LinearSolveInvocation& invocation,
                                      SolverWorkspace& workspace,
                                      ReductionEngine& reductions,
                                      ResourceCounters* resources = nullptr) noexcept {
  (void)workspace;
  const Int3 shape = invocation.rhs.interior;
  if (invocation.rhs.base == nullptr || invocation.solution.base == nullptr ||
      !detail::same_shape(shape, invocation.solution.interior) ||
      invocation.control.restart == 0U)
    return detail::refuse({StatusCode::invalid_plan, 601U}, reductions);
  hf_plan* plan = nullptr;
  std::uint8_t volume_rows = 0;
  const auto* op = dynamic_cast<const HipPressureOperator*>(&linear_operator);
  auto* mg = dynamic_cast<HipCartesianMgPlan*>(&preconditioner);
  const auto* vop = dynamic_cast<const HipVolumeRowsOperator*>(&linear_operator);
  auto* vpc = dynamic_cast<HipVolumeScaledPreconditioner*>(&preconditioner);
  if (op != nullptr && mg != nullptr && op->handle() == mg->handle()) {
    plan = op->handle();
  } else if (vop != nullptr && vpc != nullptr &&
             vop->handle() == vpc->handle()) {
    plan = vop->handle();
    volume_rows = 1;
  }
  if (plan == nullptr)
    return detail::refuse({StatusCode::invalid_plan, 603U}, reductions);
  hf_fgmres_control ctl{};
  ctl.absolute_tolerance = invocation.control.absolute_tolerance;
  ctl.relative_tolerance = invocation.control.relative_tolerance;
  ctl.maximum_iterations = invocation.control.maximum_iterations;
  ctl.true_residual_interval = invocation.control.true_residual_interval;
  ctl.restart = invocation.control.restart;
  ctl.volume_rows = volume_rows;
  hf_solve_result r{};
  hf_fgmres_solve(plan, invocation.rhs.base, invocation.solution.base, &ctl,
                  &r);
  LinearSolveResult result;
  detail::fill_result(r, result, reductions, resources);
  return result;
}

}  // namespace hundun::hip
