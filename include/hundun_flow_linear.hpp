// hundun_flow_linear.hpp -- C++ drop-in layer over the C ABI (hundun_flow.h).
//
// Mirrors the reference solver_* interface the pressure path is called
// through (versions/v0.4/include/hundun/v04_linear.hpp of the reference):
//   LinearOperator          (:396-405)   -> HipPressureOperator
//   LinearPreconditioner    (:418-452)   -> HipCartesianMgPlan
//   NativeCartesianMgPlan::compile / update_coefficients (:952-1013)
//   solve_pcg               (:1070-1075) -> solve_pcg below
// with the same names, argument meaning and Status error convention
// (include/hundun/v04_status.hpp:13-31): nothing throws, every call returns
// Status{code, detail}, failures never publish a solution.
//
// Fields are device-resident: DeviceFieldView carries a device pointer and
// the interior shape (no ghosts; the library fills halos on device).
#pragma once

#include <cstddef>
#include <cstdint>
#include <utility>

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
enum class MgStoragePrecision : std::uint8_t { fp64, fp32, bf16 };
enum class MgLineOrder : std::uint8_t { lexicographic, zebra };

struct MgHierarchyPolicy {
  double anisotropy_threshold{4.0};
  std::uint8_t pre_sweeps{2U};
  std::uint8_t post_sweeps{2U};
  std::uint8_t maximum_levels{32U};
  std::uint16_t coarse_sweeps{24U};
  std::uint8_t minimum_coarse_extent{3U};
  MgCycleKind cycle{MgCycleKind::v_cycle};
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

struct LinearOperatorCertificate {
  Int3 local_shape{};
  LinearOperatorClass operator_class{LinearOperatorClass::nonsymmetric};
};

struct LinearPreconditionerCertificate {
  LinearPreconditionerClass preconditioner_class{
      LinearPreconditionerClass::fixed_general};
};

class LinearOperator {
 public:
  virtual LinearOperatorCertificate certificate() const noexcept = 0;
  virtual Status apply(DeviceFieldView x, DeviceFieldView y) const noexcept = 0;
  virtual ~LinearOperator() = default;
};

class LinearPreconditioner {
 public:
  virtual LinearPreconditionerCertificate certificate() const noexcept = 0;
  virtual Status apply(ConstDeviceFieldView input, DeviceFieldView output,
                       std::uint32_t iteration) noexcept = 0;
  virtual ~LinearPreconditioner() = default;
};

struct LinearSolveControl {
  double absolute_tolerance{};
  double relative_tolerance{};
  std::uint32_t maximum_iterations{};
  std::uint32_t true_residual_interval{};
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
  double rhs_norm{};
  double true_residual_limit{};
};

class HipCartesianMgPlan;

// Exact IBM pressure operator over a compiled plan's level-0 coefficients.
class HipPressureOperator final : public LinearOperator {
 public:
  explicit HipPressureOperator(const HipCartesianMgPlan& plan) noexcept;
  LinearOperatorCertificate certificate() const noexcept override;
  Status apply(DeviceFieldView x, DeviceFieldView y) const noexcept override;
  hf_plan* handle() const noexcept { return plan_; }

 private:
  hf_plan* plan_{};
  Int3 shape_{};
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
        spec_(o.spec_) {}
  HipCartesianMgPlan& operator=(HipCartesianMgPlan&& o) noexcept {
    if (this != &o) {
      release();
      plan_ = std::exchange(o.plan_, nullptr);
      shape_ = o.shape_;
      spec_ = o.spec_;
    }
    return *this;
  }

  // compile: plan levels, allocate, build the hierarchy from coefficients.
  // A failed compile leaves `out` unchanged.
  static Status compile(const NativeCartesianMgSpec& spec,
                        MgCoefficientViews coefficients,
                        HipCartesianMgPlan& out) noexcept {
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
    hf_plan* plan = nullptr;
    Status st = from_c(hf_plan_create(&g, &p, spec.stream, &plan));
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
    return {};
  }

  // Numeric refresh with new coefficients (same activity and geometry).
  Status update_coefficients(MgCoefficientViews coefficients) noexcept {
    if (plan_ == nullptr) return {StatusCode::invalid_plan, 9101U};
    return from_c(hf_plan_set_coefficients(
        plan_, coefficients.diagonal, coefficients.x, coefficients.y,
        coefficients.z, spec_.activity.cells, spec_.activity.x_faces,
        spec_.activity.y_faces, spec_.activity.z_faces));
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
    return from_c(hf_mg_apply(plan_, input.base, output.base));
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
  hf_plan* handle() const noexcept { return plan_; }
  Int3 shape() const noexcept { return shape_; }

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
  if (plan_ == nullptr) return {StatusCode::invalid_plan, 9101U};
  return from_c(hf_operator_apply(plan_, x.base, y.base));
}

// solve_pcg (reference solve_pcg semantics: true residual every
// true_residual_interval iterations and at the recursive gate, direction
// restart after an unconverged check).  The device solve runs when the
// operator and preconditioner are this library's pair over one plan; any
// other pairing is rejected as invalid_plan.
inline LinearSolveResult solve_pcg(const LinearOperator& linear_operator,
                                   LinearPreconditioner& preconditioner,
                                   const LinearSolveInvocation& invocation) {
  LinearSolveResult result;
  const auto* op = dynamic_cast<const HipPressureOperator*>(&linear_operator);
  auto* mg = dynamic_cast<HipCartesianMgPlan*>(&preconditioner);
  if (op == nullptr || mg == nullptr || op->handle() != mg->handle() ||
      op->handle() == nullptr) {
    result.status = {StatusCode::invalid_plan, 9101U};
    return result;
  }
  hf_solve_control ctl{};
  ctl.absolute_tolerance = invocation.control.absolute_tolerance;
  ctl.relative_tolerance = invocation.control.relative_tolerance;
  ctl.maximum_iterations = invocation.control.maximum_iterations;
  ctl.true_residual_interval = invocation.control.true_residual_interval;
  ctl.flexible = invocation.control.flexible ? 1 : 0;
  hf_solve_result r{};
  hf_pcg_solve(op->handle(), invocation.rhs.base, invocation.solution.base,
               &ctl, &r);
  result.status = from_c(r.status);
  result.termination = static_cast<LinearTermination>(r.termination);
  result.iterations = r.iterations;
  result.initial_true_residual = r.initial_true_residual;
  result.final_true_residual = r.final_true_residual;
  result.recursive_residual = r.recursive_residual;
  result.reduction_calls = r.reduction_calls;
  result.operator_applies = r.operator_applies;
  result.preconditioner_applies = r.preconditioner_applies;
  result.rhs_norm = r.rhs_norm;
  result.true_residual_limit = r.true_residual_limit;
  return result;
}

}  // namespace hundun::hip
