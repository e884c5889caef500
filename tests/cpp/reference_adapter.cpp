// The reference-side binding of INTEGRATION.md as a translation unit that is
// compiled against the reference's own headers (versions/v0.4/include of the
// reference tree) and this repository's C++ layer.  It is only compiled
// (tests/test_linear_api.py, when the reference tree is present), never
// linked: it shows that the adapter a maintainer adds is well-formed against
// the real reference types.
//
// The adapter replaces the NativeCartesianMgPlan + solve_pcg / solve_fgmres
// pair at the call sites of src/core_product_freeze.cpp:14319-14412 and
// src/solver_piso.cpp:7024-7070, :10445-10447: it packs the interiors of the
// ghosted host fields into device arrays, runs the device plan, and unpacks
// the solution only when the solve succeeded.
#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <vector>

#include "hundun/v04_linear.hpp"
#include "hundun_flow_linear.hpp"

namespace hundun_adapter {

namespace v4 = hundun::v04;
namespace hh = hundun::hip;

inline v4::Status status_of(hh::Status s) noexcept {
  return {static_cast<v4::StatusCode>(s.code), s.detail};
}

// Interior of a ghosted cell view into a dense x-fastest device array.
inline v4::Status upload(v4::ConstFieldView v, double* device,
                         std::vector<double>& stage, hipStream_t stream) {
  stage.resize(static_cast<std::size_t>(v.interior.x) * v.interior.y *
               v.interior.z);
  std::size_t i = 0;
  for (int z = 0; z < v.interior.z; ++z)
    for (int y = 0; y < v.interior.y; ++y)
      for (int x = 0; x < v.interior.x; ++x)
        stage[i++] = v.unchecked({x, y, z}, 0U);
  if (hipMemcpyAsync(device, stage.data(), stage.size() * sizeof(double),
                     hipMemcpyHostToDevice, stream) != hipSuccess ||
      hipStreamSynchronize(stream) != hipSuccess)
    return {v4::StatusCode::invalid_plan, 9103U};
  return {};
}

inline v4::Status upload(v4::ConstFaceFieldView v, double* device,
                         std::vector<double>& stage, hipStream_t stream) {
  stage.resize(static_cast<std::size_t>(v.extents.x) * v.extents.y *
               v.extents.z);
  std::size_t i = 0;
  for (int z = 0; z < v.extents.z; ++z)
    for (int y = 0; y < v.extents.y; ++y)
      for (int x = 0; x < v.extents.x; ++x) stage[i++] = v.unchecked({x, y, z});
  if (hipMemcpyAsync(device, stage.data(), stage.size() * sizeof(double),
                     hipMemcpyHostToDevice, stream) != hipSuccess ||
      hipStreamSynchronize(stream) != hipSuccess)
    return {v4::StatusCode::invalid_plan, 9103U};
  return {};
}

inline v4::Status download(const double* device, v4::FieldView v,
                           std::vector<double>& stage, hipStream_t stream) {
  stage.resize(static_cast<std::size_t>(v.interior.x) * v.interior.y *
               v.interior.z);
  if (hipMemcpyAsync(stage.data(), device, stage.size() * sizeof(double),
                     hipMemcpyDeviceToHost, stream) != hipSuccess ||
      hipStreamSynchronize(stream) != hipSuccess)
    return {v4::StatusCode::invalid_plan, 9103U};
  std::size_t i = 0;
  for (int z = 0; z < v.interior.z; ++z)
    for (int y = 0; y < v.interior.y; ++y)
      for (int x = 0; x < v.interior.x; ++x)
        v.unchecked({x, y, z}, 0U) = stage[i++];
  return {};
}

inline hh::MgBoundaryKind kind(v4::MgBoundaryKind k) noexcept {
  return static_cast<hh::MgBoundaryKind>(static_cast<std::uint8_t>(k));
}

// The device pressure solve of one rank (one GPU per rank).
class HipPressureSolve {
 public:
  ~HipPressureSolve() { release(); }

  // where NativeCartesianMgPlan::compile is called
  v4::Status compile(const v4::NativeCartesianMgSpec& spec,
                     v4::MgCoefficientViews coefficients, hipStream_t stream) {
    release();
    stream_ = stream;
    const v4::Int3 n = spec.patch.cells;
    shape_ = {n.x, n.y, n.z};
    cells_ = static_cast<std::size_t>(n.x) * n.y * n.z;
    const std::size_t faces[3] = {
        static_cast<std::size_t>(n.x + 1) * n.y * n.z,
        static_cast<std::size_t>(n.x) * (n.y + 1) * n.z,
        static_cast<std::size_t>(n.x) * n.y * (n.z + 1)};
    if (!allocate(&d_diag_, cells_) || !allocate(&d_rhs_, cells_) ||
        !allocate(&d_x_, cells_) || !allocate(&d_fx_, faces[0]) ||
        !allocate(&d_fy_, faces[1]) || !allocate(&d_fz_, faces[2]))
      return {v4::StatusCode::allocation_failure, 9102U};
    v4::Status st = upload_coefficients(coefficients);
    if (!st) return st;
    hh::NativeCartesianMgSpec d{};
    d.geometry = {shape_, spec.geometry->x().faces().data,
                  spec.geometry->y().faces().data,
                  spec.geometry->z().faces().data};
    d.boundaries = {kind(spec.boundaries.x_min), kind(spec.boundaries.x_max),
                    kind(spec.boundaries.y_min), kind(spec.boundaries.y_max),
                    kind(spec.boundaries.z_min), kind(spec.boundaries.z_max)};
    d.null_space = spec.null_space == v4::MgNullSpace::constant
                       ? hh::MgNullSpace::constant
                       : hh::MgNullSpace::none;
    d.correction_scaling =
        spec.correction_scaling == v4::MgCorrectionScaling::unit_linear
            ? hh::MgCorrectionScaling::unit_linear
            : hh::MgCorrectionScaling::residual_minimizing;
    d.policy.anisotropy_threshold = spec.policy.anisotropy_threshold;
    d.policy.pre_sweeps = spec.policy.pre_sweeps;
    d.policy.post_sweeps = spec.policy.post_sweeps;
    d.policy.maximum_levels = spec.policy.maximum_levels;
    d.policy.coarse_sweeps = spec.policy.coarse_sweeps;
    d.policy.minimum_coarse_extent = spec.policy.minimum_coarse_extent;
    d.policy.point_smoother =
        spec.policy.point_smoother == v4::MgPointSmootherKind::chebyshev_jacobi
            ? hh::MgPointSmootherKind::chebyshev_jacobi
            : hh::MgPointSmootherKind::red_black;
    d.policy.cycle = spec.policy.cycle == v4::MgCycleKind::f_cycle
                         ? hh::MgCycleKind::f_cycle
                         : hh::MgCycleKind::v_cycle;
    d.policy.chebyshev_lower_spectrum_fraction =
        spec.policy.chebyshev_lower_spectrum_fraction;
    d.policy.diagonal_shortcut_maximum_ratio =
        spec.policy.diagonal_shortcut_maximum_ratio;
    d.policy.precision = hh::MgStoragePrecision::fp32;  // FP32 cycle
    if (spec.activity.cells.size != 0U) {
      const std::size_t counts[4] = {cells_, faces[0], faces[1], faces[2]};
      const v4::Span<const std::uint8_t> spans[4] = {
          spec.activity.cells, spec.activity.x_faces, spec.activity.y_faces,
          spec.activity.z_faces};
      for (int a = 0; a < 4; ++a) {
        if (!allocate(&d_act_[a], counts[a]) ||
            hipMemcpyAsync(d_act_[a], spans[a].data, counts[a],
                           hipMemcpyHostToDevice, stream_) != hipSuccess)
          return {v4::StatusCode::allocation_failure, 9102U};
      }
      d.activity = {d_act_[0], d_act_[1], d_act_[2], d_act_[3]};
    }
    d.stream = stream_;
    hh::MgPlanCounters counters;
    st = status_of(hh::HipCartesianMgPlan::compile(
        d, hh::MgRuntimeServices{stream_}, {d_diag_, d_fx_, d_fy_, d_fz_},
        plan_, &counters));
    if (!st) return st;
    return status_of(
        hh::SolverWorkspace::bind(shape_, 12U, stream_, workspace_));
  }

  // where update_coefficients is called
  v4::Status update_coefficients(v4::MgCoefficientViews coefficients) {
    const v4::Status st = upload_coefficients(coefficients);
    if (!st) return st;
    return status_of(plan_.update_coefficients({d_diag_, d_fx_, d_fy_, d_fz_}));
  }

  // where solve_pcg / solve_fgmres are called
  v4::LinearSolveResult solve(const v4::LinearSolveInvocation& invocation,
                              bool fgmres) {
    v4::LinearSolveResult out;
    out.status = upload(invocation.rhs, d_rhs_, stage_, stream_);
    if (out.status)
      out.status = upload(v4::as_const(invocation.solution), d_x_, stage_,
                          stream_);
    if (!out.status) return out;
    hh::HipPressureOperator op(plan_);
    hh::LinearSolveInvocation inv{};
    inv.rhs = {d_rhs_, shape_};
    inv.solution = {d_x_, shape_};
    inv.control.absolute_tolerance = invocation.control.absolute_tolerance;
    inv.control.relative_tolerance = invocation.control.relative_tolerance;
    inv.control.maximum_iterations = invocation.control.maximum_iterations;
    inv.control.true_residual_interval =
        invocation.control.true_residual_interval;
    inv.control.restart = invocation.control.restart;
    inv.control.flexible = true;  // the FP32 cycle is a flexible operator
    const hh::LinearSolveResult r =
        fgmres ? hh::solve_fgmres(op, plan_, inv, workspace_, reductions_)
               : hh::solve_pcg(op, plan_, inv, workspace_, reductions_);
    out.status = status_of(r.status);
    out.termination = static_cast<v4::LinearTermination>(
        static_cast<std::uint8_t>(r.termination));
    out.iterations = r.iterations;
    out.initial_true_residual = r.initial_true_residual;
    out.final_true_residual = r.final_true_residual;
    out.recursive_residual = r.recursive_residual;
    out.reduction_calls = r.reduction_calls;
    // publish only a converged solution, as the reference does
    if (out.status)
      out.status = download(d_x_, invocation.solution, stage_, stream_);
    return out;
  }

 private:
  v4::Status upload_coefficients(v4::MgCoefficientViews c) {
    v4::Status st = upload(c.diagonal, d_diag_, stage_, stream_);
    if (st) st = upload(c.x, d_fx_, stage_, stream_);
    if (st) st = upload(c.y, d_fy_, stage_, stream_);
    if (st) st = upload(c.z, d_fz_, stage_, stream_);
    return st;
  }
  template <class T>
  bool allocate(T** p, std::size_t count) {
    void* ptr = nullptr;
    if (hipMalloc(&ptr, count * sizeof(T)) != hipSuccess) return false;
    owned_.push_back(ptr);
    *p = static_cast<T*>(ptr);
    return true;
  }
  void release() {
    for (void* p : owned_) (void)hipFree(p);
    owned_.clear();
  }

  hipStream_t stream_{};
  hh::Int3 shape_{};
  std::size_t cells_{};
  double* d_diag_{};
  double* d_fx_{};
  double* d_fy_{};
  double* d_fz_{};
  double* d_rhs_{};
  double* d_x_{};
  std::uint8_t* d_act_[4]{};
  std::vector<void*> owned_;
  std::vector<double> stage_;
  hh::HipCartesianMgPlan plan_;
  hh::SolverWorkspace workspace_;
  hh::ReductionEngine reductions_;
};

}  // namespace hundun_adapter
