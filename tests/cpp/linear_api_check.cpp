// Test shim for include/hundun_flow_linear.hpp: drives the C++ drop-in layer
// (HipCartesianMgPlan::compile, HipPressureOperator::apply,
// LinearPreconditioner::apply, solve_pcg) from plain extern "C" entry points
// so the Python tests can compare it with the C ABI on the same device
// buffers.  Test infrastructure only; not part of the library.
#include <cstdint>

#include "hundun_flow_linear.hpp"

namespace hh = hundun::hip;

namespace {

hh::NativeCartesianMgSpec make_spec(const int32_t* n, const double* fx,
                                    const double* fy, const double* fz,
                                    const int32_t* bc, const uint8_t* const* act,
                                    const double* policy_f,
                                    const int32_t* policy_i) {
  hh::NativeCartesianMgSpec s{};
  s.geometry.cells = {n[0], n[1], n[2]};
  s.geometry.faces_x = fx;
  s.geometry.faces_y = fy;
  s.geometry.faces_z = fz;
  auto k = [](int32_t v) { return static_cast<hh::MgBoundaryKind>(v); };
  s.boundaries = {k(bc[0]), k(bc[1]), k(bc[2]), k(bc[3]), k(bc[4]), k(bc[5])};
  s.policy.anisotropy_threshold = policy_f[0];
  s.policy.pre_sweeps = static_cast<uint8_t>(policy_i[0]);
  s.policy.post_sweeps = static_cast<uint8_t>(policy_i[1]);
  s.policy.maximum_levels = static_cast<uint8_t>(policy_i[2]);
  s.policy.minimum_coarse_extent = static_cast<uint8_t>(policy_i[3]);
  s.policy.coarse_sweeps = static_cast<uint16_t>(policy_i[4]);
  s.policy.cycle = policy_i[5] ? hh::MgCycleKind::f_cycle
                               : hh::MgCycleKind::v_cycle;
  s.null_space = policy_i[6] ? hh::MgNullSpace::constant : hh::MgNullSpace::none;
  s.correction_scaling = policy_i[7] ? hh::MgCorrectionScaling::unit_linear
                                     : hh::MgCorrectionScaling::residual_minimizing;
  s.policy.precision = static_cast<hh::MgStoragePrecision>(policy_i[8]);
  s.policy.distance_corrected_coarse_faces = policy_i[9] != 0;
  s.policy.line_order = static_cast<hh::MgLineOrder>(policy_i[10]);
  s.policy.point_smoother =
      static_cast<hh::MgPointSmootherKind>(policy_i[11]);
  s.policy.chebyshev_lower_spectrum_fraction = policy_f[1];
  s.policy.diagonal_shortcut_maximum_ratio = policy_f[2];
  if (act != nullptr)
    s.activity = {act[0], act[1], act[2], act[3]};
  return s;
}

}  // namespace

extern "C" {

// Compiles a plan through the C++ layer and runs, on device buffers:
//   y = A rhs, z = M rhs, then solve_pcg(A, M, rhs -> x).
// out_i: [status code, status detail, termination, iterations, level count,
//         preconditioner class]; out_f: [final true residual, rhs norm].
int32_t hf_linear_check(const int32_t* n, const double* fx, const double* fy,
                        const double* fz, const int32_t* bc,
                        const uint8_t* const* activity, const double* policy_f,
                        const int32_t* policy_i, const double* diag,
                        const double* bx, const double* by, const double* bz,
                        const double* rhs, double* y, double* z, double* x,
                        const double* control_f, const int32_t* control_i,
                        int32_t* out_i, double* out_f) {
  const hh::NativeCartesianMgSpec spec =
      make_spec(n, fx, fy, fz, bc, activity, policy_f, policy_i);
  hh::HipCartesianMgPlan plan;
  hh::Status st = hh::HipCartesianMgPlan::compile(spec, {diag, bx, by, bz}, plan);
  out_i[0] = static_cast<int32_t>(st.code);
  out_i[1] = static_cast<int32_t>(st.detail);
  if (!st) return out_i[0];
  const hh::Int3 shape{n[0], n[1], n[2]};
  hh::HipPressureOperator op(plan);
  st = op.apply({const_cast<double*>(rhs), shape}, {y, shape});
  if (!st) {
    out_i[0] = static_cast<int32_t>(st.code);
    out_i[1] = static_cast<int32_t>(st.detail);
    return out_i[0];
  }
  hh::LinearPreconditioner& pre = plan;
  st = pre.apply({rhs, shape}, {z, shape}, 0U);
  if (!st) {
    out_i[0] = static_cast<int32_t>(st.code);
    out_i[1] = static_cast<int32_t>(st.detail);
    return out_i[0];
  }
  hh::LinearSolveInvocation inv{};
  inv.rhs = {rhs, shape};
  inv.solution = {x, shape};
  inv.control.absolute_tolerance = control_f[0];
  inv.control.relative_tolerance = control_f[1];
  inv.control.maximum_iterations = static_cast<uint32_t>(control_i[0]);
  inv.control.true_residual_interval = static_cast<uint32_t>(control_i[1]);
  inv.control.flexible = control_i[2] != 0;
  hh::SolverWorkspace ws;
  hh::ReductionEngine reductions;
  const hh::LinearSolveResult r = hh::solve_pcg(op, pre, inv, ws, reductions);
  out_i[0] = static_cast<int32_t>(r.status.code);
  out_i[1] = static_cast<int32_t>(r.status.detail);
  out_i[2] = static_cast<int32_t>(r.termination);
  out_i[3] = static_cast<int32_t>(r.iterations);
  out_i[4] = static_cast<int32_t>(plan.level_count());
  out_i[5] = static_cast<int32_t>(pre.certificate().preconditioner_class);
  out_f[0] = r.final_true_residual;
  out_f[1] = r.rhs_norm;
  return out_i[0];
}

// Host-only error paths (no device work): a compile with missing face
// coordinates, and solve_pcg with an operator/preconditioner pair that is not
// one plan.  Returns the two status codes in out[0], out[1].
void hf_linear_check_errors(int32_t* out) {
  hh::NativeCartesianMgSpec spec{};
  spec.geometry.cells = {4, 4, 4};
  hh::HipCartesianMgPlan plan;
  out[0] = static_cast<int32_t>(
      hh::HipCartesianMgPlan::compile(spec, {}, plan).code);
  hh::HipPressureOperator op(plan);
  hh::LinearSolveInvocation inv{};
  hh::SolverWorkspace ws;
  hh::ReductionEngine reductions;
  out[1] = static_cast<int32_t>(
      hh::solve_pcg(op, plan, inv, ws, reductions).status.code);
}

namespace {

// An operator that is not the library's type: the same exact IBM operator
// reached through the C ABI, so solve_pcg must take the host-driven path.
class ForeignOperator final : public hh::LinearOperator {
 public:
  explicit ForeignOperator(const hh::HipCartesianMgPlan& plan)
      : plan_(plan.handle()), shape_(plan.shape()) {}
  hh::LinearOperatorCertificate certificate() const noexcept override {
    return {shape_, hh::LinearOperatorClass::spd};
  }
  hh::Status apply(hh::DeviceFieldView x,
                   hh::DeviceFieldView y) const noexcept override {
    return hh::from_c(hf_operator_apply(plan_, x.base, y.base));
  }

 private:
  hf_plan* plan_;
  hh::Int3 shape_;
};

}  // namespace

// The rest of the drop-in surface on one plan (device buffers):
//  * solve_pcg over a foreign operator with the plan as preconditioner (the
//    host-driven path) -> xg, with the device path -> xd;
//  * apply_prepared against apply (zp, z), and the refusals of a stale
//    ticket (after a coefficient refresh) and of a foreign one;
//  * the failure provenance of an operator apply refused for a null field;
//  * solve_fgmres through the layer -> xf.
// out_i: [generic status, generic iterations, device iterations,
//         prepared status, stale ticket code, foreign ticket code,
//         provenance code, fgmres status, fgmres iterations, applications,
//         refreshes, reductions counted]
int32_t hf_linear_check_surface(const int32_t* n, const double* fx,
                                const double* fy, const double* fz,
                                const int32_t* bc,
                                const uint8_t* const* activity,
                                const double* policy_f, const int32_t* policy_i,
                                const double* diag, const double* bx,
                                const double* by, const double* bz,
                                const double* rhs, double* xg, double* xd,
                                double* z, double* zp, double* xf,
                                int32_t* out_i) {
  const hh::NativeCartesianMgSpec spec =
      make_spec(n, fx, fy, fz, bc, activity, policy_f, policy_i);
  hh::HipCartesianMgPlan plan;
  hh::MgPlanCounters counters;
  hh::Status st = hh::HipCartesianMgPlan::compile(
      spec, hh::MgRuntimeServices{}, {diag, bx, by, bz}, plan, &counters);
  if (!st) return static_cast<int32_t>(st.code);
  const hh::Int3 shape{n[0], n[1], n[2]};
  hh::SolverWorkspace ws;
  st = hh::SolverWorkspace::bind(shape, 12U, plan.stream(), ws);
  if (!st) return 100 + static_cast<int32_t>(st.code);
  hh::ReductionEngine reductions;
  hh::ResourceCounters resources;
  hh::LinearSolveInvocation inv{};
  inv.rhs = {rhs, shape};
  inv.control.relative_tolerance = 1e-10;
  inv.control.maximum_iterations = 400;
  inv.control.true_residual_interval = 4;
  ForeignOperator foreign(plan);
  inv.solution = {xg, shape};
  const hh::LinearSolveResult g =
      hh::solve_pcg(foreign, plan, inv, ws, reductions, &resources);
  hh::HipPressureOperator op(plan);
  inv.solution = {xd, shape};
  const hh::LinearSolveResult d = hh::solve_pcg(op, plan, inv, ws, reductions);
  out_i[0] = static_cast<int32_t>(g.status.code);
  out_i[1] = static_cast<int32_t>(g.iterations);
  out_i[2] = static_cast<int32_t>(d.iterations);
  // prepared batch
  hh::LinearPreconditionerBatchTicket ticket{};
  st = plan.prepare_batch({shape, 4U}, ticket);
  hh::Status applied = plan.apply({rhs, shape}, {z, shape}, 0U);
  hh::Status prepared =
      st ? plan.apply_prepared({rhs, shape}, {zp, shape}, 0U, ticket) : st;
  out_i[3] = applied && prepared ? 0 : 1;
  hh::LinearPreconditionerBatchTicket other = ticket;
  other.owner += 8;
  out_i[5] = static_cast<int32_t>(
      plan.apply_prepared({rhs, shape}, {zp, shape}, 1U, other).code);
  // a refresh with the same coefficients publishes a new generation
  st = plan.update_coefficients({diag, bx, by, bz}, &counters);
  out_i[4] = st ? static_cast<int32_t>(
                      plan.apply_prepared({rhs, shape}, {zp, shape}, 1U, ticket)
                          .code)
                : -1;
  // failure provenance
  st = op.apply({nullptr, shape}, {xd, shape});
  out_i[6] = static_cast<int32_t>(op.failure_provenance().status.code);
  // solve_fgmres through the layer (the plan pair)
  inv.solution = {xf, shape};
  inv.control.restart = 12;
  const hh::LinearSolveResult f = hh::solve_fgmres(op, plan, inv, ws, reductions);
  out_i[7] = static_cast<int32_t>(f.status.code);
  out_i[8] = static_cast<int32_t>(f.iterations);
  out_i[9] = static_cast<int32_t>(plan.counters().applications);
  out_i[10] = static_cast<int32_t>(counters.numeric_refreshes);
  out_i[11] = static_cast<int32_t>(resources.reductions);
  return 0;
}

}  // extern "C"
