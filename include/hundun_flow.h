/*
 * hundun_flow.h -- C ABI of the MI355X (gfx950) pressure-projection solver.
 *
 * The solver replaces the CPU hot path of Hundun-Flow's low-Mach pressure
 * step: the variable-coefficient (1/rho) 7-point cell-centred Poisson
 * operator with immersed-boundary (IBM) masked cells, preconditioned by the
 * native Cartesian geometric multigrid, inside a conjugate-gradient solve.
 * Reference interfaces are cited per entry point (paths relative to
 * versions/v0.4/ of the reference tree).
 *
 * Conventions
 *  - Every array argument of a compute entry point is a DEVICE pointer
 *    (HBM resident).  Host pointers are only taken by hf_plan_create (grid
 *    face coordinates) and by result/out structs.
 *  - Cell arrays are x-fastest, no ghosts: index i + nx*(j + ny*k).
 *    x-faces: i + (nx+1)*(j + ny*k); y-faces: i + nx*(j + (ny+1)*k);
 *    z-faces: i + nx*(j + ny*k) with k in [0, nz].
 *  - Every call returns hf_status{code, detail}; code values equal the
 *    reference StatusCode (include/hundun/v04_status.hpp:13-22).  Calls never
 *    throw and never publish a solution on failure.
 *  - One plan = one HIP stream (passed at creation, NULL = default stream).
 *    Plans are not thread safe; calls on one plan are serialised on its
 *    stream.
 */
#ifndef HUNDUN_FLOW_H
#define HUNDUN_FLOW_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* StatusCode, include/hundun/v04_status.hpp:13-22 */
enum hf_status_code {
  HF_OK = 0,
  HF_INVALID_CASE = 1,
  HF_INVALID_PLAN = 2,
  HF_ALLOCATION_FAILURE = 3,
  HF_MPI_FAILURE = 4, /* kept for numbering parity; never produced */
  HF_NUMERICAL_FAILURE = 5,
  HF_REJECTED_STEP = 6,
  HF_IO_FAILURE = 7
};

typedef struct hf_status {
  int32_t code;
  uint32_t detail;
} hf_status;

/* MgBoundaryKind, include/hundun/v04_linear.hpp:100-104 */
enum hf_boundary_kind {
  HF_BC_DIRICHLET = 0,
  HF_BC_NEUMANN = 1,
  HF_BC_PERIODIC = 2
};

/* Storage precision of the multigrid V-cycle.  The outer Krylov loop, its
 * residual, its inner products and the coarsest-level solve are FP64 in
 * every mode. */
enum hf_precision {
  HF_PRECISION_FP64 = 0,
  HF_PRECISION_FP32 = 1,
  HF_PRECISION_BF16 = 2 /* BF16 storage, FP32 arithmetic */
};

/* LinearTermination, include/hundun/v04_linear.hpp:75-86 */
enum hf_termination {
  HF_TERM_CONVERGED = 0,
  HF_TERM_ZERO_RHS = 1,
  HF_TERM_MAXIMUM_ITERATIONS = 2,
  HF_TERM_BREAKDOWN = 3,
  HF_TERM_NON_FINITE = 4,
  HF_TERM_OPERATOR_FAILURE = 5,
  HF_TERM_PRECONDITIONER_FAILURE = 6,
  HF_TERM_COLLECTIVE_FAILURE = 7,
  HF_TERM_INVALID_PLAN = 8
};

/* Geometry: CartesianGeometryPlan axis metrics from explicit face
 * coordinates (runtime_axes_v1, include/hundun/v04_case.hpp:28-32) and the
 * six physical boundary kinds (MgBoundarySet, v04_linear.hpp:136-143). */
typedef struct hf_grid_desc {
  int32_t nx, ny, nz;
  const double* faces_x; /* host, nx+1 increasing coordinates */
  const double* faces_y; /* host, ny+1 */
  const double* faces_z; /* host, nz+1 */
  int32_t boundary[6];   /* x_min, x_max, y_min, y_max, z_min, z_max */
} hf_grid_desc;

/* MgHierarchyPolicy (v04_linear.hpp:115-131) plus the spec fields of
 * NativeCartesianMgSpec (v04_linear.hpp:329-342) that change arithmetic. */
typedef struct hf_policy {
  double anisotropy_threshold;  /* 4.0; >= max/min spacing ratio keeps point RB */
  uint8_t pre_sweeps;           /* 2 */
  uint8_t post_sweeps;          /* 2 */
  uint8_t maximum_levels;       /* 32 */
  uint8_t minimum_coarse_extent;/* 3 */
  uint16_t coarse_sweeps;       /* 24 */
  uint8_t f_cycle;              /* 0 V-cycle, 1 F-cycle (MgCycleKind) */
  uint8_t null_space_constant;  /* MgNullSpace::constant */
  uint8_t unit_linear;          /* MgCorrectionScaling::unit_linear */
  uint8_t precision;            /* enum hf_precision */
  /* 0: reference aggregation, coarse face = sum of the fine faces tiling it
   *    (src/solver_mg.cpp:1706-1962).
   * 1: distance-corrected aggregation (not in the reference): a coarse face
   *    whose normal axis was coarsened is additionally scaled by 1/2, the
   *    ratio of fine to coarse centre distance, which makes the coarse rows a
   *    finite-volume rediscretisation and the V-cycle grid independent. */
  uint8_t coarse_face_scaling;
  /* Line relaxation order on anisotropic levels (src/solver_mg.cpp:
   * 3352-3536):
   * 0: lexicographic over lines, the reference's order, bit for bit (one
   *    workgroup walks the line hyperplanes: exact but latency bound);
   * 1: zebra (not in the reference): lines with (a+b) even, then odd
   *    (reverse sweeps odd, even), each colour one parallel pass; periodic
   *    ghosts are refreshed per colour pass. */
  uint8_t line_order;
  /* MgPointSmootherKind (v04_linear.hpp:104-107): 0 red-black (omega 0.72),
   * 1 Chebyshev-Jacobi on every point level but the coarsest
   * (src/solver_mg.cpp:2916-3090). */
  uint8_t point_smoother;
  /* chebyshev_lower_spectrum_fraction, 0.3: the smoother targets the
   * spectrum of D^-1 A in [2 fraction, 2]. */
  double chebyshev_lower_fraction;
  /* diagonal_shortcut_maximum_ratio, 0 = off: with no null space, the
   * preconditioner is z = D^-1 r when every finest row's face sum over its
   * diagonal is at most this ratio (< 1; src/solver_mg.cpp:1573-1611). */
  double diagonal_shortcut_ratio;
} hf_policy;

/* LinearSolveControl (v04_linear.hpp:454-466) restricted to PCG. */
typedef struct hf_solve_control {
  double absolute_tolerance;
  double relative_tolerance;
  uint32_t maximum_iterations;
  uint32_t true_residual_interval;
  uint8_t flexible; /* 0: beta = rho'/rho (reference); 1: Polak-Ribiere */
} hf_solve_control;

/* LinearSolveControl (v04_linear.hpp:454-466) for FGMRES. */
typedef struct hf_fgmres_control {
  double absolute_tolerance;
  double relative_tolerance;
  uint32_t maximum_iterations;
  uint32_t true_residual_interval;
  uint32_t restart;     /* basis length, 1..64 (12 in the reference's
                           production pressure solve) */
  /* 1: operator rows per unit volume (A divided by the cell volume w_x w_y
   * w_z) with the multigrid applied to V r: the cold pressure system and
   * its VolumeScaledPreconditioner (src/solver_cold.hpp:378-415). */
  uint8_t volume_rows;
} hf_fgmres_control;

/* LinearSolveResult subset (v04_linear.hpp:530-579). */
typedef struct hf_solve_result {
  hf_status status;
  int32_t termination; /* enum hf_termination */
  uint32_t iterations;
  double initial_true_residual;
  double final_true_residual;
  double recursive_residual;
  double rhs_norm;
  double true_residual_limit;
  uint64_t operator_applies;
  uint64_t preconditioner_applies;
  uint64_t reduction_calls;
  double solve_milliseconds; /* device time of the whole solve */
  uint64_t norm_breakdown_restarts; /* FGMRES: unsafe-recurrence and
                                       happy-breakdown restarts */
} hf_solve_result;

typedef struct hf_level_info {
  int32_t nx, ny, nz;
  uint8_t coarsen_mask; /* bit0 x, bit1 y, bit2 z */
  uint8_t line_mask;
  uint8_t precision;    /* storage precision of this level */
  uint8_t on_tail;      /* level handled by the single-workgroup tail */
} hf_level_info;

typedef struct hf_plan hf_plan;

/* Defaults of MgHierarchyPolicy (v04_linear.hpp:115-131), unit_linear,
 * constant null space, FP64. */
hf_status hf_policy_default(hf_policy* policy);

/* Level planning + device allocation.  Replaces make_mg_workspace_requirements
 * (src/solver_mg_workspace.cpp:317-453) and the symbolic half of
 * NativeCartesianMgPlan::compile (src/solver_mg.cpp:4361-4602).
 * stream: hipStream_t or NULL. */
hf_status hf_plan_create(const hf_grid_desc* grid, const hf_policy* policy,
                         void* stream, hf_plan** out);
hf_status hf_plan_destroy(hf_plan* plan);

/* Face coefficient rule beta_f = A_f / (d_L/m_L + d_R/m_R) and D = sum of
 * faces from a per-cell mobility m (= 1/rho).  Replaces face_pressure_
 * coefficient / fill_pressure_coefficients (src/solver_piso.cpp:1060-1160).
 * Outputs are device arrays sized as in the layout note. */
hf_status hf_assemble_coefficients(hf_plan* plan, const double* mobility,
                                   double* diag, double* bx, double* by,
                                   double* bz);

/* Pressure-correction system of the low-Mach step: replaces
 * assemble_pressure_system_impl (src/solver_pressure.cpp:110-332),
 *   D = a0 V drho_dp + (six face coefficients of the cell),
 *   b = -(V (a0 rho + a1 rho_n + a2 rho_nm1 - source) + outward flux of phi),
 * V = w_x w_y w_z of the plan's grid.  Faces bx, by, bz (>= 0) and fluxes
 * phi_x, phi_y, phi_z use the face layout; rho_nm1 NULL selects BDF1 (a2 is
 * then unused), source NULL no mass source.  Every input and every D and b
 * is checked first (positive densities, drho_dp, V and D; finite values);
 * on a failure nothing is written (numerical_failure, detail 9111). */
hf_status hf_assemble_pressure_system(
    hf_plan* plan, const double* bx, const double* by, const double* bz,
    const double* phi_x, const double* phi_y, const double* phi_z,
    const double* rho, const double* rho_n, const double* rho_nm1,
    const double* drho_dp, const double* source, double a0, double a1,
    double a2, double* diag, double* rhs);

/* Side kinds of the cold pressure coefficients. */
enum hf_cold_side {
  HF_COLD_OTHER = 0,          /* wall, inlet: exterior face coefficient 0 */
  HF_COLD_PERIODIC = 1,
  HF_COLD_PRESSURE_OUTLET = 2 /* exterior face adds 2 projection to D */
};

/* Cold-path pressure coefficients: replaces the coefficient loop of the
 * product step (src/core_product_freeze.cpp:14279-14317) over the face
 * records of src/solver_cold.hpp:990-1110, given as face arrays
 * (projection and limiter_diffusion per face):
 *   face = 0 on a non-periodic exterior face or next to an inactive cell,
 *          else projection + limiter_diffusion / 2;
 *   D    = V rho / ((p_ref + p) dt) (0 with fixed_eos) + the six faces
 *          + 2 projection per pressure-outlet face (active cells), 1 on
 *          inactive cells.
 * fluid: cell activity bytes, NULL = all active.  kinds: x_min .. z_max,
 * enum hf_cold_side (host array). */
hf_status hf_cold_coefficients(hf_plan* plan, const double* proj_x,
                               const double* proj_y, const double* proj_z,
                               const double* lim_x, const double* lim_y,
                               const double* lim_z, const uint8_t* fluid,
                               const double* rho, const double* pressure,
                               double pressure_reference, double dt,
                               int32_t fixed_eos, const int32_t* kinds,
                               double* diag, double* cx, double* cy,
                               double* cz);

/* Numeric setup: IBM masking of level 0 (src/solver_mg.cpp:1469-1570),
 * coarse aggregation (src/solver_mg.cpp:1706-1962) and the certified
 * hierarchy check (src/solver_mg.cpp:1965-2019).  Replaces the numeric half of
 * NativeCartesianMgPlan::compile and update_coefficients
 * (src/solver_mg.cpp:4604-4795).  Activity arrays are 0/1 bytes; pass all
 * four or all NULL (full Cartesian domain). */
hf_status hf_plan_set_coefficients(hf_plan* plan, const double* diag,
                                   const double* bx, const double* by,
                                   const double* bz,
                                   const uint8_t* cell_active,
                                   const uint8_t* x_face_active,
                                   const uint8_t* y_face_active,
                                   const uint8_t* z_face_active);

/* y = A x with the exact IBM pressure operator (solid rows identity, link
 * faces removed).  Replaces IbmPressureOperator::apply
 * (src/solver_ibm_pressure.cpp:256-273) over PressureLinearOperator::apply
 * (src/solver_pressure.cpp:969-991). */
hf_status hf_operator_apply(hf_plan* plan, const double* x, double* y);

/* z = M r, one multigrid application.  Replaces NativeCartesianMgPlan::apply
 * (src/solver_mg.cpp:5124-5131 -> apply_impl :4814-5122). */
hf_status hf_mg_apply(hf_plan* plan, const double* residual,
                      double* correction);

/* PCG solve, x holds the initial guess on entry.  Replaces solve_pcg
 * (src/solver_krylov.cpp:1299-1555) with the MG plan as preconditioner. */
hf_status hf_pcg_solve(hf_plan* plan, const double* rhs, double* solution,
                       const hf_solve_control* control,
                       hf_solve_result* result);

/* hf_pcg_solve that allocates nothing.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory or create HIP objects
 * inside their loop (tight serving loops, harnesses that forbid hipMalloc,
 * stream work a caller captures around the solve).  The plan's own buffers
 * are used as by hf_pcg_solve, but the reduction scratch (per-block
 * partials, folded partials, the FP64 scalar block) lives in one static
 * device array of 8 MiB per GPU, held until the process ends; a grid whose
 * reductions need more (about 88M cells and up) is refused with
 * HF_INVALID_PLAN, detail 9110.  Every call on one GPU shares that array:
 * calls that overlap on one GPU (two streams, two host threads) race.  No
 * HIP events are created: solve_milliseconds is host wall time up to the
 * final stream synchronisation, and hf_plan_profile records nothing. */
hf_status hf_pcg_solve_static(hf_plan* plan, const double* rhs,
                              double* solution,
                              const hf_solve_control* control,
                              hf_solve_result* result);

/* Restarted single-reduction FGMRES (shifted Arnoldi: one reduction per
 * iteration gives the Hessenberg column and the new vector's norm), x holds
 * the initial guess on entry.  Replaces solve_fgmres
 * (src/solver_krylov.cpp:1557-2650) with the MG plan as the flexible
 * preconditioner; with volume_rows, the operator and preconditioner pair of
 * the cold pressure solve (src/core_product_freeze.cpp:14405-14412).  The
 * basis (2 restart + 1 FP64 vectors) is allocated on the first call that
 * needs it and kept by the plan. */
hf_status hf_fgmres_solve(hf_plan* plan, const double* rhs, double* solution,
                          const hf_fgmres_control* control,
                          hf_solve_result* result);

/* Adjoint gradient of a scalar loss L(p) with p the gauge-fixed pressure
 * solution of A(rho) p = b (zero fluid mean).  Given dL/dp (grad_p) and the
 * forward solution p, solves A lambda = P grad_p (A is symmetric, so the
 * transpose solve is the same solve) and returns
 *   grad_b      = P lambda                      (zero on solid cells)
 *   grad_logrho = -sum_f (lambda_L-lambda_R)(p_L-p_R) d beta_f / d log rho
 * using the mobility m = 1/rho that built the current coefficients.
 * Either output may be NULL.  The hierarchy is recomputed, not stored. */
hf_status hf_adjoint_gradients(hf_plan* plan, const double* mobility,
                               const double* pressure, const double* grad_p,
                               const hf_solve_control* control,
                               double* grad_b, double* grad_logrho,
                               hf_solve_result* result);

/* Device vector engine for Krylov solves over operators and
 * preconditioners that are not a plan's own pair (any LinearOperator /
 * LinearPreconditioner of the C++ layer): the reference's Krylov vector
 * kernels and ReductionEngine sums (src/solver_krylov.cpp:176-399, :631-711;
 * src/solver_reduction.cpp:651-700) on one stream, with `slots` device
 * vectors of `capacity` FP64 elements owned by the engine.  Sums are
 * deterministic two-stage block reductions.  A non-finite value produced by
 * hf_vector_axpby or hf_vector_copy is reported by the next hf_vector_dot
 * (numerical_failure, detail 9107). */
typedef struct hf_vector_engine hf_vector_engine;
hf_status hf_vector_engine_create(int64_t capacity, int32_t slots,
                                  void* stream, hf_vector_engine** out);
hf_status hf_vector_engine_destroy(hf_vector_engine* engine);
hf_status hf_vector_engine_slot(hf_vector_engine* engine, int32_t slot,
                                double** vector);
/* result = a . b (host value; waits for the stream). */
hf_status hf_vector_dot(hf_vector_engine* engine, int64_t n, const double* a,
                        const double* b, double* result);
/* dest = a left + b right (dest may alias left or right). */
hf_status hf_vector_axpby(hf_vector_engine* engine, int64_t n, double a,
                          const double* left, double b, const double* right,
                          double* dest);
hf_status hf_vector_copy(hf_vector_engine* engine, int64_t n,
                         const double* source, double* destination);

/* Gauge: subtract the Euclidean mean over fluid cells (solids stay 0). */
hf_status hf_project_mean(hf_plan* plan, double* field);

hf_status hf_plan_level_count(const hf_plan* plan, int32_t* levels);
hf_status hf_plan_level_info(const hf_plan* plan, int32_t level,
                             hf_level_info* info);
/* Copies one level array of the FP64 hierarchy to a device buffer:
 * which = 0 D, 1 x-faces, 2 y-faces, 3 z-faces, 4 fluid volume. */
hf_status hf_plan_level_array(const hf_plan* plan, int32_t level,
                              int32_t which, double* out);
/* Device bytes held by the plan (hierarchy + workspaces). */
hf_status hf_plan_device_bytes(const hf_plan* plan, uint64_t* bytes);

/* Per-kernel-class device timing of the last hf_pcg_solve when enabled
 * (HIP events on the plan stream). */
hf_status hf_plan_set_profiling(hf_plan* plan, int32_t enabled);
hf_status hf_plan_profile(const hf_plan* plan, double* milliseconds,
                          uint64_t* launches, int32_t capacity,
                          int32_t* count);

const char* hf_status_message(hf_status status);

#ifdef __cplusplus
}
#endif
#endif
