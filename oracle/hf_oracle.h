/*
 * CPU oracle for the pressure-projection hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (Hundun-Flow v0.4 solver_* pressure path) used as the checker for
 * the HIP implementation.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product library never links
 * or calls it.
 *
 * Each function cites the reference file:line it restates (paths relative to
 * versions/v0.4/ of the reference tree).
 *
 * Layout convention (identical to the product C ABI):
 *   cell arrays   : x + nx*(y + ny*z), no ghosts
 *   x-face arrays : (nx+1)*ny*nz, index i + (nx+1)*(j + ny*k)
 *   y-face arrays : nx*(ny+1)*nz, index i + nx*(j + (ny+1)*k)
 *   z-face arrays : nx*ny*(nz+1), index i + nx*(j + ny*k)
 */
#ifndef HF_ORACLE_H
#define HF_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_DIRICHLET = 0, OR_NEUMANN = 1, OR_PERIODIC = 2 };

typedef struct or_policy {
  double anisotropy_threshold;  /* 4.0  */
  int pre_sweeps;               /* 2    */
  int post_sweeps;              /* 2    */
  int maximum_levels;           /* 32   */
  int coarse_sweeps;            /* 24   */
  int minimum_coarse_extent;    /* 3    */
  int f_cycle;                  /* 0 = V, 1 = F */
  int null_space_constant;      /* 0 none, 1 constant */
  int unit_linear;              /* 1 unit_linear, 0 residual_minimizing */
  int coarse_face_scaling;      /* 0 reference aggregation (sum of fine
                                   faces); 1 distance-corrected: a coarse face
                                   whose normal axis was coarsened is scaled
                                   by 1/2 (the coarse centre distance doubles) */
  int line_order;               /* line relaxation order: 0 lexicographic over
                                   lines (the reference); 1 zebra: lines with
                                   (a+b) even, then odd (reverse: odd, even),
                                   periodic ghosts refreshed per colour pass */
  int point_smoother;           /* 0 red-black, 1 Chebyshev-Jacobi
                                   (MgPointSmootherKind) */
  double chebyshev_lower_fraction;  /* 0.3 */
  double diagonal_shortcut_ratio;   /* 0 off; else D^-1 when every finest
                                       row has face sum / D <= this */
} or_policy;

typedef struct or_fgmres_control {
  double absolute_tolerance;
  double relative_tolerance;
  uint32_t maximum_iterations;
  uint32_t true_residual_interval;
  uint32_t restart;
  /* 1: operator rows per unit volume (V^-1 A) with the multigrid applied to
   * V r (VolumeScaledPreconditioner) */
  int volume_rows;
} or_fgmres_control;

typedef struct or_control {
  double absolute_tolerance;
  double relative_tolerance;
  uint32_t maximum_iterations;
  uint32_t true_residual_interval;
  int flexible; /* 0: reference beta = rho'/rho; 1: Polak-Ribiere beta */
} or_control;

typedef struct or_result {
  int status;       /* 0 ok, otherwise StatusCode value */
  int termination;  /* LinearTermination value */
  uint32_t iterations;
  double initial_true_residual;
  double final_true_residual;
  double rhs_norm;
  uint64_t operator_applies;
  uint64_t preconditioner_applies;
  double breakdown_value; /* the last gated scalar (r.z or p.Ap) */
} or_result;

typedef struct or_mg or_mg;

/* Face coefficient rule, src/solver_piso.cpp:1060-1118 with mobility m per
 * cell, and the diagonal D = sum of the six faces (pure Poisson rows). */
int or_assemble(const int n[3], const double* faces_x, const double* faces_y,
                const double* faces_z, const int bc[6], const double* mobility,
                double* bx, double* by, double* bz, double* diag);

/* Exact IBM pressure operator: regular 7-point action
 * (src/solver_pressure.cpp:746-811 with src/solver_piso.cpp:1553-1595) then
 * IBM decoration (src/solver_ibm_pressure.cpp:347-385).  cell_active may be
 * NULL (no IBM). */
void or_operator_apply(const int n[3], const int bc[6], const double* diag,
                       const double* bx, const double* by, const double* bz,
                       const uint8_t* cell_active, const double* x,
                       double* y);

/* Masked level-0 operator action used by the multigrid (same matrix as the
 * IBM operator, src/solver_mg.cpp:1469-1570 + :2405-2411). */
void or_masked_apply(const or_mg* mg, const double* x, double* y);

or_mg* or_mg_create(const int n[3], const double* widths_x,
                    const double* widths_y, const double* widths_z,
                    const int bc[6], const or_policy* policy,
                    const double* diag, const double* bx, const double* by,
                    const double* bz, const uint8_t* cell_active,
                    const uint8_t* fx_active, const uint8_t* fy_active,
                    const uint8_t* fz_active, int* status);
void or_mg_destroy(or_mg* mg);
int or_mg_apply(or_mg* mg, const double* residual, double* correction);
int or_mg_level_count(const or_mg* mg);
/* shape[3], coarsen_mask, line_mask */
int or_mg_level_info(const or_mg* mg, int level, int* shape, int* coarsen_mask,
                     int* line_mask);
/* pointers to level storage: 0 D, 1 fx, 2 fy, 3 fz, 4 vol, 5 x, 6 b, 7 r */
const double* or_mg_level_array(const or_mg* mg, int level, int which);
/* One smoother call on a level with the level's current x/b (test hook). */
int or_mg_smooth(or_mg* mg, int level, int sweeps, int reverse,
                 int retain_defect);
int or_mg_restrict(or_mg* mg, int fine_level);
int or_mg_prolongate(or_mg* mg, int fine_level);
double* or_mg_level_slot(or_mg* mg, int level, int which);

/* solve_pcg restatement, src/solver_krylov.cpp:1299-1555.  Operator is the
 * exact IBM operator (or_operator_apply); the preconditioner is the MG plan
 * (mg may be NULL: identity). */
int or_pcg_solve(const int n[3], const int bc[6], const double* diag,
                 const double* bx, const double* by, const double* bz,
                 const uint8_t* cell_active, or_mg* mg, const double* rhs,
                 double* solution, const or_control* control,
                 or_result* result);

/* The same PCG recurrence over dense row-major n x n matrices (test form):
 * operator a, preconditioner m (NULL: identity).  trace, when not NULL,
 * receives x after each of the first trace_capacity iterations. */
int or_pcg_dense(int n, const double* a, const double* m, const double* rhs,
                 double* solution, const or_control* control,
                 or_result* result, double* trace, uint32_t trace_capacity);

/* solve_fgmres restatement (single-reduction shifted Arnoldi, restarted),
 * src/solver_krylov.cpp:1557-2650, without the recycle and convergence-audit
 * options.  Operator: the exact IBM operator (rows divided by the cell
 * volume with volume_rows); preconditioner: the MG plan (NULL: identity),
 * applied to V r with volume_rows.  norm_breakdown_restarts is reported in
 * result->breakdown_value. */
int or_fgmres_solve(const int n[3], const double* widths_x,
                    const double* widths_y, const double* widths_z,
                    const int bc[6], const double* diag, const double* bx,
                    const double* by, const double* bz,
                    const uint8_t* cell_active, or_mg* mg, const double* rhs,
                    double* solution, const or_fgmres_control* control,
                    or_result* result);

/* The same FGMRES over dense row-major matrices (test form); trace receives
 * the candidate x0 + Z y after each of the first trace_capacity iterations
 * of the first Arnoldi cycle. */
int or_fgmres_dense(int n, const double* a, const double* m,
                    const double* rhs, double* solution,
                    const or_fgmres_control* control, or_result* result,
                    double* trace, uint32_t trace_capacity);

/* 1 when the plan selected the diagonal shortcut (z = D^-1 r). */
int or_mg_diagonal_shortcut(const or_mg* mg);

/* Pressure-correction system (src/solver_pressure.cpp:110-332):
 * D = a0 V drho_dp + face sum, b = -(V (a0 rho + a1 rho_n + a2 rho_nm1 - S)
 * + outward flux of phi).  rho_nm1 NULL selects BDF1 (a2 unused), source
 * NULL no mass source.  Nothing is written unless every input and result
 * passes the reference's checks. */
int or_assemble_pressure_system(const int n[3], const double* widths_x,
                                const double* widths_y, const double* widths_z,
                                const double* bx, const double* by,
                                const double* bz, const double* phi_x,
                                const double* phi_y, const double* phi_z,
                                const double* rho, const double* rho_n,
                                const double* rho_nm1, const double* drho_dp,
                                const double* source, double a0, double a1,
                                double a2, double* diag, double* rhs);

/* Cold-path pressure coefficients (src/core_product_freeze.cpp:
 * 14279-14317): face coefficient = 0 on non-periodic exterior faces and
 * faces next to an inactive cell, else projection + 0.5 limiter_diffusion;
 * D = V rho / ((p_ref + p) dt) (0 with fixed_eos) + the six faces
 * + 2 projection per pressure-outlet face for active cells, 1 for inactive
 * ones.  kinds per side: 0 other, 1 periodic, 2 pressure outlet.  fluid may
 * be NULL (every cell active). */
int or_cold_coefficients(const int n[3], const double* widths_x,
                         const double* widths_y, const double* widths_z,
                         const double* proj_x, const double* proj_y,
                         const double* proj_z, const double* lim_x,
                         const double* lim_y, const double* lim_z,
                         const uint8_t* fluid, const double* rho,
                         const double* pressure, double pressure_reference,
                         double dt, int fixed_eos, const int kinds[6],
                         double* diag, double* cx, double* cy, double* cz);

/* Vector kernels, src/solver_krylov.cpp:176-204 and :650-711. */
double or_dot(const double* a, const double* b, size_t n);
double or_norm2(const double* v, size_t n);

#ifdef __cplusplus
}
#endif
#endif
