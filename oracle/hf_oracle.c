/*
 * CPU oracle for the pressure-projection hot path -- TEST INFRASTRUCTURE.
 *
 * A single-rank plain-C restatement of the reference algorithm (paths are
 * relative to versions/v0.4/ of the reference tree).  It is the checker for
 * the HIP product path and the "port" CPU baseline; it is never shipped in or
 * called by the product library.  Build with -ffp-contract=off so that every
 * arithmetic step rounds exactly like the reference release build.
 *
 * Single rank means the halo engine reduces to periodic wrap-around and the
 * reduction engine to a local sum; both are restated as index arithmetic.
 */
#include "hf_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

enum { ST_OK = 0, ST_INVALID_PLAN = 2, ST_NUMERICAL = 5, ST_REJECTED = 6 };
enum {
  TERM_CONVERGED = 0,
  TERM_ZERO_RHS = 1,
  TERM_MAXIMUM_ITERATIONS = 2,
  TERM_BREAKDOWN = 3,
  TERM_NON_FINITE = 4
};

#define OR_MAX_LEVELS 32

typedef struct or_level {
  int n[3];
  int coarsen_mask; /* bit0 x, bit1 y, bit2 z: axes coarsened toward next */
  int line_mask;
  size_t cells, nfx, nfy, nfz;
  double *D, *fx, *fy, *fz, *vol;
  double *x, *b, *r, *t;
  int chebyshev; /* Chebyshev-Jacobi point smoother on this level */
} or_level;

struct or_mg {
  int bc[6];
  or_policy policy;
  int levels;
  or_level level[OR_MAX_LEVELS];
  uint8_t* cell_active; /* copy, level-0 cells (NULL = all active) */
  int line_failed;      /* a Thomas pivot vanished during the last apply */
  int shortcut;         /* diagonal shortcut selected: z = D^-1 r */
};

static size_t cidx(const int n[3], int i, int j, int k) {
  return (size_t)i + (size_t)n[0] * ((size_t)j + (size_t)n[1] * (size_t)k);
}
static size_t xfidx(const int n[3], int i, int j, int k) {
  return (size_t)i + (size_t)(n[0] + 1) * ((size_t)j + (size_t)n[1] * (size_t)k);
}
static size_t yfidx(const int n[3], int i, int j, int k) {
  return (size_t)i + (size_t)n[0] * ((size_t)j + (size_t)(n[1] + 1) * (size_t)k);
}
static size_t zfidx(const int n[3], int i, int j, int k) {
  return (size_t)i + (size_t)n[0] * ((size_t)j + (size_t)n[1] * (size_t)k);
}

/* ------------------------------------------------------------------ */
/* Face coefficient rule: src/solver_piso.cpp:1060-1118.               */
/* Geometry metrics: src/mesh_focus.cpp:598-608 (centre = f + 0.5 w).  */
/* Face area: src/solver_cartesian.cpp:319-331.                        */
/* ------------------------------------------------------------------ */
static double face_rule(const double* faces, int count, int normal, int kind,
                        double area, double m_left, double m_right) {
  /* m_left / m_right are the mobilities of the cells on either side; for a
   * periodic physical face they are the wrapped cells (halo-filled). */
  if ((normal == 0 || normal == count) && kind == OR_NEUMANN) return 0.0;
  if ((normal == 0 || normal == count) && kind == OR_DIRICHLET) {
    /* owner = inside cell; ghost state mobility taken equal to the owner's
     * (homogeneous EOS-closed ghost). */
    const double owner = normal == 0 ? m_right : m_left;
    double distance;
    if (normal == 0) {
      const double w0 = faces[1] - faces[0];
      distance = (faces[0] + 0.5 * w0) - faces[0];
    } else {
      const double wl = faces[count] - faces[count - 1];
      distance = faces[count] - (faces[count - 1] + 0.5 * wl);
    }
    return 2.0 * area / (distance / owner + distance / owner);
  }
  double left_distance, right_distance;
  if (normal == 0 || normal == count) { /* periodic physical face */
    left_distance = 0.5 * (faces[count] - faces[count - 1]);
    right_distance = 0.5 * (faces[1] - faces[0]);
  } else {
    const double location = faces[normal];
    const double wl = faces[normal] - faces[normal - 1];
    const double wr = faces[normal + 1] - faces[normal];
    left_distance = location - (faces[normal - 1] + 0.5 * wl);
    right_distance = (faces[normal] + 0.5 * wr) - location;
  }
  return area / (left_distance / m_left + right_distance / m_right);
}

int or_assemble(const int n[3], const double* faces_x, const double* faces_y,
                const double* faces_z, const int bc[6], const double* m,
                double* bx, double* by, double* bz, double* diag) {
  const int nx = n[0], ny = n[1], nz = n[2];
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i <= nx; ++i) {
        const double area = (faces_y[j + 1] - faces_y[j]) *
                            (faces_z[k + 1] - faces_z[k]);
        const int il = i == 0 ? nx - 1 : i - 1;
        const int ir = i == nx ? 0 : i;
        const int kind = i == 0 ? bc[0] : bc[1];
        const double v = face_rule(faces_x, nx, i, kind, area,
                                   m[cidx(n, il, j, k)], m[cidx(n, ir, j, k)]);
        if (!isfinite(v) || v < 0.0) return ST_NUMERICAL;
        bx[xfidx(n, i, j, k)] = v;
      }
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const double area = (faces_x[i + 1] - faces_x[i]) *
                            (faces_z[k + 1] - faces_z[k]);
        const int jl = j == 0 ? ny - 1 : j - 1;
        const int jr = j == ny ? 0 : j;
        const int kind = j == 0 ? bc[2] : bc[3];
        const double v = face_rule(faces_y, ny, j, kind, area,
                                   m[cidx(n, i, jl, k)], m[cidx(n, i, jr, k)]);
        if (!isfinite(v) || v < 0.0) return ST_NUMERICAL;
        by[yfidx(n, i, j, k)] = v;
      }
  for (int k = 0; k <= nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const double area = (faces_x[i + 1] - faces_x[i]) *
                            (faces_y[j + 1] - faces_y[j]);
        const int kl = k == 0 ? nz - 1 : k - 1;
        const int kr = k == nz ? 0 : k;
        const int kind = k == 0 ? bc[4] : bc[5];
        const double v = face_rule(faces_z, nz, k, kind, area,
                                   m[cidx(n, i, j, kl)], m[cidx(n, i, j, kr)]);
        if (!isfinite(v) || v < 0.0) return ST_NUMERICAL;
        bz[zfidx(n, i, j, k)] = v;
      }
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i)
        diag[cidx(n, i, j, k)] =
            bx[xfidx(n, i, j, k)] + bx[xfidx(n, i + 1, j, k)] +
            by[yfidx(n, i, j, k)] + by[yfidx(n, i, j + 1, k)] +
            bz[zfidx(n, i, j, k)] + bz[zfidx(n, i, j, k + 1)];
  return ST_OK;
}

/* ------------------------------------------------------------------ */
/* Out-of-patch neighbour: src/solver_piso.cpp:1553-1595 (operator)     */
/* and src/solver_mg.cpp:2237-2272 / :2356-2368 (MG point rows).        */
/* ------------------------------------------------------------------ */
static double neighbour(const double* x, const int n[3], const int bc[6],
                        int i, int j, int k, int axis, int dir) {
  int c[3] = {i, j, k};
  c[axis] += dir;
  if (c[axis] >= 0 && c[axis] < n[axis]) return x[cidx(n, c[0], c[1], c[2])];
  const int kind = bc[2 * axis + (dir > 0 ? 1 : 0)];
  if (kind == OR_DIRICHLET) return 0.0;
  if (kind == OR_NEUMANN) return x[cidx(n, i, j, k)];
  c[axis] = dir < 0 ? n[axis] - 1 : 0;
  return x[cidx(n, c[0], c[1], c[2])];
}

/* Stencil in the reference evaluation order (src/solver_mg.cpp:2405-2411,
 * src/solver_pressure.cpp:757-798): D*x, then x-lo, x-hi, y-lo, y-hi, z-lo,
 * z-hi, each subtracted in turn. */
static double row_value(const int n[3], const int bc[6], const double* D,
                        const double* fx, const double* fy, const double* fz,
                        const double* x, int i, int j, int k) {
  double v = D[cidx(n, i, j, k)] * x[cidx(n, i, j, k)];
  v -= fx[xfidx(n, i, j, k)] * neighbour(x, n, bc, i, j, k, 0, -1);
  v -= fx[xfidx(n, i + 1, j, k)] * neighbour(x, n, bc, i, j, k, 0, 1);
  v -= fy[yfidx(n, i, j, k)] * neighbour(x, n, bc, i, j, k, 1, -1);
  v -= fy[yfidx(n, i, j + 1, k)] * neighbour(x, n, bc, i, j, k, 1, 1);
  v -= fz[zfidx(n, i, j, k)] * neighbour(x, n, bc, i, j, k, 2, -1);
  v -= fz[zfidx(n, i, j, k + 1)] * neighbour(x, n, bc, i, j, k, 2, 1);
  return v;
}

void or_operator_apply(const int n[3], const int bc[6], const double* diag,
                       const double* bx, const double* by, const double* bz,
                       const uint8_t* act, const double* x, double* y) {
  for (int k = 0; k < n[2]; ++k)
    for (int j = 0; j < n[1]; ++j)
      for (int i = 0; i < n[0]; ++i)
        y[cidx(n, i, j, k)] = row_value(n, bc, diag, bx, by, bz, x, i, j, k);
  if (act == NULL) return;
  /* IBM decoration, src/solver_ibm_pressure.cpp:347-385: solid rows become
   * identity; each fluid/solid link adds beta*(x_solid - x_fluid) to the
   * fluid row.  Links are enumerated per solid cell and axis direction. */
  for (int k = 0; k < n[2]; ++k)
    for (int j = 0; j < n[1]; ++j)
      for (int i = 0; i < n[0]; ++i) {
        const size_t s = cidx(n, i, j, k);
        if (act[s]) continue;
        y[s] = x[s];
      }
  for (int k = 0; k < n[2]; ++k)
    for (int j = 0; j < n[1]; ++j)
      for (int i = 0; i < n[0]; ++i) {
        const size_t f = cidx(n, i, j, k);
        if (!act[f]) continue;
        for (int axis = 0; axis < 3; ++axis)
          for (int dir = -1; dir <= 1; dir += 2) {
            int c[3] = {i, j, k};
            c[axis] += dir;
            if (c[axis] < 0 || c[axis] >= n[axis]) {
              if (bc[2 * axis + (dir > 0)] != OR_PERIODIC) continue;
              c[axis] = dir < 0 ? n[axis] - 1 : 0;
            }
            const size_t s = cidx(n, c[0], c[1], c[2]);
            if (act[s]) continue;
            double face;
            if (axis == 0)
              face = bx[xfidx(n, dir < 0 ? i : i + 1, j, k)];
            else if (axis == 1)
              face = by[yfidx(n, i, dir < 0 ? j : j + 1, k)];
            else
              face = bz[zfidx(n, i, j, dir < 0 ? k : k + 1)];
            y[f] += face * (x[s] - x[f]);
          }
      }
}

/* ------------------------------------------------------------------ */
/* Level planner: src/solver_mg_workspace.cpp:113-143 (strategy mask),  */
/* :181-199 (coarsen_axis), :317-453 (level loop).                      */
/* ------------------------------------------------------------------ */
static void spacing(const double* w, int count, double* mn, double* mx,
                    double* rep) {
  *mn = INFINITY;
  *mx = 0.0;
  *rep = 0.0;
  for (int i = 0; i < count; ++i) {
    if (w[i] < *mn) *mn = w[i];
    if (w[i] > *mx) *mx = w[i];
    *rep += w[i];
  }
  *rep /= (double)count;
}

static int strategy_mask(const int n[3], const double* const w[3],
                         double threshold, int* line_mask) {
  double mn[3], mx[3], rep[3];
  for (int a = 0; a < 3; ++a) spacing(w[a], n[a], &mn[a], &mx[a], &rep[a]);
  double largest = rep[0];
  if (rep[1] > largest) largest = rep[1];
  if (rep[2] > largest) largest = rep[2];
  *line_mask = 0;
  for (int a = 0; a < 3; ++a) {
    const int internally = mn[a] > 0.0 && mx[a] / mn[a] >= threshold;
    const int cross = rep[a] > 0.0 && largest / rep[a] >= threshold;
    if (internally || cross) *line_mask |= 1 << a;
  }
  if ((*line_mask & (*line_mask - 1)) != 0) return 7;
  const int candidate = 7 & ~*line_mask;
  return candidate == 0 ? 7 : candidate;
}

/* ------------------------------------------------------------------ */
/* Hierarchy coefficients.                                              */
/* ------------------------------------------------------------------ */
static double faces_sum(const or_level* L, const double* fx, const double* fy,
                        const double* fz, int i, int j, int k) {
  return fx[xfidx(L->n, i, j, k)] + fx[xfidx(L->n, i + 1, j, k)] +
         fy[yfidx(L->n, i, j, k)] + fy[yfidx(L->n, i, j + 1, k)] +
         fz[zfidx(L->n, i, j, k)] + fz[zfidx(L->n, i, j, k + 1)];
}

/* src/solver_mg.cpp:1469-1570 */
static void copy_finest(or_mg* mg, const double* const w[3], const double* D,
                        const double* bx, const double* by, const double* bz,
                        const uint8_t* ca, const uint8_t* fxa,
                        const uint8_t* fya, const uint8_t* fza) {
  or_level* L = &mg->level[0];
  const int* n = L->n;
  for (int k = 0; k < n[2]; ++k)
    for (int j = 0; j < n[1]; ++j)
      for (int i = 0; i < n[0]; ++i) {
        const size_t c = cidx(n, i, j, k);
        L->D[c] = D[c];
        L->vol[c] = w[0][i] * w[1][j] * w[2][k];
        if (ca != NULL && ca[c] == 0) L->vol[c] = 0.0;
      }
  for (size_t f = 0; f < L->nfx; ++f)
    L->fx[f] = bx[f] * (fxa == NULL ? 1.0 : (double)fxa[f]);
  for (size_t f = 0; f < L->nfy; ++f)
    L->fy[f] = by[f] * (fya == NULL ? 1.0 : (double)fya[f]);
  for (size_t f = 0; f < L->nfz; ++f)
    L->fz[f] = bz[f] * (fza == NULL ? 1.0 : (double)fza[f]);
  if (ca == NULL) return;
  for (int k = 0; k < n[2]; ++k)
    for (int j = 0; j < n[1]; ++j)
      for (int i = 0; i < n[0]; ++i) {
        const size_t c = cidx(n, i, j, k);
        if (ca[c] == 0) {
          L->D[c] = 1.0;
          continue;
        }
        const double original =
            bx[xfidx(n, i, j, k)] + bx[xfidx(n, i + 1, j, k)] +
            by[yfidx(n, i, j, k)] + by[yfidx(n, i, j + 1, k)] +
            bz[zfidx(n, i, j, k)] + bz[zfidx(n, i, j, k + 1)];
        const double active = faces_sum(L, L->fx, L->fy, L->fz, i, j, k);
        L->D[c] = L->D[c] - original + active;
      }
}

/* Channel aggregation, src/solver_mg.cpp:1706-1856 (single rank):
 * x -> y -> z over the coarsened axes; cells and tangential faces sum their
 * two children, normal faces take the aligned fine face min(2c, n). */
static void aggregate(const or_level* F, int face_axis, const double* src,
                      double* dst, double* s1, double* s2) {
  int ext[3] = {F->n[0], F->n[1], F->n[2]};
  int cells[3] = {F->n[0], F->n[1], F->n[2]};
  if (face_axis >= 0) ext[face_axis] += 1;
  size_t count = (size_t)ext[0] * ext[1] * ext[2];
  memcpy(s1, src, count * sizeof(double));
  double* cur = s1;
  double* nxt = s2;
  for (int axis = 0; axis < 3; ++axis) {
    if ((F->coarsen_mask & (1 << axis)) == 0) continue;
    const int global_extent = cells[axis];
    int ncells[3] = {cells[0], cells[1], cells[2]};
    ncells[axis] = (cells[axis] + 1) / 2;
    int next[3] = {ncells[0], ncells[1], ncells[2]};
    if (face_axis >= 0) next[face_axis] += 1;
    for (int k = 0; k < next[2]; ++k)
      for (int j = 0; j < next[1]; ++j)
        for (int i = 0; i < next[0]; ++i) {
          int o[3] = {i, j, k};
          double value = 0.0;
          if (face_axis == axis) {
            int s[3] = {i, j, k};
            s[axis] = 2 * o[axis] < global_extent ? 2 * o[axis] : global_extent;
            value = cur[(size_t)s[0] + (size_t)ext[0] * ((size_t)s[1] + (size_t)ext[1] * s[2])];
          } else {
            for (int child = 0; child < 2; ++child) {
              const int fg = 2 * o[axis] + child;
              if (fg >= global_extent) continue;
              int s[3] = {i, j, k};
              s[axis] = fg;
              value += cur[(size_t)s[0] + (size_t)ext[0] * ((size_t)s[1] + (size_t)ext[1] * s[2])];
            }
          }
          nxt[(size_t)i + (size_t)next[0] * ((size_t)j + (size_t)next[1] * k)] = value;
        }
    double* t = cur;
    cur = nxt;
    nxt = t;
    memcpy(cells, ncells, sizeof cells);
    memcpy(ext, next, sizeof ext);
  }
  count = (size_t)ext[0] * ext[1] * ext[2];
  memcpy(dst, cur, count * sizeof(double));
}

/* src/solver_mg.cpp:1875-1962 */
static int build_coarse(or_mg* mg) {
  size_t maxv = 0;
  for (int l = 0; l < mg->levels; ++l) {
    const or_level* L = &mg->level[l];
    if (L->cells > maxv) maxv = L->cells;
    if (L->nfx > maxv) maxv = L->nfx;
    if (L->nfy > maxv) maxv = L->nfy;
    if (L->nfz > maxv) maxv = L->nfz;
  }
  double* s1 = (double*)malloc(maxv * sizeof(double));
  double* s2 = (double*)malloc(maxv * sizeof(double));
  double* tmp = (double*)malloc(maxv * sizeof(double));
  int status = ST_OK;
  for (int l = 1; l < mg->levels && status == ST_OK; ++l) {
    const or_level* F = &mg->level[l - 1];
    or_level* C = &mg->level[l];
    aggregate(F, -1, F->vol, C->vol, s1, s2);
    for (int k = 0; k < F->n[2]; ++k)
      for (int j = 0; j < F->n[1]; ++j)
        for (int i = 0; i < F->n[0]; ++i) {
          const size_t c = cidx(F->n, i, j, k);
          if (F->vol[c] == 0.0) {
            tmp[c] = 0.0;
            continue;
          }
          const double faces = faces_sum(F, F->fx, F->fy, F->fz, i, j, k);
          const double reaction = F->D[c] - faces;
          tmp[c] = reaction > 0.0 ? reaction : 0.0;
        }
    aggregate(F, -1, tmp, C->D, s1, s2);
    aggregate(F, 0, F->fx, C->fx, s1, s2);
    aggregate(F, 1, F->fy, C->fy, s1, s2);
    aggregate(F, 2, F->fz, C->fz, s1, s2);
    if (mg->policy.coarse_face_scaling) {
      /* Not in the reference: rediscretise the aggregated transmissibility
       * over the doubled centre distance of a coarsened normal axis. */
      if (F->coarsen_mask & 1)
        for (size_t f = 0; f < C->nfx; ++f) C->fx[f] *= 0.5;
      if (F->coarsen_mask & 2)
        for (size_t f = 0; f < C->nfy; ++f) C->fy[f] *= 0.5;
      if (F->coarsen_mask & 4)
        for (size_t f = 0; f < C->nfz; ++f) C->fz[f] *= 0.5;
    }
    for (int k = 0; k < C->n[2]; ++k)
      for (int j = 0; j < C->n[1]; ++j)
        for (int i = 0; i < C->n[0]; ++i) {
          const size_t c = cidx(C->n, i, j, k);
          C->D[c] += faces_sum(C, C->fx, C->fy, C->fz, i, j, k);
          if (C->vol[c] == 0.0) C->D[c] = 1.0;
          if (C->vol[c] < 0.0 || !isfinite(C->vol[c]) || !(C->D[c] > 0.0) ||
              !isfinite(C->D[c]))
            status = ST_NUMERICAL;
        }
  }
  free(s1);
  free(s2);
  free(tmp);
  return status;
}

/* select_diagonal_shortcut, src/solver_mg.cpp:1573-1611: with no null
 * space, use D^-1 when the largest finest-row face sum over the diagonal
 * (masked coefficients, physical boundary faces included) is at most the
 * policy ratio. */
static int select_diagonal_shortcut(const or_mg* mg) {
  if (mg->policy.diagonal_shortcut_ratio == 0.0 ||
      mg->policy.null_space_constant)
    return 0;
  const or_level* L = &mg->level[0];
  double maximum = 0.0;
  for (int k = 0; k < L->n[2]; ++k)
    for (int j = 0; j < L->n[1]; ++j)
      for (int i = 0; i < L->n[0]; ++i) {
        const double sum = faces_sum(L, L->fx, L->fy, L->fz, i, j, k);
        const double ratio = sum / L->D[cidx(L->n, i, j, k)];
        if (ratio > maximum) maximum = ratio; /* std::max(maximum, ratio) */
      }
  return maximum <= mg->policy.diagonal_shortcut_ratio;
}

int or_mg_diagonal_shortcut(const or_mg* mg) { return mg->shortcut; }

or_mg* or_mg_create(const int n[3], const double* wx, const double* wy,
                    const double* wz, const int bc[6], const or_policy* policy,
                    const double* diag, const double* bx, const double* by,
                    const double* bz, const uint8_t* ca, const uint8_t* fxa,
                    const uint8_t* fya, const uint8_t* fza, int* status) {
  *status = ST_INVALID_PLAN;
  if (n[0] <= 0 || n[1] <= 0 || n[2] <= 0) return NULL;
  for (int a = 0; a < 3; ++a)
    if ((bc[2 * a] == OR_PERIODIC) != (bc[2 * a + 1] == OR_PERIODIC))
      return NULL;
  if (policy->pre_sweeps <= 0 || policy->post_sweeps <= 0 ||
      policy->coarse_sweeps <= 0 || policy->maximum_levels < 2 ||
      policy->maximum_levels > OR_MAX_LEVELS ||
      policy->minimum_coarse_extent < 2 || !(policy->anisotropy_threshold > 1.0) ||
      policy->point_smoother < 0 || policy->point_smoother > 1 ||
      !(policy->chebyshev_lower_fraction > 0.0) ||
      !(policy->chebyshev_lower_fraction < 1.0) ||
      !(policy->diagonal_shortcut_ratio >= 0.0) ||
      !(policy->diagonal_shortcut_ratio < 1.0))
    return NULL;
  or_mg* mg = (or_mg*)calloc(1, sizeof(or_mg));
  memcpy(mg->bc, bc, sizeof mg->bc);
  mg->policy = *policy;
  const double* w[3] = {wx, wy, wz};
  int line_mask = 0;
  const int preferred = strategy_mask(n, w, policy->anisotropy_threshold,
                                      &line_mask);
  int g[3] = {n[0], n[1], n[2]};
  for (int l = 0; l < policy->maximum_levels; ++l) {
    or_level* L = &mg->level[l];
    memcpy(L->n, g, sizeof g);
    L->line_mask = line_mask;
    int next[3] = {g[0], g[1], g[2]};
    int mask = 0;
    for (int a = 0; a < 3; ++a) {
      if ((preferred & (1 << a)) == 0) continue;
      if (g[a] <= policy->minimum_coarse_extent) continue;
      next[a] = (g[a] + 1) / 2;
      mask |= 1 << a;
    }
    L->coarsen_mask = mask;
    mg->levels = l + 1;
    if (mask == 0 || mg->levels == policy->maximum_levels) break;
    memcpy(g, next, sizeof g);
  }
  if (mg->levels < 2) {
    free(mg);
    return NULL;
  }
  for (int l = 0; l < mg->levels; ++l) {
    or_level* L = &mg->level[l];
    L->cells = (size_t)L->n[0] * L->n[1] * L->n[2];
    L->nfx = (size_t)(L->n[0] + 1) * L->n[1] * L->n[2];
    L->nfy = (size_t)L->n[0] * (L->n[1] + 1) * L->n[2];
    L->nfz = (size_t)L->n[0] * L->n[1] * (L->n[2] + 1);
    L->D = (double*)calloc(L->cells, sizeof(double));
    L->vol = (double*)calloc(L->cells, sizeof(double));
    L->fx = (double*)calloc(L->nfx, sizeof(double));
    L->fy = (double*)calloc(L->nfy, sizeof(double));
    L->fz = (double*)calloc(L->nfz, sizeof(double));
    L->x = (double*)calloc(L->cells, sizeof(double));
    L->b = (double*)calloc(L->cells, sizeof(double));
    L->r = (double*)calloc(L->cells, sizeof(double));
    L->t = (double*)calloc(L->cells, sizeof(double));
  }
  if (ca != NULL) {
    mg->cell_active = (uint8_t*)malloc(mg->level[0].cells);
    memcpy(mg->cell_active, ca, mg->level[0].cells);
  }
  copy_finest(mg, w, diag, bx, by, bz, ca, fxa, fya, fza);
  /* the Chebyshev-Jacobi smoother on every point level but the coarsest
   * (src/solver_mg.cpp:1048-1062) */
  for (int l = 0; l < mg->levels; ++l) {
    or_level* L = &mg->level[l];
    L->chebyshev = l + 1 < mg->levels && L->line_mask == 0 &&
                   policy->point_smoother == 1;
  }
  mg->shortcut = select_diagonal_shortcut(mg);
  if (mg->shortcut) {
    /* the coarse hierarchy of a diagonal plan is never read
     * (src/solver_mg.cpp:4556-4566) */
    *status = ST_OK;
    return mg;
  }
  *status = build_coarse(mg);
  if (*status != ST_OK) {
    or_mg_destroy(mg);
    return NULL;
  }
  return mg;
}

void or_mg_destroy(or_mg* mg) {
  if (mg == NULL) return;
  for (int l = 0; l < mg->levels; ++l) {
    or_level* L = &mg->level[l];
    free(L->D); free(L->vol); free(L->fx); free(L->fy); free(L->fz);
    free(L->x); free(L->b); free(L->r); free(L->t);
  }
  free(mg->cell_active);
  free(mg);
}

int or_mg_level_count(const or_mg* mg) { return mg->levels; }

int or_mg_level_info(const or_mg* mg, int level, int* shape, int* cm,
                     int* lm) {
  if (level < 0 || level >= mg->levels) return ST_INVALID_PLAN;
  memcpy(shape, mg->level[level].n, 3 * sizeof(int));
  *cm = mg->level[level].coarsen_mask;
  *lm = mg->level[level].line_mask;
  return ST_OK;
}

const double* or_mg_level_array(const or_mg* mg, int level, int which) {
  const or_level* L = &mg->level[level];
  switch (which) {
    case 0: return L->D;
    case 1: return L->fx;
    case 2: return L->fy;
    case 3: return L->fz;
    case 4: return L->vol;
    case 5: return L->x;
    case 6: return L->b;
    case 7: return L->r;
    default: return NULL;
  }
}

double* or_mg_level_slot(or_mg* mg, int level, int which) {
  return (double*)or_mg_level_array(mg, level, which);
}

static void level_matrix(const or_mg* mg, const or_level* L, const double* x,
                         double* out) {
  for (int k = 0; k < L->n[2]; ++k)
    for (int j = 0; j < L->n[1]; ++j)
      for (int i = 0; i < L->n[0]; ++i)
        out[cidx(L->n, i, j, k)] =
            row_value(L->n, mg->bc, L->D, L->fx, L->fy, L->fz, x, i, j, k);
}

static void level_defect(const or_mg* mg, const or_level* L, const double* x,
                         const double* b, double* out) {
  for (int k = 0; k < L->n[2]; ++k)
    for (int j = 0; j < L->n[1]; ++j)
      for (int i = 0; i < L->n[0]; ++i) {
        const size_t c = cidx(L->n, i, j, k);
        out[c] = b[c] - row_value(L->n, mg->bc, L->D, L->fx, L->fy, L->fz, x,
                                  i, j, k);
      }
}

void or_masked_apply(const or_mg* mg, const double* x, double* y) {
  level_matrix(mg, &mg->level[0], x, y);
}

/* src/solver_mg.cpp:3112-3126: the fused colour update is allowed unless a
 * periodic axis has an odd extent on this level. */
static int fused_allowed(const or_mg* mg, const or_level* L) {
  for (int a = 0; a < 3; ++a)
    if (mg->bc[2 * a] == OR_PERIODIC && (L->n[a] & 1)) return 0;
  return 1;
}

/* Red-black point smoother, src/solver_mg.cpp:3149-3349 (omega 0.72). */
static void point_smooth(or_mg* mg, int level, int sweeps, int reverse,
                         int retain) {
  or_level* L = &mg->level[level];
  const double omega = 0.72;
  const int fused = fused_allowed(mg, L);
  const int* n = L->n;
  if (!fused || sweeps == 0) {
    if (retain && sweeps == 0)
      level_defect(mg, L, L->x, L->b, L->r);
    else
      level_matrix(mg, L, L->x, L->r);
  }
  for (int sweep = 0; sweep < sweeps; ++sweep) {
    const int first = reverse ? 1 : 0;
    for (int pass = 0; pass < 2; ++pass) {
      const int color = (first + pass) & 1;
      /* Same-colour cells never neighbour under the fused-allowed rule, so
       * the update is order independent; the reference loop order is kept. */
      for (int kk = 0; kk < n[2]; ++kk) {
        const int k = reverse ? n[2] - 1 - kk : kk;
        for (int jj = 0; jj < n[1]; ++jj) {
          const int j = reverse ? n[1] - 1 - jj : jj;
          for (int ii = 0; ii < n[0]; ++ii) {
            const int i = reverse ? n[0] - 1 - ii : ii;
            if (((i + j + k) & 1) != color) continue;
            const size_t c = cidx(n, i, j, k);
            const double value =
                fused ? row_value(n, mg->bc, L->D, L->fx, L->fy, L->fz, L->x,
                                  i, j, k)
                      : L->r[c];
            L->x[c] += omega * (L->b[c] - value) / L->D[c];
          }
        }
      }
      const int final_update = sweep + 1 == sweeps && pass == 1;
      if (!final_update || retain) {
        if (!fused || (final_update && retain)) {
          if (retain && final_update)
            level_defect(mg, L, L->x, L->b, L->r);
          else
            level_matrix(mg, L, L->x, L->r);
        }
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* Line relaxation, src/solver_mg.cpp:3352-3536 (line_smooth_axis) and  */
/* :3540-3590 (line_smooth, smooth).                                    */
/* ------------------------------------------------------------------ */
/* MgHierarchyPolicy::line_relaxation_maximum_extent default
 * (include/hundun/v04_linear.hpp:123): longer lines use the point smoother. */
#define OR_LINE_MAXIMUM_EXTENT 4096

/* Cell m of line (a, b) along axis ax: x-lines run over (y=a, z=b), y-lines
 * over (x=a, z=b), z-lines over (x=a, y=b). */
static void line_cell(int ax, int a, int b, int m, int c[3]) {
  if (ax == 0) {
    c[0] = m; c[1] = a; c[2] = b;
  } else if (ax == 1) {
    c[0] = a; c[1] = m; c[2] = b;
  } else {
    c[0] = a; c[1] = b; c[2] = m;
  }
}

static double face_low(const or_level* L, int axis, const int c[3]) {
  if (axis == 0) return L->fx[xfidx(L->n, c[0], c[1], c[2])];
  if (axis == 1) return L->fy[yfidx(L->n, c[0], c[1], c[2])];
  return L->fz[zfidx(L->n, c[0], c[1], c[2])];
}

static double face_high(const or_level* L, int axis, const int c[3]) {
  if (axis == 0) return L->fx[xfidx(L->n, c[0] + 1, c[1], c[2])];
  if (axis == 1) return L->fy[yfidx(L->n, c[0], c[1] + 1, c[2])];
  return L->fz[zfidx(L->n, c[0], c[1], c[2] + 1)];
}

/* level_neighbor (src/solver_mg.cpp:2229-2265) on one rank: in-range reads
 * the current iterate, a periodic wrap reads the ghost copy made when the
 * sweep (zebra: the colour pass) began, Neumann clamps to the cell itself,
 * Dirichlet is 0. */
static double line_neighbour(const or_mg* mg, const or_level* L,
                             const double* ghost, const int c[3], int axis,
                             int dir) {
  int s[3] = {c[0], c[1], c[2]};
  s[axis] += dir;
  if (s[axis] >= 0 && s[axis] < L->n[axis])
    return L->x[cidx(L->n, s[0], s[1], s[2])];
  const int kind = mg->bc[2 * axis + (dir > 0 ? 1 : 0)];
  if (kind == OR_DIRICHLET) return 0.0;
  if (kind == OR_NEUMANN) return L->x[cidx(L->n, c[0], c[1], c[2])];
  s[axis] = dir < 0 ? L->n[axis] - 1 : 0;
  return ghost[cidx(L->n, s[0], s[1], s[2])];
}

/* One Thomas solve (solve_one, src/solver_mg.cpp:3390-3502): c' in the
 * temporary slot, d' in the residual slot, then back substitution into x.
 * Returns 0 (x of the line untouched) when a pivot is not above DBL_MIN. */
static int solve_line(const or_mg* mg, or_level* L, const double* ghost,
                      int ax, int a, int b) {
  const int E = L->n[ax];
  int c[3], p[3];
  for (int m = 0; m < E; ++m) {
    line_cell(ax, a, b, m, c);
    const size_t ci = cidx(L->n, c[0], c[1], c[2]);
    const double lo = face_low(L, ax, c);
    const double hi = face_high(L, ax, c);
    const double sub = m == 0 ? 0.0 : -lo;
    const double dia = L->D[ci];
    const double sup = m + 1 == E ? 0.0 : -hi;
    double d = L->b[ci];
    for (int t = 0; t < 3; ++t) {
      if (t == ax) continue;
      d += face_low(L, t, c) * line_neighbour(mg, L, ghost, c, t, -1);
      d += face_high(L, t, c) * line_neighbour(mg, L, ghost, c, t, 1);
    }
    if (m == 0) d += lo * line_neighbour(mg, L, ghost, c, ax, -1);
    if (m + 1 == E) d += hi * line_neighbour(mg, L, ghost, c, ax, 1);
    if (m > 0) {
      line_cell(ax, a, b, m - 1, p);
      const size_t pi = cidx(L->n, p[0], p[1], p[2]);
      const double den = dia - sub * L->t[pi];
      if (!(fabs(den) > DBL_MIN)) return 0;
      L->t[ci] = sup / den;
      L->r[ci] = (d - sub * L->r[pi]) / den;
    } else {
      if (!(fabs(dia) > DBL_MIN)) return 0;
      L->t[ci] = sup / dia;
      L->r[ci] = d / dia;
    }
  }
  for (int m = E; m-- > 0;) {
    line_cell(ax, a, b, m, c);
    const size_t ci = cidx(L->n, c[0], c[1], c[2]);
    if (m + 1 == E) {
      L->x[ci] = L->r[ci];
    } else {
      line_cell(ax, a, b, m + 1, p);
      L->x[ci] = L->r[ci] - L->t[ci] * L->x[cidx(L->n, p[0], p[1], p[2])];
    }
  }
  return 1;
}


static void line_smooth_axis(or_mg* mg, int level, int ax, int sweeps,
                             int reverse) {
  or_level* L = &mg->level[level];
  const int E = L->n[ax];
  if (E < 1 || E > OR_LINE_MAXIMUM_EXTENT) {
    point_smooth(mg, level, sweeps, reverse, 0);
    return;
  }
  const int A = ax == 0 ? L->n[1] : L->n[0];
  const int B = ax == 2 ? L->n[1] : L->n[2];
  double* ghost = (double*)malloc(L->cells * sizeof(double));
  int ok = 1;
  for (int sweep = 0; sweep < sweeps; ++sweep) {
    if (mg->policy.line_order == 0) {
      memcpy(ghost, L->x, L->cells * sizeof(double));
      if (reverse) {
        for (int b = B; b-- > 0;)
          for (int a = A; a-- > 0;) ok = solve_line(mg, L, ghost, ax, a, b) && ok;
      } else {
        for (int b = 0; b < B; ++b)
          for (int a = 0; a < A; ++a) ok = solve_line(mg, L, ghost, ax, a, b) && ok;
      }
    } else {
      for (int pass = 0; pass < 2; ++pass) {
        const int colour = ((reverse ? 1 : 0) + pass) & 1;
        memcpy(ghost, L->x, L->cells * sizeof(double));
        for (int b = 0; b < B; ++b)
          for (int a = 0; a < A; ++a)
            if (((a + b) & 1) == colour)
              ok = solve_line(mg, L, ghost, ax, a, b) && ok;
      }
    }
  }
  free(ghost);
  if (!ok) mg->line_failed = 1;
}

/* Chebyshev-Jacobi smoothing, src/solver_mg.cpp:2916-3090 with the
 * streamed stage of :2440-2560: the Jacobi-preconditioned spectrum is
 * bounded by [fraction * 2, 2]; stage s updates the direction
 * d = f_prev * d + f_res * (b - A x) / D and x += d from the previous
 * stage's x (Jacobi: all neighbours from the previous stage).  The retained
 * defect is b - A x after the last stage. */
static void chebyshev_smooth(or_mg* mg, int level, int stages, int retain) {
  or_level* L = &mg->level[level];
  const int* n = L->n;
  const double beta = 2.0;
  const double alpha = mg->policy.chebyshev_lower_fraction * beta;
  const double theta = (beta + alpha) * 0.5;
  const double delta = (beta - alpha) * 0.5;
  const double sigma = theta / delta;
  double previous_rho = 1.0 / sigma;
  double* current = L->x;
  double* next = L->t;
  for (int stage = 0; stage < stages; ++stage) {
    const double rho = stage == 0 ? 0.0 : 1.0 / (2.0 * sigma - previous_rho);
    const double previous_factor = stage == 0 ? 0.0 : rho * previous_rho;
    const double residual_factor = stage == 0 ? 1.0 / theta : 2.0 * rho / delta;
    if (stage != 0) previous_rho = rho;
    for (int k = 0; k < n[2]; ++k)
      for (int j = 0; j < n[1]; ++j)
        for (int i = 0; i < n[0]; ++i) {
          const size_t c = cidx(n, i, j, k);
          const double value = row_value(n, mg->bc, L->D, L->fx, L->fy, L->fz,
                                         current, i, j, k);
          const double defect = L->b[c] - value;
          const double scaled = residual_factor * defect / L->D[c];
          const double direction =
              stage == 0 ? scaled : previous_factor * L->r[c] + scaled;
          L->r[c] = direction;
          next[c] = current[c] + direction;
        }
    double* swap = current;
    current = next;
    next = swap;
  }
  if (current != L->x) memcpy(L->x, current, L->cells * sizeof(double));
  if (retain) level_defect(mg, L, L->x, L->b, L->r);
}

/* smooth(), src/solver_mg.cpp:3557-3590: point levels use red-black (or
 * Chebyshev-Jacobi, :3240-3244), line levels relax each line axis in x, y,
 * z order (reverse: z, y, x). */
static void smooth(or_mg* mg, int level, int sweeps, int reverse, int retain) {
  const int mask = mg->level[level].line_mask;
  if (mg->level[level].chebyshev) {
    chebyshev_smooth(mg, level, sweeps, retain);
    return;
  }
  if (mask == 0) {
    point_smooth(mg, level, sweeps, reverse, retain);
    return;
  }
  for (int i = 0; i < 3; ++i) {
    const int ax = reverse ? 2 - i : i;
    if (mask & (1 << ax)) line_smooth_axis(mg, level, ax, sweeps, reverse);
  }
}

int or_mg_smooth(or_mg* mg, int level, int sweeps, int reverse, int retain) {
  if (level < 0 || level >= mg->levels) return ST_INVALID_PLAN;
  smooth(mg, level, sweeps, reverse, retain);
  return ST_OK;
}

/* src/solver_mg.cpp:1412-1454: volume-weighted mean removal. */
static int project_constant(const or_level* L, double* field) {
  double weighted = 0.0, volume = 0.0;
  for (size_t c = 0; c < L->cells; ++c) {
    weighted += L->vol[c] * field[c];
    volume += L->vol[c];
  }
  if (!(volume > 0.0) || !isfinite(weighted) || !isfinite(volume))
    return ST_NUMERICAL;
  const double mean = weighted / volume;
  for (size_t c = 0; c < L->cells; ++c)
    field[c] = L->vol[c] > 0.0 ? field[c] - mean : 0.0;
  return ST_OK;
}

/* src/solver_mg.cpp:3633-3720: coarse rhs = sum of children (dz, dy, dx
 * order), coarse x := 0, optional constant projection. */
int or_mg_restrict(or_mg* mg, int fine) {
  const or_level* F = &mg->level[fine];
  or_level* C = &mg->level[fine + 1];
  const int mx = F->coarsen_mask & 1, my = (F->coarsen_mask >> 1) & 1,
            mz = (F->coarsen_mask >> 2) & 1;
  for (int ck = 0; ck < C->n[2]; ++ck)
    for (int cj = 0; cj < C->n[1]; ++cj)
      for (int ci = 0; ci < C->n[0]; ++ci) {
        const int xb = mx ? 2 * ci : ci, yb = my ? 2 * cj : cj,
                  zb = mz ? 2 * ck : ck;
        double sum = 0.0;
        for (int dz = 0; dz < (mz ? 2 : 1); ++dz)
          for (int dy = 0; dy < (my ? 2 : 1); ++dy)
            for (int dx = 0; dx < (mx ? 2 : 1); ++dx) {
              const int fx = xb + dx, fy = yb + dy, fz = zb + dz;
              if (fx >= F->n[0] || fy >= F->n[1] || fz >= F->n[2]) continue;
              sum += F->r[cidx(F->n, fx, fy, fz)];
            }
        const size_t c = cidx(C->n, ci, cj, ck);
        C->b[c] = sum;
        C->x[c] = 0.0;
      }
  if (mg->policy.null_space_constant) return project_constant(C, C->b);
  return ST_OK;
}

static void interpolation(int fine_global, int coarsened, int coarse_extent,
                          int periodic, int* left, int* right, double* rw) {
  if (!coarsened) {
    *left = fine_global;
    *right = fine_global;
    *rw = 0.0;
    return;
  }
  const int parent = fine_global / 2;
  if ((fine_global & 1) == 0) {
    *left = parent - 1;
    *right = parent;
    *rw = 0.75;
  } else {
    *left = parent;
    *right = parent + 1;
    *rw = 0.25;
  }
  if (!periodic && *left < 0) {
    *left = 0;
    *right = 0;
    *rw = 0.0;
  } else if (!periodic && *right >= coarse_extent) {
    *left = coarse_extent - 1;
    *right = coarse_extent - 1;
    *rw = 0.0;
  }
}

static int wrap(int v, int extent) {
  if (v < 0) return v + extent;
  if (v >= extent) return v - extent;
  return v;
}

/* src/solver_mg.cpp:3893-3993: trilinear (0.75/0.25) add, periodic via the
 * wrapped halo, clamp at physical faces, inactive fine cells set to 0. */
int or_mg_prolongate(or_mg* mg, int fine) {
  or_level* F = &mg->level[fine];
  const or_level* C = &mg->level[fine + 1];
  const int m[3] = {F->coarsen_mask & 1, (F->coarsen_mask >> 1) & 1,
                    (F->coarsen_mask >> 2) & 1};
  const int per[3] = {mg->bc[0] == OR_PERIODIC, mg->bc[2] == OR_PERIODIC,
                      mg->bc[4] == OR_PERIODIC};
  for (int k = 0; k < F->n[2]; ++k)
    for (int j = 0; j < F->n[1]; ++j)
      for (int i = 0; i < F->n[0]; ++i) {
        const size_t c = cidx(F->n, i, j, k);
        if (F->vol[c] == 0.0) {
          F->x[c] = 0.0;
          continue;
        }
        int xl, xr, yl, yr, zl, zr;
        double xw, yw, zw;
        interpolation(i, m[0], C->n[0], per[0], &xl, &xr, &xw);
        interpolation(j, m[1], C->n[1], per[1], &yl, &yr, &yw);
        interpolation(k, m[2], C->n[2], per[2], &zl, &zr, &zw);
        double value = 0.0;
        for (int az = 0; az < (zl == zr ? 1 : 2); ++az) {
          const int gz = az == 0 ? zl : zr;
          const double wz = az == 0 ? 1.0 - zw : zw;
          for (int ay = 0; ay < (yl == yr ? 1 : 2); ++ay) {
            const int gy = ay == 0 ? yl : yr;
            const double wy = ay == 0 ? 1.0 - yw : yw;
            for (int ax = 0; ax < (xl == xr ? 1 : 2); ++ax) {
              const int gx = ax == 0 ? xl : xr;
              const double wx = ax == 0 ? 1.0 - xw : xw;
              value += wx * wy * wz *
                       C->x[cidx(C->n, wrap(gx, C->n[0]), wrap(gy, C->n[1]),
                                 wrap(gz, C->n[2]))];
            }
          }
        }
        F->x[c] += value;
      }
  return ST_OK;
}

/* src/solver_mg.cpp:4192-4338 */
static int mg_cycle(or_mg* mg, int level, int counter) {
  if (level + 1 == mg->levels) {
    smooth(mg, level, mg->policy.coarse_sweeps, 0, 0);
    return ST_OK;
  }
  smooth(mg, level, mg->policy.pre_sweeps, 0, 1);
  /* line levels recompute the defect (compute_residual, :4256-4262) */
  if (mg->level[level].line_mask != 0)
    level_defect(mg, &mg->level[level], mg->level[level].x,
                 mg->level[level].b, mg->level[level].r);
  int st = or_mg_restrict(mg, level);
  if (st != ST_OK) return st;
  st = mg_cycle(mg, level + 1, counter);
  if (st != ST_OK) return st;
  if (counter > 1) {
    st = mg_cycle(mg, level + 1, counter - 1);
    if (st != ST_OK) return st;
  }
  or_mg_prolongate(mg, level);
  smooth(mg, level, mg->policy.post_sweeps, 1, 0);
  return ST_OK;
}

/* NativeCartesianMgPlan::apply_impl, src/solver_mg.cpp:4814-5122. */
int or_mg_apply(or_mg* mg, const double* residual, double* correction) {
  or_level* L = &mg->level[0];
  const uint8_t* act = mg->cell_active;
  for (size_t c = 0; c < L->cells; ++c) {
    L->x[c] = 0.0;
    L->b[c] = (act == NULL || act[c]) ? residual[c] : 0.0;
  }
  int st;
  if (mg->policy.null_space_constant) {
    st = project_constant(L, L->b);
    if (st != ST_OK) return st;
  }
  mg->line_failed = 0;
  if (mg->shortcut) {
    /* src/solver_mg.cpp:4926-4936 */
    for (size_t c = 0; c < L->cells; ++c) L->x[c] = L->b[c] / L->D[c];
  } else {
    st = mg_cycle(mg, 0, mg->policy.f_cycle ? 2 : 1);
    if (st != ST_OK) return st;
  }
  if (mg->line_failed) return ST_NUMERICAL;
  if (mg->policy.null_space_constant) {
    st = project_constant(L, L->x);
    if (st != ST_OK) return st;
  }
  level_defect(mg, L, L->x, L->b, L->r);
  double p0 = 0.0, p1 = 0.0, p2 = 0.0;
  for (size_t c = 0; c < L->cells; ++c) {
    const double r = (act == NULL || act[c]) ? residual[c] : 0.0;
    const double az = r - L->r[c];
    p0 += r * r;
    p1 += r * az;
    p2 += az * az;
  }
  if (!isfinite(p0) || !isfinite(p1) || !isfinite(p2) || p0 < 0.0 || p2 < 0.0)
    return ST_NUMERICAL;
  const double tiny = DBL_EPSILON * p0;
  const double alpha = mg->policy.unit_linear ? 1.0
                                              : (p2 > tiny ? p1 / p2 : 0.0);
  if (!isfinite(alpha)) return ST_NUMERICAL;
  for (size_t c = 0; c < L->cells; ++c) {
    const int a = act == NULL || act[c];
    L->x[c] = a ? L->x[c] * alpha : 0.0;
    const double r = a ? residual[c] : 0.0;
    L->r[c] = r - alpha * (r - L->r[c]);
  }
  memcpy(correction, L->x, L->cells * sizeof(double));
  return ST_OK;
}

/* ------------------------------------------------------------------ */
/* Krylov vector kernels, src/solver_krylov.cpp:176-204 and :650-711.   */
/* ------------------------------------------------------------------ */
double or_dot(const double* a, const double* b, size_t n) {
  double sum = 0.0, correction = 0.0;
  for (size_t i = 0; i < n; ++i) {
    const double product = a[i] * b[i];
    const double next = sum + product;
    if (fabs(sum) >= fabs(product))
      correction += (sum - next) + product;
    else
      correction += (product - next) + sum;
    sum = next;
  }
  return sum + correction;
}

double or_norm2(const double* v, size_t n) {
  double scale = 0.0;
  for (size_t i = 0; i < n; ++i)
    if (fabs(v[i]) > scale) scale = fabs(v[i]);
  if (scale == 0.0) return 0.0;
  double squares = 0.0, correction = 0.0;
  for (size_t i = 0; i < n; ++i) {
    const double s = v[i] / scale;
    const double sq = s * s;
    const double next = squares + sq;
    if (fabs(squares) >= fabs(sq))
      correction += (squares - next) + sq;
    else
      correction += (sq - next) + squares;
    squares = next;
  }
  squares += correction;
  return scale * sqrt(squares);
}

/* Operator and preconditioner of one Krylov solve. */
typedef struct or_krylov_ops {
  size_t n;
  void (*apply_a)(void* ctx, const double* x, double* y);
  void* a_ctx;
  /* NULL: identity preconditioner */
  int (*apply_m)(void* ctx, const double* r, double* z);
  void* m_ctx;
  /* optional: x after every iteration, row-major [iteration][n] */
  double* trace;
  uint32_t trace_capacity;
} or_krylov_ops;

/* solve_pcg, src/solver_krylov.cpp:1299-1555: true residual every
 * true_residual_interval iterations (and at the recursive gate); an
 * unconverged verification restarts the direction p = M r. */
static int pcg_core(const or_krylov_ops* ops, const double* rhs,
                    double* solution, const or_control* ctl, or_result* res) {
  const size_t N = ops->n;
  memset(res, 0, sizeof *res);
  double* x = (double*)malloc(N * sizeof(double));
  double* r = (double*)malloc(N * sizeof(double));
  double* z = (double*)malloc(N * sizeof(double));
  double* p = (double*)malloc(N * sizeof(double));
  double* ap = (double*)malloc(N * sizeof(double));
  int st = ST_OK;
#define OR_FAIL(code, term) \
  do {                      \
    st = (code);            \
    res->termination = (term); \
    goto done;              \
  } while (0)
  const double rhs_norm = or_norm2(rhs, N);
  res->rhs_norm = rhs_norm;
  const double tol = fmax(ctl->absolute_tolerance,
                          ctl->relative_tolerance * rhs_norm);
  if (rhs_norm == 0.0) {
    memset(solution, 0, N * sizeof(double));
    res->termination = TERM_ZERO_RHS;
    goto done;
  }
  memcpy(x, solution, N * sizeof(double));
  ops->apply_a(ops->a_ctx, x, ap);
  res->operator_applies++;
  for (size_t i = 0; i < N; ++i) r[i] = 1.0 * rhs[i] + -1.0 * ap[i];
  res->initial_true_residual = or_norm2(r, N);
  res->final_true_residual = res->initial_true_residual;
  if (res->initial_true_residual <= tol) {
    memcpy(solution, x, N * sizeof(double));
    res->termination = TERM_CONVERGED;
    goto done;
  }
#define OR_PRECONDITION()                                             \
  do {                                                                \
    if (ops->apply_m != NULL) {                                       \
      const int pst = ops->apply_m(ops->m_ctx, r, z);                 \
      if (pst != ST_OK) OR_FAIL(pst, TERM_NON_FINITE);                \
    } else {                                                          \
      memcpy(z, r, N * sizeof(double));                               \
    }                                                                 \
    res->preconditioner_applies++;                                    \
  } while (0)
  OR_PRECONDITION();
  double rho = or_dot(r, z, N);
  res->breakdown_value = rho;
  if (!(rho > 0.0) || !isfinite(rho)) OR_FAIL(ST_NUMERICAL, TERM_BREAKDOWN);
  memcpy(p, z, N * sizeof(double));
  while (res->iterations < ctl->maximum_iterations) {
    ops->apply_a(ops->a_ctx, p, ap);
    res->operator_applies++;
    const double denominator = or_dot(p, ap, N);
    res->breakdown_value = denominator;
    if (!(denominator > 0.0) || !isfinite(denominator))
      OR_FAIL(ST_NUMERICAL, TERM_BREAKDOWN);
    const double alpha = rho / denominator;
    if (!isfinite(alpha)) OR_FAIL(ST_NUMERICAL, TERM_NON_FINITE);
    for (size_t i = 0; i < N; ++i) x[i] = x[i] + alpha * p[i];
    for (size_t i = 0; i < N; ++i) r[i] = r[i] + -alpha * ap[i];
    if (ops->trace != NULL && res->iterations < ops->trace_capacity)
      memcpy(ops->trace + (size_t)res->iterations * N, x, N * sizeof(double));
    res->iterations++;
    const double recursive = or_norm2(r, N);
    const int verify = recursive <= tol ||
                       res->iterations % ctl->true_residual_interval == 0 ||
                       res->iterations == ctl->maximum_iterations;
    if (verify) {
      ops->apply_a(ops->a_ctx, x, ap);
      res->operator_applies++;
      for (size_t i = 0; i < N; ++i) r[i] = 1.0 * rhs[i] + -1.0 * ap[i];
      res->final_true_residual = or_norm2(r, N);
      if (res->final_true_residual <= tol) {
        memcpy(solution, x, N * sizeof(double));
        res->termination = TERM_CONVERGED;
        goto done;
      }
      if (res->iterations == ctl->maximum_iterations)
        OR_FAIL(ST_REJECTED, TERM_MAXIMUM_ITERATIONS);
      OR_PRECONDITION();
      rho = or_dot(r, z, N);
      res->breakdown_value = rho;
      if (!(rho > 0.0) || !isfinite(rho)) OR_FAIL(ST_NUMERICAL, TERM_BREAKDOWN);
      memcpy(p, z, N * sizeof(double));
      continue;
    }
    OR_PRECONDITION();
    const double rho_next = or_dot(r, z, N);
    res->breakdown_value = rho_next;
    if (!(rho_next > 0.0) || !isfinite(rho_next))
      OR_FAIL(ST_NUMERICAL, TERM_BREAKDOWN);
    double beta;
    if (ctl->flexible) {
      /* Polak-Ribiere: z+ . (r+ - r) / rho, with r+ - r = -alpha * Ap. */
      beta = -alpha * or_dot(z, ap, N) / rho;
    } else {
      beta = rho_next / rho;
    }
    for (size_t i = 0; i < N; ++i) p[i] = 1.0 * z[i] + beta * p[i];
    rho = rho_next;
  }
  OR_FAIL(ST_REJECTED, TERM_MAXIMUM_ITERATIONS);
done:
  res->status = st;
  free(x);
  free(r);
  free(z);
  free(p);
  free(ap);
  return st;
#undef OR_FAIL
#undef OR_PRECONDITION
}

typedef struct ibm_operator_ctx {
  const int* n;
  const int* bc;
  const double *diag, *bx, *by, *bz;
  const uint8_t* act;
} ibm_operator_ctx;

static void ibm_operator_cb(void* ctx, const double* x, double* y) {
  const ibm_operator_ctx* c = (const ibm_operator_ctx*)ctx;
  or_operator_apply(c->n, c->bc, c->diag, c->bx, c->by, c->bz, c->act, x, y);
}

static int mg_precondition_cb(void* ctx, const double* r, double* z) {
  return or_mg_apply((or_mg*)ctx, r, z);
}

int or_pcg_solve(const int n[3], const int bc[6], const double* diag,
                 const double* bx, const double* by, const double* bz,
                 const uint8_t* act, or_mg* mg, const double* rhs,
                 double* solution, const or_control* ctl, or_result* res) {
  ibm_operator_ctx a = {n, bc, diag, bx, by, bz, act};
  or_krylov_ops ops;
  memset(&ops, 0, sizeof ops);
  ops.n = (size_t)n[0] * n[1] * n[2];
  ops.apply_a = ibm_operator_cb;
  ops.a_ctx = &a;
  if (mg != NULL) {
    ops.apply_m = mg_precondition_cb;
    ops.m_ctx = mg;
  }
  return pcg_core(&ops, rhs, solution, ctl, res);
}

/* Dense test form of the same recurrence: y = A x and z = M r with row-major
 * n x n matrices (M NULL: identity). */
typedef struct dense_ctx {
  size_t n;
  const double* m;
} dense_ctx;

static void dense_apply_cb(void* ctx, const double* x, double* y) {
  const dense_ctx* d = (const dense_ctx*)ctx;
  for (size_t i = 0; i < d->n; ++i) {
    double v = 0.0;
    for (size_t j = 0; j < d->n; ++j) v += d->m[i * d->n + j] * x[j];
    y[i] = v;
  }
}

static int dense_precondition_cb(void* ctx, const double* r, double* z) {
  dense_apply_cb(ctx, r, z);
  return ST_OK;
}

int or_pcg_dense(int n, const double* a, const double* m, const double* rhs,
                 double* solution, const or_control* ctl, or_result* res,
                 double* trace, uint32_t trace_capacity) {
  dense_ctx ac = {(size_t)n, a};
  dense_ctx mc = {(size_t)n, m};
  or_krylov_ops ops;
  memset(&ops, 0, sizeof ops);
  ops.n = (size_t)n;
  ops.apply_a = dense_apply_cb;
  ops.a_ctx = &ac;
  if (m != NULL) {
    ops.apply_m = dense_precondition_cb;
    ops.m_ctx = &mc;
  }
  ops.trace = trace;
  ops.trace_capacity = trace_capacity;
  return pcg_core(&ops, rhs, solution, ctl, res);
}

/* ------------------------------------------------------------------ */
/* FGMRES, src/solver_krylov.cpp:1557-2650 (single rank, no recycle, no */
/* convergence audit).                                                  */
/* ------------------------------------------------------------------ */
/* kSingleReductionFgmresShift, src/solver_krylov.cpp:35 */
#define OR_FGMRES_SHIFT 1.0

/* single_reduction_fgmres_norm, src/solver_krylov.cpp:565-604: returns 0 ok,
 * 1 unsafe recurrence (t < 0), -1 non-finite. */
static int fgmres_norm(const double* h, size_t count, double bar, double* t,
                       double* next) {
  if (!isfinite(bar)) return -1;
  *t = bar;
  for (size_t row = 0; row < count; ++row) {
    const double projection = h[row];
    if (!isfinite(projection)) return -1;
    const double square = projection * projection;
    if (!isfinite(square)) return -1;
    *t = *t - square;
    if (!isfinite(*t)) return -1;
  }
  if (*t < 0.0) {
    *next = 0.0;
    return 1;
  }
  *next = sqrt(*t);
  return isfinite(*next) ? 0 : -1;
}

/* add_scaled_basis, src/solver_krylov.cpp:510-555: destination += sum of
 * scales[i] * sources[i], added in row order per element. */
static void add_scaled_basis(double* dst, double* const* src,
                             const double* scales, size_t count, size_t n) {
  for (size_t e = 0; e < n; ++e) {
    double v = dst[e];
    for (size_t row = 0; row < count; ++row) v = v + scales[row] * src[row][e];
    dst[e] = v;
  }
}

/* Upper-triangular back substitution of the rotated Hessenberg columns
 * (column-major, `rows` per column). */
static int fgmres_back_substitute(const double* h, size_t rows,
                                  const double* g, size_t count, double* y) {
  for (size_t reverse = count; reverse > 0; --reverse) {
    const size_t row = reverse - 1;
    double value = g[row];
    for (size_t upper = row + 1; upper < count; ++upper)
      value -= h[upper * rows + row] * y[upper];
    const double divisor = h[row * rows + row];
    if (divisor == 0.0 || !isfinite(divisor)) return TERM_BREAKDOWN;
    y[row] = value / divisor;
    if (!isfinite(y[row])) return TERM_NON_FINITE;
  }
  return 0;
}

static int fgmres_core(const or_krylov_ops* ops, const double* rhs,
                       double* solution, const or_fgmres_control* ctl,
                       or_result* res) {
  const size_t N = ops->n;
  const size_t R = ctl->restart;
  memset(res, 0, sizeof *res);
  if (R == 0 || R > 64 || ctl->maximum_iterations == 0 ||
      ctl->true_residual_interval == 0) {
    res->status = ST_INVALID_PLAN;
    res->termination = 8;
    return ST_INVALID_PLAN;
  }
  const size_t rows = R + 1;
  double* x = (double*)malloc(N * sizeof(double));
  double* ax = (double*)malloc(N * sizeof(double));
  double** v = (double**)malloc((R + 1) * sizeof(double*));
  double** z = (double**)malloc(R * sizeof(double*));
  for (size_t i = 0; i <= R; ++i) v[i] = (double*)malloc(N * sizeof(double));
  for (size_t i = 0; i < R; ++i) z[i] = (double*)malloc(N * sizeof(double));
  double* h = (double*)calloc(rows * R, sizeof(double));
  double* cs = (double*)calloc(R, sizeof(double));
  double* sn = (double*)calloc(R, sizeof(double));
  double* g = (double*)calloc(rows, sizeof(double));
  double* y = (double*)calloc(R, sizeof(double));
  double* scales = (double*)calloc(R + 1, sizeof(double));
  double restarts = 0.0;
  int st = ST_OK;
  const double eps = 2.220446049250313080847e-16;
#define FG_FAIL(code, term) \
  do {                      \
    st = (code);            \
    res->termination = (term); \
    goto done;              \
  } while (0)
#define FG_TRUE_RESIDUAL(vec, out, norm)                                  \
  do {                                                                    \
    ops->apply_a(ops->a_ctx, (vec), ax);                                  \
    res->operator_applies++;                                              \
    for (size_t e = 0; e < N; ++e) (out)[e] = 1.0 * rhs[e] + -1.0 * ax[e]; \
    (norm) = or_norm2((out), N);                                          \
    if (!isfinite(norm)) FG_FAIL(ST_NUMERICAL, 5);                        \
  } while (0)
  const double rhs_norm = or_norm2(rhs, N);
  res->rhs_norm = rhs_norm;
  if (!isfinite(rhs_norm)) FG_FAIL(ST_NUMERICAL, TERM_NON_FINITE);
  const double tol = fmax(ctl->absolute_tolerance,
                          ctl->relative_tolerance * rhs_norm);
  if (rhs_norm == 0.0) {
    memset(solution, 0, N * sizeof(double));
    res->termination = TERM_ZERO_RHS;
    goto done;
  }
  memcpy(x, solution, N * sizeof(double));
  FG_TRUE_RESIDUAL(x, v[0], res->initial_true_residual);
  res->final_true_residual = res->initial_true_residual;
  if (!isfinite(tol)) FG_FAIL(ST_NUMERICAL, TERM_NON_FINITE);
  if (res->initial_true_residual <= tol) {
    memcpy(solution, x, N * sizeof(double));
    res->termination = TERM_CONVERGED;
    goto done;
  }
  while (res->iterations < ctl->maximum_iterations) {
    const double beta = res->final_true_residual;
    const double inv = 1.0 / beta;
    for (size_t e = 0; e < N; ++e) {
      v[0][e] = inv * v[0][e];
      if (!isfinite(v[0][e])) FG_FAIL(ST_NUMERICAL, TERM_NON_FINITE);
    }
    for (size_t i = 0; i < rows; ++i) g[i] = 0.0;
    g[0] = beta;
    int restarted = 0;
    for (size_t column = 0;
         column < R && res->iterations < ctl->maximum_iterations; ++column) {
      if (ops->apply_m != NULL) {
        const int pst = ops->apply_m(ops->m_ctx, v[column], z[column]);
        res->preconditioner_applies++;
        if (pst != ST_OK) FG_FAIL(pst, 6);
      } else {
        memcpy(z[column], v[column], N * sizeof(double));
        res->preconditioner_applies++;
      }
      ops->apply_a(ops->a_ctx, z[column], ax);
      res->operator_applies++;
      for (size_t e = 0; e < N; ++e) ax[e] = ax[e] + -OR_FGMRES_SHIFT * v[column][e];
      const size_t dots = column + 1;
      double* hc = h + column * rows;
      for (size_t row = 0; row < dots; ++row) hc[row] = or_dot(v[row], ax, N);
      hc[dots] = or_dot(ax, ax, N);
      for (size_t row = 0; row <= dots; ++row)
        if (!isfinite(hc[row])) FG_FAIL(ST_NUMERICAL, TERM_NON_FINITE);
      const double bar = hc[dots];
      for (size_t row = 0; row < dots; ++row) scales[row] = -hc[row];
      double t = 0.0, next_norm = 0.0;
      const int nst = fgmres_norm(hc, dots, bar, &t, &next_norm);
      if (nst < 0) FG_FAIL(ST_NUMERICAL, TERM_NON_FINITE);
      const int unsafe = nst == 1;
      if (unsafe && column != 0) {
        /* discard the column, publish the shorter cycle, restart
         * (src/solver_krylov.cpp:2232-2340) */
        res->iterations++;
        const int term = fgmres_back_substitute(h, rows, g, column, y);
        if (term) FG_FAIL(ST_NUMERICAL, term);
        add_scaled_basis(x, z, y, column, N);
        FG_TRUE_RESIDUAL(x, v[0], res->final_true_residual);
        if (res->final_true_residual <= tol) {
          memcpy(solution, x, N * sizeof(double));
          res->termination = TERM_CONVERGED;
          goto done;
        }
        if (res->iterations == ctl->maximum_iterations)
          FG_FAIL(ST_REJECTED, TERM_MAXIMUM_ITERATIONS);
        restarts += 1.0;
        memset(h, 0, rows * R * sizeof(double));
        memset(cs, 0, R * sizeof(double));
        memset(sn, 0, R * sizeof(double));
        memset(g, 0, rows * sizeof(double));
        memset(y, 0, R * sizeof(double));
        restarted = 1;
        break;
      }
      int explicit_pass = 0;
      if (unsafe || next_norm == 0.0) {
        /* second orthogonalisation pass (src/solver_krylov.cpp:2342-2400) */
        add_scaled_basis(ax, v, scales, dots, N);
        for (size_t row = 0; row < dots; ++row) y[row] = or_dot(v[row], ax, N);
        for (size_t row = 0; row < dots; ++row) {
          hc[row] += y[row];
          if (!isfinite(hc[row])) FG_FAIL(ST_NUMERICAL, TERM_NON_FINITE);
          scales[row] = -y[row];
        }
        add_scaled_basis(ax, v, scales, dots, N);
        next_norm = or_norm2(ax, N);
        if (!isfinite(next_norm)) FG_FAIL(ST_NUMERICAL, TERM_NON_FINITE);
        explicit_pass = 1;
      }
      hc[column] += OR_FGMRES_SHIFT;
      if (!isfinite(hc[column])) FG_FAIL(ST_NUMERICAL, TERM_NON_FINITE);
      hc[column + 1] = next_norm;
      for (size_t row = 0; row < column; ++row) {
        const double upper = hc[row], lower = hc[row + 1];
        const double rotated = cs[row] * upper + sn[row] * lower;
        hc[row + 1] = -sn[row] * upper + cs[row] * lower;
        hc[row] = rotated;
      }
      const double rn = hypot(hc[column], hc[column + 1]);
      if (!(rn > 0.0) || !isfinite(rn)) FG_FAIL(ST_NUMERICAL, TERM_BREAKDOWN);
      cs[column] = hc[column] / rn;
      sn[column] = hc[column + 1] / rn;
      hc[column] = rn;
      hc[column + 1] = 0.0;
      g[column + 1] = -sn[column] * g[column];
      g[column] = cs[column] * g[column];
      const double recursive = fabs(g[column + 1]);
      res->iterations++;
      const int happy = next_norm == 0.0;
      const int cycle_end = happy || column + 1 == R ||
                            res->iterations == ctl->maximum_iterations;
      const int verify = happy || recursive <= tol ||
                         res->iterations % ctl->true_residual_interval == 0 ||
                         cycle_end;
      if (!cycle_end) {
        if (!explicit_pass) add_scaled_basis(ax, v, scales, dots, N);
        const double inv_norm = 1.0 / next_norm;
        for (size_t e = 0; e < N; ++e) {
          v[column + 1][e] = inv_norm * ax[e];
          if (!isfinite(v[column + 1][e])) FG_FAIL(ST_NUMERICAL, TERM_NON_FINITE);
        }
      }
      if (ops->trace != NULL && res->iterations <= ops->trace_capacity &&
          restarts == 0.0 && res->iterations == column + 1) {
        double* tr = ops->trace + (size_t)(res->iterations - 1) * N;
        double* yt = (double*)calloc(dots, sizeof(double));
        if (fgmres_back_substitute(h, rows, g, dots, yt) == 0) {
          memcpy(tr, x, N * sizeof(double));
          add_scaled_basis(tr, z, yt, dots, N);
        }
        free(yt);
      }
      if (!verify) continue;
      const int term = fgmres_back_substitute(h, rows, g, dots, y);
      if (term) FG_FAIL(ST_NUMERICAL, term);
      if (!cycle_end) {
        /* audit a trial solution in the unused last basis slot */
        double* trial = v[R];
        memcpy(trial, x, N * sizeof(double));
        add_scaled_basis(trial, z, y, dots, N);
        double audited = 0.0;
        FG_TRUE_RESIDUAL(trial, ax, audited);
        res->final_true_residual = audited;
        if (audited <= tol) {
          memcpy(solution, trial, N * sizeof(double));
          res->termination = TERM_CONVERGED;
          goto done;
        }
        continue;
      }
      add_scaled_basis(x, z, y, dots, N);
      FG_TRUE_RESIDUAL(x, v[0], res->final_true_residual);
      if (res->final_true_residual <= tol) {
        memcpy(solution, x, N * sizeof(double));
        res->termination = TERM_CONVERGED;
        goto done;
      }
      if (res->iterations == ctl->maximum_iterations)
        FG_FAIL(ST_REJECTED, TERM_MAXIMUM_ITERATIONS);
      if (happy) {
        const double guard = 64.0 * eps * fmax(beta, tol);
        if (!(res->final_true_residual + guard < beta))
          FG_FAIL(ST_NUMERICAL, TERM_BREAKDOWN);
        restarts += 1.0;
      }
      restarted = 1;
      break;
    }
    if (!restarted) break;
  }
  FG_FAIL(ST_REJECTED, TERM_MAXIMUM_ITERATIONS);
done:
  res->status = st;
  res->breakdown_value = restarts;
  free(x);
  free(ax);
  for (size_t i = 0; i <= R; ++i) free(v[i]);
  for (size_t i = 0; i < R; ++i) free(z[i]);
  free(v);
  free(z);
  free(h);
  free(cs);
  free(sn);
  free(g);
  free(y);
  free(scales);
  return st;
#undef FG_FAIL
#undef FG_TRUE_RESIDUAL
}

/* Operator rows per unit volume and the volume-scaled preconditioner
 * (VolumeScaledPreconditioner, src/solver_cold.hpp:378-415): the rows are
 * the integrated operator divided by the cell volume (w_x w_y) w_z. */
typedef struct volume_ctx {
  ibm_operator_ctx a;
  or_mg* mg;
  const double *wx, *wy, *wz;
  double* scratch;
} volume_ctx;

static double cell_volume(const volume_ctx* v, size_t c) {
  const int* n = v->a.n;
  const size_t i = c % (size_t)n[0];
  const size_t j = (c / (size_t)n[0]) % (size_t)n[1];
  const size_t k = c / ((size_t)n[0] * (size_t)n[1]);
  return v->wx[i] * v->wy[j] * v->wz[k];
}

static void volume_operator_cb(void* ctx, const double* x, double* y) {
  const volume_ctx* v = (const volume_ctx*)ctx;
  ibm_operator_cb((void*)&v->a, x, y);
  const size_t N = (size_t)v->a.n[0] * v->a.n[1] * v->a.n[2];
  for (size_t c = 0; c < N; ++c) y[c] = y[c] / cell_volume(v, c);
}

static int volume_precondition_cb(void* ctx, const double* r, double* z) {
  volume_ctx* v = (volume_ctx*)ctx;
  const size_t N = (size_t)v->a.n[0] * v->a.n[1] * v->a.n[2];
  for (size_t c = 0; c < N; ++c) v->scratch[c] = r[c] * cell_volume(v, c);
  if (v->mg == NULL) {
    memcpy(z, v->scratch, N * sizeof(double));
    return ST_OK;
  }
  return or_mg_apply(v->mg, v->scratch, z);
}

int or_fgmres_solve(const int n[3], const double* wx, const double* wy,
                    const double* wz, const int bc[6], const double* diag,
                    const double* bx, const double* by, const double* bz,
                    const uint8_t* act, or_mg* mg, const double* rhs,
                    double* solution, const or_fgmres_control* ctl,
                    or_result* res) {
  const size_t N = (size_t)n[0] * n[1] * n[2];
  volume_ctx v;
  memset(&v, 0, sizeof v);
  v.a.n = n;
  v.a.bc = bc;
  v.a.diag = diag;
  v.a.bx = bx;
  v.a.by = by;
  v.a.bz = bz;
  v.a.act = act;
  v.mg = mg;
  v.wx = wx;
  v.wy = wy;
  v.wz = wz;
  or_krylov_ops ops;
  memset(&ops, 0, sizeof ops);
  ops.n = N;
  if (ctl->volume_rows) {
    v.scratch = (double*)malloc(N * sizeof(double));
    ops.apply_a = volume_operator_cb;
    ops.a_ctx = &v;
    ops.apply_m = volume_precondition_cb;
    ops.m_ctx = &v;
  } else {
    ops.apply_a = ibm_operator_cb;
    ops.a_ctx = &v.a;
    if (mg != NULL) {
      ops.apply_m = mg_precondition_cb;
      ops.m_ctx = mg;
    }
  }
  const int st = fgmres_core(&ops, rhs, solution, ctl, res);
  free(v.scratch);
  return st;
}

int or_fgmres_dense(int n, const double* a, const double* m,
                    const double* rhs, double* solution,
                    const or_fgmres_control* ctl, or_result* res,
                    double* trace, uint32_t trace_capacity) {
  dense_ctx ac = {(size_t)n, a};
  dense_ctx mc = {(size_t)n, m};
  or_krylov_ops ops;
  memset(&ops, 0, sizeof ops);
  ops.n = (size_t)n;
  ops.apply_a = dense_apply_cb;
  ops.a_ctx = &ac;
  if (m != NULL) {
    ops.apply_m = dense_precondition_cb;
    ops.m_ctx = &mc;
  }
  ops.trace = trace;
  ops.trace_capacity = trace_capacity;
  return fgmres_core(&ops, rhs, solution, ctl, res);
}

/* ------------------------------------------------------------------ */
/* Pressure-correction system, src/solver_pressure.cpp:110-332          */
/* (assemble_pressure_system_impl) with face_sum / outward_flux of      */
/* :80-94: D = a0 V drho/dp + sum of the six face coefficients,         */
/* b = -(V (a0 rho + a1 rho_n + a2 rho_nm1 - S) + outward flux).        */
/* ------------------------------------------------------------------ */
int or_assemble_pressure_system(const int n[3], const double* wx,
                                const double* wy, const double* wz,
                                const double* bx, const double* by,
                                const double* bz, const double* phix,
                                const double* phiy, const double* phiz,
                                const double* rho, const double* rho_n,
                                const double* rho_nm1, const double* drho_dp,
                                const double* source, double a0, double a1,
                                double a2, double* diag, double* rhs) {
  const int bdf2 = rho_nm1 != NULL;
  if (!(a0 > 0.0) || !isfinite(a0) || !isfinite(a1) || !isfinite(a2))
    return ST_INVALID_PLAN;
  const size_t nfx = (size_t)(n[0] + 1) * n[1] * n[2];
  const size_t nfy = (size_t)n[0] * (n[1] + 1) * n[2];
  const size_t nfz = (size_t)n[0] * n[1] * (n[2] + 1);
  const size_t N = (size_t)n[0] * n[1] * n[2];
  /* input finiteness (finite_field_box / finite_nonnegative_faces) */
  for (size_t c = 0; c < N; ++c)
    if (!isfinite(rho[c]) || !isfinite(rho_n[c]) ||
        (bdf2 && !isfinite(rho_nm1[c])) || !isfinite(drho_dp[c]) ||
        (source != NULL && !isfinite(source[c])))
      return ST_NUMERICAL;
  for (size_t f = 0; f < nfx; ++f)
    if (!isfinite(bx[f]) || bx[f] < 0.0 || !isfinite(phix[f])) return ST_NUMERICAL;
  for (size_t f = 0; f < nfy; ++f)
    if (!isfinite(by[f]) || by[f] < 0.0 || !isfinite(phiy[f])) return ST_NUMERICAL;
  for (size_t f = 0; f < nfz; ++f)
    if (!isfinite(bz[f]) || bz[f] < 0.0 || !isfinite(phiz[f])) return ST_NUMERICAL;
  /* two passes: the preflight publishes nothing on failure */
  for (int pass = 0; pass < 2; ++pass)
    for (int k = 0; k < n[2]; ++k)
      for (int j = 0; j < n[1]; ++j)
        for (int i = 0; i < n[0]; ++i) {
          const size_t c = cidx(n, i, j, k);
          const double volume = wx[i] * wy[j] * wz[k];
          const double r = rho[c], rn = rho_n[c];
          const double rnm1 = bdf2 ? rho_nm1[c] : 0.0;
          const double s = source != NULL ? source[c] : 0.0;
          const double defect = volume * (a0 * r + a1 * rn + a2 * rnm1 - s);
          const double faces =
              bx[xfidx(n, i, j, k)] + bx[xfidx(n, i + 1, j, k)] +
              by[yfidx(n, i, j, k)] + by[yfidx(n, i, j + 1, k)] +
              bz[zfidx(n, i, j, k)] + bz[zfidx(n, i, j, k + 1)];
          const double d = a0 * volume * drho_dp[c] + faces;
          const double flux = phix[xfidx(n, i + 1, j, k)] - phix[xfidx(n, i, j, k)] +
                              phiy[yfidx(n, i, j + 1, k)] - phiy[yfidx(n, i, j, k)] +
                              phiz[zfidx(n, i, j, k + 1)] - phiz[zfidx(n, i, j, k)];
          const double b = -(defect + flux);
          if (pass == 0) {
            if (!(r > 0.0) || !(rn > 0.0) || (bdf2 && !(rnm1 > 0.0)) ||
                !(drho_dp[c] > 0.0) || !(volume > 0.0) || !(d > 0.0) ||
                !isfinite(defect) || !isfinite(d) || !isfinite(b))
              return ST_NUMERICAL;
          } else {
            diag[c] = d;
            rhs[c] = b;
          }
        }
  return ST_OK;
}

/* ------------------------------------------------------------------ */
/* Cold-path coefficient loop, src/core_product_freeze.cpp:14279-14317, */
/* with the face records of src/solver_cold.hpp:990-1110 given as face   */
/* arrays: projection and limiter_diffusion per face (a fluid cell's     */
/* record of a face holds the face's values; an inactive cell's records  */
/* are zero), solid_neighbour = the cell across the face is not fluid    */
/* (periodic wrap across a periodic side).  The loop runs over cells in  */
/* x-fastest order and, as the reference, writes each of the six face    */
/* coefficients of every cell, so a shared face keeps the value of the   */
/* later cell.  kinds: 0 other, 1 periodic, 2 pressure outlet.           */
/* ------------------------------------------------------------------ */
static int or_wrap_fluid(const int n[3], const uint8_t* fluid, const int kinds[6],
                         int i, int j, int k) {
  const int c[3] = {i, j, k};
  int w[3] = {i, j, k};
  for (int a = 0; a < 3; ++a) {
    if (c[a] >= 0 && c[a] < n[a]) continue;
    if (kinds[2 * a] != 1) return 0; /* outside a non-periodic side */
    w[a] = (c[a] + n[a]) % n[a];
  }
  return fluid == NULL || fluid[cidx(n, w[0], w[1], w[2])] != 0;
}

int or_cold_coefficients(const int n[3], const double* wx, const double* wy,
                         const double* wz, const double* proj_x,
                         const double* proj_y, const double* proj_z,
                         const double* lim_x, const double* lim_y,
                         const double* lim_z, const uint8_t* fluid,
                         const double* rho, const double* pressure,
                         double pressure_reference, double dt, int fixed_eos,
                         const int kinds[6], double* diag, double* cx,
                         double* cy, double* cz) {
  for (int a = 0; a < 3; ++a)
    if ((kinds[2 * a] == 1) != (kinds[2 * a + 1] == 1)) return ST_INVALID_PLAN;
  if (!(dt > 0.0) || !isfinite(dt) || !isfinite(pressure_reference))
    return ST_INVALID_PLAN;
  const double* proj[3] = {proj_x, proj_y, proj_z};
  const double* lim[3] = {lim_x, lim_y, lim_z};
  double* out[3] = {cx, cy, cz};
  for (int k = 0; k < n[2]; ++k)
    for (int j = 0; j < n[1]; ++j)
      for (int i = 0; i < n[0]; ++i) {
        const size_t c = cidx(n, i, j, k);
        const int cell_fluid = fluid == NULL || fluid[c] != 0;
        const double volume = wx[i] * wy[j] * wz[k];
        double d = cell_fluid
                       ? (fixed_eos ? 0.0
                                    : volume * rho[c] /
                                          ((pressure_reference + pressure[c]) * dt))
                       : 1.0;
        for (int f = 0; f < 6; ++f) {
          const int a = f / 2;
          int at[3] = {i, j, k};
          if (f % 2) at[a]++;
          const size_t fi = a == 0   ? xfidx(n, at[0], at[1], at[2])
                            : a == 1 ? yfidx(n, at[0], at[1], at[2])
                                     : zfidx(n, at[0], at[1], at[2]);
          const int coordinate = at[a];
          const int exterior =
              (coordinate == 0 || coordinate == n[a]) && kinds[f] != 1;
          int nb[3] = {i, j, k};
          nb[a] += f % 2 ? 1 : -1;
          /* an inactive cell's records are zero: projection 0, limiter 0,
           * no solid neighbour */
          const double p = cell_fluid ? proj[a][fi] : 0.0;
          const double l = cell_fluid ? lim[a][fi] : 0.0;
          const int solid =
              cell_fluid && !or_wrap_fluid(n, fluid, kinds, nb[0], nb[1], nb[2]) &&
              !exterior;
          const double coefficient = exterior || solid ? 0.0 : p + 0.5 * l;
          if (cell_fluid) {
            d += coefficient;
            if (exterior && kinds[f] == 2) d += 2.0 * p;
          }
          out[a][fi] = coefficient;
        }
        diag[c] = d;
      }
  return ST_OK;
}
