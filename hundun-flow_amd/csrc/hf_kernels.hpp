// gfx950 kernels of the pressure solver.  Every kernel is HBM-stream bound
// (the 7-point operator is ~0.27 flop/byte in FP64), so the design goals are
// coalesced x-fastest streams, one pass per array where the algorithm allows
// it, and an XCD-aware block order so neighbour rows/planes hit the same L2.
#pragma once

#include "hf_device.hpp"

namespace hf {

// ---------------------------------------------------------------------------
// Setup: face coefficient rule (src/solver_piso.cpp:1060-1118), mobility
// m = 1/rho, beta = A_f / (d_L/m_L + d_R/m_R); Neumann physical faces are 0;
// periodic physical faces use half widths of the wrapped cells.
// ---------------------------------------------------------------------------
struct AxisGeom {
  const double* faces[3];  // n+1 coordinates per axis
  int n[3];
  int8_t kind[6];          // hf_boundary_kind per physical face
};

__device__ __forceinline__ double face_rule(const double* f, int count,
                                            int normal, int kind, double area,
                                            double ml, double mr) {
  if ((normal == 0 || normal == count) && kind == 1) return 0.0;
  if ((normal == 0 || normal == count) && kind == 0) {
    const double owner = normal == 0 ? mr : ml;
    double distance;
    if (normal == 0) {
      const double w0 = f[1] - f[0];
      distance = (f[0] + 0.5 * w0) - f[0];
    } else {
      const double wl = f[count] - f[count - 1];
      distance = f[count] - (f[count - 1] + 0.5 * wl);
    }
    return 2.0 * area / (distance / owner + distance / owner);
  }
  double dl, dr;
  if (normal == 0 || normal == count) {
    dl = 0.5 * (f[count] - f[count - 1]);
    dr = 0.5 * (f[1] - f[0]);
  } else {
    const double loc = f[normal];
    const double wl = f[normal] - f[normal - 1];
    const double wr = f[normal + 1] - f[normal];
    dl = loc - (f[normal - 1] + 0.5 * wl);
    dr = (f[normal] + 0.5 * wr) - loc;
  }
  return area / (dl / ml + dr / mr);
}

// d beta / d m_L and d beta / d m_R for the same rule (adjoint).
__device__ __forceinline__ void face_rule_grad(const double* f, int count,
                                               int normal, int kind,
                                               double area, double ml,
                                               double mr, double* gl,
                                               double* gr) {
  *gl = 0.0;
  *gr = 0.0;
  if ((normal == 0 || normal == count) && kind == 1) return;
  if ((normal == 0 || normal == count) && kind == 0) {
    // beta = area*m_o/d: linear in the owner mobility
    const double owner = normal == 0 ? mr : ml;
    const double beta = face_rule(f, count, normal, kind, area, ml, mr);
    if (normal == 0)
      *gr = beta / owner;
    else
      *gl = beta / owner;
    return;
  }
  double dl, dr;
  if (normal == 0 || normal == count) {
    dl = 0.5 * (f[count] - f[count - 1]);
    dr = 0.5 * (f[1] - f[0]);
  } else {
    const double loc = f[normal];
    const double wl = f[normal] - f[normal - 1];
    const double wr = f[normal + 1] - f[normal];
    dl = loc - (f[normal - 1] + 0.5 * wl);
    dr = (f[normal] + 0.5 * wr) - loc;
  }
  const double s = dl / ml + dr / mr;
  const double beta = area / s;
  // d beta/d m = beta^2/area * d/m^2
  *gl = beta * beta / area * (dl / (ml * ml));
  *gr = beta * beta / area * (dr / (mr * mr));
}

__global__ void k_assemble_faces(AxisGeom g, int axis, const double* m,
                                 double* out, int64_t count) {
  const int64_t t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= count) return;
  int e[3] = {g.n[0], g.n[1], g.n[2]};
  e[axis] += 1;
  const int i = static_cast<int>(t % e[0]);
  const int j = static_cast<int>((t / e[0]) % e[1]);
  const int k = static_cast<int>(t / ((int64_t)e[0] * e[1]));
  int c[3] = {i, j, k};
  const int nrm = c[axis];
  const int a1 = axis == 0 ? 1 : 0;
  const int a2 = axis == 2 ? 1 : 2;
  const double area = (g.faces[a1][c[a1] + 1] - g.faces[a1][c[a1]]) *
                      (g.faces[a2][c[a2] + 1] - g.faces[a2][c[a2]]);
  int cl[3] = {i, j, k};
  int cr[3] = {i, j, k};
  cl[axis] = nrm == 0 ? g.n[axis] - 1 : nrm - 1;
  cr[axis] = nrm == g.n[axis] ? 0 : nrm;
  const int64_t nx = g.n[0], ny = g.n[1];
  const double ml = m[cl[0] + nx * (cl[1] + ny * cl[2])];
  const double mr = m[cr[0] + nx * (cr[1] + ny * cr[2])];
  const int kind = nrm == 0 ? g.kind[2 * axis] : g.kind[2 * axis + 1];
  out[t] = face_rule(g.faces[axis], g.n[axis], nrm, kind, area, ml, mr);
}

__global__ void k_face_sum_diag(int nx, int ny, int nz, const double* bx,
                                const double* by, const double* bz,
                                double* diag) {
  const int64_t cells = (int64_t)nx * ny * nz;
  const int64_t c = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (c >= cells) return;
  const int j = static_cast<int>((c / nx) % ny);
  const int k = static_cast<int>(c / ((int64_t)nx * ny));
  const int64_t row = j + (int64_t)ny * k;
  const int64_t fyo = c + (int64_t)nx * k;
  diag[c] = bx[c + row] + bx[c + row + 1] + by[fyo] + by[fyo + nx] + bz[c] +
            bz[c + (int64_t)nx * ny];
}

// Level-0 masking (src/solver_mg.cpp:1469-1570): faces times activity,
// solid diagonal 1, fluid diagonal D - original faces + active faces, fluid
// volume w_x*w_y*w_z (0 on solids).
struct Finest {
  int nx, ny, nz;
  const double* D;
  const double* bx;
  const double* by;
  const double* bz;
  const uint8_t* ca;
  const uint8_t* fxa;
  const uint8_t* fya;
  const uint8_t* fza;
  const double* wx;
  const double* wy;
  const double* wz;
  double* oD;
  double* ofx;
  double* ofy;
  double* ofz;
  double* vol;
  uint8_t* act;
};

__global__ void k_mask_faces(const double* raw, const uint8_t* active,
                             double* out, int64_t count) {
  const int64_t t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= count) return;
  out[t] = raw[t] * (active == nullptr ? 1.0 : static_cast<double>(active[t]));
}

__global__ void k_mask_cells(Finest f) {
  const int64_t nx = f.nx, ny = f.ny;
  const int64_t cells = nx * ny * f.nz;
  const int64_t c = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (c >= cells) return;
  const int i = static_cast<int>(c % nx);
  const int j = static_cast<int>((c / nx) % ny);
  const int k = static_cast<int>(c / (nx * ny));
  double v = f.wx[i] * f.wy[j] * f.wz[k];
  const bool solid = f.ca != nullptr && f.ca[c] == 0;
  if (solid) v = 0.0;
  f.vol[c] = v;
  f.act[c] = solid ? 0 : 1;
  double d = f.D[c];
  if (f.ca != nullptr) {
    if (solid) {
      d = 1.0;
    } else {
      const int64_t row = j + ny * k;
      const int64_t fyo = c + nx * k;
      const int64_t zhi = c + nx * ny;
      const double original = f.bx[c + row] + f.bx[c + row + 1] + f.by[fyo] +
                              f.by[fyo + nx] + f.bz[c] + f.bz[zhi];
      const double active = f.ofx[c + row] + f.ofx[c + row + 1] + f.ofy[fyo] +
                            f.ofy[fyo + nx] + f.ofz[c] + f.ofz[zhi];
      d = d - original + active;
    }
  }
  f.oD[c] = d;
}

// Coarse aggregation (src/solver_mg.cpp:1706-1962) evaluated per coarse
// entry in the reference's association order: the x pass sums (or, for a
// normal face, selects) first, then y, then z.
struct Agg {
  int fn[3];  // fine shape
  int cn[3];  // coarse shape
  int mask;   // coarsened axes
};

__device__ __forceinline__ int64_t fidx(const int e[3], int i, int j, int k) {
  return i + (int64_t)e[0] * (j + (int64_t)e[1] * k);
}

// Value of one fine channel entry (cell channel: face_axis < 0).
template <class Loader>
__device__ __forceinline__ double aggregate_entry(const Agg& a, int face_axis,
                                                  int ci, int cj, int ck,
                                                  Loader&& load) {
  // Per axis: list of up to 2 fine indices (selection or pair).
  int idx[3][2];
  int cnt[3];
  const int co[3] = {ci, cj, ck};
  for (int ax = 0; ax < 3; ++ax) {
    if ((a.mask & (1 << ax)) == 0) {
      idx[ax][0] = co[ax];
      cnt[ax] = 1;
    } else if (face_axis == ax) {
      const int g = 2 * co[ax];
      idx[ax][0] = g < a.fn[ax] ? g : a.fn[ax];
      cnt[ax] = 1;
    } else {
      idx[ax][0] = 2 * co[ax];
      idx[ax][1] = 2 * co[ax] + 1;
      cnt[ax] = idx[ax][1] < a.fn[ax] ? 2 : 1;
    }
  }
  // x pass innermost, z pass outermost: z(y(x)).
  double vz = 0.0;
  for (int dz = 0; dz < cnt[2]; ++dz) {
    double vy = 0.0;
    for (int dy = 0; dy < cnt[1]; ++dy) {
      double vx = 0.0;
      for (int dx = 0; dx < cnt[0]; ++dx)
        vx = (dx == 0 && cnt[0] == 1 && ((a.mask & 1) == 0 || face_axis == 0))
                 ? load(idx[0][dx], idx[1][dy], idx[2][dz])
                 : vx + load(idx[0][dx], idx[1][dy], idx[2][dz]);
      vy = (cnt[1] == 1 && ((a.mask & 2) == 0 || face_axis == 1)) ? vx : vy + vx;
    }
    vz = (cnt[2] == 1 && ((a.mask & 4) == 0 || face_axis == 2)) ? vy : vz + vy;
  }
  return vz;
}

struct Fine64 {
  const double* D;
  const double* fx;
  const double* fy;
  const double* fz;
  const double* vol;
};
struct Coarse64 {
  double* D;
  double* fx;
  double* fy;
  double* fz;
  double* vol;
  uint8_t* act;
};

__global__ void k_coarse_cells(Agg a, Fine64 f, Coarse64 c) {
  const int64_t cells = (int64_t)a.cn[0] * a.cn[1] * a.cn[2];
  const int64_t t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= cells) return;
  const int ci = static_cast<int>(t % a.cn[0]);
  const int cj = static_cast<int>((t / a.cn[0]) % a.cn[1]);
  const int ck = static_cast<int>(t / ((int64_t)a.cn[0] * a.cn[1]));
  const int* fn = a.fn;
  const int fxe[3] = {fn[0] + 1, fn[1], fn[2]};
  const int fye[3] = {fn[0], fn[1] + 1, fn[2]};
  const int fze[3] = {fn[0], fn[1], fn[2] + 1};
  c.vol[t] = aggregate_entry(a, -1, ci, cj, ck, [&](int i, int j, int k) {
    return f.vol[fidx(fn, i, j, k)];
  });
  c.D[t] = aggregate_entry(a, -1, ci, cj, ck, [&](int i, int j, int k) {
    const int64_t cc = fidx(fn, i, j, k);
    if (f.vol[cc] == 0.0) return 0.0;
    const double faces = f.fx[fidx(fxe, i, j, k)] + f.fx[fidx(fxe, i + 1, j, k)] +
                         f.fy[fidx(fye, i, j, k)] + f.fy[fidx(fye, i, j + 1, k)] +
                         f.fz[fidx(fze, i, j, k)] + f.fz[fidx(fze, i, j, k + 1)];
    const double reaction = f.D[cc] - faces;
    return reaction > 0.0 ? reaction : 0.0;
  });
}

__global__ void k_coarse_faces(Agg a, int axis, const double* fine,
                               double* coarse) {
  int ce[3] = {a.cn[0], a.cn[1], a.cn[2]};
  ce[axis] += 1;
  int fe[3] = {a.fn[0], a.fn[1], a.fn[2]};
  fe[axis] += 1;
  const int64_t count = (int64_t)ce[0] * ce[1] * ce[2];
  const int64_t t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const int ci = static_cast<int>(t % ce[0]);
  const int cj = static_cast<int>((t / ce[0]) % ce[1]);
  const int ck = static_cast<int>(t / ((int64_t)ce[0] * ce[1]));
  coarse[t] = aggregate_entry(a, axis, ci, cj, ck, [&](int i, int j, int k) {
    return fine[fidx(fe, i, j, k)];
  });
}

__global__ void k_coarse_diag(int nx, int ny, int nz, Coarse64 c,
                              int* bad) {
  const int64_t cells = (int64_t)nx * ny * nz;
  const int64_t t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= cells) return;
  const int j = static_cast<int>((t / nx) % ny);
  const int k = static_cast<int>(t / ((int64_t)nx * ny));
  const int64_t row = j + (int64_t)ny * k;
  const int64_t fyo = t + (int64_t)nx * k;
  double d = c.D[t];
  d += c.fx[t + row] + c.fx[t + row + 1] + c.fy[fyo] + c.fy[fyo + nx] +
       c.fz[t] + c.fz[t + (int64_t)nx * ny];
  const double v = c.vol[t];
  if (v == 0.0) d = 1.0;
  c.D[t] = d;
  c.act[t] = v > 0.0 ? 1 : 0;
  if (v < 0.0 || !isfinite(v) || !(d > 0.0) || !isfinite(d)) atomicOr(bad, 1);
}

// Certified SDD M-matrix check of one level (src/solver_mg.cpp:1965-2019).
__global__ void k_validate(int nx, int ny, int nz, const double* D,
                           const double* fx, const double* fy,
                           const double* fz, const double* vol, int* bad) {
  const int64_t cells = (int64_t)nx * ny * nz;
  const int64_t t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= cells) return;
  const int j = static_cast<int>((t / nx) % ny);
  const int k = static_cast<int>(t / ((int64_t)nx * ny));
  const int64_t row = j + (int64_t)ny * k;
  const int64_t fyo = t + (int64_t)nx * k;
  const double faces = fx[t + row] + fx[t + row + 1] + fy[fyo] + fy[fyo + nx] +
                       fz[t] + fz[t + (int64_t)nx * ny];
  const double d = D[t];
  bool fail = !isfinite(d) || !(d > 0.0) || !isfinite(faces) || faces < 0.0;
  if (vol[t] == 0.0)
    fail = fail || d != 1.0 || faces != 0.0;
  else
    fail = fail || faces > d;
  if (fail) atomicOr(bad, 2);
}

template <class T>
__global__ void k_convert(const double* src, T* dst, int64_t count) {
  const int64_t t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t < count) dst[t] = st<T>(src[t]);
}

// ---------------------------------------------------------------------------
// Reductions: per-block partials (deterministic shuffle tree) and a single
// block finaliser that also performs the scalar Krylov algebra on device so
// the host only reads back what it must test.
// ---------------------------------------------------------------------------
enum ScalarSlot : int {
  S_RHO = 0,
  S_PAP = 1,
  S_ALPHA = 2,
  S_RNORM2 = 3,
  S_BETA = 4,
  S_RHONEXT = 5,
  S_ZQ = 6,
  S_MEAN0 = 7,  // 7..7+kMaxLevels: projection means per level
  S_PROJ0 = 48, // three dots of the residual-minimising scaling
  S_PROJ1 = 49,
  S_PROJ2 = 50,
  S_MGALPHA = 51,
  S_SUM = 52,
  S_SUM2 = 53,
  S_COUNT = 64
};

enum FinalOp : int {
  F_STORE = 0,     // scalars[a] = sum0
  F_ALPHA = 1,     // pap = sum0, alpha = rho / pap
  F_RHO = 2,       // rho = sum0
  F_BETA_FR = 3,   // beta = sum0/rho, rho = sum0
  F_BETA_PR = 4,   // rho_next = sum0, zq = sum1, beta = -alpha*zq/rho
  F_MEAN = 5,      // scalars[a] = sum0 / volume
  F_MGALPHA = 6,   // three sums -> residual-minimising alpha
  F_STORE2 = 7     // scalars[a] = sum0, scalars[a+1] = sum1
};

__global__ void __launch_bounds__(1024)
k_finalize(const double* partials, int64_t blocks, int64_t stride, int nacc,
           double* scalars, int op, int slot, double volume, int unit_linear) {
  __shared__ double scratch[16];
  double sums[3] = {0.0, 0.0, 0.0};
  for (int a = 0; a < nacc; ++a) {
    double v = 0.0;
    for (int64_t b = threadIdx.x; b < blocks; b += blockDim.x)
      v += partials[a * stride + b];
    sums[a] = block_sum(v, scratch);
  }
  if (threadIdx.x != 0) return;
  switch (op) {
    case F_STORE:
      scalars[slot] = sums[0];
      break;
    case F_STORE2:
      scalars[slot] = sums[0];
      scalars[slot + 1] = sums[1];
      break;
    case F_ALPHA:
      scalars[S_PAP] = sums[0];
      scalars[S_ALPHA] = scalars[S_RHO] / sums[0];
      break;
    case F_RHO:
      scalars[S_RHO] = sums[0];
      break;
    case F_BETA_FR:
      scalars[S_RHONEXT] = sums[0];
      scalars[S_BETA] = sums[0] / scalars[S_RHO];
      scalars[S_RHO] = sums[0];
      break;
    case F_BETA_PR:
      scalars[S_RHONEXT] = sums[0];
      scalars[S_ZQ] = sums[1];
      scalars[S_BETA] = -scalars[S_ALPHA] * sums[1] / scalars[S_RHO];
      scalars[S_RHO] = sums[0];
      break;
    case F_MEAN:
      scalars[slot] = sums[0] / volume;
      break;
    case F_MGALPHA: {
      scalars[S_PROJ0] = sums[0];
      scalars[S_PROJ1] = sums[1];
      scalars[S_PROJ2] = sums[2];
      const double tiny = 2.220446049250313080847e-16 * sums[0];
      scalars[S_MGALPHA] =
          unit_linear ? 1.0 : (sums[2] > tiny ? sums[1] / sums[2] : 0.0);
      break;
    }
    default:
      break;
  }
}

template <class T>
__global__ void k_vol_dot(const T* field, const double* vol, int64_t cells,
                          double* partials) {
  __shared__ double scratch[16];
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t c = blk * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (c < cells) v = vol[c] * static_cast<double>(ld(field[c]));
  v = block_sum(v, scratch);
  if (threadIdx.x == 0) partials[blk] = v;
}

__global__ void k_sum(const double* field, int64_t cells, double* partials,
                      const uint8_t* act) {
  __shared__ double scratch[16];
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t c = blk * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (c < cells && (act == nullptr || act[c])) v = field[c];
  v = block_sum(v, scratch);
  if (threadIdx.x == 0) partials[blk] = v;
}

__global__ void k_count(const uint8_t* act, int64_t cells, double* partials) {
  __shared__ double scratch[16];
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t c = blk * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (c < cells) v = act == nullptr ? 1.0 : static_cast<double>(act[c] != 0);
  v = block_sum(v, scratch);
  if (threadIdx.x == 0) partials[blk] = v;
}

__global__ void k_subtract_mean(double* field, int64_t cells,
                                const double* mean, const uint8_t* act) {
  const int64_t c = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (c >= cells) return;
  field[c] = (act == nullptr || act[c]) ? field[c] - *mean : 0.0;
}

// f := vol > 0 ? f - mean : 0 (src/solver_mg.cpp:1445-1452)
template <class T>
__global__ void k_project_apply(T* field, const double* vol, int64_t cells,
                                const double* mean) {
  const int64_t c = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (c >= cells) return;
  field[c] = vol[c] > 0.0
                 ? st<T>(static_cast<double>(ld(field[c])) - *mean)
                 : st<T>(0.0);
}

// ---------------------------------------------------------------------------
// Multigrid grid-level kernels.
// ---------------------------------------------------------------------------

// apply_impl entry (src/solver_mg.cpp:4867-4879): x := 0, b := r on fluid.
template <class T>
__global__ void k_mg_entry(const double* r, const uint8_t* act, T* b, T* x,
                           int64_t cells) {
  const int64_t c = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (c >= cells) return;
  x[c] = st<T>(0.0);
  b[c] = act[c] ? st<T>(r[c]) : st<T>(0.0);
}

// One colour pass of the fused red-black update (src/solver_mg.cpp:
// 3149-3222): x += 0.72 (b - A x)/D on cells with (i+j+k)&1 == color.  One
// thread per cell of the colour; same-colour cells never neighbour, so the
// in-place update is order independent.
template <class T>
__global__ void k_rb(Lv<T> L, Bc bc, int color, int64_t work) {
  using C = typename Compute<T>::type;
  const int64_t t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= work) return;
  const int hx = (L.nx + 1) >> 1;
  const int ti = static_cast<int>(t % hx);
  const int64_t rest = t / hx;
  const int j = static_cast<int>(rest % L.ny);
  const int k = static_cast<int>(rest / L.ny);
  const int i = 2 * ti + ((color + j + k) & 1);
  if (i >= L.nx) return;
  const int64_t c = i + (int64_t)L.nx * (j + (int64_t)L.ny * k);
  const C xc = ld(L.x[c]);
  const C v = stencil<C>(L, bc, L.x, i, j, k, c, xc);
  const C omega = C(0.72);
  L.x[c] = st<T>(xc + omega * (C(ld(L.b[c])) - v) / C(ld(L.D[c])));
}

// Separated route for levels with an odd periodic extent
// (src/solver_mg.cpp:3262-3307): the colour update uses the A x computed
// before the pass.
template <class T>
__global__ void k_rb_separated(Lv<T> L, int color, int64_t work) {
  using C = typename Compute<T>::type;
  const int64_t t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= work) return;
  const int hx = (L.nx + 1) >> 1;
  const int ti = static_cast<int>(t % hx);
  const int64_t rest = t / hx;
  const int j = static_cast<int>(rest % L.ny);
  const int k = static_cast<int>(rest / L.ny);
  const int i = 2 * ti + ((color + j + k) & 1);
  if (i >= L.nx) return;
  const int64_t c = i + (int64_t)L.nx * (j + (int64_t)L.ny * k);
  const C omega = C(0.72);
  L.x[c] = st<T>(C(ld(L.x[c])) +
                 omega * (C(ld(L.b[c])) - C(ld(L.r[c]))) / C(ld(L.D[c])));
}

// r := A x (defect = 0) or r := b - A x (defect = 1).
template <class T>
__global__ void k_level_apply(Lv<T> L, Bc bc, int defect) {
  using C = typename Compute<T>::type;
  const int64_t c = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (c >= L.cells) return;
  const int i = static_cast<int>(c % L.nx);
  const int j = static_cast<int>((c / L.nx) % L.ny);
  const int k = static_cast<int>(c / ((int64_t)L.nx * L.ny));
  const C v = stencil<C>(L, bc, L.x, i, j, k, c, C(ld(L.x[c])));
  L.r[c] = defect ? st<T>(C(ld(L.b[c])) - v) : st<T>(v);
}

// Restriction (src/solver_mg.cpp:3633-3720): coarse b = sum of children in
// dz, dy, dx order, coarse x := 0.  Optional partial sums of vol*b for the
// constant projection.
template <class Tf, class Tc>
__global__ void k_restrict(Lv<Tf> F, Lv<Tc> Cl, int mask, double* partials) {
  using C = typename Compute<Tc>::type;
  __shared__ double scratch[16];
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t t = blk * blockDim.x + threadIdx.x;
  double pv = 0.0;
  if (t < Cl.cells) {
    const int ci = static_cast<int>(t % Cl.nx);
    const int cj = static_cast<int>((t / Cl.nx) % Cl.ny);
    const int ck = static_cast<int>(t / ((int64_t)Cl.nx * Cl.ny));
    const int mx = mask & 1, my = (mask >> 1) & 1, mz = (mask >> 2) & 1;
    const int xb = mx ? 2 * ci : ci, yb = my ? 2 * cj : cj, zb = mz ? 2 * ck : ck;
    C sum = C(0);
    for (int dz = 0; dz < (mz ? 2 : 1); ++dz)
      for (int dy = 0; dy < (my ? 2 : 1); ++dy)
        for (int dx = 0; dx < (mx ? 2 : 1); ++dx) {
          const int fx = xb + dx, fy = yb + dy, fz = zb + dz;
          if (fx >= F.nx || fy >= F.ny || fz >= F.nz) continue;
          sum += C(ld(F.r[fx + (int64_t)F.nx * (fy + (int64_t)F.ny * fz)]));
        }
    Cl.b[t] = st<Tc>(sum);
    Cl.x[t] = st<Tc>(0.0);
    if (partials != nullptr) pv = Cl.vol[t] * static_cast<double>(ld(Cl.b[t]));
  }
  if (partials != nullptr) {
    pv = block_sum(pv, scratch);
    if (threadIdx.x == 0) partials[blk] = pv;
  }
}

__device__ __forceinline__ void interp(int fine, int coarsened, int cext,
                                       int periodic, int* l, int* r,
                                       int* rw4) {
  // rw4: right weight in quarters (0, 1 or 3)
  if (!coarsened) {
    *l = fine;
    *r = fine;
    *rw4 = 0;
    return;
  }
  const int parent = fine / 2;
  if ((fine & 1) == 0) {
    *l = parent - 1;
    *r = parent;
    *rw4 = 3;
  } else {
    *l = parent;
    *r = parent + 1;
    *rw4 = 1;
  }
  if (!periodic && *l < 0) {
    *l = 0;
    *r = 0;
    *rw4 = 0;
  } else if (!periodic && *r >= cext) {
    *l = cext - 1;
    *r = cext - 1;
    *rw4 = 0;
  }
}

__device__ __forceinline__ int wrapi(int v, int extent) {
  return v < 0 ? v + extent : (v >= extent ? v - extent : v);
}

// Prolongation (src/solver_mg.cpp:3893-3993): x_f += sum of ((wx*wy)*wz)*x_c
// over the (az, ay, ax) corners; fine solids := 0.
template <class Tc, class Tf>
__device__ __forceinline__ void prolong_cell(const Lv<Tc>& Cl, const Lv<Tf>& F,
                                             int mask, int per, int64_t c) {
  using C = typename Compute<Tf>::type;
  if (!F.act[c]) {
    F.x[c] = st<Tf>(0.0);
    return;
  }
  const int i = static_cast<int>(c % F.nx);
  const int j = static_cast<int>((c / F.nx) % F.ny);
  const int k = static_cast<int>(c / ((int64_t)F.nx * F.ny));
  int xl, xr, yl, yr, zl, zr, xw, yw, zw;
  interp(i, mask & 1, Cl.nx, per & 1, &xl, &xr, &xw);
  interp(j, (mask >> 1) & 1, Cl.ny, (per >> 1) & 1, &yl, &yr, &yw);
  interp(k, (mask >> 2) & 1, Cl.nz, (per >> 2) & 1, &zl, &zr, &zw);
  C value = C(0);
  for (int az = 0; az < (zl == zr ? 1 : 2); ++az) {
    const int gz = wrapi(az == 0 ? zl : zr, Cl.nz);
    const C wz = az == 0 ? C(1) - C(zw) * C(0.25) : C(zw) * C(0.25);
    for (int ay = 0; ay < (yl == yr ? 1 : 2); ++ay) {
      const int gy = wrapi(ay == 0 ? yl : yr, Cl.ny);
      const C wy = ay == 0 ? C(1) - C(yw) * C(0.25) : C(yw) * C(0.25);
      for (int ax = 0; ax < (xl == xr ? 1 : 2); ++ax) {
        const int gx = wrapi(ax == 0 ? xl : xr, Cl.nx);
        const C wx = ax == 0 ? C(1) - C(xw) * C(0.25) : C(xw) * C(0.25);
        value += wx * wy * wz *
                 C(ld(Cl.x[gx + (int64_t)Cl.nx * (gy + (int64_t)Cl.ny * gz)]));
      }
    }
  }
  F.x[c] = st<Tf>(C(ld(F.x[c])) + value);
}

template <class Tc, class Tf>
__global__ void k_prolong(Lv<Tc> Cl, Lv<Tf> F, int mask, int per) {
  const int64_t c = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (c >= F.cells) return;
  prolong_cell(Cl, F, mask, per, c);
}

// apply_impl exit (src/solver_mg.cpp:5082-5118): correction = alpha * x on
// fluid, 0 on solids; fused with the Krylov inner products r.z (and z.q for
// the flexible beta).
template <class T>
__global__ void k_mg_exit(const T* x, const uint8_t* act, const double* alpha,
                          double* z, const double* r, const double* q,
                          int64_t cells, double* partials, int64_t stride) {
  __shared__ double scratch[16];
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t c = blk * blockDim.x + threadIdx.x;
  double rz = 0.0, zq = 0.0;
  if (c < cells) {
    const double a = alpha == nullptr ? 1.0 : *alpha;
    const double v = act[c] ? static_cast<double>(ld(x[c])) * a : 0.0;
    z[c] = v;
    if (r != nullptr) rz = r[c] * v;
    if (q != nullptr) zq = v * q[c];
  }
  if (partials != nullptr) {
    rz = block_sum(rz, scratch);
    if (threadIdx.x == 0) partials[blk] = rz;
    if (q != nullptr) {
      zq = block_sum(zq, scratch);
      if (threadIdx.x == 0) partials[stride + blk] = zq;
    }
  }
}

// Three dots of the residual-minimising scaling (src/solver_mg.cpp:
// 4988-5001): r = masked input, az = r - (b - A x).
template <class T>
__global__ void k_mg_proj3(const double* rin, const uint8_t* act, const T* rec,
                           int64_t cells, double* partials, int64_t stride) {
  __shared__ double scratch[16];
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t c = blk * blockDim.x + threadIdx.x;
  double p0 = 0.0, p1 = 0.0, p2 = 0.0;
  if (c < cells) {
    const double r = act[c] ? rin[c] : 0.0;
    const double az = r - static_cast<double>(ld(rec[c]));
    p0 = r * r;
    p1 = r * az;
    p2 = az * az;
  }
  p0 = block_sum(p0, scratch);
  p1 = block_sum(p1, scratch);
  p2 = block_sum(p2, scratch);
  if (threadIdx.x == 0) {
    partials[blk] = p0;
    partials[stride + blk] = p1;
    partials[2 * stride + blk] = p2;
  }
}

// ---------------------------------------------------------------------------
// Single-workgroup tail: every level below the tail threshold (FP64) runs as
// one kernel that interprets the host-built cycle program, removing ~10
// launches per coarse level.
// ---------------------------------------------------------------------------
enum TailKind : uint8_t { T_SMOOTH = 0, T_RESTRICT = 1, T_PROLONG = 2 };
struct TailOp {
  uint8_t kind;
  uint8_t level;
  uint8_t reverse;
  uint8_t retain;
  uint16_t sweeps;
  uint16_t pad;
};

struct TailLevel {
  Lv<double> L;
  int coarsen_mask;
  int fused;
  double volume;
};

__device__ __forceinline__ void tail_apply(const Lv<double>& L, const Bc& bc,
                                           int defect) {
  for (int64_t c = threadIdx.x; c < L.cells; c += blockDim.x) {
    const int i = static_cast<int>(c % L.nx);
    const int j = static_cast<int>((c / L.nx) % L.ny);
    const int k = static_cast<int>(c / ((int64_t)L.nx * L.ny));
    const double v = stencil<double>(L, bc, L.x, i, j, k, c, L.x[c]);
    L.r[c] = defect ? L.b[c] - v : v;
  }
  __syncthreads();
}

__device__ void tail_smooth(const TailLevel& T, const Bc& bc, int sweeps,
                            int reverse, int retain) {
  const Lv<double>& L = T.L;
  const double omega = 0.72;
  if (!T.fused) tail_apply(L, bc, 0);
  for (int sweep = 0; sweep < sweeps; ++sweep) {
    const int first = reverse ? 1 : 0;
    for (int pass = 0; pass < 2; ++pass) {
      const int color = (first + pass) & 1;
      for (int64_t c = threadIdx.x; c < L.cells; c += blockDim.x) {
        const int i = static_cast<int>(c % L.nx);
        const int j = static_cast<int>((c / L.nx) % L.ny);
        const int k = static_cast<int>(c / ((int64_t)L.nx * L.ny));
        if (((i + j + k) & 1) != color) continue;
        const double xc = L.x[c];
        const double v =
            T.fused ? stencil<double>(L, bc, L.x, i, j, k, c, xc) : L.r[c];
        L.x[c] = xc + omega * (L.b[c] - v) / L.D[c];
      }
      __syncthreads();
      const bool final_update = sweep + 1 == sweeps && pass == 1;
      if (!final_update || retain) {
        if (!T.fused || (final_update && retain))
          tail_apply(L, bc, retain && final_update);
      }
    }
  }
}

__device__ void tail_project(const Lv<double>& L, double volume, double* field,
                             double* scratch) {
  double v = 0.0;
  for (int64_t c = threadIdx.x; c < L.cells; c += blockDim.x)
    v += L.vol[c] * field[c];
  __shared__ double mean;
  v = block_sum(v, scratch);
  if (threadIdx.x == 0) mean = v / volume;
  __syncthreads();
  const double m = mean;
  for (int64_t c = threadIdx.x; c < L.cells; c += blockDim.x)
    field[c] = L.vol[c] > 0.0 ? field[c] - m : 0.0;
  __syncthreads();
}

__global__ void __launch_bounds__(1024)
k_tail(const TailLevel* levels, int first_level, const TailOp* ops, int nops,
       Bc bc, int periodic_mask, int null_space) {
  __shared__ double scratch[16];
  for (int o = 0; o < nops; ++o) {
    const TailOp op = ops[o];
    const TailLevel& T = levels[op.level - first_level];
    if (op.kind == T_SMOOTH) {
      tail_smooth(T, bc, op.sweeps, op.reverse, op.retain);
    } else if (op.kind == T_RESTRICT) {
      const TailLevel& Cn = levels[op.level + 1 - first_level];
      const Lv<double>& F = T.L;
      const Lv<double>& Cl = Cn.L;
      const int mask = T.coarsen_mask;
      const int mx = mask & 1, my = (mask >> 1) & 1, mz = (mask >> 2) & 1;
      for (int64_t t = threadIdx.x; t < Cl.cells; t += blockDim.x) {
        const int ci = static_cast<int>(t % Cl.nx);
        const int cj = static_cast<int>((t / Cl.nx) % Cl.ny);
        const int ck = static_cast<int>(t / ((int64_t)Cl.nx * Cl.ny));
        const int xb = mx ? 2 * ci : ci, yb = my ? 2 * cj : cj,
                  zb = mz ? 2 * ck : ck;
        double sum = 0.0;
        for (int dz = 0; dz < (mz ? 2 : 1); ++dz)
          for (int dy = 0; dy < (my ? 2 : 1); ++dy)
            for (int dx = 0; dx < (mx ? 2 : 1); ++dx) {
              const int fx = xb + dx, fy = yb + dy, fz = zb + dz;
              if (fx >= F.nx || fy >= F.ny || fz >= F.nz) continue;
              sum += F.r[fx + (int64_t)F.nx * (fy + (int64_t)F.ny * fz)];
            }
        Cl.b[t] = sum;
        Cl.x[t] = 0.0;
      }
      __syncthreads();
      if (null_space) tail_project(Cl, Cn.volume, Cl.b, scratch);
    } else {
      const TailLevel& Cn = levels[op.level + 1 - first_level];
      for (int64_t c = threadIdx.x; c < T.L.cells; c += blockDim.x)
        prolong_cell(Cn.L, T.L, T.coarsen_mask, periodic_mask, c);
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------------------
// Krylov kernels (FP64).  The operator arrays are the masked level-0
// coefficients, which equal the exact IBM operator.
// ---------------------------------------------------------------------------

// q = A p fused with the partial sum of p.q.
__global__ void k_op_dot(Lv<double> L, Bc bc, const double* p, double* q,
                         double* partials) {
  __shared__ double scratch[16];
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t c = blk * blockDim.x + threadIdx.x;
  double d = 0.0;
  if (c < L.cells) {
    const int i = static_cast<int>(c % L.nx);
    const int j = static_cast<int>((c / L.nx) % L.ny);
    const int k = static_cast<int>(c / ((int64_t)L.nx * L.ny));
    const double pc = p[c];
    const double v = stencil<double>(L, bc, p, i, j, k, c, pc);
    q[c] = v;
    d = pc * v;
  }
  if (partials != nullptr) {
    d = block_sum(d, scratch);
    if (threadIdx.x == 0) partials[blk] = d;
  }
}

// r = b - A x fused with the partial sum of r.r.
__global__ void k_true_residual(Lv<double> L, Bc bc, const double* b,
                                const double* x, double* r, double* partials) {
  __shared__ double scratch[16];
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t c = blk * blockDim.x + threadIdx.x;
  double d = 0.0;
  if (c < L.cells) {
    const int i = static_cast<int>(c % L.nx);
    const int j = static_cast<int>((c / L.nx) % L.ny);
    const int k = static_cast<int>(c / ((int64_t)L.nx * L.ny));
    const double v = stencil<double>(L, bc, x, i, j, k, c, x[c]);
    const double rc = 1.0 * b[c] + -1.0 * v;
    r[c] = rc;
    d = rc * rc;
  }
  d = block_sum(d, scratch);
  if (threadIdx.x == 0) partials[blk] = d;
}

// x += alpha p; r += (-alpha) q; partial sum of r.r.
__global__ void k_update_xr(double* x, const double* p, double* r,
                            const double* q, const double* alpha,
                            int64_t cells, double* partials) {
  __shared__ double scratch[16];
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t c = blk * blockDim.x + threadIdx.x;
  double d = 0.0;
  if (c < cells) {
    const double a = *alpha;
    x[c] = x[c] + a * p[c];
    const double rc = r[c] + -a * q[c];
    r[c] = rc;
    d = rc * rc;
  }
  d = block_sum(d, scratch);
  if (threadIdx.x == 0) partials[blk] = d;
}

// p = 1.0*z + beta*p
__global__ void k_update_p(double* p, const double* z, const double* beta,
                           int64_t cells) {
  const int64_t c = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (c >= cells) return;
  p[c] = 1.0 * z[c] + *beta * p[c];
}

__global__ void k_copy(double* dst, const double* src, int64_t cells) {
  const int64_t c = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (c < cells) dst[c] = src[c];
}

__global__ void k_sumsq(const double* v, int64_t cells, double* partials) {
  __shared__ double scratch[16];
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t c = blk * blockDim.x + threadIdx.x;
  double d = 0.0;
  if (c < cells) d = v[c] * v[c];
  d = block_sum(d, scratch);
  if (threadIdx.x == 0) partials[blk] = d;
}

__global__ void k_dot(const double* a, const double* b, int64_t cells,
                      double* partials) {
  __shared__ double scratch[16];
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t c = blk * blockDim.x + threadIdx.x;
  double d = 0.0;
  if (c < cells) d = a[c] * b[c];
  d = block_sum(d, scratch);
  if (threadIdx.x == 0) partials[blk] = d;
}

// Adjoint: grad_logrho(c) = -sum over the six faces f of c of
//   (lam_L - lam_R)(p_L - p_R) * d beta_f / d m_c * d m_c / d log rho_c
// with m = 1/rho, d m / d log rho = -m.  Masked faces carry no gradient.
__global__ void k_grad_logrho(AxisGeom g, const double* m, const double* p,
                              const double* lam, const uint8_t* fxa,
                              const uint8_t* fya, const uint8_t* fza,
                              double* out) {
  const int64_t nx = g.n[0], ny = g.n[1], nz = g.n[2];
  const int64_t cells = nx * ny * nz;
  const int64_t c = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (c >= cells) return;
  const int ijk[3] = {static_cast<int>(c % nx), static_cast<int>((c / nx) % ny),
                      static_cast<int>(c / (nx * ny))};
  const uint8_t* fa[3] = {fxa, fya, fza};
  double grad = 0.0;
  for (int axis = 0; axis < 3; ++axis) {
    const int a1 = axis == 0 ? 1 : 0;
    const int a2 = axis == 2 ? 1 : 2;
    const double area =
        (g.faces[a1][ijk[a1] + 1] - g.faces[a1][ijk[a1]]) *
        (g.faces[a2][ijk[a2] + 1] - g.faces[a2][ijk[a2]]);
    int e[3] = {g.n[0], g.n[1], g.n[2]};
    e[axis] += 1;
    for (int side = 0; side < 2; ++side) {
      const int normal = ijk[axis] + side;  // face index along axis
      int fc[3] = {ijk[0], ijk[1], ijk[2]};
      fc[axis] = normal;
      const int64_t fi = fc[0] + (int64_t)e[0] * (fc[1] + (int64_t)e[1] * fc[2]);
      if (fa[axis] != nullptr && fa[axis][fi] == 0) continue;
      int cl[3] = {fc[0], fc[1], fc[2]};
      int cr[3] = {fc[0], fc[1], fc[2]};
      cl[axis] = normal == 0 ? g.n[axis] - 1 : normal - 1;
      cr[axis] = normal == g.n[axis] ? 0 : normal;
      const int kind = normal == 0 ? g.kind[2 * axis]
                                   : (normal == g.n[axis] ? g.kind[2 * axis + 1]
                                                          : 2);
      const bool physical = normal == 0 || normal == g.n[axis];
      if (physical && kind == 1) continue;
      const int64_t il = cl[0] + nx * (cl[1] + ny * cl[2]);
      const int64_t ir = cr[0] + nx * (cr[1] + ny * cr[2]);
      double gl, gr;
      face_rule_grad(g.faces[axis], g.n[axis], normal, kind, area, m[il], m[ir],
                     &gl, &gr);
      double jump;
      if (physical && kind == 0) {
        const int64_t owner = normal == 0 ? ir : il;
        jump = lam[owner] * p[owner];  // ghost value 0
      } else {
        jump = (lam[il] - lam[ir]) * (p[il] - p[ir]);
      }
      // this cell is the left cell of its high face and the right cell of
      // its low face (a periodic self-face counts both)
      double dbeta = 0.0;
      if (il == c) dbeta += gl;
      if (ir == c) dbeta += gr;
      grad -= jump * dbeta * (-m[c]);
    }
  }
  out[c] = grad;
}

}  // namespace hf
