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
                               double* coarse, double scale) {
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
  const double v = aggregate_entry(a, axis, ci, cj, ck,
                                   [&](int i, int j, int k) {
                                     return fine[fidx(fe, i, j, k)];
                                   });
  coarse[t] = scale == 1.0 ? v : v * scale;
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
  S_LINE_FAIL = 54,  // nonzero after a line relaxation pivot vanished
  S_NSCALE = 55,     // scaled norm: largest magnitude
  S_NSUM = 56,       // scaled norm: sum of scaled squares
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

// First fold of many block partials: block (s, a) sums span s of
// accumulator a in index order.
__global__ void k_fold_partials(const double* partials, int64_t blocks,
                                int64_t stride, double* out,
                                int64_t out_stride) {
  __shared__ double scratch[16];
  const int64_t a = blockIdx.y;
  const int64_t lo = static_cast<int64_t>(blockIdx.x) * 2048;
  const int64_t hi = lo + 2048 < blocks ? lo + 2048 : blocks;
  double v = 0.0;
  for (int64_t b = lo + threadIdx.x; b < hi; b += blockDim.x)
    v += partials[a * stride + b];
  v = block_sum(v, scratch);
  if (threadIdx.x == 0) out[a * out_stride + blockIdx.x] = v;
}

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

}  // namespace hf

#include "hf_cycle.hpp"

namespace hf {

// Scaled 2-norm (global_norm, src/solver_krylov.cpp:650-711): the largest
// magnitude, then the compensated sum of squares of the elements divided by
// it.  The solvers take it where a plain sum of squares overflowed, so a
// norm overflows only where the reference's does.  A non-finite element
// makes its block's maximum +inf.
__device__ __forceinline__ double block_max(double v, double* scratch) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
#pragma unroll
  for (int offset = kWave / 2; offset > 0; offset >>= 1)
    v = fmax(v, __shfl_down(v, offset, kWave));
  if (lane == 0) scratch[wave] = v;
  __syncthreads();
  double m = 0.0;
  if (threadIdx.x == 0) {
    const int waves = (blockDim.x + kWave - 1) / kWave;
    for (int w = 0; w < waves; ++w) m = fmax(m, scratch[w]);
  }
  __syncthreads();
  return m;
}

__global__ void k_absmax(Chunk ch, const double* v, double* partials) {
  __shared__ double scratch[16];
  const int lb = xcd_block32(blockIdx.x, gridDim.x);
  const int lo = lb * ch.chunk;
  const int hi = min(ch.n, lo + ch.chunk);
  double m = 0.0;
  for (int c = lo + threadIdx.x; c < hi; c += blockDim.x) {
    const double a = v[c];
    m = isfinite(a) ? fmax(m, fabs(a)) : INFINITY;
  }
  m = block_max(m, scratch);
  if (threadIdx.x == 0) partials[lb] = m;
}

__global__ void __launch_bounds__(1024)
k_max_fold(const double* partials, int64_t count, double* scalars, int slot) {
  __shared__ double scratch[16];
  double m = 0.0;
  for (int64_t b = threadIdx.x; b < count; b += blockDim.x)
    m = fmax(m, partials[b]);
  m = block_max(m, scratch);
  if (threadIdx.x == 0) scalars[slot] = m;
}

// Per block: the Neumaier-compensated sum of (v / scale)^2 of each thread's
// elements, then the block's sum of those.
__global__ void k_scaled_squares(Chunk ch, const double* v,
                                 const double* scale, double* partials) {
  __shared__ double scratch[16];
  const int lb = xcd_block32(blockIdx.x, gridDim.x);
  const int lo = lb * ch.chunk;
  const int hi = min(ch.n, lo + ch.chunk);
  const double s = *scale;
  double sum = 0.0, correction = 0.0;
  for (int c = lo + threadIdx.x; c < hi; c += blockDim.x) {
    const double q = v[c] / s;
    const double square = q * q;
    const double next = sum + square;
    correction += fabs(sum) >= fabs(square) ? (sum - next) + square
                                            : (square - next) + sum;
    sum = next;
  }
  sum = block_sum(sum + correction, scratch);
  if (threadIdx.x == 0) partials[lb] = sum;
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
