// Kernels of the reference's production pressure pairing on gfx950:
// Chebyshev-Jacobi smoothing, the diagonal shortcut of the preconditioner,
// and the single-reduction FGMRES (Arnoldi step, basis combinations and the
// volume-scaled rows of the cold pressure system).
#pragma once

#include "hf_kernels.hpp"

namespace hf {

// Geometric cell volume w_x w_y w_z, multiplied left to right as the
// reference's cell_volume (src/solver_cold.hpp:378-415 scales by it).
__device__ __forceinline__ double geom_volume(const Widths& w, int i, int j,
                                              int k) {
  return w.x[i] * w.y[j] * w.z[k];
}

// ---------------------------------------------------------------------------
// Chebyshev-Jacobi smoothing (src/solver_mg.cpp:2916-3090).  One launch is
// one stage: every neighbour comes from the stage's input x (Jacobi), the
// direction lives in the level's r slot (read from L.r, stored to dout; a
// smoothing's last stage passes dout == nullptr, as no stage reads its
// direction), and
//   d = rf (b - A x) / D            (first stage)
//   d = pf d + rf (b - A x) / D     (later stages)
//   x_next = x + d.
// z-marching as k_stencil_z; x and x_next are distinct buffers.
// ---------------------------------------------------------------------------
template <class T>
__global__ void k_cheb_stage(Tiles t, Lv<T> L, Bc bc, const T* x, T* xn,
                             typename Compute<T>::type rf,
                             typename Compute<T>::type pf, int first,
                             T* dout) {
  using C = typename Compute<T>::type;
  const int lb = xcd_block32(blockIdx.x, gridDim.x);
  const int gxy = t.gx * t.gy;
  const int zb = lb / gxy;
  int i, j, k_unused;
  tile_xyz(t, lb - zb * gxy, i, j, k_unused);
  const int nx = t.nx, ny = t.ny, nz = t.nz;
  if (i >= nx || j >= ny) return;
  const int k0 = zb * t.per;
  const int k1 = min(nz, k0 + t.per);
  const int plane = nx * ny;
  int c = cell_index(nx, ny, i, j, k0);
  C xc = C(ld(x[c]));
  C zl = k0 > 0 ? C(ld(x[c - plane]))
                : neighbour_or_boundary<C>(bc.m[4], xc, x, c + (nz - 1) * plane);
  C fzl = C(ld(L.fz[c]));
  for (int k = k0; k < k1; ++k) {
    const C zh = k + 1 < nz ? C(ld(x[c + plane]))
                            : neighbour_or_boundary<C>(bc.m[5], xc, x,
                                                       c - (nz - 1) * plane);
    const C fzh = C(ld(L.fz[c + plane]));
    const int row = j + ny * k;
    const C xl = i > 0 ? C(ld(x[c - 1]))
                       : neighbour_or_boundary<C>(bc.m[0], xc, x, c + (nx - 1));
    const C xh = i + 1 < nx ? C(ld(x[c + 1]))
                            : neighbour_or_boundary<C>(bc.m[1], xc, x,
                                                       c - (nx - 1));
    const C yl = j > 0 ? C(ld(x[c - nx]))
                       : neighbour_or_boundary<C>(bc.m[2], xc, x,
                                                  c + (ny - 1) * nx);
    const C yh = j + 1 < ny ? C(ld(x[c + nx]))
                            : neighbour_or_boundary<C>(bc.m[3], xc, x,
                                                       c - (ny - 1) * nx);
    const int fyo = c + nx * k;
    const C d = C(ld(L.D[c]));
    C v = d * xc;
    v -= C(ld(L.fx[c + row])) * xl;
    v -= C(ld(L.fx[c + row + 1])) * xh;
    v -= C(ld(L.fy[fyo])) * yl;
    v -= C(ld(L.fy[fyo + nx])) * yh;
    v -= fzl * zl;
    v -= fzh * zh;
    const C defect = C(ld(L.b[c])) - v;
    const C scaled = rf * defect / d;
    const C dir = first ? scaled : pf * C(ld(L.r[c])) + scaled;
    if (dout != nullptr) dout[c] = st<T>(dir);
    xn[c] = st<T>(xc + dir);
    zl = xc;
    xc = zh;
    fzl = fzh;
    c += plane;
  }
}

// k_cheb_stage over pairs of cells (2m, 2m + 1) of a row, for an even x
// extent: x, D, b, r, the y and z faces and both stores move two cells per
// access (8-byte words on FP32 levels); the x faces of a pair are three
// loads.  Each cell's operands and order of operations are those of
// k_cheb_stage, the pair's inner x neighbours being each other's x.  t tiles
// (nx / 2) x ny columns in z-chunks.
template <class T>
__global__ void k_cheb_stage2(Tiles t, Lv<T> L, Bc bc, const T* x, T* xn,
                              typename Compute<T>::type rf,
                              typename Compute<T>::type pf, int first,
                              T* dout) {
  using C = typename Compute<T>::type;
  struct alignas(2 * sizeof(T)) Pair {
    T v[2];
  };
  const int lb = xcd_block32(blockIdx.x, gridDim.x);
  const int gxy = t.gx * t.gy;
  const int zb = lb / gxy;
  int ip, j, k_unused;
  tile_xyz(t, lb - zb * gxy, ip, j, k_unused);
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  if (ip >= t.nx || j >= ny) return;
  const int i0 = 2 * ip;
  const int k0 = zb * t.per;
  const int k1 = min(nz, k0 + t.per);
  const int plane = nx * ny;
  auto pair = [](const T* a, int c, C* out) {
    const Pair q = *reinterpret_cast<const Pair*>(a + c);
    out[0] = C(ld(q.v[0]));
    out[1] = C(ld(q.v[1]));
  };
  int c = cell_index(nx, ny, i0, j, k0);
  C xc[2], zl[2], fzl[2];
  pair(x, c, xc);
  if (k0 > 0) {
    pair(x, c - plane, zl);
  } else {
#pragma unroll
    for (int e = 0; e < 2; ++e)
      zl[e] = neighbour_or_boundary<C>(bc.m[4], xc[e], x,
                                       c + e + (nz - 1) * plane);
  }
  pair(L.fz, c, fzl);
  for (int k = k0; k < k1; ++k) {
    C zh[2], fzh[2], yl[2], yh[2], fyl[2], fyh[2], d[2], b[2], r[2];
    if (k + 1 < nz) {
      pair(x, c + plane, zh);
    } else {
#pragma unroll
      for (int e = 0; e < 2; ++e)
        zh[e] = neighbour_or_boundary<C>(bc.m[5], xc[e], x,
                                         c + e - (nz - 1) * plane);
    }
    pair(L.fz, c + plane, fzh);
    if (j > 0) {
      pair(x, c - nx, yl);
    } else {
#pragma unroll
      for (int e = 0; e < 2; ++e)
        yl[e] = neighbour_or_boundary<C>(bc.m[2], xc[e], x,
                                         c + e + (ny - 1) * nx);
    }
    if (j + 1 < ny) {
      pair(x, c + nx, yh);
    } else {
#pragma unroll
      for (int e = 0; e < 2; ++e)
        yh[e] = neighbour_or_boundary<C>(bc.m[3], xc[e], x,
                                         c + e - (ny - 1) * nx);
    }
    const int row = j + ny * k;
    const int fyo = c + nx * k;
    pair(L.fy, fyo, fyl);
    pair(L.fy, fyo + nx, fyh);
    pair(L.D, c, d);
    pair(L.b, c, b);
    if (!first) pair(L.r, c, r);
    const C fx0 = C(ld(L.fx[c + row]));
    const C fx1 = C(ld(L.fx[c + row + 1]));
    const C fx2 = C(ld(L.fx[c + row + 2]));
    const C xl0 = i0 > 0 ? C(ld(x[c - 1]))
                         : neighbour_or_boundary<C>(bc.m[0], xc[0], x,
                                                    c + (nx - 1));
    const C xh1 = i0 + 2 < nx ? C(ld(x[c + 2]))
                              : neighbour_or_boundary<C>(bc.m[1], xc[1], x,
                                                         c + 1 - (nx - 1));
    const C xl[2] = {xl0, xc[0]};
    const C xh[2] = {xc[1], xh1};
    const C fxl[2] = {fx0, fx1};
    const C fxh[2] = {fx1, fx2};
    Pair ro, xo;
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      C v = d[e] * xc[e];
      v -= fxl[e] * xl[e];
      v -= fxh[e] * xh[e];
      v -= fyl[e] * yl[e];
      v -= fyh[e] * yh[e];
      v -= fzl[e] * zl[e];
      v -= fzh[e] * zh[e];
      const C defect = b[e] - v;
      const C scaled = rf * defect / d[e];
      const C dir = first ? scaled : pf * r[e] + scaled;
      ro.v[e] = st<T>(dir);
      xo.v[e] = st<T>(xc[e] + dir);
    }
    if (dout != nullptr) *reinterpret_cast<Pair*>(dout + c) = ro;
    *reinterpret_cast<Pair*>(xn + c) = xo;
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      zl[e] = xc[e];
      xc[e] = zh[e];
      fzl[e] = fzh[e];
    }
    c += plane;
  }
}

// First Chebyshev-Jacobi stage of a level's first visit in a cycle, where x
// holds zero: A x = +0 on every row (non-negative faces, positive D), so the
// stage's d = rf (b - A x) / D is rf b / D and x_next = 0 + d, the values
// k_cheb_stage stores there, from b and D alone.  Four consecutive cells per
// thread.  dir == nullptr: the direction is not stored (a one-stage
// smoothing, whose direction no later stage reads).
template <class T>
__global__ void k_cheb_first(int cells, Lv<T> L, T* xn,
                             typename Compute<T>::type rf, T* dir) {
  using C = typename Compute<T>::type;
  struct alignas(4 * sizeof(T)) Quad {
    T v[4];
  };
  const int64_t c0 =
      4 * (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x);
  if (c0 >= cells) return;
  if (c0 + 4 <= cells) {
    const Quad b = *reinterpret_cast<const Quad*>(L.b + c0);
    const Quad d = *reinterpret_cast<const Quad*>(L.D + c0);
    Quad r, x;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const C q = rf * C(ld(b.v[e])) / C(ld(d.v[e]));
      r.v[e] = st<T>(q);
      x.v[e] = st<T>(C(0) + q);
    }
    if (dir != nullptr) *reinterpret_cast<Quad*>(dir + c0) = r;
    *reinterpret_cast<Quad*>(xn + c0) = x;
    return;
  }
  for (int64_t c = c0; c < cells; ++c) {
    const C d = rf * C(ld(L.b[c])) / C(ld(L.D[c]));
    if (dir != nullptr) dir[c] = st<T>(d);
    xn[c] = st<T>(C(0) + d);
  }
}

// ---------------------------------------------------------------------------
// Diagonal shortcut (src/solver_mg.cpp:1573-1611 selects it, :4926-4936
// applies it): x = b / D on level 0.
// ---------------------------------------------------------------------------
template <class T>
__global__ void k_diag_solve(Chunk ch, Lv<T> L) {
  using C = typename Compute<T>::type;
  int lb, lo, hi;
  chunk_range(ch, lb, lo, hi);
  for (int c = lo + threadIdx.x; c < hi; c += kBlock)
    L.x[c] = st<T>(C(ld(L.b[c])) / C(ld(L.D[c])));
}

// Largest finest-row ratio (sum of the six masked faces) / D, physical
// boundary faces included, as the non-negative bit pattern max into *out
// (ratios are >= 0 on a validated hierarchy, so integer order is float
// order; the maximum does not depend on the order of the updates).  A
// grid-stride loop over a bounded grid, then one atomic per block: one
// atomic per wave put half a million updates on one address (6 ms at
// 512 x 256 x 256).
constexpr int kRatioBlocks = 2048;

__global__ void k_row_ratio_max(int nx, int ny, int nz, const double* D,
                                const double* fx, const double* fy,
                                const double* fz,
                                unsigned long long* out) {
  __shared__ double scratch[16];
  const int64_t cells = static_cast<int64_t>(nx) * ny * nz;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  double m = 0.0;
  for (int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       c < cells; c += stride) {
    const int i = static_cast<int>(c % nx);
    const int j = static_cast<int>((c / nx) % ny);
    const int k = static_cast<int>(c / (static_cast<int64_t>(nx) * ny));
    const int64_t xf = i + static_cast<int64_t>(nx + 1) * (j + static_cast<int64_t>(ny) * k);
    const int64_t yf = i + static_cast<int64_t>(nx) * (j + static_cast<int64_t>(ny + 1) * k);
    const int64_t zf = c;
    const double sum = fx[xf] + fx[xf + 1] + fy[yf] + fy[yf + nx] + fz[zf] +
                       fz[zf + static_cast<int64_t>(nx) * ny];
    double ratio = sum / D[c];
    if (!(ratio >= 0.0)) ratio = 0.0;  // max(0, NaN) keeps 0 (std::max)
    m = fmax(m, ratio);
  }
  m = block_max(m, scratch);
  if (threadIdx.x == 0)
    atomicMax(out, static_cast<unsigned long long>(__double_as_longlong(m)));
}

// ---------------------------------------------------------------------------
// Single-reduction FGMRES (src/solver_krylov.cpp:1557-2650).
// ---------------------------------------------------------------------------
constexpr int kFgDots = 16;       // accumulators of one Arnoldi pass
constexpr int kFgMaxRestart = 64;

enum FgMode : int { kFgArnoldi = 0, kFgResidual = 1, kFgOpDot = 2 };

// Level-0 FP64 action y = A' x, A' = A or A divided by the geometric cell
// volume (the per-volume rows of the cold system), z-marching as
// k_stencil_z, with the FGMRES epilogues:
//  kFgArnoldi:  out = A' x - shift * aux (aux = v_col); accumulator a < kNd
//               holds basis_a . out, accumulator wacc holds out . out;
//  kFgResidual: out = 1 b + (-1) A' x (b = aux); accumulator 0 = out . out
//               (kNd = 0).
// Partials of accumulator a sit at partials + a * pstride.  One instance
// per dot count, so the kNd basis loads of a step are unconditional and
// issue together: at 512 x 256 x 256 a basis vector adds 56 us to a launch,
// against 85 us when a runtime count guarded each load.
template <int kMode, int kNd>
__global__ void k_fg_stencil(Tiles t, Lv<double> A, Bc bc, const double* x,
                             const double* aux, double* out,
                             const double* basis, int64_t vstride,
                             int wacc, double shift, int volume_rows,
                             Widths w, double* partials, int64_t pstride) {
  static_assert(kNd >= 0 && kNd <= kFgDots, "dots of one pass");
  __shared__ double scratch[16];
  const int lb = xcd_block32(blockIdx.x, gridDim.x);
  const int gxy = t.gx * t.gy;
  const int zb = lb / gxy;
  int i, j, k_unused;
  tile_xyz(t, lb - zb * gxy, i, j, k_unused);
  const int nx = t.nx, ny = t.ny, nz = t.nz;
  const int k0 = zb * t.per;
  const int k1 = min(nz, k0 + t.per);
  double acc[kNd > 0 ? kNd : 1];
#pragma unroll
  for (int a = 0; a < kNd; ++a) acc[a] = 0.0;
  double ww = 0.0;
  if (i < nx && j < ny) {
    const int plane = nx * ny;
    int c = cell_index(nx, ny, i, j, k0);
    double xc = x[c];
    double zl = k0 > 0 ? x[c - plane]
                       : neighbour_or_boundary<double>(bc.m[4], xc, x,
                                                       c + (nz - 1) * plane);
    double fzl = A.fz[c];
    const double wxy = volume_rows ? w.x[i] * w.y[j] : 1.0;
    for (int k = k0; k < k1; ++k) {
      const double zh = k + 1 < nz
                            ? x[c + plane]
                            : neighbour_or_boundary<double>(
                                  bc.m[5], xc, x, c - (nz - 1) * plane);
      const double fzh = A.fz[c + plane];
      const int row = j + ny * k;
      const double xl = i > 0 ? x[c - 1]
                              : neighbour_or_boundary<double>(bc.m[0], xc, x,
                                                              c + (nx - 1));
      const double xh = i + 1 < nx
                            ? x[c + 1]
                            : neighbour_or_boundary<double>(bc.m[1], xc, x,
                                                            c - (nx - 1));
      const double yl = j > 0 ? x[c - nx]
                              : neighbour_or_boundary<double>(
                                    bc.m[2], xc, x, c + (ny - 1) * nx);
      const double yh = j + 1 < ny
                            ? x[c + nx]
                            : neighbour_or_boundary<double>(
                                  bc.m[3], xc, x, c - (ny - 1) * nx);
      const int fyo = c + nx * k;
      double v = A.D[c] * xc;
      v -= A.fx[c + row] * xl;
      v -= A.fx[c + row + 1] * xh;
      v -= A.fy[fyo] * yl;
      v -= A.fy[fyo + nx] * yh;
      v -= fzl * zl;
      v -= fzh * zh;
      if (volume_rows) v = v / (wxy * w.z[k]);
      if constexpr (kMode == kFgArnoldi) {
        const double o = v + -shift * aux[c];
        out[c] = o;
#pragma unroll
        for (int a = 0; a < kNd; ++a) acc[a] += basis[a * vstride + c] * o;
        ww += o * o;
      } else {
        const double o = 1.0 * aux[c] + -1.0 * v;
        out[c] = o;
        ww += o * o;
      }
      zl = xc;
      xc = zh;
      fzl = fzh;
      c += plane;
    }
  }
  if constexpr (kMode == kFgArnoldi) {
#pragma unroll
    for (int a = 0; a < kNd; ++a) {
      const double s = block_sum(acc[a], scratch);
      if (threadIdx.x == 0) partials[a * pstride + lb] = s;
    }
  }
  const double s = block_sum(ww, scratch);
  if (threadIdx.x == 0) partials[wacc * pstride + lb] = s;
}

// More dots of the Arnoldi column when it has more rows than one pass
// holds: accumulator acc0 + a = basis_(row0 + a) . wv, a < kNd, in the
// z-marching tile layout of k_fg_stencil (same partial count; one instance
// per dot count, as there).
template <int kNd>
__global__ void k_fg_dots(Tiles t, const double* wv, const double* basis,
                          int64_t vstride, int row0, int acc0,
                          double* partials, int64_t pstride) {
  static_assert(kNd >= 1 && kNd <= kFgDots, "dots of one pass");
  __shared__ double scratch[16];
  const int lb = xcd_block32(blockIdx.x, gridDim.x);
  const int gxy = t.gx * t.gy;
  const int zb = lb / gxy;
  int i, j, k_unused;
  tile_xyz(t, lb - zb * gxy, i, j, k_unused);
  const int nx = t.nx, ny = t.ny, nz = t.nz;
  const int k0 = zb * t.per;
  const int k1 = min(nz, k0 + t.per);
  double acc[kNd];
#pragma unroll
  for (int a = 0; a < kNd; ++a) acc[a] = 0.0;
  if (i < nx && j < ny) {
    const int plane = nx * ny;
    int c = cell_index(nx, ny, i, j, k0);
    for (int k = k0; k < k1; ++k) {
      const double o = wv[c];
#pragma unroll
      for (int a = 0; a < kNd; ++a)
        acc[a] += basis[(row0 + a) * vstride + c] * o;
      c += plane;
    }
  }
#pragma unroll
  for (int a = 0; a < kNd; ++a) {
    const double s = block_sum(acc[a], scratch);
    if (threadIdx.x == 0) partials[(acc0 + a) * pstride + lb] = s;
  }
}

// k_fg_stencil over pairs of cells (2m, 2m + 1) of a row, for an even x
// extent: x, aux, the basis rows, D, the y and z faces and out move two
// cells (16 bytes) per access; the x faces of a pair are three loads.  Each
// cell's operator value is formed as in k_fg_stencil; a thread adds its two
// cells' products to its accumulators in cell order.  t tiles (nx / 2) x ny
// columns in z-chunks; partials as in k_fg_stencil, one per block of t.
// kFgOpDot (the PCG operator, k_stencil_z<double, kZOpDot> over pairs):
// out = A' x and accumulator wacc holds x . out; aux is not read.
template <int kMode, int kNd>
__global__ void k_fg_stencil2(Tiles t, Lv<double> A, Bc bc, const double* x,
                              const double* aux, double* out,
                              const double* basis, int64_t vstride,
                              int wacc, double shift, int volume_rows,
                              Widths w, double* partials, int64_t pstride) {
  static_assert(kNd >= 0 && kNd <= kFgDots, "dots of one pass");
  struct alignas(16) Pair {
    double v[2];
  };
  auto pair = [](const double* a, int64_t c, double* o2) {
    const Pair q = *reinterpret_cast<const Pair*>(a + c);
    o2[0] = q.v[0];
    o2[1] = q.v[1];
  };
  __shared__ double scratch[16];
  const int lb = xcd_block32(blockIdx.x, gridDim.x);
  const int gxy = t.gx * t.gy;
  const int zb = lb / gxy;
  int ip, j, k_unused;
  tile_xyz(t, lb - zb * gxy, ip, j, k_unused);
  const int nx = A.nx, ny = A.ny, nz = A.nz;
  const int k0 = zb * t.per;
  const int k1 = min(nz, k0 + t.per);
  double acc[kNd > 0 ? kNd : 1];
#pragma unroll
  for (int a = 0; a < kNd; ++a) acc[a] = 0.0;
  double ww = 0.0;
  if (ip < t.nx && j < ny) {
    const int i0 = 2 * ip;
    const int plane = nx * ny;
    int c = cell_index(nx, ny, i0, j, k0);
    double xc[2], zl[2], fzl[2];
    pair(x, c, xc);
    if (k0 > 0) {
      pair(x, c - plane, zl);
    } else {
#pragma unroll
      for (int e = 0; e < 2; ++e)
        zl[e] = neighbour_or_boundary<double>(bc.m[4], xc[e], x,
                                              c + e + (nz - 1) * plane);
    }
    pair(A.fz, c, fzl);
    double wxy[2] = {1.0, 1.0};
    if (volume_rows) {
      wxy[0] = w.x[i0] * w.y[j];
      wxy[1] = w.x[i0 + 1] * w.y[j];
    }
    for (int k = k0; k < k1; ++k) {
      double zh[2], fzh[2], yl[2], yh[2], fyl[2], fyh[2], d[2], av[2];
      if (k + 1 < nz) {
        pair(x, c + plane, zh);
      } else {
#pragma unroll
        for (int e = 0; e < 2; ++e)
          zh[e] = neighbour_or_boundary<double>(bc.m[5], xc[e], x,
                                                c + e - (nz - 1) * plane);
      }
      pair(A.fz, c + plane, fzh);
      if (j > 0) {
        pair(x, c - nx, yl);
      } else {
#pragma unroll
        for (int e = 0; e < 2; ++e)
          yl[e] = neighbour_or_boundary<double>(bc.m[2], xc[e], x,
                                                c + e + (ny - 1) * nx);
      }
      if (j + 1 < ny) {
        pair(x, c + nx, yh);
      } else {
#pragma unroll
        for (int e = 0; e < 2; ++e)
          yh[e] = neighbour_or_boundary<double>(bc.m[3], xc[e], x,
                                                c + e - (ny - 1) * nx);
      }
      const int row = j + ny * k;
      const int fyo = c + nx * k;
      pair(A.fy, fyo, fyl);
      pair(A.fy, fyo + nx, fyh);
      pair(A.D, c, d);
      if constexpr (kMode != kFgOpDot) pair(aux, c, av);
      const double fx0 = A.fx[c + row];
      const double fx1 = A.fx[c + row + 1];
      const double fx2 = A.fx[c + row + 2];
      const double xl0 = i0 > 0 ? x[c - 1]
                                : neighbour_or_boundary<double>(
                                      bc.m[0], xc[0], x, c + (nx - 1));
      const double xh1 = i0 + 2 < nx
                             ? x[c + 2]
                             : neighbour_or_boundary<double>(
                                   bc.m[1], xc[1], x, c + 1 - (nx - 1));
      const double xl[2] = {xl0, xc[0]};
      const double xh[2] = {xc[1], xh1};
      const double fxl[2] = {fx0, fx1};
      const double fxh[2] = {fx1, fx2};
      Pair o;
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        double v = d[e] * xc[e];
        v -= fxl[e] * xl[e];
        v -= fxh[e] * xh[e];
        v -= fyl[e] * yl[e];
        v -= fyh[e] * yh[e];
        v -= fzl[e] * zl[e];
        v -= fzh[e] * zh[e];
        if (volume_rows) v = v / (wxy[e] * w.z[k]);
        if constexpr (kMode == kFgArnoldi)
          o.v[e] = v + -shift * av[e];
        else if constexpr (kMode == kFgResidual)
          o.v[e] = 1.0 * av[e] + -1.0 * v;
        else
          o.v[e] = v;
      }
      *reinterpret_cast<Pair*>(out + c) = o;
      if constexpr (kMode == kFgArnoldi) {
#pragma unroll
        for (int a = 0; a < kNd; ++a) {
          double bv[2];
          pair(basis + a * vstride, c, bv);
          acc[a] += bv[0] * o.v[0];
          acc[a] += bv[1] * o.v[1];
        }
      }
      if constexpr (kMode == kFgOpDot) {
        ww += xc[0] * o.v[0];
        ww += xc[1] * o.v[1];
      } else {
        ww += o.v[0] * o.v[0];
        ww += o.v[1] * o.v[1];
      }
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        zl[e] = xc[e];
        xc[e] = zh[e];
        fzl[e] = fzh[e];
      }
      c += plane;
    }
  }
  if constexpr (kMode == kFgArnoldi) {
#pragma unroll
    for (int a = 0; a < kNd; ++a) {
      const double s = block_sum(acc[a], scratch);
      if (threadIdx.x == 0) partials[a * pstride + lb] = s;
    }
  }
  const double s = block_sum(ww, scratch);
  if (threadIdx.x == 0) partials[wacc * pstride + lb] = s;
}

// k_fg_dots over pairs of cells of a row (t tiles (nx / 2) x ny columns),
// in the partial layout of k_fg_stencil2.
template <int kNd>
__global__ void k_fg_dots2(Tiles t, const double* wv, const double* basis,
                           int64_t vstride, int row0, int acc0,
                           double* partials, int64_t pstride) {
  static_assert(kNd >= 1 && kNd <= kFgDots, "dots of one pass");
  struct alignas(16) Pair {
    double v[2];
  };
  __shared__ double scratch[16];
  const int lb = xcd_block32(blockIdx.x, gridDim.x);
  const int gxy = t.gx * t.gy;
  const int zb = lb / gxy;
  int ip, j, k_unused;
  tile_xyz(t, lb - zb * gxy, ip, j, k_unused);
  const int nx = 2 * t.nx, ny = t.ny, nz = t.nz;
  const int k0 = zb * t.per;
  const int k1 = min(nz, k0 + t.per);
  double acc[kNd];
#pragma unroll
  for (int a = 0; a < kNd; ++a) acc[a] = 0.0;
  if (ip < t.nx && j < ny) {
    const int plane = nx * ny;
    int c = cell_index(nx, ny, 2 * ip, j, k0);
    for (int k = k0; k < k1; ++k) {
      const Pair o = *reinterpret_cast<const Pair*>(wv + c);
#pragma unroll
      for (int a = 0; a < kNd; ++a) {
        const Pair b = *reinterpret_cast<const Pair*>(
            basis + (row0 + a) * vstride + c);
        acc[a] += b.v[0] * o.v[0];
        acc[a] += b.v[1] * o.v[1];
      }
      c += plane;
    }
  }
#pragma unroll
  for (int a = 0; a < kNd; ++a) {
    const double s = block_sum(acc[a], scratch);
    if (threadIdx.x == 0) partials[(acc0 + a) * pstride + lb] = s;
  }
}

// Scales of one basis combination, passed by value.
struct FgScales {
  double s[kFgMaxRestart + 1];
};

// dst = src + sum_{row < count} s[row] * basis_row, the terms added in row
// order (add_scaled_basis, src/solver_krylov.cpp:510-555); with scale != 0
// the sum is then multiplied by post (the normalisation of a basis vector,
// which must stay finite: a non-finite element sets *flag to 1).  pairs
// (an even vector length, so every basis row starts 16-byte aligned): a
// thread forms two consecutive elements with 16-byte accesses
// (ch.chunk = 2 kBlock); else one element per thread (ch.chunk = kBlock).
struct alignas(16) FgPair {
  double v[2];
};

__device__ __forceinline__ double fg_combine_one(const double* src,
                                                 const double* basis,
                                                 int64_t vstride,
                                                 const FgScales& s, int count,
                                                 int64_t c) {
  double v = src[c];
  for (int row = 0; row < count; ++row) v = v + s.s[row] * basis[row * vstride + c];
  return v;
}

__global__ void k_fg_combine(Chunk ch, double* dst, const double* src,
                             const double* basis, int64_t vstride,
                             FgScales s, int count, double post, int scale,
                             double* flag, int pairs) {
  int lb, lo, hi;
  chunk_range(ch, lb, lo, hi);
  bool bad = false;
  const int step = pairs ? 2 : 1;
  for (int c = lo + step * static_cast<int>(threadIdx.x); c < hi;
       c += step * kBlock) {
    if (pairs && c + 1 < hi) {
      FgPair v = *reinterpret_cast<const FgPair*>(src + c);
      for (int row = 0; row < count; ++row) {
        const FgPair b =
            *reinterpret_cast<const FgPair*>(basis + row * vstride + c);
        v.v[0] = v.v[0] + s.s[row] * b.v[0];
        v.v[1] = v.v[1] + s.s[row] * b.v[1];
      }
      if (scale) {
        v.v[0] = post * v.v[0];
        v.v[1] = post * v.v[1];
        bad = bad || !isfinite(v.v[0]) || !isfinite(v.v[1]);
      }
      *reinterpret_cast<FgPair*>(dst + c) = v;
    } else {
      double v = fg_combine_one(src, basis, vstride, s, count, c);
      if (scale) {
        v = post * v;
        bad = bad || !isfinite(v);
      }
      dst[c] = v;
    }
  }
  if (bad) *flag = 1.0;
}

// k_fg_combine with the normalisation (scale), also storing the
// multigrid's level-0 entry of the result, b = fluid ? v : 0 at level
// storage (k_mg_entry's values, without projection sums): the
// preconditioner applied to dst next then starts from b without reading dst
// again.
template <class T>
__global__ void k_fg_combine_entry(Chunk ch, double* dst, const double* src,
                                   const double* basis, int64_t vstride,
                                   FgScales s, int count, double post,
                                   double* flag, const uint8_t* act, T* b,
                                   int pairs) {
  struct alignas(2 * sizeof(T)) BPair {
    T v[2];
  };
  int lb, lo, hi;
  chunk_range(ch, lb, lo, hi);
  bool bad = false;
  const int step = pairs ? 2 : 1;
  for (int c = lo + step * static_cast<int>(threadIdx.x); c < hi;
       c += step * kBlock) {
    if (pairs && c + 1 < hi) {
      FgPair v = *reinterpret_cast<const FgPair*>(src + c);
      for (int row = 0; row < count; ++row) {
        const FgPair bs =
            *reinterpret_cast<const FgPair*>(basis + row * vstride + c);
        v.v[0] = v.v[0] + s.s[row] * bs.v[0];
        v.v[1] = v.v[1] + s.s[row] * bs.v[1];
      }
      v.v[0] = post * v.v[0];
      v.v[1] = post * v.v[1];
      bad = bad || !isfinite(v.v[0]) || !isfinite(v.v[1]);
      *reinterpret_cast<FgPair*>(dst + c) = v;
      const uint16_t fl = *reinterpret_cast<const uint16_t*>(act + c);
      BPair bo;
      bo.v[0] = (fl & 0xffu) != 0 ? st<T>(v.v[0]) : st<T>(0.0);
      bo.v[1] = (fl >> 8) != 0 ? st<T>(v.v[1]) : st<T>(0.0);
      *reinterpret_cast<BPair*>(b + c) = bo;
    } else {
      double v = post * fg_combine_one(src, basis, vstride, s, count, c);
      bad = bad || !isfinite(v);
      dst[c] = v;
      b[c] = act[c] != 0 ? st<T>(v) : st<T>(0.0);
    }
  }
  if (bad) *flag = 1.0;
}

// dst = src * (w_x w_y w_z): the residual handed to the multigrid by the
// volume-scaled preconditioner (src/solver_cold.hpp:378-415).
__global__ void k_scale_volume(Tiles t, double* dst, const double* src,
                               Widths w) {
  const int lb = xcd_block32(blockIdx.x, gridDim.x);
  int i, j, k;
  tile_xyz(t, lb, i, j, k);
  if (i >= t.nx || j >= t.ny) return;
  const int c = cell_index(t.nx, t.ny, i, j, k);
  dst[c] = src[c] * geom_volume(w, i, j, k);
}

// out[a] = sum of accumulator a's partials, a < nacc (one block; partials
// are summed in the fixed shuffle-tree order of block_sum).
__global__ void __launch_bounds__(1024)
k_finalize_many(const double* partials, int64_t blocks, int64_t stride,
                int nacc, double* out) {
  __shared__ double scratch[16];
  for (int a = 0; a < nacc; ++a) {
    double v = 0.0;
    for (int64_t b = threadIdx.x; b < blocks; b += blockDim.x)
      v += partials[a * stride + b];
    v = block_sum(v, scratch);
    if (threadIdx.x == 0) out[a] = v;
  }
}

}  // namespace hf
