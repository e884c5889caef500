// Line relaxation for strongly anisotropic levels (src/solver_mg.cpp:
// 3352-3536, line_smooth_axis): every line along the line axis is solved
// exactly by the Thomas algorithm, with the transverse couplings taken from
// the current iterate.
//
// Lines are indexed (a, b), a fastest: x-lines run over (y, z), y-lines over
// (x, z), z-lines over (x, y).  Two orders are provided:
//  * lexicographic (the reference): b outer, a inner.  A line reads its
//    (a-1, b) and (a, b-1) neighbours after their update and (a+1, b),
//    (a, b+1) before it, so every line on the hyperplane a + b = t depends
//    only on hyperplane t - 1: one workgroup walks the hyperplanes with a
//    barrier between them and produces the sequential result bit for bit.
//  * zebra: lines with (a + b) even, then odd (reverse sweeps: odd, even);
//    lines of one colour never neighbour, so a colour is one parallel pass.
// A periodic transverse wrap reads a ghost copy of the boundary plane taken
// when the sweep (zebra: the colour pass) begins -- the single-rank halo of
// the reference.  A periodic wrap along the line reads the line's own far
// end, which the line has not yet written when it is read.
#pragma once

#include <float.h>

#include "hf_device.hpp"

namespace hf {

// Longest line relaxed exactly; longer lines fall back to the point smoother
// (MgHierarchyPolicy::line_relaxation_maximum_extent default,
// include/hundun/v04_linear.hpp:123).
constexpr int kLineMaximumExtent = 4096;

template <class T>
struct LineArgs {
  using C = typename Compute<T>::type;
  Lv<T> L;
  int modes;      // boundary mode of face f (Bc encoding) in bits 2f, 2f+1
  int ax;         // line axis
  int A, B, E;    // lines per a, lines per b, cells per line
  C* cw;          // Thomas c' per cell (the reference's temporary slot)
  C* cd;          // Thomas d' per cell (the reference's residual slot)
  T* ghost_a;     // [side][b][m] of planes a = 0 (side 0), a = A-1 (side 1)
  T* ghost_b;     // [side][a][m] of planes b = 0, b = B-1
  double* flag;   // set to 1 when a pivot vanishes (scalar slot)
};

__device__ __forceinline__ void line_cell(int ax, int a, int b, int m,
                                          int& i, int& j, int& k) {
  if (ax == 0) {
    i = m; j = a; k = b;
  } else if (ax == 1) {
    i = a; j = m; k = b;
  } else {
    i = a; j = b; k = m;
  }
}

__device__ __forceinline__ double abs_c(double v) { return fabs(v); }
__device__ __forceinline__ float abs_c(float v) { return fabsf(v); }
__device__ __forceinline__ double tiny_c(double) { return DBL_MIN; }
__device__ __forceinline__ float tiny_c(float) { return FLT_MIN; }

template <class T>
__device__ __forceinline__ typename Compute<T>::type line_face(
    const Lv<T>& L, int axis, int high, int i, int j, int k, int c) {
  if (axis == 0) return ld(L.fx[c + j + L.ny * k + high]);
  if (axis == 1) return ld(L.fy[c + L.nx * k + high * L.nx]);
  return ld(L.fz[c + high * L.nx * L.ny]);
}

// level_neighbor (src/solver_mg.cpp:2229-2265) from line (a, b) at cell m.
template <class T>
__device__ __forceinline__ typename Compute<T>::type line_neighbour(
    const LineArgs<T>& g, int a, int b, int m, int i, int j, int k, int c,
    int axis, int dir) {
  using C = typename Compute<T>::type;
  const Lv<T>& L = g.L;
  const int extent = axis == 0 ? L.nx : (axis == 1 ? L.ny : L.nz);
  const int pos = axis == 0 ? i : (axis == 1 ? j : k);
  const int stride = axis == 0 ? 1 : (axis == 1 ? L.nx : L.nx * L.ny);
  const int next = pos + dir;
  if (next >= 0 && next < extent) return ld(L.x[c + dir * stride]);
  const int face = 2 * axis + (dir > 0 ? 1 : 0);
  const int mode = (g.modes >> (2 * face)) & 3;
  if (mode == 0) return C(0);
  if (mode == 1) return ld(L.x[c]);
  if (axis == g.ax) return ld(L.x[c - dir * (extent - 1) * stride]);
  const int a_axis = g.ax == 0 ? 1 : 0;
  // wrap from side 0 reads the far plane (side 1) and vice versa
  const int side = dir < 0 ? 1 : 0;
  if (axis == a_axis) return ld(g.ghost_a[(side * g.B + b) * g.E + m]);
  return ld(g.ghost_b[(side * g.A + a) * g.E + m]);
}

// solve_one (src/solver_mg.cpp:3390-3502) for line (a, b).  Returns false,
// leaving x of the line untouched, when a pivot is not above the smallest
// normal number.
template <class T>
__device__ __forceinline__ bool line_solve(const LineArgs<T>& g, int a, int b) {
  using C = typename Compute<T>::type;
  const Lv<T>& L = g.L;
  const int ax = g.ax;
  C cw_prev = C(0), cd_prev = C(0);
  for (int m = 0; m < g.E; ++m) {
    int i, j, k;
    line_cell(ax, a, b, m, i, j, k);
    const int c = i + L.nx * (j + L.ny * k);
    const C lo = line_face(L, ax, 0, i, j, k, c);
    const C hi = line_face(L, ax, 1, i, j, k, c);
    const C sub = m == 0 ? C(0) : -lo;
    const C dia = ld(L.D[c]);
    const C sup = m + 1 == g.E ? C(0) : -hi;
    C d = ld(L.b[c]);
    for (int t = 0; t < 3; ++t) {
      if (t == ax) continue;
      d += line_face(L, t, 0, i, j, k, c) *
           line_neighbour(g, a, b, m, i, j, k, c, t, -1);
      d += line_face(L, t, 1, i, j, k, c) *
           line_neighbour(g, a, b, m, i, j, k, c, t, 1);
    }
    if (m == 0) d += lo * line_neighbour(g, a, b, m, i, j, k, c, ax, -1);
    if (m + 1 == g.E) d += hi * line_neighbour(g, a, b, m, i, j, k, c, ax, 1);
    C cw, cd;
    if (m > 0) {
      const C den = dia - sub * cw_prev;
      if (!(abs_c(den) > tiny_c(den))) return false;
      cw = sup / den;
      cd = (d - sub * cd_prev) / den;
    } else {
      if (!(abs_c(dia) > tiny_c(dia))) return false;
      cw = sup / dia;
      cd = d / dia;
    }
    g.cw[c] = cw;
    g.cd[c] = cd;
    cw_prev = cw;
    cd_prev = cd;
  }
  C x_next = C(0);
  for (int m = g.E - 1; m >= 0; --m) {
    int i, j, k;
    line_cell(ax, a, b, m, i, j, k);
    const int c = i + L.nx * (j + L.ny * k);
    const C v = m + 1 == g.E ? g.cd[c] : g.cd[c] - g.cw[c] * x_next;
    const T stored = st<T>(v);
    L.x[c] = stored;
    x_next = ld(stored);
  }
  return true;
}

// Copy the periodic transverse boundary planes of x into the ghosts; entry
// e enumerates the ghost cells (a-planes first, then b-planes).
template <class T>
__device__ __forceinline__ void line_ghost_copy(const LineArgs<T>& g, int e) {
  const Lv<T>& L = g.L;
  const int na = g.ghost_a != nullptr ? 2 * g.B * g.E : 0;
  int a, b, m;
  T* dst;
  if (e < na) {
    const int side = e / (g.B * g.E);
    const int rem = e - side * g.B * g.E;
    b = rem / g.E;
    m = rem - b * g.E;
    a = side == 0 ? 0 : g.A - 1;
    dst = g.ghost_a + e;
  } else {
    const int f = e - na;
    const int side = f / (g.A * g.E);
    const int rem = f - side * g.A * g.E;
    a = rem / g.E;
    m = rem - a * g.E;
    b = side == 0 ? 0 : g.B - 1;
    dst = g.ghost_b + f;
  }
  int i, j, k;
  line_cell(g.ax, a, b, m, i, j, k);
  *dst = L.x[i + L.nx * (j + L.ny * k)];
}

template <class T>
__device__ __forceinline__ int line_ghost_count(const LineArgs<T>& g) {
  return (g.ghost_a != nullptr ? 2 * g.B * g.E : 0) +
         (g.ghost_b != nullptr ? 2 * g.A * g.E : 0);
}

// Lexicographic sweeps of one axis in one workgroup (hyperplane order).
template <class T>
__global__ void __launch_bounds__(1024)
    k_line_lex(LineArgs<T> g, int sweeps, int reverse) {
  const int tid = threadIdx.x;
  const int nthreads = blockDim.x;
  const int planes = g.A + g.B - 1;
  const int ghosts = line_ghost_count(g);
  for (int sweep = 0; sweep < sweeps; ++sweep) {
    for (int e = tid; e < ghosts; e += nthreads) line_ghost_copy(g, e);
    __syncthreads();
    for (int step = 0; step < planes; ++step) {
      const int t = reverse ? planes - 1 - step : step;
      const int a_lo = max(0, t - (g.B - 1));
      const int a_hi = min(t, g.A - 1);
      for (int a = a_lo + tid; a <= a_hi; a += nthreads) {
        if (!line_solve(g, a, t - a)) g.flag[0] = 1;
      }
      __syncthreads();
    }
  }
}

template <class T>
__global__ void k_line_ghost(LineArgs<T> g) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < line_ghost_count(g)) line_ghost_copy(g, e);
}

// One zebra colour pass: thread -> line (2q + ((colour + b) & 1), b).
template <class T>
__global__ void k_line_zebra(LineArgs<T> g, int colour) {
  const int half = (g.A + 1) / 2;
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= half * g.B) return;
  const int b = id / half;
  const int a = 2 * (id - b * half) + ((colour + b) & 1);
  if (a >= g.A) return;
  if (!line_solve(g, a, b)) g.flag[0] = 1;
}

}  // namespace hf
