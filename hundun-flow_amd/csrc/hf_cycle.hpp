// Multigrid cycle and Krylov kernels (gfx950).
//
// Grid-level kernels map one 256-thread block to a 2-D tile of one z-plane
// (Tiles in hf_device.hpp) and vector kernels one block to 256 consecutive
// elements, so every launch streams HBM in dispatch order.  Reductions write
// one partial per block; the host-side finaliser folds them in fixed order.
// Block numbers are remapped so that every XCD owns a contiguous z-slab and
// neighbour planes are served by its own L2.
#pragma once

#include "hf_device.hpp"

namespace hf {

struct Widths {
  const double* x;
  const double* y;
  const double* z;
};

__device__ __forceinline__ int cell_index(int nx, int ny, int i, int j, int k) {
  return i + nx * (j + ny * k);
}

// ---------------------------------------------------------------------------
// Vector kernels: each block covers one contiguous chunk of elements.
// ---------------------------------------------------------------------------
struct Chunk {
  int n;       // elements
  int chunk;   // elements per block (multiple of kBlock)
};

__device__ __forceinline__ void chunk_range(const Chunk& ch, int& lb, int& lo,
                                            int& hi) {
  lb = xcd_block32(blockIdx.x, gridDim.x);
  lo = lb * ch.chunk;
  hi = min(ch.n, lo + ch.chunk);
}

__global__ void k_sum(Chunk ch, const double* field, const uint8_t* act,
                      double* partials) {
  __shared__ double scratch[16];
  int lb, lo, hi;
  chunk_range(ch, lb, lo, hi);
  double v = 0.0;
  for (int c = lo + threadIdx.x; c < hi; c += kBlock)
    if (act == nullptr || act[c]) v += field[c];
  v = block_sum(v, scratch);
  if (threadIdx.x == 0) partials[lb] = v;
}

__global__ void k_count(Chunk ch, const uint8_t* act, double* partials) {
  __shared__ double scratch[16];
  int lb, lo, hi;
  chunk_range(ch, lb, lo, hi);
  double v = 0.0;
  for (int c = lo + threadIdx.x; c < hi; c += kBlock)
    v += act == nullptr ? 1.0 : static_cast<double>(act[c] != 0);
  v = block_sum(v, scratch);
  if (threadIdx.x == 0) partials[lb] = v;
}

__global__ void k_dot(Chunk ch, const double* a, const double* b,
                      double* partials) {
  __shared__ double scratch[16];
  int lb, lo, hi;
  chunk_range(ch, lb, lo, hi);
  double d = 0.0;
  for (int c = lo + threadIdx.x; c < hi; c += kBlock) d += a[c] * b[c];
  d = block_sum(d, scratch);
  if (threadIdx.x == 0) partials[lb] = d;
}

__global__ void k_subtract_mean(Chunk ch, double* field, const double* mean,
                                const uint8_t* act) {
  int lb, lo, hi;
  chunk_range(ch, lb, lo, hi);
  const double m = *mean;
  for (int c = lo + threadIdx.x; c < hi; c += kBlock)
    field[c] = (act == nullptr || act[c]) ? field[c] - m : 0.0;
}

// x += alpha p; r += (-alpha) q; partial sum of r.r
// (add_scaled, src/solver_krylov.cpp:383-399, twice).
__global__ void k_update_xr(Chunk ch, double* x, const double* p, double* r,
                            const double* q, const double* alpha,
                            double* partials) {
  __shared__ double scratch[16];
  int lb, lo, hi;
  chunk_range(ch, lb, lo, hi);
  const double a = *alpha;
  double d = 0.0;
  for (int c = lo + threadIdx.x; c < hi; c += kBlock) {
    x[c] = x[c] + a * p[c];
    const double rc = r[c] + -a * q[c];
    r[c] = rc;
    d += rc * rc;
  }
  d = block_sum(d, scratch);
  if (threadIdx.x == 0) partials[lb] = d;
}


// x += alpha p alone, for an iteration whose r the true residual replaces
// (the x update of k_update_xr, same expression).
__global__ void k_update_x(Chunk ch, double* x, const double* p,
                           const double* alpha) {
  int lb, lo, hi;
  chunk_range(ch, lb, lo, hi);
  const double a = *alpha;
  for (int c = lo + threadIdx.x; c < hi; c += kBlock) x[c] = x[c] + a * p[c];
}

__global__ void k_copy(Chunk ch, double* dst, const double* src) {
  int lb, lo, hi;
  chunk_range(ch, lb, lo, hi);
  for (int c = lo + threadIdx.x; c < hi; c += kBlock) dst[c] = src[c];
}

// dest = a * left + b * right (axpby, src/solver_krylov.cpp:344-361); a
// non-finite result sets *flag.
__global__ void k_axpby(Chunk ch, double a, const double* left, double b,
                        const double* right, double* dest, double* flag) {
  int lb, lo, hi;
  chunk_range(ch, lb, lo, hi);
  bool bad = false;
  for (int c = lo + threadIdx.x; c < hi; c += kBlock) {
    const double v = a * left[c] + b * right[c];
    bad = bad || !isfinite(v);
    dest[c] = v;
  }
  if (bad) *flag = 1.0;
}

// dst = src (copy_field, src/solver_krylov.cpp:300-316); a non-finite
// value sets *flag.
__global__ void k_copy_checked(Chunk ch, double* dst, const double* src,
                               double* flag) {
  int lb, lo, hi;
  chunk_range(ch, lb, lo, hi);
  bool bad = false;
  for (int c = lo + threadIdx.x; c < hi; c += kBlock) {
    const double v = src[c];
    bad = bad || !isfinite(v);
    dst[c] = v;
  }
  if (bad) *flag = 1.0;
}

// dst[0..n) = value: clears memory with vector stores, where a solve may run
// inside a caller's captured stream work (no memset nodes).
__global__ void k_fill(Chunk ch, double* dst, double value) {
  int lb, lo, hi;
  chunk_range(ch, lb, lo, hi);
  for (int c = lo + threadIdx.x; c < hi; c += kBlock) dst[c] = value;
}

// ---------------------------------------------------------------------------
// Projection helpers (constant null space, src/solver_mg.cpp:1412-1454).
// Level 0 recomputes its fluid volumes from the axis widths and the activity
// byte, (w_x*w_y)*w_z exactly as the hierarchy stored them.
// ---------------------------------------------------------------------------
// Reduction-only kernels take kRedPer consecutive tiles per block: the loads
// of all of them are issued before any is used and one block sum serves
// kRedPer * 256 cells (these kernels move only a few bytes per cell).
constexpr int kRedPer = 8;

// Cell of tile slot u of block lb in a kRedPer launch; ok = inside the level.
__device__ __forceinline__ int red_cell(const Tiles& t, int lb, int u, int& i,
                                        int& j, int& k, bool& ok) {
  const int tile = lb * kRedPer + u;
  tile_xyz(t, tile < t.count ? tile : 0, i, j, k);
  ok = tile < t.count && i < t.nx && j < t.ny;
  if (!ok) {
    i = 0;
    j = 0;
    k = 0;
  }
  return ok ? cell_index(t.nx, t.ny, i, j, k) : 0;
}

template <class T>
__global__ void k_vol_dot_finest(Tiles t, const T* field, const uint8_t* act,
                                 Widths w, double* partials) {
  __shared__ double scratch[16];
  const int lb = xcd_block32(blockIdx.x, gridDim.x);
  double f[kRedPer], vol[kRedPer];
  bool use[kRedPer];
#pragma unroll
  for (int u = 0; u < kRedPer; ++u) {
    int i, j, k;
    bool ok;
    const int c = red_cell(t, lb, u, i, j, k, ok);
    use[u] = ok && act[c] != 0;
    f[u] = static_cast<double>(ld(field[c]));
    vol[u] = w.x[i] * w.y[j] * w.z[k];
  }
  double v = 0.0;
#pragma unroll
  for (int u = 0; u < kRedPer; ++u)
    if (use[u]) v += vol[u] * f[u];
  v = block_sum(v, scratch);
  if (threadIdx.x == 0) partials[lb] = v;
}

template <class T>
__global__ void k_vol_dot(Tiles t, const T* field, const double* vol,
                          double* partials) {
  __shared__ double scratch[16];
  const int lb = xcd_block32(blockIdx.x, gridDim.x);
  double f[kRedPer], vv[kRedPer];
  bool use[kRedPer];
#pragma unroll
  for (int u = 0; u < kRedPer; ++u) {
    int i, j, k;
    bool ok;
    const int c = red_cell(t, lb, u, i, j, k, ok);
    use[u] = ok;
    f[u] = static_cast<double>(ld(field[c]));
    vv[u] = vol[c];
  }
  double v = 0.0;
#pragma unroll
  for (int u = 0; u < kRedPer; ++u)
    if (use[u]) v += vv[u] * f[u];
  v = block_sum(v, scratch);
  if (threadIdx.x == 0) partials[lb] = v;
}

// The PCG update x += a p, r -= a q of k_update_xr fused with the
// preconditioner's entry (k_mg_entry): b = fluid ? r : 0 at the level-0
// storage type.  Partial sums: r.r in accumulator 0 and, with the constant
// null space, vol*b in accumulator 1 (stride apart).  Tiles of the reduction
// layout (kRedPer per block).  x == nullptr: x is left for the direction
// update (k_update_p_exit), which reads p anyway.
template <class T>
__global__ void k_update_xr_entry(Tiles t, double* x, const double* pv,
                                  double* r, const double* q,
                                  const double* alpha, const uint8_t* act,
                                  Widths w, T* b, double* partials,
                                  int stride, int proj) {
  __shared__ double scratch[16];
  const int lb = xcd_block32(blockIdx.x, gridDim.x);
  const double a = *alpha;
  const bool with_x = x != nullptr;
  // two groups of four tiles: each group's loads are issued together (eight
  // tiles of five FP64 operands would not fit the register file)
  constexpr int kGroup = 4;
  static_assert(kRedPer % kGroup == 0, "tile groups");
  double d = 0.0, m = 0.0;
#pragma unroll
  for (int g = 0; g < kRedPer; g += kGroup) {
    double xv[kGroup], pp[kGroup], rv[kGroup], qv[kGroup], vol[kGroup];
    bool ok[kGroup], fluid[kGroup];
    int cc[kGroup];
#pragma unroll
    for (int u = 0; u < kGroup; ++u) {
      int i, j, k;
      cc[u] = red_cell(t, lb, g + u, i, j, k, ok[u]);
      xv[u] = with_x ? x[cc[u]] : 0.0;
      pp[u] = with_x ? pv[cc[u]] : 0.0;
      rv[u] = r[cc[u]];
      qv[u] = q[cc[u]];
      fluid[u] = act[cc[u]] != 0;
      vol[u] = w.x[i] * w.y[j] * w.z[k];
    }
#pragma unroll
    for (int u = 0; u < kGroup; ++u) {
      if (!ok[u]) continue;
      if (with_x) x[cc[u]] = xv[u] + a * pp[u];
      const double rc = rv[u] + -a * qv[u];
      r[cc[u]] = rc;
      d += rc * rc;
      const T bv = fluid[u] ? st<T>(rc) : st<T>(0.0);
      b[cc[u]] = bv;
      if (proj && fluid[u]) m += vol[u] * static_cast<double>(ld(bv));
    }
  }
  d = block_sum(d, scratch);
  if (threadIdx.x == 0) partials[lb] = d;
  if (proj) {
    m = block_sum(m, scratch);
    if (threadIdx.x == 0) partials[stride + lb] = m;
  }
}

// f := fluid ? f - mean : 0
template <class T>
__global__ void k_project_apply(Tiles t, T* field, const uint8_t* act,
                                const double* mean) {
  int i, j, k;
  tile_xyz(t, xcd_block32(blockIdx.x, gridDim.x), i, j, k);
  if (i >= t.nx || j >= t.ny) return;
  const int c = cell_index(t.nx, t.ny, i, j, k);
  field[c] = act[c] ? st<T>(static_cast<double>(ld(field[c])) - *mean)
                    : st<T>(0.0);
}

// ---------------------------------------------------------------------------
// Multigrid grid-level kernels.
// ---------------------------------------------------------------------------

// apply_impl entry (src/solver_mg.cpp:4867-4879): b := r on fluid cells and,
// unless the first smoothing pass starts from zero itself, x := 0.  With a
// constant null space the block partials of vol*b are produced here.
template <class T>
__global__ void k_mg_entry(Tiles t, const double* r, const uint8_t* act,
                           Widths w, T* b, T* x, double* partials) {
  __shared__ double scratch[16];
  const int lb = xcd_block32(blockIdx.x, gridDim.x);
  double rv[kRedPer], vol[kRedPer];
  bool ok[kRedPer], fluid[kRedPer];
  int cc[kRedPer];
#pragma unroll
  for (int u = 0; u < kRedPer; ++u) {
    int i, j, k;
    cc[u] = red_cell(t, lb, u, i, j, k, ok[u]);
    fluid[u] = act[cc[u]] != 0;
    rv[u] = r[cc[u]];
    vol[u] = w.x[i] * w.y[j] * w.z[k];
  }
  double v = 0.0;
#pragma unroll
  for (int u = 0; u < kRedPer; ++u) {
    if (!ok[u]) continue;
    const bool a = fluid[u];
    const T bv = a ? st<T>(rv[u]) : st<T>(0.0);
    b[cc[u]] = bv;
    if (x != nullptr) x[cc[u]] = st<T>(0.0);
    if (partials != nullptr && a) v += vol[u] * static_cast<double>(ld(bv));
  }
  if (partials != nullptr) {
    v = block_sum(v, scratch);
    if (threadIdx.x == 0) partials[lb] = v;
  }
}

// First colour pass of a smoothing that starts from x = 0: A x = 0, so
// colour-0 cells get x = 0 + 0.72 b / D and colour-1 cells x = 0, reading no
// neighbours and no faces.  A pending constant projection of b (mean != null)
// is applied and stored here: b := fluid ? b - mean : 0.
template <class T>
__global__ void k_rb_first(Tiles t, Lv<T> L, const double* mean) {
  using C = typename Compute<T>::type;
  int i, j, k;
  tile_xyz(t, xcd_block32(blockIdx.x, gridDim.x), i, j, k);
  if (i >= t.nx || j >= t.ny) return;
  const int c = cell_index(t.nx, t.ny, i, j, k);
  C bv;
  if (mean != nullptr) {
    const T stored = L.act[c] ? st<T>(static_cast<double>(ld(L.b[c])) - *mean)
                              : st<T>(0.0);
    L.b[c] = stored;
    bv = C(ld(stored));
  } else {
    bv = C(ld(L.b[c]));
  }
  if (((i + j + k) & 1) == 0) {
    const C omega = C(0.72);
    const C zero = C(0);
    L.x[c] = st<T>(zero + omega * (bv - zero) / C(ld(L.D[c])));
  } else {
    L.x[c] = st<T>(0.0);
  }
}

// One colour pass of the fused red-black update (src/solver_mg.cpp:
// 3149-3222): x += 0.72 (b - A x)/D on cells with (i+j+k)&1 == color.  One
// thread per cell of the colour; same-colour cells never neighbour, so the
// in-place update is order independent.
template <class T>
__global__ void k_rb(Tiles t, Lv<T> L, Bc bc, int color) {
  using C = typename Compute<T>::type;
  int ic, j, k;
  tile_xyz(t, xcd_block32(blockIdx.x, gridDim.x), ic, j, k);
  if (ic >= t.nx || j >= t.ny) return;
  const int i = 2 * ic + ((color + j + k) & 1);
  if (i >= L.nx) return;
  const int c = cell_index(L.nx, L.ny, i, j, k);
  const C xc = ld(L.x[c]);
  const C v = stencil<C>(L, bc, L.x, i, j, k, c, xc);
  const C omega = C(0.72);
  L.x[c] = st<T>(xc + omega * (C(ld(L.b[c])) - v) / C(ld(L.D[c])));
}

// Separated route for levels with an odd periodic extent
// (src/solver_mg.cpp:3262-3307): the colour update uses the A x computed
// before the pass.
template <class T>
__global__ void k_rb_separated(Tiles t, Lv<T> L, int color) {
  using C = typename Compute<T>::type;
  int ic, j, k;
  tile_xyz(t, xcd_block32(blockIdx.x, gridDim.x), ic, j, k);
  if (ic >= t.nx || j >= t.ny) return;
  const int i = 2 * ic + ((color + j + k) & 1);
  if (i >= L.nx) return;
  const int c = cell_index(L.nx, L.ny, i, j, k);
  const C omega = C(0.72);
  L.x[c] = st<T>(C(ld(L.x[c])) +
                 omega * (C(ld(L.b[c])) - C(ld(L.r[c]))) / C(ld(L.D[c])));
}


// Restriction (src/solver_mg.cpp:3633-3720): coarse b = sum of children in
// dz, dy, dx order, coarse x := 0 (skipped when the next smoothing starts
// from zero by itself).  With a constant null space the block partials of
// vol*b are produced (chunked launch).
template <class Tf, class Tc>
__global__ void k_restrict(Tiles t, Lv<Tf> F, Lv<Tc> Cl, int mask,
                           int zero_x, double* partials) {
  using C = typename Compute<Tc>::type;
  __shared__ double scratch[16];
  const int lb = xcd_block32(blockIdx.x, gridDim.x);
  const int t0 = lb * t.per;
  const int t1 = min(t.count, t0 + t.per);
  const int mx = mask & 1, my = (mask >> 1) & 1, mz = (mask >> 2) & 1;
  double pv = 0.0;
  for (int tile = t0; tile < t1; ++tile) {
    int ci, cj, ck;
    tile_xyz(t, tile, ci, cj, ck);
    if (ci >= t.nx || cj >= t.ny) continue;
    const int xb = mx ? 2 * ci : ci, yb = my ? 2 * cj : cj,
              zb = mz ? 2 * ck : ck;
    C sum = C(0);
    for (int dz = 0; dz < (mz ? 2 : 1); ++dz)
      for (int dy = 0; dy < (my ? 2 : 1); ++dy)
        for (int dx = 0; dx < (mx ? 2 : 1); ++dx) {
          const int fx = xb + dx, fy = yb + dy, fz = zb + dz;
          if (fx >= F.nx || fy >= F.ny || fz >= F.nz) continue;
          sum += C(ld(F.r[cell_index(F.nx, F.ny, fx, fy, fz)]));
        }
    const int c = cell_index(Cl.nx, Cl.ny, ci, cj, ck);
    const Tc stored = st<Tc>(sum);
    Cl.b[c] = stored;
    if (zero_x) Cl.x[c] = st<Tc>(0.0);
    if (partials != nullptr) pv += Cl.vol[c] * static_cast<double>(ld(stored));
  }
  if (partials != nullptr) {
    pv = block_sum(pv, scratch);
    if (threadIdx.x == 0) partials[lb] = pv;
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
                                             int mask, int per, int i, int j,
                                             int k, int c) {
  using C = typename Compute<Tf>::type;
  // the fluid flag, x and the corners are loaded together (one memory
  // round trip); solid cells store zero
  const bool fluid = F.act[c] != 0;
  int xl, xr, yl, yr, zl, zr, xw, yw, zw;
  interp(i, mask & 1, Cl.nx, per & 1, &xl, &xr, &xw);
  interp(j, (mask >> 1) & 1, Cl.ny, (per >> 1) & 1, &yl, &yr, &yw);
  interp(k, (mask >> 2) & 1, Cl.nz, (per >> 2) & 1, &zl, &zr, &zw);
  // all eight corner values are loaded before any is used (a corner that
  // does not exist on an axis repeats the first one and is skipped below)
  const int gx[2] = {wrapi(xl, Cl.nx), wrapi(xr, Cl.nx)};
  const int gy[2] = {wrapi(yl, Cl.ny), wrapi(yr, Cl.ny)};
  const int gz[2] = {wrapi(zl, Cl.nz), wrapi(zr, Cl.nz)};
  C corner[2][2][2];
#pragma unroll
  for (int az = 0; az < 2; ++az)
#pragma unroll
    for (int ay = 0; ay < 2; ++ay)
#pragma unroll
      for (int ax = 0; ax < 2; ++ax)
        corner[az][ay][ax] = C(ld(
            Cl.x[cell_index(Cl.nx, Cl.ny, gx[ax], gy[ay], gz[az])]));
  const C xf = C(ld(F.x[c]));
  const int nz2 = zl == zr ? 1 : 2, ny2 = yl == yr ? 1 : 2,
            nx2 = xl == xr ? 1 : 2;
  // the reference order: az, ay, ax; value += ((wx*wy)*wz)*x_c
  C value = C(0);
#pragma unroll
  for (int az = 0; az < 2; ++az) {
    const C wz = az == 0 ? C(1) - C(zw) * C(0.25) : C(zw) * C(0.25);
#pragma unroll
    for (int ay = 0; ay < 2; ++ay) {
      const C wy = ay == 0 ? C(1) - C(yw) * C(0.25) : C(yw) * C(0.25);
#pragma unroll
      for (int ax = 0; ax < 2; ++ax) {
        const C wx = ax == 0 ? C(1) - C(xw) * C(0.25) : C(xw) * C(0.25);
        if (az < nz2 && ay < ny2 && ax < nx2)
          value += wx * wy * wz * corner[az][ay][ax];
      }
    }
  }
  F.x[c] = fluid ? st<Tf>(xf + value) : st<Tf>(0.0);
}

// Retained defect fused with the restriction that consumes it
// (src/solver_mg.cpp:3633-3720).  The z-marching layout of k_stencil_z with
// full tiles of even width and height and even z-chunks, on a level with
// even extents coarsened along every axis: each step's defects, stored and
// reloaded at the fine storage type, go through LDS to the thread owning the
// coarse cell (even i, even j), which adds its eight children in the order
// of k_restrict (dz, dy, dx) and stores the coarse b at the end of the odd
// plane.  The fine r is not stored.  zero_x and partials as in k_restrict.
//
// kFirstStage: the level was just smoothed by one Chebyshev-Jacobi stage
// from x = 0 (the pre-smoothing of a from-zero visit), which is not run:
// every x this kernel reads is formed from that cell's b and D as the
// stage stores it, st(0 + rf0 b / D), and the kernel also stores x of its
// own cells into xs (the stage's direction is dead: the next stage on the
// level starts a new recurrence).
template <class Tf, class Tc, bool kFirstStage = false>
__global__ void k_defect_restrict(Tiles t, Lv<Tf> L, Bc bc, Lv<Tc> Cl,
                                  int zero_x, double* partials,
                                  typename Compute<Tf>::type rf0 = 0,
                                  Tf* xs = nullptr) {
  using C = typename Compute<Tf>::type;
  using Cc = typename Compute<Tc>::type;
  __shared__ double scratch[16];
  __shared__ Cc defect[kBlock];
  const int lb = xcd_block32(blockIdx.x, gridDim.x);
  const int gxy = t.gx * t.gy;
  const int zb = lb / gxy;
  int i, j, k_unused;
  tile_xyz(t, lb - zb * gxy, i, j, k_unused);
  const int nx = t.nx, ny = t.ny, nz = t.nz;
  const int k0 = zb * t.per;
  const int k1 = min(nz, k0 + t.per);
  const int plane = nx * ny;
  const int tw = 1 << t.txl;
  const int tid = threadIdx.x;
  const bool owner = ((i | j) & 1) == 0;
  const Tf* x = L.x;
  // x at cell q: stored, or formed from b and D as the first stage stores it
  auto xat = [&](int q) -> C {
    if constexpr (kFirstStage)
      return C(ld(st<Tf>(C(0) + rf0 * C(ld(L.b[q])) / C(ld(L.D[q])))));
    else
      return C(ld(x[q]));
  };
  // neighbour_or_boundary through xat
  auto nob = [&](int mode, C self, int wrapped) -> C {
    if (mode == 0) return C(0);
    if (mode == 1) return self;
    return xat(wrapped);
  };
  int c = cell_index(nx, ny, i, j, k0);
  C xc = xat(c);
  C zl = k0 > 0 ? xat(c - plane) : nob(bc.m[4], xc, c + (nz - 1) * plane);
  C fzl = C(ld(L.fz[c]));
  Cc sum = Cc(0);
  double pv = 0.0;
  for (int k = k0; k < k1; ++k) {
    const C zh = k + 1 < nz ? xat(c + plane)
                            : nob(bc.m[5], xc, c - (nz - 1) * plane);
    const C fzh = C(ld(L.fz[c + plane]));
    const int row = j + ny * k;
    const C xl = i > 0 ? xat(c - 1) : nob(bc.m[0], xc, c + (nx - 1));
    const C xh = i + 1 < nx ? xat(c + 1) : nob(bc.m[1], xc, c - (nx - 1));
    const C yl = j > 0 ? xat(c - nx) : nob(bc.m[2], xc, c + (ny - 1) * nx);
    const C yh = j + 1 < ny ? xat(c + nx)
                            : nob(bc.m[3], xc, c - (ny - 1) * nx);
    if constexpr (kFirstStage) xs[c] = st<Tf>(xc);
    const int fyo = c + nx * k;
    C v = C(ld(L.D[c])) * xc;
    v -= C(ld(L.fx[c + row])) * xl;
    v -= C(ld(L.fx[c + row + 1])) * xh;
    v -= C(ld(L.fy[fyo])) * yl;
    v -= C(ld(L.fy[fyo + nx])) * yh;
    v -= fzl * zl;
    v -= fzh * zh;
    const Tf r = st<Tf>(C(ld(L.b[c])) - v);
    __syncthreads();  // the previous step's children have been read
    defect[tid] = Cc(ld(r));
    __syncthreads();
    if (owner) {
      // children (dx, dy) of this plane: dx fastest, as k_restrict adds them
      sum += defect[tid];
      sum += defect[tid + 1];
      sum += defect[tid + tw];
      sum += defect[tid + tw + 1];
      if (k & 1) {
        const int cc = cell_index(Cl.nx, Cl.ny, i >> 1, j >> 1, k >> 1);
        const Tc stored = st<Tc>(sum);
        Cl.b[cc] = stored;
        if (zero_x) Cl.x[cc] = st<Tc>(0.0);
        if (partials != nullptr)
          pv += Cl.vol[cc] * static_cast<double>(ld(stored));
        sum = Cc(0);
      }
    }
    zl = xc;
    xc = zh;
    fzl = fzh;
    c += plane;
  }
  if (partials != nullptr) {
    pv = block_sum(pv, scratch);
    if (threadIdx.x == 0) partials[lb] = pv;
  }
}

// z-marching prolongation: a block owns one 2-D tile of the fine level and
// t.per consecutive fine planes.  Each thread walks its fine column; the x
// and y interpolation is set up once, and the four corners of a coarse
// plane are loaded once and reused while consecutive fine planes need that
// plane (fine planes 2K+1 and 2K+2 both interpolate between coarse planes K
// and K+1).  Terms are added in the order of prolong_cell (the tail's
// per-cell form).
template <class Tc, class Tf>
__global__ void k_prolong_z(Tiles t, Lv<Tc> Cl, Lv<Tf> F, int mask, int per) {
  using C = typename Compute<Tf>::type;
  const int lb = xcd_block32(blockIdx.x, gridDim.x);
  const int gxy = t.gx * t.gy;
  const int zb = lb / gxy;
  int i, j, k_unused;
  tile_xyz(t, lb - zb * gxy, i, j, k_unused);
  if (i >= t.nx || j >= t.ny) return;
  const int k0 = zb * t.per;
  const int k1 = min(t.nz, k0 + t.per);
  int xl, xr, xw, yl, yr, yw;
  interp(i, mask & 1, Cl.nx, per & 1, &xl, &xr, &xw);
  interp(j, (mask >> 1) & 1, Cl.ny, (per >> 1) & 1, &yl, &yr, &yw);
  const int cx[2] = {wrapi(xl, Cl.nx), wrapi(xr, Cl.nx)};
  const int cy[2] = {wrapi(yl, Cl.ny), wrapi(yr, Cl.ny)};
  const int ny2 = yl == yr ? 1 : 2, nx2 = xl == xr ? 1 : 2;
  const int cplane = Cl.nx * Cl.ny;
  int off[4];  // in-plane offsets of the corners, index ay * 2 + ax
#pragma unroll
  for (int ay = 0; ay < 2; ++ay)
#pragma unroll
    for (int ax = 0; ax < 2; ++ax) off[ay * 2 + ax] = cy[ay] * Cl.nx + cx[ax];
  // the two coarse planes held (index, corner values)
  int held0 = -1, held1 = -1;
  C h0[4] = {C(0), C(0), C(0), C(0)}, h1[4] = {C(0), C(0), C(0), C(0)};
  int c = cell_index(t.nx, t.ny, i, j, k0);
  const int plane = t.nx * t.ny;
  for (int k = k0; k < k1; ++k, c += plane) {
    int zl, zr, zw;
    interp(k, (mask >> 2) & 1, Cl.nz, (per >> 2) & 1, &zl, &zr, &zw);
    const int gz[2] = {wrapi(zl, Cl.nz), wrapi(zr, Cl.nz)};
    const bool fluid = F.act[c] != 0;
    const C xf = C(ld(F.x[c]));
    C v[2][4];
#pragma unroll
    for (int az = 0; az < 2; ++az) {
      const int g = gz[az];
      if (g == held0) {
#pragma unroll
        for (int e = 0; e < 4; ++e) v[az][e] = h0[e];
      } else if (g == held1) {
#pragma unroll
        for (int e = 0; e < 4; ++e) v[az][e] = h1[e];
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) v[az][e] = C(ld(Cl.x[g * cplane + off[e]]));
      }
    }
    held0 = gz[0];
    held1 = gz[1];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      h0[e] = v[0][e];
      h1[e] = v[1][e];
    }
    const int nz2 = zl == zr ? 1 : 2;
    // the reference order: az, ay, ax; value += ((wx*wy)*wz)*x_c
    C value = C(0);
#pragma unroll
    for (int az = 0; az < 2; ++az) {
      const C wz = az == 0 ? C(1) - C(zw) * C(0.25) : C(zw) * C(0.25);
#pragma unroll
      for (int ay = 0; ay < 2; ++ay) {
        const C wy = ay == 0 ? C(1) - C(yw) * C(0.25) : C(yw) * C(0.25);
#pragma unroll
        for (int ax = 0; ax < 2; ++ax) {
          const C wx = ax == 0 ? C(1) - C(xw) * C(0.25) : C(xw) * C(0.25);
          if (az < nz2 && ay < ny2 && ax < nx2)
            value += wx * wy * wz * v[az][ay * 2 + ax];
        }
      }
    }
    F.x[c] = fluid ? st<Tf>(xf + value) : st<Tf>(0.0);
  }
}

// k_prolong_z over pairs of fine cells (2m, 2m + 1) of a row, for an even
// fine x extent: the pair's fluid flags, x and stored x move as one access
// of 2 and 2 sizeof(Tf) bytes (one cell per lane moved FP32 levels in 4-byte
// words, which stream at about 3.8 TB/s against 6 TB/s for 8-byte words).
// Each cell's
// value is formed exactly as in k_prolong_z.  t tiles (nx / 2) x ny columns.
template <class Tc, class Tf>
__global__ void k_prolong_z2(Tiles t, Lv<Tc> Cl, Lv<Tf> F, int mask, int per) {
  using C = typename Compute<Tf>::type;
  struct alignas(2 * sizeof(Tf)) Pair {
    Tf v[2];
  };
  const int lb = xcd_block32(blockIdx.x, gridDim.x);
  const int gxy = t.gx * t.gy;
  const int zb = lb / gxy;
  int ip, j, k_unused;
  tile_xyz(t, lb - zb * gxy, ip, j, k_unused);
  if (ip >= t.nx || j >= t.ny) return;
  const int k0 = zb * t.per;
  const int k1 = min(t.nz, k0 + t.per);
  int yl, yr, yw;
  interp(j, (mask >> 1) & 1, Cl.ny, (per >> 1) & 1, &yl, &yr, &yw);
  const int cy[2] = {wrapi(yl, Cl.ny), wrapi(yr, Cl.ny)};
  const int ny2 = yl == yr ? 1 : 2;
  int xw[2], nx2[2], off[2][4];
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    int xl, xr;
    interp(2 * ip + e, mask & 1, Cl.nx, per & 1, &xl, &xr, &xw[e]);
    const int cx[2] = {wrapi(xl, Cl.nx), wrapi(xr, Cl.nx)};
    nx2[e] = xl == xr ? 1 : 2;
#pragma unroll
    for (int ay = 0; ay < 2; ++ay)
#pragma unroll
      for (int ax = 0; ax < 2; ++ax)
        off[e][ay * 2 + ax] = cy[ay] * Cl.nx + cx[ax];
  }
  const int cplane = Cl.nx * Cl.ny;
  int held0 = -1, held1 = -1;
  C h0[2][4], h1[2][4];
#pragma unroll
  for (int e = 0; e < 2; ++e)
#pragma unroll
    for (int q = 0; q < 4; ++q) h0[e][q] = h1[e][q] = C(0);
  const int plane = F.nx * F.ny;
  int c = cell_index(F.nx, F.ny, 2 * ip, j, k0);
  for (int k = k0; k < k1; ++k, c += plane) {
    int zl, zr, zw;
    interp(k, (mask >> 2) & 1, Cl.nz, (per >> 2) & 1, &zl, &zr, &zw);
    const int gz[2] = {wrapi(zl, Cl.nz), wrapi(zr, Cl.nz)};
    const uint16_t flags = *reinterpret_cast<const uint16_t*>(F.act + c);
    const Pair xp = *reinterpret_cast<const Pair*>(F.x + c);
    C v[2][2][4];  // [az][cell][corner]
#pragma unroll
    for (int az = 0; az < 2; ++az) {
      const int g = gz[az];
#pragma unroll
      for (int e = 0; e < 2; ++e)
#pragma unroll
        for (int q = 0; q < 4; ++q)
          v[az][e][q] = g == held0   ? h0[e][q]
                        : g == held1 ? h1[e][q]
                                     : C(ld(Cl.x[g * cplane + off[e][q]]));
    }
    held0 = gz[0];
    held1 = gz[1];
#pragma unroll
    for (int e = 0; e < 2; ++e)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        h0[e][q] = v[0][e][q];
        h1[e][q] = v[1][e][q];
      }
    const int nz2 = zl == zr ? 1 : 2;
    Pair out;
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      // the reference order: az, ay, ax; value += ((wx*wy)*wz)*x_c
      C value = C(0);
#pragma unroll
      for (int az = 0; az < 2; ++az) {
        const C wz = az == 0 ? C(1) - C(zw) * C(0.25) : C(zw) * C(0.25);
#pragma unroll
        for (int ay = 0; ay < 2; ++ay) {
          const C wy = ay == 0 ? C(1) - C(yw) * C(0.25) : C(yw) * C(0.25);
#pragma unroll
          for (int ax = 0; ax < 2; ++ax) {
            const C wx =
                ax == 0 ? C(1) - C(xw[e]) * C(0.25) : C(xw[e]) * C(0.25);
            if (az < nz2 && ay < ny2 && ax < nx2[e])
              value += wx * wy * wz * v[az][e][ay * 2 + ax];
          }
        }
      }
      const bool fluid = ((flags >> (8 * e)) & 0xffu) != 0;
      out.v[e] = fluid ? st<Tf>(C(ld(xp.v[e])) + value) : st<Tf>(0.0);
    }
    *reinterpret_cast<Pair*>(F.x + c) = out;
  }
}

// apply_impl exit (src/solver_mg.cpp:4954-4976, :5082-5118): the candidate's
// constant projection (mean != null), scale alpha (null = unit) and
// correction = fluid ? (x - mean) * alpha : 0, fused with the Krylov inner
// products r.z (and z.q for the flexible beta).
template <class T>
__global__ void k_mg_exit(Tiles t, const T* x, const uint8_t* act,
                          const double* mean, const double* alpha, double* z,
                          const double* r, const double* q, double* partials,
                          int stride) {
  __shared__ double scratch[16];
  const int lb = xcd_block32(blockIdx.x, gridDim.x);
  const double a = alpha == nullptr ? 1.0 : *alpha;
  const double m = mean == nullptr ? 0.0 : *mean;
  T xs[kRedPer];
  double rv[kRedPer], qv[kRedPer];
  bool ok[kRedPer], fluid[kRedPer];
  int cc[kRedPer];
#pragma unroll
  for (int u = 0; u < kRedPer; ++u) {
    int i, j, k;
    cc[u] = red_cell(t, lb, u, i, j, k, ok[u]);
    fluid[u] = act[cc[u]] != 0;
    xs[u] = x[cc[u]];
    rv[u] = r != nullptr ? r[cc[u]] : 0.0;
    qv[u] = q != nullptr ? q[cc[u]] : 0.0;
  }
  double rz = 0.0, zq = 0.0;
#pragma unroll
  for (int u = 0; u < kRedPer; ++u) {
    if (!ok[u]) continue;
    double v = 0.0;
    if (fluid[u]) {
      double xv = static_cast<double>(ld(xs[u]));
      if (mean != nullptr) xv = static_cast<double>(ld(st<T>(xv - m)));
      v = xv * a;
    }
    if (z != nullptr) z[cc[u]] = v;
    if (r != nullptr) rz += rv[u] * v;
    if (q != nullptr) zq += v * qv[u];
  }
  if (partials != nullptr) {
    rz = block_sum(rz, scratch);
    if (threadIdx.x == 0) partials[lb] = rz;
    if (q != nullptr) {
      zq = block_sum(zq, scratch);
      if (threadIdx.x == 0) partials[stride + lb] = zq;
    }
  }
}

// p = 1.0*z + beta*p (axpby, src/solver_krylov.cpp:321-339) with z formed
// from the cycle's level-0 candidate as
// k_mg_exit forms it (fluid ? (x - mean, stored and reloaded) * alpha : 0),
// so the exit need not store z.  With sol: first the iteration's solution
// update sol += step p (add_scaled, src/solver_krylov.cpp:383-399) with the
// p it replaces, which k_update_xr_entry left out.  Tiles of the reduction
// layout.
template <class T>
__global__ void k_update_p_exit(Tiles t, const T* x, const uint8_t* act,
                                const double* mean, const double* alpha,
                                double* p, const double* beta, double* sol,
                                const double* step) {
  const int lb = xcd_block32(blockIdx.x, gridDim.x);
  const double a = alpha == nullptr ? 1.0 : *alpha;
  const double m = mean == nullptr ? 0.0 : *mean;
  const double be = *beta;
  const double sa = sol != nullptr ? *step : 0.0;
  T xs[kRedPer];
  double pv[kRedPer], sv[kRedPer];
  bool ok[kRedPer], fluid[kRedPer];
  int cc[kRedPer];
#pragma unroll
  for (int u = 0; u < kRedPer; ++u) {
    int i, j, k;
    cc[u] = red_cell(t, lb, u, i, j, k, ok[u]);
    fluid[u] = act[cc[u]] != 0;
    xs[u] = x[cc[u]];
    pv[u] = p[cc[u]];
    sv[u] = sol != nullptr ? sol[cc[u]] : 0.0;
  }
#pragma unroll
  for (int u = 0; u < kRedPer; ++u) {
    if (!ok[u]) continue;
    if (sol != nullptr) sol[cc[u]] = sv[u] + sa * pv[u];
    double v = 0.0;
    if (fluid[u]) {
      double xv = static_cast<double>(ld(xs[u]));
      if (mean != nullptr) xv = static_cast<double>(ld(st<T>(xv - m)));
      v = xv * a;
    }
    p[cc[u]] = 1.0 * v + be * pv[u];
  }
}

// Three dots of the residual-minimising scaling (src/solver_mg.cpp:
// 4988-5001): r = masked input, az = r - (b - A x).
template <class T>
__global__ void k_mg_proj3(Tiles t, const double* rin, const uint8_t* act,
                           const T* rec, double* partials, int stride) {
  __shared__ double scratch[16];
  const int lb = xcd_block32(blockIdx.x, gridDim.x);
  double rv[kRedPer];
  T recv[kRedPer];
  bool ok[kRedPer], fluid[kRedPer];
#pragma unroll
  for (int u = 0; u < kRedPer; ++u) {
    int i, j, k;
    const int c = red_cell(t, lb, u, i, j, k, ok[u]);
    fluid[u] = act[c] != 0;
    rv[u] = rin[c];
    recv[u] = rec[c];
  }
  double p0 = 0.0, p1 = 0.0, p2 = 0.0;
#pragma unroll
  for (int u = 0; u < kRedPer; ++u) {
    if (!ok[u]) continue;
    const double r = fluid[u] ? rv[u] : 0.0;
    const double az = r - static_cast<double>(ld(recv[u]));
    p0 += r * r;
    p1 += r * az;
    p2 += az * az;
  }
  p0 = block_sum(p0, scratch);
  p1 = block_sum(p1, scratch);
  p2 = block_sum(p2, scratch);
  if (threadIdx.x == 0) {
    partials[lb] = p0;
    partials[stride + lb] = p1;
    partials[2 * stride + lb] = p2;
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
  Lv<double> L;      // coefficients in global memory; x/b/r live in LDS
  int coarsen_mask;
  int fused;
  int cheb;          // Chebyshev-Jacobi smoothing (else red-black)
  double volume;
  int lds;           // offset (doubles) of this level's x, b, r in LDS
};

constexpr int kTailLdsBytes = 64 * 1024;

// Chebyshev-Jacobi stage factors of the tail's Chebyshev levels, stage s:
// direction d = pf[s] d + rf[s] (b - A x) / D (d = rf[0] (b - A x) / D at
// s = 0).  The host fills them with the recurrence of the grid levels.
constexpr int kTailChebStages = 16;
struct TailCheb {
  double rf[kTailChebStages];
  double pf[kTailChebStages];
};

__device__ __forceinline__ void tail_ijk(const Lv<double>& L, int c, int& i,
                                         int& j, int& k) {
  const int plane = L.nx * L.ny;
  k = c / plane;
  const int r = c - k * plane;
  j = r / L.nx;
  i = r - j * L.nx;
}

// Level view with the work slots redirected into LDS.
__device__ __forceinline__ Lv<double> tail_view(const TailLevel& T,
                                                double* lds) {
  Lv<double> L = T.L;
  L.x = lds + T.lds;
  L.b = L.x + L.cells;
  L.r = L.b + L.cells;
  return L;
}

__device__ __forceinline__ void tail_apply(const Lv<double>& L, const Bc& bc,
                                           int defect) {
  for (int c = threadIdx.x; c < L.cells; c += blockDim.x) {
    int i, j, k;
    tail_ijk(L, c, i, j, k);
    const double v = stencil<double>(L, bc, L.x, i, j, k, c, L.x[c]);
    L.r[c] = defect ? L.b[c] - v : v;
  }
  __syncthreads();
}

// Where a tail cell's six neighbours (x-lo, x-hi, y-lo, y-hi, z-lo, z-hi)
// sit in its level's x: the neighbour cell, the cell itself on a Neumann
// face, the wrapped cell on a periodic face, -1 on a Dirichlet face (read
// as +0).  The values are those neighbour_or_boundary gives; set up once
// per smoothing call, they spare every pass the boundary-mode branches.
struct TailNb {
  int n[6];
};

__device__ __forceinline__ TailNb tail_neighbours(const Lv<double>& L,
                                                  const Bc& bc, int i, int j,
                                                  int k, int c) {
  const int nx = L.nx, ny = L.ny, nz = L.nz, plane = nx * ny;
  auto face = [&](bool inside, int next, int mode, int wrapped) {
    if (inside) return next;
    return mode == 0 ? -1 : (mode == 1 ? c : wrapped);
  };
  TailNb nb;
  nb.n[0] = face(i > 0, c - 1, bc.m[0], c + (nx - 1));
  nb.n[1] = face(i + 1 < nx, c + 1, bc.m[1], c - (nx - 1));
  nb.n[2] = face(j > 0, c - nx, bc.m[2], c + (ny - 1) * nx);
  nb.n[3] = face(j + 1 < ny, c + nx, bc.m[3], c - (ny - 1) * nx);
  nb.n[4] = face(k > 0, c - plane, bc.m[4], c + (nz - 1) * plane);
  nb.n[5] = face(k + 1 < nz, c + plane, bc.m[5], c - (nz - 1) * plane);
  return nb;
}

// stencil_with on a tail cell through its neighbour positions: the same
// operands in the same order.
__device__ __forceinline__ double tail_stencil(const Lv<double>& L,
                                               const TailNb& nb, int c,
                                               double xc,
                                               const StencilCoef<double>& f) {
  double v6[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const int at = nb.n[q];
    const double v = L.x[at < 0 ? c : at];
    v6[q] = at < 0 ? 0.0 : v;
  }
  double v = f.d * xc;
  v -= f.xl * v6[0];
  v -= f.xh * v6[1];
  v -= f.yl * v6[2];
  v -= f.yh * v6[3];
  v -= f.zl * v6[4];
  v -= f.zh * v6[5];
  return v;
}

// tail_smooth on a level of at most one cell per thread: the cell's
// coefficients are loaded once and kept in registers across the passes (the
// same values and arithmetic as the general form below).
__device__ void tail_smooth_small(const Lv<double>& L, int fused, const Bc& bc,
                                  int sweeps, int reverse, int retain) {
  const double omega = 0.72;
  const int c = threadIdx.x;
  const bool mine = c < L.cells;
  int i = 0, j = 0, k = 0;
  StencilCoef<double> f{};
  TailNb nb{};
  if (mine) {
    tail_ijk(L, c, i, j, k);
    f = stencil_coef<double>(L, j, k, c);
    nb = tail_neighbours(L, bc, i, j, k, c);
  }
  auto apply = [&](int defect) {
    if (mine) {
      const double v = tail_stencil(L, nb, c, L.x[c], f);
      L.r[c] = defect ? L.b[c] - v : v;
    }
    __syncthreads();
  };
  if (!fused) apply(0);
  for (int sweep = 0; sweep < sweeps; ++sweep) {
    const int first = reverse ? 1 : 0;
    for (int pass = 0; pass < 2; ++pass) {
      const int color = (first + pass) & 1;
      if (mine && ((i + j + k) & 1) == color) {
        const double xc = L.x[c];
        const double v =
            fused ? tail_stencil(L, nb, c, xc, f) : L.r[c];
        L.x[c] = xc + omega * (L.b[c] - v) / f.d;
      }
      __syncthreads();
      const bool final_update = sweep + 1 == sweeps && pass == 1;
      if (!final_update || retain) {
        if (!fused || (final_update && retain)) apply(retain && final_update);
      }
    }
  }
}

__device__ void tail_smooth(const Lv<double>& L, int fused, const Bc& bc,
                            int sweeps, int reverse, int retain) {
  if (L.cells <= static_cast<int>(blockDim.x)) {
    tail_smooth_small(L, fused, bc, sweeps, reverse, retain);
    return;
  }
  const double omega = 0.72;
  if (!fused) tail_apply(L, bc, 0);
  for (int sweep = 0; sweep < sweeps; ++sweep) {
    const int first = reverse ? 1 : 0;
    for (int pass = 0; pass < 2; ++pass) {
      const int color = (first + pass) & 1;
      for (int c = threadIdx.x; c < L.cells; c += blockDim.x) {
        int i, j, k;
        tail_ijk(L, c, i, j, k);
        if (((i + j + k) & 1) != color) continue;
        const double xc = L.x[c];
        const double v =
            fused ? stencil<double>(L, bc, L.x, i, j, k, c, xc) : L.r[c];
        L.x[c] = xc + omega * (L.b[c] - v) / L.D[c];
      }
      __syncthreads();
      const bool final_update = sweep + 1 == sweeps && pass == 1;
      if (!final_update || retain) {
        if (!fused || (final_update && retain))
          tail_apply(L, bc, retain && final_update);
      }
    }
  }
}

// Chebyshev-Jacobi smoothing of a tail level (src/solver_mg.cpp:2916-3090):
// the stages of k_cheb_stage with x updated in place.  A stage stores its
// direction in r, and x += r follows a barrier, so every neighbour the
// stencil reads is the stage's input x (Jacobi), the operands and order of
// operations being those of the grid kernel.  retain: then r = b - A x.
__device__ void tail_cheb(const Lv<double>& L, const Bc& bc, int stages,
                          int retain, const TailCheb& ch) {
  if (L.cells <= static_cast<int>(blockDim.x)) {
    // at most one cell per thread: its coefficients stay in registers
    const int c = threadIdx.x;
    const bool mine = c < L.cells;
    int i = 0, j = 0, k = 0;
    StencilCoef<double> f{};
    TailNb nb{};
    if (mine) {
      tail_ijk(L, c, i, j, k);
      f = stencil_coef<double>(L, j, k, c);
      nb = tail_neighbours(L, bc, i, j, k, c);
    }
    for (int s = 0; s < stages; ++s) {
      double xn = 0.0;
      if (mine) {
        const double xc = L.x[c];
        const double v = tail_stencil(L, nb, c, xc, f);
        const double defect = L.b[c] - v;
        const double scaled = ch.rf[s] * defect / f.d;
        const double dir = s == 0 ? scaled : ch.pf[s] * L.r[c] + scaled;
        L.r[c] = dir;
        xn = xc + dir;
      }
      __syncthreads();
      if (mine) L.x[c] = xn;
      __syncthreads();
    }
    if (retain) {
      if (mine)
        L.r[c] = L.b[c] - tail_stencil(L, nb, c, L.x[c], f);
      __syncthreads();
    }
    return;
  }
  for (int s = 0; s < stages; ++s) {
    for (int c = threadIdx.x; c < L.cells; c += blockDim.x) {
      int i, j, k;
      tail_ijk(L, c, i, j, k);
      const double xc = L.x[c];
      const double v = stencil<double>(L, bc, L.x, i, j, k, c, xc);
      const double defect = L.b[c] - v;
      const double scaled = ch.rf[s] * defect / L.D[c];
      L.r[c] = s == 0 ? scaled : ch.pf[s] * L.r[c] + scaled;
    }
    __syncthreads();
    for (int c = threadIdx.x; c < L.cells; c += blockDim.x)
      L.x[c] = L.x[c] + L.r[c];
    __syncthreads();
  }
  if (retain) tail_apply(L, bc, 1);
}

__device__ void tail_project(const Lv<double>& L, double volume, double* field,
                             double* scratch, double* mean) {
  double v = 0.0;
  for (int c = threadIdx.x; c < L.cells; c += blockDim.x)
    v += L.vol[c] * field[c];
  v = block_sum(v, scratch);
  if (threadIdx.x == 0) *mean = v / volume;
  __syncthreads();
  const double m = *mean;
  for (int c = threadIdx.x; c < L.cells; c += blockDim.x)
    field[c] = L.vol[c] > 0.0 ? field[c] - m : 0.0;
  __syncthreads();
}

// The tail's first level receives b (already projected) and x = 0 from the
// grid-level restriction; its x is stored back for the grid prolongation.
__global__ void __launch_bounds__(1024)
k_tail(const TailLevel* levels, int first_level, const TailOp* ops, int nops,
       Bc bc, int periodic_mask, int null_space, TailCheb cheb) {
  extern __shared__ double lds[];
  __shared__ double scratch[16];
  __shared__ double mean;
  {
    const TailLevel& T0 = levels[0];
    const Lv<double> L0 = tail_view(T0, lds);
    for (int c = threadIdx.x; c < L0.cells; c += blockDim.x) {
      L0.x[c] = T0.L.x[c];
      L0.b[c] = T0.L.b[c];
    }
    __syncthreads();
  }
  for (int o = 0; o < nops; ++o) {
    const TailOp op = ops[o];
    const TailLevel& T = levels[op.level - first_level];
    const Lv<double> L = tail_view(T, lds);
    if (op.kind == T_SMOOTH) {
      if (T.cheb)
        tail_cheb(L, bc, op.sweeps, op.retain, cheb);
      else
        tail_smooth(L, T.fused, bc, op.sweeps, op.reverse, op.retain);
    } else if (op.kind == T_RESTRICT) {
      const TailLevel& Cn = levels[op.level + 1 - first_level];
      const Lv<double> Cl = tail_view(Cn, lds);
      const int mask = T.coarsen_mask;
      const int mx = mask & 1, my = (mask >> 1) & 1, mz = (mask >> 2) & 1;
      for (int t = threadIdx.x; t < Cl.cells; t += blockDim.x) {
        int ci, cj, ck;
        tail_ijk(Cl, t, ci, cj, ck);
        const int xb = mx ? 2 * ci : ci, yb = my ? 2 * cj : cj,
                  zb = mz ? 2 * ck : ck;
        double sum = 0.0;
        for (int dz = 0; dz < (mz ? 2 : 1); ++dz)
          for (int dy = 0; dy < (my ? 2 : 1); ++dy)
            for (int dx = 0; dx < (mx ? 2 : 1); ++dx) {
              const int fx = xb + dx, fy = yb + dy, fz = zb + dz;
              if (fx >= L.nx || fy >= L.ny || fz >= L.nz) continue;
              sum += L.r[cell_index(L.nx, L.ny, fx, fy, fz)];
            }
        Cl.b[t] = sum;
        Cl.x[t] = 0.0;
      }
      __syncthreads();
      if (null_space) tail_project(Cl, Cn.volume, Cl.b, scratch, &mean);
    } else {
      const TailLevel& Cn = levels[op.level + 1 - first_level];
      const Lv<double> Cl = tail_view(Cn, lds);
      for (int c = threadIdx.x; c < L.cells; c += blockDim.x) {
        int i, j, k;
        tail_ijk(L, c, i, j, k);
        prolong_cell(Cl, L, T.coarsen_mask, periodic_mask, i, j, k, c);
      }
      __syncthreads();
    }
  }
  {
    const TailLevel& T0 = levels[0];
    const Lv<double> L0 = tail_view(T0, lds);
    for (int c = threadIdx.x; c < L0.cells; c += blockDim.x)
      T0.L.x[c] = L0.x[c];
  }
}

// ---------------------------------------------------------------------------
// z-marching stencil kernels.  A block owns one 2-D tile of the plane (the
// Tiles layout with k = 0) and t.per consecutive planes of it; each thread
// walks its cell column upwards, carrying x at k-1 and k and the z-face
// coefficient between steps, so a step loads x and the coefficients of one
// new plane and the in-plane neighbours (cache hits from the tile's rows).
// Operand values and the order of operations are those of stencil():
//   kZOpDot    q = A p and the partial sum of p.q (Krylov operator)
//   kZResidual r = 1 b + (-1) A x and the partial sum of r.r (the true
//              residual, axpby(r, 1, b, -1, Ax): src/solver_krylov.cpp:
//              1082-1091); aux > 0: also the partial sum of b.b, aux
//              partials after those of r.r
//   kZApply    r = b - A x (aux != 0, defect) or r = A x at level storage
// ---------------------------------------------------------------------------
constexpr int kZOpDot = 0, kZResidual = 1, kZApply = 2;

template <class T, int kMode>
__global__ void k_stencil_z(Tiles t, Lv<T> L, Bc bc, const T* x, const T* b,
                            T* out, double* partials, int aux) {
  using C = typename Compute<T>::type;
  __shared__ double scratch[16];
  const int lb = xcd_block32(blockIdx.x, gridDim.x);
  const int gxy = t.gx * t.gy;
  const int zb = lb / gxy;
  int i, j, k_unused;
  tile_xyz(t, lb - zb * gxy, i, j, k_unused);
  const int nx = t.nx, ny = t.ny, nz = t.nz;
  const int k0 = zb * t.per;
  const int k1 = min(nz, k0 + t.per);
  double d = 0.0, bb = 0.0;
  if (i < nx && j < ny) {
    const int plane = nx * ny;
    int c = cell_index(nx, ny, i, j, k0);
    C xc = C(ld(x[c]));
    C zl = k0 > 0 ? C(ld(x[c - plane]))
                  : neighbour_or_boundary<C>(bc.m[4], xc, x,
                                             c + (nz - 1) * plane);
    C fzl = C(ld(L.fz[c]));
    for (int k = k0; k < k1; ++k) {
      const C zh = k + 1 < nz
                       ? C(ld(x[c + plane]))
                       : neighbour_or_boundary<C>(bc.m[5], xc, x,
                                                  c - (nz - 1) * plane);
      const C fzh = C(ld(L.fz[c + plane]));
      const int row = j + ny * k;
      const C xl = i > 0 ? C(ld(x[c - 1]))
                         : neighbour_or_boundary<C>(bc.m[0], xc, x,
                                                    c + (nx - 1));
      const C xh = i + 1 < nx
                       ? C(ld(x[c + 1]))
                       : neighbour_or_boundary<C>(bc.m[1], xc, x,
                                                  c - (nx - 1));
      const C yl = j > 0 ? C(ld(x[c - nx]))
                         : neighbour_or_boundary<C>(bc.m[2], xc, x,
                                                    c + (ny - 1) * nx);
      const C yh = j + 1 < ny
                       ? C(ld(x[c + nx]))
                       : neighbour_or_boundary<C>(bc.m[3], xc, x,
                                                  c - (ny - 1) * nx);
      const int fyo = c + nx * k;
      C v = C(ld(L.D[c])) * xc;
      v -= C(ld(L.fx[c + row])) * xl;
      v -= C(ld(L.fx[c + row + 1])) * xh;
      v -= C(ld(L.fy[fyo])) * yl;
      v -= C(ld(L.fy[fyo + nx])) * yh;
      v -= fzl * zl;
      v -= fzh * zh;
      if constexpr (kMode == kZOpDot) {
        out[c] = v;
        d += xc * v;
      } else if constexpr (kMode == kZResidual) {
        const double bv = b[c];
        const double rc = 1.0 * bv + -1.0 * v;
        out[c] = rc;
        d += rc * rc;
        bb += bv * bv;
      } else {
        out[c] = aux ? st<T>(C(ld(b[c])) - v) : st<T>(v);
      }
      zl = xc;
      xc = zh;
      fzl = fzh;
      c += plane;
    }
  }
  if (kMode != kZApply && partials != nullptr) {
    d = block_sum(d, scratch);
    if (threadIdx.x == 0) partials[lb] = d;
    if (kMode == kZResidual && aux > 0) {
      bb = block_sum(bb, scratch);
      if (threadIdx.x == 0) partials[lb + aux] = bb;
    }
  }
}

// The true residual of k_stencil_z<double, kZResidual> (same operands and
// order) fused with the preconditioner's entry of k_mg_entry, which the
// restart after an unconverged check runs on that residual: b_mg = fluid ?
// r : 0 at the level-0 storage precision.  Accumulators `stride` apart:
// r.r, then (proj) vol * b_mg over the fluid cells, then (with_bb) b.b.
template <class Te>
__global__ void k_residual_entry(Tiles t, Lv<double> L, Bc bc, const double* x,
                                 const double* rhs, double* r,
                                 const uint8_t* act, Widths w, Te* bmg,
                                 double* partials, int stride, int proj,
                                 int with_bb) {
  __shared__ double scratch[16];
  const int lb = xcd_block32(blockIdx.x, gridDim.x);
  const int gxy = t.gx * t.gy;
  const int zb = lb / gxy;
  int i, j, k_unused;
  tile_xyz(t, lb - zb * gxy, i, j, k_unused);
  const int nx = t.nx, ny = t.ny, nz = t.nz;
  const int k0 = zb * t.per;
  const int k1 = min(nz, k0 + t.per);
  double d = 0.0, m = 0.0, bb = 0.0;
  if (i < nx && j < ny) {
    const int plane = nx * ny;
    const double wxy = w.x[i] * w.y[j];
    int c = cell_index(nx, ny, i, j, k0);
    double xc = x[c];
    double zl = k0 > 0 ? x[c - plane]
                       : neighbour_or_boundary<double>(bc.m[4], xc, x,
                                                       c + (nz - 1) * plane);
    double fzl = L.fz[c];
    for (int k = k0; k < k1; ++k) {
      const double zh = k + 1 < nz
                            ? x[c + plane]
                            : neighbour_or_boundary<double>(bc.m[5], xc, x,
                                                            c - (nz - 1) * plane);
      const double fzh = L.fz[c + plane];
      const int row = j + ny * k;
      const double xl = i > 0 ? x[c - 1]
                              : neighbour_or_boundary<double>(bc.m[0], xc, x,
                                                              c + (nx - 1));
      const double xh = i + 1 < nx
                            ? x[c + 1]
                            : neighbour_or_boundary<double>(bc.m[1], xc, x,
                                                            c - (nx - 1));
      const double yl = j > 0 ? x[c - nx]
                              : neighbour_or_boundary<double>(bc.m[2], xc, x,
                                                              c + (ny - 1) * nx);
      const double yh = j + 1 < ny
                            ? x[c + nx]
                            : neighbour_or_boundary<double>(bc.m[3], xc, x,
                                                            c - (ny - 1) * nx);
      const int fyo = c + nx * k;
      double v = L.D[c] * xc;
      v -= L.fx[c + row] * xl;
      v -= L.fx[c + row + 1] * xh;
      v -= L.fy[fyo] * yl;
      v -= L.fy[fyo + nx] * yh;
      v -= fzl * zl;
      v -= fzh * zh;
      const double bv = rhs[c];
      const double rc = 1.0 * bv + -1.0 * v;
      r[c] = rc;
      d += rc * rc;
      bb += bv * bv;
      const bool fluid = act[c] != 0;
      const Te be = fluid ? st<Te>(rc) : st<Te>(0.0);
      bmg[c] = be;
      if (fluid) m += wxy * w.z[k] * static_cast<double>(ld(be));
      zl = xc;
      xc = zh;
      fzl = fzh;
      c += plane;
    }
  }
  d = block_sum(d, scratch);
  if (threadIdx.x == 0) partials[lb] = d;
  if (proj) {
    m = block_sum(m, scratch);
    if (threadIdx.x == 0) partials[lb + stride] = m;
  }
  if (with_bb) {
    bb = block_sum(bb, scratch);
    if (threadIdx.x == 0) partials[lb + 2 * stride] = bb;
  }
}


}  // namespace hf

namespace hf {

// ---------------------------------------------------------------------------
// Fused red-black sweep (both colour passes of src/solver_mg.cpp:3149-3222 in
// one launch).  A block owns a 128 x 8 tile of a z-chunk and marches through
// its planes with five LDS plane buffers holding x with a two-cell ring.  At
// step k the first colour is updated on plane k over the tile plus a one-cell
// ring (recomputing what the neighbouring tiles own, from the same inputs),
// then the second colour on plane k-1 over the tile, which then has every
// first-colour neighbour it needs.  x is read from L.x and written to xout
// (out of place), so tiles never see each other's updates; every update uses
// the reference arithmetic, so the result equals two colour passes bit for
// bit.  With `from_zero` the input x is zero (no loads) and a pending
// constant projection (`mean`) of b is applied on the fly and stored to bout.
//
// Boundaries live in the LDS halo: a position outside the domain holds 0
// (Dirichlet), the wrapped cell (periodic) or the mirrored boundary cell as
// loaded (Neumann: the reference reads the cell's own value, which is still
// its value at sweep start when the cell is updated, each cell being updated
// once per sweep), so every update reads its six neighbours unconditionally.
//
// Work is owned by cell pairs: lane l of a wave owns cells x0 + 2l and
// x0 + 2l + 1 of a row, one of each colour.  At step k it loads the pair's
// coefficients on plane k with one two-element load per array (whole cache
// lines per instruction), updates the first-colour cell at once and keeps the
// second-colour cell's coefficients in registers for step k + 1, where that
// cell is updated.  Wave w of W owns tile rows w and w + W; waves 0 and 1 also
// own the ring rows above and below the tile (first colour only) and wave 2
// the ring cells left and right of it (one per tile row and plane, the side
// alternating with the parity).  Each load is a raw buffer load: scalar resource, 32-bit
// lane offset within the plane, scalar plane offset.  Rows hold an even
// number of cells (the host plans this kernel for even nx only).
// ---------------------------------------------------------------------------
constexpr int kSwTX = 128;
constexpr int kSwTY = 8;
constexpr int kSwThreads = 256;  // threads per block of the fused sweep
constexpr int kSwPX = kSwTX + 4;
constexpr int kSwPY = kSwTY + 4;
constexpr int kSwPlane = kSwPX * kSwPY;
// LDS planes k - 3 .. k + 1: the x load of step k + 1 overwrites plane k - 3,
// which no thread still working on step k reads (its reads are planes
// k - 2 .. k + 1), so one barrier per step suffices
constexpr int kSwPlanes = 5;
__device__ __forceinline__ int sw_slot(int kk) {  // kk >= -2
  return (kk + 2 * kSwPlanes) % kSwPlanes;
}
constexpr int kSwRingRow = kSwTX / 2 + 1;  // first-colour cells per ring row

template <class T>
struct SweepArgs {
  Lv<T> L;
  Bc bc;
  T* xout;
  T* bout;             // projected b (from_zero with mean), else null
  const double* mean;  // pending projection mean, else null
  int first;           // colour updated first: 0 forward, 1 reverse
  int from_zero;
  int zc;              // planes per block
  int gx, gy;          // tiles along x and y
  // with kMean: per-block partial sums of vol * x over fluid cells of the
  // output (the constant projection's mean), cell widths per axis
  double* partials;
  const double* wx;
  const double* wy;
  const double* wz;
};

// A value updated in LDS rounded to the storage type, as if it had been
// stored and loaded again (the two colour passes read the first colour's
// values from storage); exact for 32- and 64-bit storage.
template <class T, class C>
__device__ __forceinline__ C sw_round(C v) {
  return C(ld(st<T>(v)));
}

// Coefficients of one cell update, fetched ahead of the LDS phase.
template <class C>
struct SwCoef {
  C d, xl, xh, yl, yh, zl, zh, b;
};

// LDS positions of the x plane per thread.
constexpr int kSwLoadSlots = (kSwPlane + kSwThreads - 1) / kSwThreads;

// one pair per lane covers a tile row; each wave owns two tile rows, and the
// ring column cells of the tile rows fit one wave
constexpr int kSwWaves = kSwThreads / 64;
static_assert(kSwTX == 128 && kSwThreads == 32 * kSwTY && kSwTY <= 64 &&
                  kSwWaves >= 3,
              "fused sweep pair mapping");

// Source index along one axis of a halo position g (-2 <= g <= n + 1), or -1
// where the position holds zero.
__device__ __forceinline__ int sw_source(int g, int n, int mlo, int mhi) {
  if (g >= 0 && g < n) return g;
  if (g < -2 || g > n + 1) return -1;  // never read
  if (g < 0) {
    if (mlo == 2) return g + n;
    return (g == -1 && mlo != 0) ? 0 : -1;
  }
  if (mhi == 2) return g - n;
  return (g == n && mhi != 0) ? n - 1 : -1;
}

// Raw buffer access: a scalar resource per array, a 32-bit per-lane byte
// offset within a plane and a scalar byte offset of the plane, so a load
// carries no 64-bit address arithmetic.  The host plans the fused sweep only
// where every array of the level spans less than 4 GiB.
using SwRsrc = __amdgpu_buffer_rsrc_t;
constexpr int kBufferDword3 = 0x00020000;  // gfx9 raw buffer descriptor

template <class T>
__device__ __forceinline__ SwRsrc sw_rsrc(const T* p, uint64_t count) {
  const uint64_t bytes = count * sizeof(T);
  return __builtin_amdgcn_make_buffer_rsrc(
      const_cast<T*>(p), 0,
      static_cast<int>(bytes > 0xffffffffull ? 0xffffffffu
                                             : static_cast<uint32_t>(bytes)),
      kBufferDword3);
}

template <class T>
__device__ __forceinline__ T sw_load(SwRsrc r, uint32_t voff, uint32_t soff);
template <>
__device__ __forceinline__ float sw_load<float>(SwRsrc r, uint32_t voff,
                                                uint32_t soff) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(
      r, static_cast<int>(voff), static_cast<int>(soff), 0));
}
template <>
__device__ __forceinline__ double sw_load<double>(SwRsrc r, uint32_t voff,
                                                  uint32_t soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(
                                        r, static_cast<int>(voff),
                                        static_cast<int>(soff), 0));
}
template <>
__device__ __forceinline__ bf16 sw_load<bf16>(SwRsrc r, uint32_t voff,
                                              uint32_t soff) {
  bf16 v;
  v.bits = __builtin_amdgcn_raw_buffer_load_b16(r, static_cast<int>(voff),
                                                static_cast<int>(soff), 0);
  return v;
}

// Two consecutive elements (the first at an even element offset).
template <class T>
struct SwPair {
  T lo, hi;
};
template <class T>
__device__ __forceinline__ SwPair<T> sw_load2(SwRsrc r, uint32_t voff,
                                              uint32_t soff);
template <>
__device__ __forceinline__ SwPair<float> sw_load2<float>(SwRsrc r,
                                                         uint32_t voff,
                                                         uint32_t soff) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b64(
      r, static_cast<int>(voff), static_cast<int>(soff), 0);
  return {__uint_as_float(v[0]), __uint_as_float(v[1])};
}
template <>
__device__ __forceinline__ SwPair<double> sw_load2<double>(SwRsrc r,
                                                           uint32_t voff,
                                                           uint32_t soff) {
  return __builtin_bit_cast(
      SwPair<double>, __builtin_amdgcn_raw_buffer_load_b128(
                          r, static_cast<int>(voff), static_cast<int>(soff), 0));
}
template <>
__device__ __forceinline__ SwPair<bf16> sw_load2<bf16>(SwRsrc r,
                                                       uint32_t voff,
                                                       uint32_t soff) {
  const uint32_t v = __builtin_amdgcn_raw_buffer_load_b32(
      r, static_cast<int>(voff), static_cast<int>(soff), 0);
  SwPair<bf16> p;
  p.lo.bits = static_cast<uint16_t>(v & 0xffffu);
  p.hi.bits = static_cast<uint16_t>(v >> 16);
  return p;
}

__device__ __forceinline__ void sw_store2(SwRsrc r, uint32_t voff,
                                          uint32_t soff, float lo, float hi) {
  using u2 = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
  const u2 v = {__float_as_uint(lo), __float_as_uint(hi)};
  __builtin_amdgcn_raw_buffer_store_b64(v, r, static_cast<int>(voff),
                                        static_cast<int>(soff), 0);
}
__device__ __forceinline__ void sw_store2(SwRsrc r, uint32_t voff,
                                          uint32_t soff, double lo,
                                          double hi) {
  using u4 = decltype(__builtin_amdgcn_raw_buffer_load_b128(r, 0, 0, 0));
  const SwPair<double> p = {lo, hi};
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u4, p), r,
                                         static_cast<int>(voff),
                                         static_cast<int>(soff), 0);
}
__device__ __forceinline__ void sw_store2(SwRsrc r, uint32_t voff,
                                          uint32_t soff, bf16 lo, bf16 hi) {
  const uint32_t v = static_cast<uint32_t>(lo.bits) |
                     (static_cast<uint32_t>(hi.bits) << 16);
  __builtin_amdgcn_raw_buffer_store_b32(v, r, static_cast<int>(voff),
                                        static_cast<int>(soff), 0);
}

// Coefficients of one cell pair on one plane.
template <class T>
struct SwPairCoef {
  SwPair<T> d, yl, yh, zl, zh, b;
  T fx0, fx1, fx2;  // x faces of the pair: low, middle, high
  uint32_t act;     // fluid flags of the pair (first sweep with projection)
};

// kZero: the first sweep of a smoothing from x = 0 (a.from_zero), which loads
// no x and applies a pending projection of b; otherwise a.from_zero is 0 and
// a.mean is null.  kMean: the block also sums vol * x over the fluid cells
// it stores (the terms of k_vol_dot_finest) into a.partials.
template <class T, bool kZero, bool kMean = false>
__global__ void __launch_bounds__(kSwThreads) k_rb_sweep(SweepArgs<T> a) {
  using C = typename Compute<T>::type;
  __shared__ C buf[kSwPlanes][kSwPlane];
  __shared__ double red[16];
  double msum = 0.0;
  const Lv<T>& L = a.L;
  int blk = xcd_block32(blockIdx.x, gridDim.x);
  // the tile coordinates are block-uniform: held in scalar registers, every
  // per-plane offset derived from them is scalar too (buffer soffsets)
  const int tx = __builtin_amdgcn_readfirstlane(blk % a.gx);
  blk /= a.gx;
  const int ty = __builtin_amdgcn_readfirstlane(blk % a.gy);
  const int tz = __builtin_amdgcn_readfirstlane(blk / a.gy);
  const int x0 = tx * kSwTX, y0 = ty * kSwTY;
  const int z0 = tz * a.zc;
  const int z1 = min(L.nz, z0 + a.zc);
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  const int mx0 = a.bc.m[0], mx1 = a.bc.m[1], my0 = a.bc.m[2],
            my1 = a.bc.m[3], mz0 = a.bc.m[4], mz1 = a.bc.m[5];
  const bool perx = mx0 == 2, pery = my0 == 2, perz = mz0 == 2;
  const uint32_t nxy = static_cast<uint32_t>(nx) * ny;
  constexpr uint32_t kS = sizeof(T);
  const C omega = C(0.72);
  const double mean = a.mean != nullptr ? *a.mean : 0.0;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int c0 = a.first;
  const uint64_t cells = static_cast<uint64_t>(nxy) * nz;
  const SwRsrc rx = sw_rsrc(L.x, cells);
  const SwRsrc rd = sw_rsrc(L.D, cells);
  const SwRsrc rfx = sw_rsrc(L.fx, cells + static_cast<uint64_t>(ny) * nz);
  const SwRsrc rfy = sw_rsrc(L.fy, cells + static_cast<uint64_t>(nx) * nz);
  const SwRsrc rfz = sw_rsrc(L.fz, cells + nxy);
  const SwRsrc rb = sw_rsrc(L.b, cells);
  const SwRsrc rout = sw_rsrc(a.xout, cells);
  const SwRsrc rbout = sw_rsrc(kZero ? a.bout : a.xout, cells);

  // ---- per-block set-up
  // x plane: in-plane source byte offset of every LDS position this thread
  // fills
  uint32_t xo[kSwLoadSlots];
  uint32_t xok = 0;
#pragma unroll
  for (int s = 0; s < kSwLoadSlots; ++s) {
    const int p = tid + s * kSwThreads;
    const int iy = p / kSwPX;
    const int ix = p - iy * kSwPX - 2;
    const int si = sw_source(x0 + ix, nx, mx0, mx1);
    const int sj = sw_source(y0 + iy - 2, ny, my0, my1);
    const bool ok = p < kSwPlane && si >= 0 && sj >= 0;
    xo[s] = ok ? static_cast<uint32_t>(sj * nx + si) * kS : 0u;
    xok |= ok ? (1u << s) : 0u;
  }
  // pair rows of this wave: tile rows wave and wave + kSwWaves, and the ring row
  // above the tile (wave 0) or below its last row inside the domain (wave 1).
  // Per row: tile row index r, unwrapped row gj (parity), whether the row
  // exists (tile rows: inside the domain; ring rows: inside, or wrapped
  // across a periodic face), and the pair's byte offsets (cells, x faces).
  constexpr int kRows = 3;  // two tile rows, one ring row (waves 0 and 1)
  int rr[kRows], rgj[kRows];
  bool rvalid[kRows];
  uint32_t pob[kRows], pfb[kRows];
  const bool lane_in = x0 + 2 * lane < nx;  // nx even: both cells inside
#pragma unroll
  for (int s = 0; s < kRows; ++s) {
    const int r = s == 0   ? wave
                  : s == 1 ? wave + kSwWaves
                           : (wave == 0 ? -1 : min(kSwTY, ny - y0));
    int gj = y0 + r;
    rr[s] = r;
    rgj[s] = gj;
    bool ok;
    if (s < 2) {
      ok = gj < ny;
    } else {
      ok = wave < 2;
      if (gj < 0 || gj >= ny) {
        if (pery) gj = wrapi(gj, ny); else ok = false;
      }
    }
    rvalid[s] = ok;
    const bool lok = ok && lane_in;
    pob[s] = lok ? static_cast<uint32_t>(gj * nx + x0 + 2 * lane) * kS : 0u;
    pfb[s] = lok ? static_cast<uint32_t>(gj * (nx + 1) + x0 + 2 * lane) * kS
                 : 0u;
  }
  // kMean: widths of the pair's cells and of the two tile rows (constant
  // over the block's planes)
  double mwx0 = 0.0, mwx1 = 0.0, mwy[2] = {0.0, 0.0};
  if (kMean) {
    if (lane_in) {
      mwx0 = a.wx[x0 + 2 * lane];
      mwx1 = a.wx[x0 + 2 * lane + 1];
    }
#pragma unroll
    for (int s = 0; s < 2; ++s)
      if (rvalid[s]) mwy[s] = a.wy[rgj[s]];
  }
  // ring columns (wave 2, lane = tile row): byte offsets of the cell
  // left of the tile (sel 0) and right of its last column inside the domain
  // (sel 1, at an even x as nx is even); ring cells beyond a periodic face
  // are the wrapped cells
  uint32_t cob[2], cfb[2];
  uint32_t cok = 0;
  const int crow = lane % kSwTY;
  const int cgj = y0 + crow;
  const int cright = min(x0 + kSwTX, nx);
#pragma unroll
  for (int sel = 0; sel < 2; ++sel) {
    int gi = sel == 0 ? x0 - 1 : cright;
    bool ok = wave == 2 && lane < kSwTY && cgj < ny;
    if (gi < 0 || gi >= nx) {
      if (perx) gi = wrapi(gi, nx); else ok = false;
    }
    cob[sel] = ok ? static_cast<uint32_t>(cgj * nx + gi) * kS : 0u;
    cfb[sel] = ok ? static_cast<uint32_t>(cgj * (nx + 1) + gi) * kS : 0u;
    cok |= ok ? (1u << sel) : 0u;
  }

  auto plane_of = [&](int kk, int& kg) -> bool {
    kg = kk;
    if (kk >= 0 && kk < nz) return true;
    if (!perz) return false;
    kg = kk < 0 ? kk + nz : kk - nz;
    return true;
  };
  // Every global load of one plane step is issued before any of its values
  // is used, so a step waits on memory once rather than once per cell.
  auto fetch_plane = [&](int kk, C (&v)[kSwLoadSlots]) {
    const int kz = kZero ? -1 : sw_source(kk, nz, mz0, mz1);
    if (kz < 0) {
#pragma unroll
      for (int s = 0; s < kSwLoadSlots; ++s) v[s] = C(0);
      return;
    }
    const uint32_t sx = static_cast<uint32_t>(kz) * nxy * kS;
#pragma unroll
    for (int s = 0; s < kSwLoadSlots; ++s) {
      const C x = C(ld(sw_load<T>(rx, xo[s], sx)));
      v[s] = ((xok >> s) & 1u) ? x : C(0);
    }
  };
  auto store_plane = [&](int kk, const C (&v)[kSwLoadSlots]) {
    C* P = buf[sw_slot(kk)];
#pragma unroll
    for (int s = 0; s < kSwLoadSlots; ++s) {
      const int p = tid + s * kSwThreads;
      if (p < kSwPlane) P[p] = v[s];
    }
  };
  // byte offsets of plane kg in the coefficient arrays (cells, x faces,
  // y faces and the high y face, the high z face)
  struct PlaneOff {
    uint32_t c, fx, fy, fyh, fzh;
  };
  auto plane_off = [&](int kg) -> PlaneOff {
    const uint32_t k = static_cast<uint32_t>(kg);
    PlaneOff pl;
    pl.c = k * nxy * kS;
    pl.fx = k * (nxy + ny) * kS;
    pl.fy = k * (nxy + nx) * kS;
    pl.fyh = pl.fy + nx * kS;
    pl.fzh = pl.c + nxy * kS;
    return pl;
  };
  // b at a cell (byte offset ob in plane pl), projected on the first sweep
  auto bval = [&](const PlaneOff& pl, uint32_t ob, T braw) -> C {
    if (!kZero || a.mean == nullptr) return C(ld(braw));
    const T projected = st<T>(static_cast<double>(ld(braw)) - mean);
    return L.act[(pl.c + ob) / kS] ? C(ld(projected)) : C(0);
  };
  auto fetch_pair = [&](const PlaneOff& pl, uint32_t ob,
                        uint32_t fb) -> SwPairCoef<T> {
    SwPairCoef<T> c;
    c.d = sw_load2<T>(rd, ob, pl.c);
    c.fx0 = sw_load<T>(rfx, fb, pl.fx);
    c.fx1 = sw_load<T>(rfx, fb + kS, pl.fx);
    c.fx2 = sw_load<T>(rfx, fb + 2 * kS, pl.fx);
    c.yl = sw_load2<T>(rfy, ob, pl.fy);
    c.yh = sw_load2<T>(rfy, ob, pl.fyh);
    c.zl = sw_load2<T>(rfz, ob, pl.c);
    c.zh = sw_load2<T>(rfz, ob, pl.fzh);
    c.b = sw_load2<T>(rb, ob, pl.c);
    c.act = 0u;
    if ((kZero && a.mean != nullptr) || kMean)
      c.act = *reinterpret_cast<const uint16_t*>(L.act + (pl.c + ob) / kS);
    return c;
  };
  // b of the pair's two cells, projected on the first sweep
  auto pair_b = [&](const SwPairCoef<T>& c, C& blo, C& bhi) {
    if (!kZero || a.mean == nullptr) {
      blo = C(ld(c.b.lo));
      bhi = C(ld(c.b.hi));
      return;
    }
    const T plo = st<T>(static_cast<double>(ld(c.b.lo)) - mean);
    const T phi = st<T>(static_cast<double>(ld(c.b.hi)) - mean);
    blo = (c.act & 0xffu) ? C(ld(plo)) : C(0);
    bhi = (c.act & 0xff00u) ? C(ld(phi)) : C(0);
  };
  // coefficients of the pair's cell `hi` (0: x0 + 2l, 1: x0 + 2l + 1);
  // values are widened before the select, so it picks between registers
  auto pick = [&](const SwPairCoef<T>& c, C blo, C bhi,
                  bool hi) -> SwCoef<C> {
    auto sel = [hi](C lo, C h) -> C { return hi ? h : lo; };
    SwCoef<C> f;
    f.d = sel(C(ld(c.d.lo)), C(ld(c.d.hi)));
    const C x0v = C(ld(c.fx0)), x1v = C(ld(c.fx1)), x2v = C(ld(c.fx2));
    f.xl = sel(x0v, x1v);
    f.xh = sel(x1v, x2v);
    f.yl = sel(C(ld(c.yl.lo)), C(ld(c.yl.hi)));
    f.yh = sel(C(ld(c.yh.lo)), C(ld(c.yh.hi)));
    f.zl = sel(C(ld(c.zl.lo)), C(ld(c.zl.hi)));
    f.zh = sel(C(ld(c.zh.lo)), C(ld(c.zh.hi)));
    f.b = sel(blo, bhi);
    return f;
  };
  auto fetch_cell = [&](const PlaneOff& pl, uint32_t ob,
                        uint32_t fb) -> SwCoef<C> {
    SwCoef<C> f;
    f.d = C(ld(sw_load<T>(rd, ob, pl.c)));
    f.xl = C(ld(sw_load<T>(rfx, fb, pl.fx)));
    f.xh = C(ld(sw_load<T>(rfx, fb + kS, pl.fx)));
    f.yl = C(ld(sw_load<T>(rfy, ob, pl.fy)));
    f.yh = C(ld(sw_load<T>(rfy, ob, pl.fyh)));
    f.zl = C(ld(sw_load<T>(rfz, ob, pl.c)));
    f.zh = C(ld(sw_load<T>(rfz, ob, pl.fzh)));
    f.b = bval(pl, ob, sw_load<T>(rb, ob, pl.c));
    return f;
  };
  // x <- x + omega (b - A x) / D at LDS position q, in the reference
  // evaluation order; the halo supplies boundary neighbours
  auto update = [&](const SwCoef<C>& f, const C* Pm, C* P0, const C* Pp,
                    int q) {
    const C xc = P0[q];
    C v = f.d * xc;
    v -= f.xl * P0[q - 1];
    v -= f.xh * P0[q + 1];
    v -= f.yl * P0[q - kSwPX];
    v -= f.yh * P0[q + kSwPX];
    v -= f.zl * Pm[q];
    v -= f.zh * Pp[q];
    P0[q] = sw_round<T>(xc + omega * (f.b - v) / f.d);
  };
  // first-colour cell of a pair in row gj on plane kg: 0 (x0 + 2l) or 1
  auto first_hi = [&](int gj, int kg) -> bool {
    return ((c0 + gj + kg) & 1) != 0;
  };

  const int kstart = z0 - 1;
  {
    C v[kSwLoadSlots];
    fetch_plane(kstart - 1, v);
    store_plane(kstart - 1, v);
    fetch_plane(kstart, v);
    store_plane(kstart, v);
  }
  // second-colour coefficients of the two tile rows, carried one step (with
  // kMean, the fluid flags of the pair too)
  SwCoef<C> carry[2]{};
  uint32_t act_carry[2] = {0u, 0u};
  for (int k = kstart; k <= z1; ++k) {
    // ---- issue: x of plane k+1 and the coefficients of plane k
    C vx[kSwLoadSlots];
    fetch_plane(k + 1, vx);
    int kg;
    const bool first_plane = plane_of(k, kg);
    const PlaneOff pk = plane_off(first_plane ? kg : 0);
    SwPairCoef<T> pc[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      if (first_plane && rvalid[s]) pc[s] = fetch_pair(pk, pob[s], pfb[s]);
    }
    // ring cell of this lane (first colour only): in the ring row (waves 0
    // and 1) or column (wave 2; column -1 when sel is 0)
    const bool rhi = first_hi(rgj[2], kg);
    const int sel = (c0 + 1 + cgj + kg) & 1;
    bool ring_ok;
    uint32_t ring_ob, ring_fb;
    int ring_q;
    if (wave < 2) {
      ring_ok = rvalid[2] && lane_in;
      ring_ob = pob[2] + (rhi ? kS : 0u);
      ring_fb = pfb[2] + (rhi ? kS : 0u);
      ring_q = (2 * lane + 2 + (rhi ? 1 : 0)) + kSwPX * (rr[2] + 2);
    } else {
      ring_ok = (cok >> sel) & 1u;
      ring_ob = sel ? cob[1] : cob[0];
      ring_fb = sel ? cfb[1] : cfb[0];
      ring_q = (sel ? cright - x0 + 2 : 1) + kSwPX * (crow + 2);
    }
    ring_ok = ring_ok && first_plane;
    SwCoef<C> fr{};
    if (wave < 3 && first_plane)
      fr = fetch_cell(pk, ring_ok ? ring_ob : 0u, ring_ok ? ring_fb : 0u);
    const int kb = k - 1;
    const bool second_plane = kb >= z0 && kb < z1;
    // ---- x of plane k+1 into LDS
    store_plane(k + 1, vx);
    __syncthreads();
    SwCoef<C> next[2]{};
    uint32_t act_next[2] = {0u, 0u};
    if (first_plane) {
      const C* Pm = buf[sw_slot(k - 1)];
      C* P0 = buf[sw_slot(k)];
      const C* Pp = buf[sw_slot(k + 1)];
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        if (!rvalid[s]) continue;
        const bool hi = first_hi(rgj[s], kg);
        const int q = (2 * lane + 2 + (hi ? 1 : 0)) + kSwPX * (rr[s] + 2);
        C blo, bhi;
        pair_b(pc[s], blo, bhi);
        if (lane_in) update(pick(pc[s], blo, bhi, hi), Pm, P0, Pp, q);
        next[s] = pick(pc[s], blo, bhi, !hi);
        if (kMean) act_next[s] = pc[s].act;
        // the projected b of this block's planes, stored once
        if (kZero && a.bout != nullptr && lane_in && k >= z0 && k < z1)
          sw_store2(rbout, pob[s], pk.c, st<T>(blo), st<T>(bhi));
      }
      if (ring_ok) update(fr, Pm, P0, Pp, ring_q);
    }
    // no barrier before the second colour: it reads first-colour cells of
    // plane kb written in the previous step, and plane k only at its own
    // pair, written by this lane above
    if (second_plane) {
      const C* Pm = buf[sw_slot(kb - 1)];
      C* P0 = buf[sw_slot(kb)];
      const C* Pp = buf[sw_slot(kb + 1)];
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        if (!rvalid[s]) continue;
        const bool hi = !first_hi(rgj[s], kb);
        const int q = (2 * lane + 2 + (hi ? 1 : 0)) + kSwPX * (rr[s] + 2);
        if (lane_in) update(carry[s], Pm, P0, Pp, q);
      }
      // both cells of a stored pair were written by this lane
      const PlaneOff pb = plane_off(kb);
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        if (!rvalid[s] || !lane_in) continue;
        const int q = (2 * lane + 2) + kSwPX * (rr[s] + 2);
        const T lo = st<T>(P0[q]), hi = st<T>(P0[q + 1]);
        sw_store2(rout, pob[s], pb.c, lo, hi);
        if (kMean) {
          // the terms of k_vol_dot_finest: ((wx wy) wz) x on fluid cells
          const double wz = a.wz[kb];
          const double vlo = mwx0 * mwy[s] * wz;
          const double vhi = mwx1 * mwy[s] * wz;
          if (act_carry[s] & 0xffu) msum += vlo * static_cast<double>(ld(lo));
          if (act_carry[s] & 0xff00u) msum += vhi * static_cast<double>(ld(hi));
        }
      }
    }
    carry[0] = next[0];
    carry[1] = next[1];
    act_carry[0] = act_next[0];
    act_carry[1] = act_next[1];
  }
  if (kMean) {
    msum = block_sum(msum, red);
    if (threadIdx.x == 0) a.partials[xcd_block32(blockIdx.x, gridDim.x)] = msum;
  }
}

}  // namespace hf
