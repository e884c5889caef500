// Device-side building blocks of the gfx950 pressure solver: storage types,
// level descriptors, the 7-point stencil in the reference evaluation order,
// and block reductions built on 64-lane wavefront shuffles.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hf {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kXcds = 8;  // MI355X: 256 CUs in 8 XCDs, each with its own L2

// BF16 storage only; arithmetic always widens to FP32.
struct bf16 {
  uint16_t bits;
};

template <class T> struct Compute { using type = float; };
template <> struct Compute<double> { using type = double; };

__device__ __forceinline__ double ld(double v) { return v; }
__device__ __forceinline__ float ld(float v) { return v; }
__device__ __forceinline__ float ld(bf16 v) {
  return __uint_as_float(static_cast<uint32_t>(v.bits) << 16);
}

template <class T> __device__ __forceinline__ T st(double v);
template <> __device__ __forceinline__ double st<double>(double v) { return v; }
template <> __device__ __forceinline__ float st<float>(double v) {
  return static_cast<float>(v);
}
template <> __device__ __forceinline__ bf16 st<bf16>(double v) {
  // round-to-nearest-even on the FP32 value, quiet NaN preserved
  const float f = static_cast<float>(v);
  uint32_t u = __float_as_uint(f);
  bf16 out;
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu) != 0u) {
    out.bits = 0x7fc0u;
    return out;
  }
  u += 0x7fffu + ((u >> 16) & 1u);
  out.bits = static_cast<uint16_t>(u >> 16);
  return out;
}
template <class T> __device__ __forceinline__ T st(float v);
template <> __device__ __forceinline__ double st<double>(float v) { return v; }
template <> __device__ __forceinline__ float st<float>(float v) { return v; }
template <> __device__ __forceinline__ bf16 st<bf16>(float v) {
  return st<bf16>(static_cast<double>(v));
}

// Boundary handling of one physical face of the point row
// (src/solver_mg.cpp:938-963 and :2356-2368): 0 -> zero (Dirichlet),
// 1 -> self (Neumann), 2 -> wrapped neighbour (periodic halo).
struct Bc {
  int8_t m[6];
};

// One multigrid level in storage type T.  Coefficient arrays hold the
// masked operator; act marks fluid cells (fluid volume > 0).
template <class T>
struct Lv {
  int nx, ny, nz;
  int cells;  // every level fits 32-bit indexing (< 2^31 cells and faces)
  const T* D;
  const T* fx;
  const T* fy;
  const T* fz;
  const uint8_t* act;
  const double* vol;
  T* x;
  T* b;
  T* r;
};

template <class C, class T>
__device__ __forceinline__ C neighbour_or_boundary(int mode, C self,
                                                   const T* x, int wrapped) {
  if (mode == 0) return C(0);
  if (mode == 1) return self;
  return ld(x[wrapped]);
}

// The seven coefficients of the row of cell c = (i, j, k) of a level.
template <class C>
struct StencilCoef {
  C d, xl, xh, yl, yh, zl, zh;
};

template <class C, class T>
__device__ __forceinline__ StencilCoef<C> stencil_coef(const Lv<T>& L, int j,
                                                       int k, int c) {
  const int row = j + L.ny * k;
  const int fyo = c + L.nx * k;
  StencilCoef<C> f;
  f.d = C(ld(L.D[c]));
  f.xl = C(ld(L.fx[c + row]));
  f.xh = C(ld(L.fx[c + row + 1]));
  f.yl = C(ld(L.fy[fyo]));
  f.yh = C(ld(L.fy[fyo + L.nx]));
  f.zl = C(ld(L.fz[c]));
  f.zh = C(ld(L.fz[c + L.nx * L.ny]));
  return f;
}

// A x at cell (i,j,k) in the reference order: D*x, then x-lo, x-hi, y-lo,
// y-hi, z-lo, z-hi, each product subtracted in turn (src/solver_mg.cpp:
// 2405-2411), with the row's coefficients given.  Built with
// -ffp-contract=off so FP64 results are bitwise equal to the sequential
// reference.
template <class C, class T>
__device__ __forceinline__ C stencil_with(const Lv<T>& L, const Bc& bc,
                                          const T* x, int i, int j, int k,
                                          int c, C xc,
                                          const StencilCoef<C>& f) {
  const int nx = L.nx;
  const int plane = nx * L.ny;
  const C xl = i > 0 ? ld(x[c - 1])
                     : neighbour_or_boundary<C>(bc.m[0], xc, x, c + (nx - 1));
  const C xh = i + 1 < nx
                   ? ld(x[c + 1])
                   : neighbour_or_boundary<C>(bc.m[1], xc, x, c - (nx - 1));
  const C yl = j > 0 ? ld(x[c - nx])
                     : neighbour_or_boundary<C>(bc.m[2], xc, x,
                                                c + (L.ny - 1) * nx);
  const C yh = j + 1 < L.ny
                   ? ld(x[c + nx])
                   : neighbour_or_boundary<C>(bc.m[3], xc, x,
                                              c - (L.ny - 1) * nx);
  const C zl = k > 0 ? ld(x[c - plane])
                     : neighbour_or_boundary<C>(bc.m[4], xc, x,
                                                c + (L.nz - 1) * plane);
  const C zh = k + 1 < L.nz
                   ? ld(x[c + plane])
                   : neighbour_or_boundary<C>(bc.m[5], xc, x,
                                              c - (L.nz - 1) * plane);
  C v = f.d * xc;
  v -= f.xl * xl;
  v -= f.xh * xh;
  v -= f.yl * yl;
  v -= f.yh * yh;
  v -= f.zl * zl;
  v -= f.zh * zh;
  return v;
}

// A x at cell (i,j,k), coefficients loaded from the level.
template <class C, class T>
__device__ __forceinline__ C stencil(const Lv<T>& L, const Bc& bc, const T* x,
                                     int i, int j, int k, int c, C xc) {
  return stencil_with<C>(L, bc, x, i, j, k, c, xc, stencil_coef<C>(L, j, k, c));
}

// Block index -> logical tile such that each XCD owns one contiguous slab of
// the domain: neighbouring rows and planes are then served by the same L2.
__device__ __forceinline__ int64_t xcd_block(int64_t block, int64_t blocks) {
  const int64_t q = blocks / kXcds;
  const int64_t rem = blocks % kXcds;
  const int64_t xcd = block % kXcds;
  const int64_t local = block / kXcds;
  return xcd < rem ? xcd * (q + 1) + local
                   : rem * (q + 1) + (xcd - rem) * q + local;
}

__device__ __forceinline__ int xcd_block32(int block, int blocks) {
  const int q = blocks / kXcds;
  const int rem = blocks % kXcds;
  const int xcd = block % kXcds;
  const int local = block / kXcds;
  return xcd < rem ? xcd * (q + 1) + local
                   : rem * (q + 1) + (xcd - rem) * q + local;
}

// 2-D tiling of one z-plane: a 256-thread block covers 2^txl cells in x by
// 256 >> txl rows in y, so (i, j, k) come from uniform per-tile arithmetic
// and lane bit fields instead of per-thread divisions.  Tiles are numbered
// x-fastest, then y, then z; chunked kernels give each block `per`
// consecutive tiles.
struct Tiles {
  int nx, ny, nz;  // iteration extents (x counts colour cells for RB passes)
  int txl;         // log2 of the tile width
  int gx, gy;      // tiles along x and y
  int count;       // gx * gy * nz
  int per;         // tiles per block of a chunked launch
  int blocks;      // blocks of a chunked launch
};

__device__ __forceinline__ void tile_xyz(const Tiles& t, int tile, int& i,
                                         int& j, int& k) {
  const int gxy = t.gx * t.gy;
  k = tile / gxy;
  const int r = tile - k * gxy;
  const int ty = r / t.gx;
  i = ((r - ty * t.gx) << t.txl) +
      static_cast<int>(threadIdx.x & ((1u << t.txl) - 1u));
  j = ty * (kBlock >> t.txl) + static_cast<int>(threadIdx.x >> t.txl);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int offset = kWave / 2; offset > 0; offset >>= 1)
    v += __shfl_down(v, offset, kWave);
  return v;
}

// Deterministic block sum (fixed shuffle tree, fixed wave order).  Result is
// valid in thread 0.  `scratch` needs blockDim/64 doubles of LDS.
__device__ __forceinline__ double block_sum(double v, double* scratch) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  v = wave_sum(v);
  if (lane == 0) scratch[wave] = v;
  __syncthreads();
  double total = 0.0;
  if (threadIdx.x == 0) {
    const int waves = (blockDim.x + kWave - 1) / kWave;
    for (int w = 0; w < waves; ++w) total += scratch[w];
  }
  __syncthreads();
  return total;
}

}  // namespace hf
