// On-device assembly of the pressure-correction system (gfx950): the
// compressible system D = a0 V drho/dp + sum of faces, b = -(V density defect
// + outward flux) of src/solver_pressure.cpp:110-332, and the cold path's
// coefficient loop of src/core_product_freeze.cpp:14279-14317.  Both are
// elementwise passes over cells (and faces) in the reference's operation
// order, so with -ffp-contract=off every value equals the sequential loop's.
#pragma once

#include "hf_cycle.hpp"

namespace hf {

// ---------------------------------------------------------------------------
// Compressible pressure-correction system.
// ---------------------------------------------------------------------------
struct PressureSystem {
  int nx, ny, nz;
  Widths w;
  const double* bx;
  const double* by;
  const double* bz;
  const double* phix;
  const double* phiy;
  const double* phiz;
  const double* rho;
  const double* rho_n;
  const double* rho_nm1;  // null: BDF1
  const double* drho_dp;
  const double* source;   // null: no mass source
  double a0, a1, a2;
  double* diag;
  double* rhs;
};

// Face inputs must be finite, the coefficients also non-negative
// (finite_nonnegative_faces / finite_faces); a violation sets bit 0 of *bad.
__global__ void k_check_faces(const double* beta, const double* phi,
                              int64_t count, int* bad) {
  const int64_t f = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (f >= count) return;
  const double b = beta[f];
  if (!isfinite(b) || b < 0.0 || !isfinite(phi[f])) atomicOr(bad, 1);
}

// One cell per thread.  write == 0: the preflight, which only checks every
// value the reference checks (a failure sets bit 1 of *bad); write == 1:
// stores D and b.
__global__ void k_pressure_system(PressureSystem s, int write, int* bad) {
  const int64_t c = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const int64_t plane = static_cast<int64_t>(s.nx) * s.ny;
  if (c >= plane * s.nz) return;
  const int i = static_cast<int>(c % s.nx);
  const int j = static_cast<int>((c / s.nx) % s.ny);
  const int k = static_cast<int>(c / plane);
  const double volume = s.w.x[i] * s.w.y[j] * s.w.z[k];
  const double r = s.rho[c], rn = s.rho_n[c];
  const double rnm1 = s.rho_nm1 != nullptr ? s.rho_nm1[c] : 0.0;
  const double src = s.source != nullptr ? s.source[c] : 0.0;
  const double drho = s.drho_dp[c];
  const double defect = volume * (s.a0 * r + s.a1 * rn + s.a2 * rnm1 - src);
  const int64_t xf = i + static_cast<int64_t>(s.nx + 1) * (j + static_cast<int64_t>(s.ny) * k);
  const int64_t yf = i + static_cast<int64_t>(s.nx) * (j + static_cast<int64_t>(s.ny + 1) * k);
  const int64_t zf = c;
  const double faces = s.bx[xf] + s.bx[xf + 1] + s.by[yf] + s.by[yf + s.nx] +
                       s.bz[zf] + s.bz[zf + plane];
  const double d = s.a0 * volume * drho + faces;
  const double flux = s.phix[xf + 1] - s.phix[xf] + s.phiy[yf + s.nx] -
                      s.phiy[yf] + s.phiz[zf + plane] - s.phiz[zf];
  const double b = -(defect + flux);
  if (write) {
    s.diag[c] = d;
    s.rhs[c] = b;
    return;
  }
  const bool inputs = isfinite(r) && isfinite(rn) && isfinite(drho) &&
                      (s.rho_nm1 == nullptr || isfinite(rnm1)) &&
                      (s.source == nullptr || isfinite(src));
  const bool ok = inputs && r > 0.0 && rn > 0.0 &&
                  (s.rho_nm1 == nullptr || rnm1 > 0.0) && drho > 0.0 &&
                  volume > 0.0 && d > 0.0 && isfinite(defect) && isfinite(d) &&
                  isfinite(b);
  if (!ok) atomicOr(bad, 2);
}

// ---------------------------------------------------------------------------
// Cold-path coefficients.  The reference loops over cells and writes all six
// face coefficients of each cell from that cell's face records; both cells
// of a face always write the same value (an inactive cell's records are
// zero, an active cell drops a face whose other cell is inactive), so the
// faces are computed here once each, from the face:
//   coefficient = 0 on a non-periodic exterior face or next to an inactive
//                 cell (periodic sides wrap), else projection + limiter / 2.
// D of an active cell = V rho / ((p_ref + p) dt) (0 with a fixed EOS), then
// per face x-, x+, y-, y+, z-, z+ the coefficient and, on a pressure-outlet
// exterior face, 2 projection; D of an inactive cell = 1.
// ---------------------------------------------------------------------------
struct ColdSystem {
  int n[3];
  Widths w;
  const double* proj[3];
  const double* lim[3];
  const uint8_t* fluid;  // null: every cell active
  const double* rho;
  const double* pressure;
  double pressure_reference;
  double dt;
  int fixed_eos;
  int8_t kind[6];  // 0 other, 1 periodic, 2 pressure outlet
  double* diag;
  double* coef[3];
};

__device__ __forceinline__ bool cold_fluid(const ColdSystem& s, int i, int j,
                                           int k) {
  int c[3] = {i, j, k};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (c[a] >= 0 && c[a] < s.n[a]) continue;
    if (s.kind[2 * a] != 1) return false;
    c[a] = (c[a] + s.n[a]) % s.n[a];
  }
  if (s.fluid == nullptr) return true;
  return s.fluid[c[0] + static_cast<int64_t>(s.n[0]) *
                             (c[1] + static_cast<int64_t>(s.n[1]) * c[2])] != 0;
}

// Coefficient of face (i, j, k) on axis a (the face below cell (i, j, k)).
__device__ __forceinline__ double cold_face(const ColdSystem& s, int a, int i,
                                            int j, int k, int64_t f) {
  const int at[3] = {i, j, k};
  const bool boundary = at[a] == 0 || at[a] == s.n[a];
  const bool periodic = s.kind[2 * a] == 1;
  if (boundary && !periodic) return 0.0;
  int lo[3] = {i, j, k};
  lo[a] -= 1;
  if (!cold_fluid(s, lo[0], lo[1], lo[2]) || !cold_fluid(s, i, j, k))
    return 0.0;
  return s.proj[a][f] + 0.5 * s.lim[a][f];
}

__global__ void k_cold_faces(ColdSystem s, int a, int64_t count) {
  const int64_t f = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (f >= count) return;
  const int ex = s.n[0] + (a == 0 ? 1 : 0);
  const int ey = s.n[1] + (a == 1 ? 1 : 0);
  const int i = static_cast<int>(f % ex);
  const int j = static_cast<int>((f / ex) % ey);
  const int k = static_cast<int>(f / (static_cast<int64_t>(ex) * ey));
  s.coef[a][f] = cold_face(s, a, i, j, k, f);
}

__global__ void k_cold_diag(ColdSystem s) {
  const int64_t c = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const int64_t plane = static_cast<int64_t>(s.n[0]) * s.n[1];
  if (c >= plane * s.n[2]) return;
  const int i = static_cast<int>(c % s.n[0]);
  const int j = static_cast<int>((c / s.n[0]) % s.n[1]);
  const int k = static_cast<int>(c / plane);
  if (s.fluid != nullptr && s.fluid[c] == 0) {
    s.diag[c] = 1.0;
    return;
  }
  const double volume = s.w.x[i] * s.w.y[j] * s.w.z[k];
  double d = s.fixed_eos
                 ? 0.0
                 : volume * s.rho[c] /
                       ((s.pressure_reference + s.pressure[c]) * s.dt);
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    const int a = f / 2;
    int at[3] = {i, j, k};
    if (f & 1) at[a] += 1;
    const int ex = s.n[0] + (a == 0 ? 1 : 0);
    const int ey = s.n[1] + (a == 1 ? 1 : 0);
    const int64_t fi = at[0] + static_cast<int64_t>(ex) *
                                   (at[1] + static_cast<int64_t>(ey) * at[2]);
    d += cold_face(s, a, at[0], at[1], at[2], fi);
    const bool exterior =
        (at[a] == 0 || at[a] == s.n[a]) && s.kind[f] != 1;
    if (exterior && s.kind[f] == 2) d += 2.0 * s.proj[a][fi];
  }
  s.diag[c] = d;
}

}  // namespace hf
