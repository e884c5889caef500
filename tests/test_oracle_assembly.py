"""Pins of the oracle's pressure-correction system assembly
(or_assemble_pressure_system, restating src/solver_pressure.cpp:110-332).

Each check is a property the finite-volume system must have, chosen so that
a plausible slip fails one of them:
* the right-hand sides telescope: summed over the grid, every interior face
  flux cancels and only the boundary outflux and the volume-integrated
  density defect remain (a wrong face index, a sign of the flux or a dropped
  axis breaks it);
* with no flux, b / V is the negated density defect per unit volume, the
  same in every cell when the densities are uniform, on a grid whose three
  axes have different, non-uniform widths (pins V = w_x w_y w_z);
* on all-Neumann walls the operator applied to a constant leaves
  a0 V drho/dp: the diagonal holds the six faces of the row (pins the face
  sum and the compressibility term against the operator, not the formula);
* the BDF1 form equals the BDF2 form with a2 = 0, and b is affine in the
  mass source (b(S) - b(0) = V S);
* an inadmissible input (a non-positive density, a negative face) is
  refused with numerical_failure and nothing is written.
"""
import numpy as np
import pytest

import hf_oracle as O

N = (7, 5, 6)


def _faces(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for a, count in enumerate(n):
        w = (1.0 + a) * (0.5 + rng.random(count))
        out.append(np.concatenate([[0.0], np.cumsum(w)]))
    return out


def _inputs(n, seed):
    rng = np.random.default_rng(seed)
    cells = n[0] * n[1] * n[2]
    nf = ((n[0] + 1) * n[1] * n[2], n[0] * (n[1] + 1) * n[2],
          n[0] * n[1] * (n[2] + 1))
    beta = [rng.random(c) for c in nf]
    phi = [rng.standard_normal(c) for c in nf]
    rho = 0.5 + rng.random(cells)
    rho_n = 0.5 + rng.random(cells)
    rho_nm1 = 0.5 + rng.random(cells)
    drho = 0.1 + rng.random(cells)
    src = rng.standard_normal(cells)
    return beta, phi, rho, rho_n, rho_nm1, drho, src


def _volumes(n, faces):
    w = [np.diff(f) for f in faces]
    return ((w[0][None, None, :] * w[1][None, :, None]) *
            w[2][:, None, None]).ravel()


def _boundary_outflux(n, phi):
    nx, ny, nz = n
    fx = phi[0].reshape(nz, ny, nx + 1)
    fy = phi[1].reshape(nz, ny + 1, nx)
    fz = phi[2].reshape(nz + 1, ny, nx)
    return ((fx[:, :, -1].sum() - fx[:, :, 0].sum()) +
            (fy[:, -1, :].sum() - fy[:, 0, :].sum()) +
            (fz[-1].sum() - fz[0].sum()))


A = (1.5, -2.0, 0.5)


def test_rhs_telescopes_to_boundary_outflux():
    faces = _faces(N, 1)
    beta, phi, rho, rho_n, rho_nm1, drho, src = _inputs(N, 2)
    st, d, b = O.assemble_pressure_system(N, faces, *beta, phi, rho, rho_n,
                                          drho, A, rho_nm1=rho_nm1,
                                          source=src)
    assert st == 0
    V = _volumes(N, faces)
    defect = V * (A[0] * rho + A[1] * rho_n + A[2] * rho_nm1 - src)
    want = -(defect.sum() + _boundary_outflux(N, phi))
    scale = np.abs(defect).sum() + sum(np.abs(p).sum() for p in phi)
    assert abs(b.sum() - want) <= 1e-13 * scale


def test_rhs_per_volume_without_flux():
    faces = _faces(N, 3)
    beta, phi, _, _, _, drho, _ = _inputs(N, 4)
    cells = N[0] * N[1] * N[2]
    zero = [np.zeros_like(p) for p in phi]
    rho = np.full(cells, 1.25)
    rho_n = np.full(cells, 0.75)
    st, d, b = O.assemble_pressure_system(N, faces, *beta, zero, rho, rho_n,
                                          drho, (A[0], A[1], 0.0))
    assert st == 0
    V = _volumes(N, faces)
    per_volume = b / V
    expect = -(A[0] * 1.25 + A[1] * 0.75)
    assert np.allclose(per_volume, expect, rtol=1e-14, atol=0.0)
    # the widths really differ from cell to cell
    assert V.max() / V.min() > 2.0


def test_diagonal_against_operator_on_neumann_walls():
    faces = _faces(N, 5)
    beta, phi, rho, rho_n, _, drho, _ = _inputs(N, 6)
    st, d, b = O.assemble_pressure_system(N, faces, *beta, phi, rho, rho_n,
                                          drho, (A[0], A[1], 0.0))
    assert st == 0
    cells = N[0] * N[1] * N[2]
    ones = np.ones(cells)
    y = O.operator_apply(N, [1] * 6, d, beta[0], beta[1], beta[2], None,
                         ones)
    V = _volumes(N, faces)
    want = A[0] * V * drho
    assert np.allclose(y, want, rtol=1e-12, atol=1e-14 * np.abs(d).max())


def test_bdf1_is_bdf2_with_zero_a2_and_source_is_affine():
    faces = _faces(N, 7)
    beta, phi, rho, rho_n, rho_nm1, drho, src = _inputs(N, 8)
    a = (A[0], A[1], 0.0)
    s1, d1, b1 = O.assemble_pressure_system(N, faces, *beta, phi, rho, rho_n,
                                            drho, a)
    s2, d2, b2 = O.assemble_pressure_system(N, faces, *beta, phi, rho, rho_n,
                                            drho, a, rho_nm1=rho_nm1)
    assert s1 == s2 == 0
    assert np.array_equal(d1, d2) and np.array_equal(b1, b2)
    s3, d3, b3 = O.assemble_pressure_system(N, faces, *beta, phi, rho, rho_n,
                                            drho, a, source=src)
    assert s3 == 0 and np.array_equal(d3, d1)
    V = _volumes(N, faces)
    assert np.allclose(b3 - b1, V * src, rtol=1e-10,
                       atol=1e-12 * np.abs(b1).max())


@pytest.mark.parametrize("poison", ["rho", "rho_nm1", "drho", "face", "a0"])
def test_inadmissible_input_publishes_nothing(poison):
    faces = _faces(N, 9)
    beta, phi, rho, rho_n, rho_nm1, drho, _ = _inputs(N, 10)
    a = list(A)
    if poison == "rho":
        rho = rho.copy()
        rho[17] = 0.0
    elif poison == "rho_nm1":
        rho_nm1 = rho_nm1.copy()
        rho_nm1[3] = -1.0
    elif poison == "drho":
        drho = drho.copy()
        drho[-1] = np.nan
    elif poison == "face":
        beta = [v.copy() for v in beta]
        beta[1][5] = -1e-3
    else:
        a[0] = 0.0
    st, d, b = O.assemble_pressure_system(N, faces, *beta, phi, rho, rho_n,
                                          drho, a, rho_nm1=rho_nm1)
    assert st != 0
    assert np.isnan(d).all() and np.isnan(b).all()


# ---------------------------------------------------------------------------
# Cold-path coefficient loop (or_cold_coefficients, restating
# src/core_product_freeze.cpp:14279-14317).  Pinned through the operator it
# builds, not its formula:
# * every row of A applied to a constant, with Neumann closure of the
#   exterior faces, leaves the compressibility term V rho / ((p_ref + p) dt)
#   plus 2 projection per pressure-outlet face, and 1 on a solid row (the
#   identity);
# * faces on a non-periodic exterior side and faces next to an inactive
#   cell are zero; an interior fluid-fluid face is projection + limiter / 2;
# * periodic sides couple the first and last cells through the face that
#   both write, and the two copies of that face agree;
# * fixed_eos drops the compressibility term (a closed fluid domain is then
#   singular: A 1 = 0).
# ---------------------------------------------------------------------------
def _cold_inputs(n, seed):
    rng = np.random.default_rng(seed)
    fc = O.face_counts(n)
    proj = [0.2 + rng.random(c) for c in fc]
    lim = [rng.random(c) for c in fc]
    cells = n[0] * n[1] * n[2]
    rho = 0.5 + rng.random(cells)
    pres = rng.standard_normal(cells) * 0.1
    return proj, lim, rho, pres


def _sphere_fluid(n, radius=0.3):
    nx, ny, nz = n
    z, y, x = np.meshgrid((np.arange(nz) + 0.5) / nz,
                          (np.arange(ny) + 0.5) / ny,
                          (np.arange(nx) + 0.5) / nx, indexing="ij")
    solid = (x - 0.5) ** 2 + (y - 0.5) ** 2 + (z - 0.5) ** 2 < radius ** 2
    return (~solid).astype(np.uint8).ravel()


@pytest.mark.parametrize("kinds", [[0] * 6, [2, 0, 1, 1, 0, 2],
                                   [1, 1, 1, 1, 1, 1]])
def test_cold_rows_against_operator(kinds):
    n = (9, 7, 8)
    faces = _faces(n, 11)
    proj, lim, rho, pres = _cold_inputs(n, 12)
    fluid = _sphere_fluid(n)
    pref, dt = 2.0, 0.05
    st, d, cx, cy, cz = O.cold_coefficients(n, faces, proj, lim, fluid, rho,
                                            pres, pref, dt, kinds)
    assert st == 0
    bc = [2 if k == 1 else 1 for k in kinds]  # exterior faces: Neumann
    y = O.operator_apply(n, bc, d, cx, cy, cz, None, np.ones(d.size))
    V = _volumes(n, faces)
    want = V * rho / ((pref + pres) * dt)
    # pressure-outlet faces of fluid cells add 2 projection to the diagonal
    nx, ny, nz = n
    P = [p.copy() for p in proj]
    out = np.zeros(d.size).reshape(nz, ny, nx)
    px = P[0].reshape(nz, ny, nx + 1)
    py = P[1].reshape(nz, ny + 1, nx)
    pz = P[2].reshape(nz + 1, ny, nx)
    if kinds[0] == 2: out[:, :, 0] += 2 * px[:, :, 0]
    if kinds[1] == 2: out[:, :, -1] += 2 * px[:, :, -1]
    if kinds[2] == 2: out[:, 0, :] += 2 * py[:, 0, :]
    if kinds[3] == 2: out[:, -1, :] += 2 * py[:, -1, :]
    if kinds[4] == 2: out[0] += 2 * pz[0]
    if kinds[5] == 2: out[-1] += 2 * pz[-1]
    want = want + out.ravel()
    solid = fluid == 0
    want[solid] = 1.0
    # a solid row is the identity: its faces are all zero
    assert np.allclose(y, want, rtol=1e-12, atol=1e-12)
    assert np.all(d[solid] == 1.0)


def test_cold_face_values():
    n = (6, 5, 4)
    faces = _faces(n, 13)
    proj, lim, rho, pres = _cold_inputs(n, 14)
    fluid = np.ones(n[0] * n[1] * n[2], dtype=np.uint8)
    fluid[2 + n[0] * (2 + n[1] * 1)] = 0  # one inactive cell (2, 2, 1)
    kinds = [0, 2, 1, 1, 0, 0]
    st, d, cx, cy, cz = O.cold_coefficients(n, faces, proj, lim, fluid, rho,
                                            pres, 1.0, 0.1, kinds)
    assert st == 0
    nx, ny, nz = n
    full = [p + 0.5 * l for p, l in zip(proj, lim)]
    X = cx.reshape(nz, ny, nx + 1)
    FX = full[0].reshape(nz, ny, nx + 1)
    # non-periodic exterior x sides are zero, interior x faces are full
    assert np.all(X[:, :, 0] == 0) and np.all(X[:, :, -1] == 0)
    mask = np.ones_like(X, dtype=bool)
    mask[:, :, [0, -1]] = False
    mask[1, 2, 2] = mask[1, 2, 3] = False  # faces of the inactive cell
    assert np.array_equal(X[mask], FX[mask])
    assert X[1, 2, 2] == 0 and X[1, 2, 3] == 0
    # periodic y: the face both edge cells write is the full coefficient,
    # and its two copies (index 0 and ny) agree
    Y = cy.reshape(nz, ny + 1, nx)
    FY = full[1].reshape(nz, ny + 1, nx)
    assert np.array_equal(Y[:, 0, :], FY[:, 0, :])
    assert np.array_equal(Y[:, -1, :], FY[:, -1, :])


def test_cold_fixed_eos_is_singular_on_closed_fluid():
    n = (5, 6, 4)
    faces = _faces(n, 15)
    proj, lim, rho, pres = _cold_inputs(n, 16)
    st, d, cx, cy, cz = O.cold_coefficients(n, faces, proj, lim, None, rho,
                                            pres, 1.0, 0.1, [0] * 6,
                                            fixed_eos=True)
    assert st == 0
    y = O.operator_apply(n, [1] * 6, d, cx, cy, cz, None, np.ones(d.size))
    assert np.abs(y).max() <= 1e-13 * np.abs(d).max()
