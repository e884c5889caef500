"""CPU pins of the oracle beyond the reference-closure fixtures.

The golden fixtures pin the multigrid and PCG bit for bit.  The face
coefficient rule and the exact IBM operator live in reference objects outside
the compiled closure (solver_piso / solver_pressure), so they are pinned here
by what the mathematics fixes: closed-form coefficients, flux continuity of
the harmonic rule, symmetry / positive semi-definiteness / null space of the
dense operator, identity rows on solids, agreement with the reference MG's own
level-0 matrix action recorded in the fixtures, and a dense least-squares
solve.
"""
import glob
import math
import os

import numpy as np
import pytest

import hf_cases
import hf_oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


def small_case(n=(5, 4, 3), seed=3, bc=(1, 1, 1, 1, 2, 2), solid=True,
               stretched=False):
    rng = np.random.default_rng(seed)
    nx, ny, nz = n
    faces = [hf_cases.uniform_faces(nx),
             hf_cases.tanh_faces(ny, 4.0) if stretched
             else hf_cases.uniform_faces(ny),
             hf_cases.uniform_faces(nz)]
    rho = np.exp(0.5 * rng.standard_normal((nz, ny, nx)))
    act = np.ones((nz, ny, nx), dtype=np.uint8)
    if solid:
        act[1, 1, 2] = 0
        act[1, 2, 2] = 0
    fx, fy, fz = hf_cases.face_activity(act, list(bc))
    case = {"n": n, "faces": faces, "bc": list(bc),
            "mobility": (1.0 / rho).ravel(), "cell_active": act.ravel(),
            "x_active": fx, "y_active": fy, "z_active": fz}
    return case


def dense_operator(case):
    n = case["n"]
    N = n[0] * n[1] * n[2]
    d, bx, by, bz = O.assemble(n, case["faces"], case["bc"], case["mobility"])
    A = np.zeros((N, N))
    for c in range(N):
        e = np.zeros(N)
        e[c] = 1.0
        A[:, c] = O.operator_apply(n, case["bc"], d, bx, by, bz,
                                   case["cell_active"], e)
    return A, (d, bx, by, bz)


def test_face_rule_closed_form_uniform():
    """Constant rho on a uniform grid: beta = area / (h * rho); Neumann walls
    carry 0; periodic faces use the same value; D = sum of faces.  Extents
    are powers of two so the REAL4 face coordinates are exact."""
    n = (4, 2, 8)
    h = [1.0 / v for v in n]
    rho = 2.5
    faces = [hf_cases.uniform_faces(v) for v in n]
    bc = [1, 1, 1, 1, 2, 2]
    m = np.full(n[0] * n[1] * n[2], 1.0 / rho)
    d, bx, by, bz = O.assemble(n, faces, bc, m)
    ex = h[1] * h[2] / (h[0] * rho)
    bx = bx.reshape(n[2], n[1], n[0] + 1)
    assert np.allclose(bx[:, :, 1:-1], ex, rtol=1e-14)
    assert np.all(bx[:, :, 0] == 0.0) and np.all(bx[:, :, -1] == 0.0)
    bz = bz.reshape(n[2] + 1, n[1], n[0])
    ez = h[0] * h[1] / (h[2] * rho)
    assert np.allclose(bz, ez, rtol=1e-14)  # periodic faces included
    assert np.array_equal(bz[0], bz[-1])


def test_face_rule_flux_continuity():
    """The harmonic coefficient is the one whose flux beta*(pL - pR) equals
    both half-cell fluxes m*(p - p_face)/d for the interface pressure that
    balances them (two cells of different mobility and width)."""
    faces = [np.array([0.0, 0.25, 1.0]), hf_cases.uniform_faces(1),
             hf_cases.uniform_faces(1)]
    mL, mR = 0.37, 2.9
    d, bx, by, bz = O.assemble((2, 1, 1), faces, [1, 1, 1, 1, 1, 1],
                               np.array([mL, mR]))
    beta = bx[1]
    dL, dR = 0.125, 0.375
    pL, pR = 1.7, -0.4
    # interface pressure from flux balance
    pf = (mL * pL / dL + mR * pR / dR) / (mL / dL + mR / dR)
    flux = mL * (pL - pf) / dL  # per unit area; face area is 1
    assert math.isclose(beta * (pL - pR), flux, rel_tol=1e-14)
    assert math.isclose(flux, mR * (pf - pR) / dR, rel_tol=1e-14)


@pytest.mark.parametrize("bc", [(1, 1, 1, 1, 2, 2), (0, 0, 1, 1, 2, 2),
                                (2, 2, 2, 2, 2, 2), (1, 1, 1, 1, 1, 1)])
@pytest.mark.parametrize("stretched", [False, True])
def test_dense_operator_structure(bc, stretched):
    case = small_case(bc=bc, stretched=stretched)
    A, _ = dense_operator(case)
    act = case["cell_active"] != 0
    # solid rows and columns are the identity
    solid = np.where(~act)[0]
    for s in solid:
        e = np.zeros(len(act))
        e[s] = 1.0
        assert np.array_equal(A[s], e)
        assert np.array_equal(A[:, s], e)
    F = A[np.ix_(act, act)]
    assert np.allclose(F, F.T, rtol=0, atol=1e-15 * np.abs(F).max())
    w = np.linalg.eigvalsh(0.5 * (F + F.T))
    tol = 1e-12 * w.max()
    if 0 in bc:
        assert w.min() > tol  # Dirichlet walls: SPD
    else:
        # pure Neumann/periodic: one constant null vector on the fluid
        assert abs(w[0]) <= tol and w[1] > tol
        assert np.allclose(F @ np.ones(F.shape[0]), 0.0,
                           atol=1e-13 * np.abs(F).max())


def test_masked_form_equals_decorated_operator():
    """The multigrid's masked level-0 matrix (faces x activity, fluid
    diagonal D - original + active faces) is the IBM-decorated operator."""
    case = small_case(n=(6, 5, 4), seed=9)
    (d, bx, by, bz), mg = O.from_case(case, O.policy(maximum_levels=2))
    x = np.random.default_rng(1).standard_normal(6 * 5 * 4)
    y1 = O.operator_apply(case["n"], case["bc"], d, bx, by, bz,
                          case["cell_active"], x)
    y2 = mg.masked_apply(x)
    assert np.allclose(y1, y2, rtol=0, atol=4e-16 * np.abs(y1).max() * 8)


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(
    HERE, "golden", "ref_*.npz"))), ids=lambda p: os.path.basename(p)[4:-4])
def test_operator_against_reference_level0_action(path):
    """After one reference MG application the level-0 slots hold the
    candidate x, the projected rhs b and (unit scaling) the defect b - A0 x,
    all computed by the reference code.  The oracle operator applied to x must
    reproduce A0 x = b - defect to rounding."""
    with np.load(path, allow_pickle=False) as g:
        if int(g["policy"][8]) != 1:
            pytest.skip("residual-minimising scaling rescales the slots")
        n = tuple(int(v) for v in g["n"])
        ax_ref = g["slot_b_0"] - g["slot_r_0"]
        ax = O.operator_apply(n, [int(v) for v in g["bc"]], g["diag"],
                              g["bx"], g["by"], g["bz"], g["cell_active"],
                              g["slot_x_0"])
        act = g["cell_active"] != 0
        scale = np.abs(ax_ref).max()
        assert np.max(np.abs(ax[act] - ax_ref[act])) <= 1e-12 * scale


def test_pcg_against_dense_least_squares():
    case = small_case(n=(6, 5, 4), seed=5)
    A, (d, bx, by, bz) = dense_operator(case)
    act = case["cell_active"] != 0
    rng = np.random.default_rng(7)
    b = np.where(act, rng.standard_normal(act.size), 0.0)
    b[act] -= b[act].mean()
    (_, _, _, _), mg = O.from_case(case, O.policy(maximum_levels=3))
    x, res = O.pcg(case["n"], case["bc"], d, bx, by, bz, case["cell_active"],
                   mg, b, rtol=1e-13)
    assert res.status == 0
    xd = np.linalg.lstsq(A, b, rcond=None)[0]
    g = lambda v: np.where(act, v - v[act].mean(), 0.0)
    assert np.linalg.norm(g(x) - g(xd)) <= 1e-9 * np.linalg.norm(g(xd))
    assert np.all(x[~act] == 0.0)


def test_dot_and_norm_exact():
    rng = np.random.default_rng(11)
    a = rng.standard_normal(10007) * np.exp(rng.uniform(-20, 20, 10007))
    b = rng.standard_normal(10007)
    exact = math.fsum(float(u) * float(v) for u, v in zip(a, b))
    assert math.isclose(O.dot(a, b), exact, rel_tol=1e-15, abs_tol=1e-300) or \
        abs(O.dot(a, b) - exact) <= 1e-15 * math.fsum(abs(a * b))
    s = np.abs(a).max()
    exact_norm = s * math.sqrt(math.fsum((float(v) / s) ** 2 for v in a))
    assert math.isclose(O.norm2(a), exact_norm, rel_tol=1e-15)
    assert O.norm2(np.zeros(5)) == 0.0


def test_restriction_conserves_and_prolongation_reproduces_constants():
    case = small_case(n=(8, 6, 4), solid=False, bc=(1, 1, 1, 1, 2, 2))
    pol = O.policy(null_space_constant=0, maximum_levels=2)
    (_, _, _, _), mg = O.from_case(case, pol)
    fine_r = mg.slot(0, 7)
    fine_r[:] = np.random.default_rng(2).standard_normal(fine_r.size)
    mg_r = fine_r.copy()
    assert O.lib().or_mg_restrict(mg.h, 0) == 0
    coarse_b = mg.slot(1, 6)
    assert math.isclose(coarse_b.sum(), mg_r.sum(), rel_tol=1e-12)
    coarse_x = mg.slot(1, 5)
    coarse_x[:] = 3.25
    fine_x = mg.slot(0, 5)
    fine_x[:] = 0.0
    assert O.lib().or_mg_prolongate(mg.h, 0) == 0
    assert np.allclose(mg.slot(0, 5), 3.25, rtol=0, atol=1e-15)


def test_coarse_operator_preserves_constants():
    """Face aggregation keeps A_c 1 = 0 on pure Neumann/periodic levels and
    inactive coarse rows exactly identity."""
    case = small_case(n=(8, 8, 8), seed=4, bc=(1, 1, 1, 1, 2, 2))
    (_, _, _, _), mg = O.from_case(case, O.policy())
    for level in range(1, mg.level_count()):
        nx, ny, nz = mg.level_info(level)["shape"]
        D = mg.level_array(level, 0)
        fx = mg.level_array(level, 1).reshape(nz, ny, nx + 1)
        fy = mg.level_array(level, 2).reshape(nz, ny + 1, nx)
        fz = mg.level_array(level, 3).reshape(nz + 1, ny, nx)
        vol = mg.level_array(level, 4)
        faces = (fx[:, :, :-1] + fx[:, :, 1:] + fy[:, :-1, :] + fy[:, 1:, :]
                 + fz[:-1] + fz[1:]).ravel()
        fluid = vol > 0
        assert np.allclose(D[fluid], faces[fluid], rtol=1e-13)
        assert np.all(D[~fluid] == 1.0) and np.all(faces[~fluid] == 0.0)


@pytest.mark.parametrize("order", [0, 1])
@pytest.mark.parametrize("reverse", [0, 1])
def test_line_smoother_is_block_gauss_seidel(order, reverse):
    """One line sweep of the oracle equals block Gauss-Seidel over the y-lines
    of the dense level-0 matrix (each block solved by LAPACK), visited in
    lexicographic order (b outer, a inner; reverse: both descending) or in
    zebra colour passes ((a+b) even first; reverse: odd first)."""
    case = hf_cases.flame_cylinder(n=(6, 10, 5), seed=61)
    n = case["n"]
    bc = [0, 0, 1, 1, 0, 0]
    act = case["cell_active"].reshape(n[2], n[1], n[0])
    fx, fy, fz = hf_cases.face_activity(act, bc)
    case.update(bc=bc, x_active=fx, y_active=fy, z_active=fz)
    _, mg = O.from_case(case, O.policy(line_order=order,
                                       null_space_constant=0))
    assert mg.level_info(0)["line_mask"] == 2
    N = n[0] * n[1] * n[2]
    A = np.zeros((N, N))
    for c in range(N):
        e = np.zeros(N)
        e[c] = 1.0
        A[:, c] = mg.masked_apply(e)
    rng = np.random.default_rng(62)
    x0 = rng.standard_normal(N)
    b = rng.standard_normal(N)
    mg.slot(0, 5)[:] = x0
    mg.slot(0, 6)[:] = b
    mg.smooth(0, 1, reverse, 0)
    got = mg.slot(0, 5).copy()
    nx, ny, nz = n
    lines = [(a, bb) for bb in range(nz) for a in range(nx)]
    if order == 0:
        seq = lines[::-1] if reverse else lines
    else:
        first = 1 if reverse else 0
        seq = ([l for l in lines if (l[0] + l[1]) % 2 == first] +
               [l for l in lines if (l[0] + l[1]) % 2 != first])
    x = x0.copy()
    for a, bb in seq:
        idx = np.array([a + nx * (j + ny * bb) for j in range(ny)])
        rest = np.setdiff1d(np.arange(N), idx)
        rhs = b[idx] - A[np.ix_(idx, rest)] @ x[rest]
        x[idx] = np.linalg.solve(A[np.ix_(idx, idx)], rhs)
    assert np.allclose(got, x, rtol=1e-12, atol=1e-12 * np.abs(x).max())
    # the two orders really differ
    if order == 0 and reverse == 0:
        _, mz = O.from_case(case, O.policy(line_order=1,
                                           null_space_constant=0))
        mz.slot(0, 5)[:] = x0
        mz.slot(0, 6)[:] = b
        mz.smooth(0, 1, 0, 0)
        assert not np.allclose(mz.slot(0, 5), got, rtol=1e-6)
