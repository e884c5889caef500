"""Pins of the two oracle functions that are not in the reference.

* The distance-corrected coarse faces (policy coarse_face_scaling = 1): on a
  uniform grid whose mobility is constant over the cells that coarsen
  together, the finite-volume rediscretisation of the coarse grid is known in
  closed form -- the face rule of the fine grid applied to the coarse grid
  (pinned on its own by the face-rule tests in test_oracle.py).  Each
  distance-corrected coarse level must equal it, and the reference
  aggregation (scaling 0) must equal it times 2 for every coarsening of the
  face's normal axis so far (tangential coarsening leaves it unchanged).
* The flexible (Polak-Ribiere) beta of the PCG: with a symmetric positive
  definite preconditioner every PCG iterate is the A-norm minimiser over the
  preconditioned Krylov space, for both beta forms (checked against a dense
  Galerkin projection, not against the recurrence); with a non-symmetric
  preconditioner the flexible form keeps consecutive directions A-conjugate,
  which the reference form does not.
"""
import numpy as np
import pytest

import hf_cases
import hf_oracle as O


def _block_mobility(n, block, seed):
    """Mobility constant over blocks of `block` cells (x, y, z)."""
    rng = np.random.default_rng(seed)
    nb = [n[a] // block[a] for a in range(3)]
    coarse = np.exp(0.7 * rng.standard_normal((nb[2], nb[1], nb[0])))
    fine = np.repeat(np.repeat(np.repeat(coarse, block[2], axis=0),
                               block[1], axis=1), block[0], axis=2)
    return fine.ravel(), coarse


def _coarse_mobility(fine, n, factor):
    m = fine.reshape(n[2], n[1], n[0])
    return m[::factor[2], ::factor[1], ::factor[0]].ravel()


CASES = [
    # (fine extents, boundary kinds, levels checked)
    ((16, 8, 16), (1, 1, 1, 1, 2, 2), 2),
    ((16, 16, 8), (0, 0, 1, 1, 2, 2), 2),
    ((8, 16, 16), (2, 2, 2, 2, 2, 2), 2),
    ((16, 16, 3), (1, 1, 0, 1, 1, 1), 2),  # z never coarsened (extent 3)
]


@pytest.mark.parametrize("n,bc,levels", CASES,
                         ids=["neumann_periodic", "dirichlet_x",
                              "all_periodic", "semi_xy"])
@pytest.mark.parametrize("scaling", [0, 1])
def test_coarse_faces_against_rediscretisation(n, bc, levels, scaling):
    n = tuple(n)
    block = tuple(4 if n[a] > 3 else 1 for a in range(3))
    mob, _ = _block_mobility(n, block, seed=sum(n) + 7 * scaling)
    faces = [hf_cases.uniform_faces(v) for v in n]
    d, bx, by, bz = O.assemble(n, faces, list(bc), mob)
    mg = O.MG(n, faces, list(bc), O.policy(coarse_face_scaling=scaling,
                                           null_space_constant=0),
              d, bx, by, bz)
    assert mg.level_count() > levels
    factor = [1, 1, 1]
    for level in range(1, levels + 1):
        info = mg.level_info(level - 1)
        mask = info["coarsen_mask"]
        for a in range(3):
            if mask & (1 << a):
                factor[a] *= 2
        cn = mg.level_info(level)["shape"]
        assert all(cn[a] * factor[a] == n[a] for a in range(3))
        cm = _coarse_mobility(mob, n, factor)
        cfaces = [hf_cases.uniform_faces(v) for v in cn]
        rd, rbx, rby, rbz = O.assemble(cn, cfaces, list(bc), cm)
        # the reference sums faces: each coarsening of the normal axis
        # doubles the face relative to the rediscretisation (2^levels after
        # several); scaling 1 removes exactly that factor each time
        want = []
        for a, r in enumerate((rbx, rby, rbz)):
            scale = 1.0 if scaling else float(factor[a])
            want.append(r * scale)
        for a in range(3):
            got = mg.level_array(level, 1 + a)
            assert np.allclose(got, want[a], rtol=1e-13, atol=0.0), (level, a)
        # the coarse diagonal is the sum of its faces (no reaction term)
        got_d = mg.level_array(level, 0)
        cnx, cny, cnz = cn
        fx = want[0].reshape(cnz, cny, cnx + 1)
        fy = want[1].reshape(cnz, cny + 1, cnx)
        fz = want[2].reshape(cnz + 1, cny, cnx)
        face_sum = (fx[:, :, :-1] + fx[:, :, 1:] + fy[:, :-1, :] +
                    fy[:, 1:, :] + fz[:-1] + fz[1:]).ravel()
        assert np.allclose(got_d, face_sum, rtol=1e-13, atol=0.0)
        if scaling:
            assert np.allclose(got_d, rd, rtol=1e-13, atol=0.0)


@pytest.mark.parametrize("scaling", [0, 1])
def test_coarse_operator_null_space_and_dominance(scaling):
    """Every coarse level of a Neumann/periodic case with solids annihilates
    constants on its fluid rows and is a diagonally dominant M-matrix."""
    case = hf_cases.lognormal_sphere(n=(16, 16, 16), seed=21)
    (_, _, _, _), mg = O.from_case(case, O.policy(coarse_face_scaling=scaling))
    for level in range(1, mg.level_count()):
        nx, ny, nz = mg.level_info(level)["shape"]
        D = mg.level_array(level, 0)
        fx = mg.level_array(level, 1).reshape(nz, ny, nx + 1)
        fy = mg.level_array(level, 2).reshape(nz, ny + 1, nx)
        fz = mg.level_array(level, 3).reshape(nz + 1, ny, nx)
        vol = mg.level_array(level, 4)
        assert np.all(fx >= 0) and np.all(fy >= 0) and np.all(fz >= 0)
        faces = (fx[:, :, :-1] + fx[:, :, 1:] + fy[:, :-1, :] + fy[:, 1:, :]
                 + fz[:-1] + fz[1:]).ravel()
        fluid = vol > 0
        assert np.allclose(D[fluid], faces[fluid], rtol=1e-13)
        assert np.all(D[~fluid] == 1.0)


def _spd(n, seed, cond=50.0):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    w = np.geomspace(1.0, cond, n)
    return (q * w) @ q.T


def _krylov_minimiser(a, m, r0, x0, k):
    """x0 + argmin over K_k(MA, M r0) of the A-norm error (dense Galerkin)."""
    n = a.shape[0]
    basis = np.zeros((n, k))
    v = m @ r0
    for j in range(k):
        basis[:, j] = v
        v = m @ (a @ v)
    q, _ = np.linalg.qr(basis)
    y = np.linalg.solve(q.T @ a @ q, q.T @ r0)
    return x0 + q @ y


@pytest.mark.parametrize("flexible", [False, True])
@pytest.mark.parametrize("precond", ["identity", "spd"])
def test_pcg_iterates_are_krylov_minimisers(flexible, precond):
    n = 24
    a = _spd(n, seed=3)
    m = None if precond == "identity" else np.linalg.inv(_spd(n, seed=4, cond=8))
    m = None if m is None else 0.5 * (m + m.T)
    rng = np.random.default_rng(5)
    b = rng.standard_normal(n)
    x0 = rng.standard_normal(n)
    _, res, it = O.pcg_dense(a, m, b, x0=x0, rtol=1e-13, flexible=flexible,
                             trace=8)
    assert res.status == 0 and res.iterations >= 8
    mm = np.eye(n) if m is None else m
    r0 = b - a @ x0
    xs = np.linalg.solve(a, b)
    for k in range(1, 9):
        want = _krylov_minimiser(a, mm, r0, x0, k)
        err = np.sqrt((it[k - 1] - want) @ a @ (it[k - 1] - want))
        scale = np.sqrt((x0 - xs) @ a @ (x0 - xs))
        assert err <= 1e-9 * scale, (k, err / scale)


def test_flexible_beta_keeps_directions_conjugate():
    """A non-symmetric preconditioner: the flexible beta makes each new
    direction A-orthogonal to the previous one (the Polak-Ribiere
    condition); the reference beta does not."""
    n = 24
    a = _spd(n, seed=6)
    rng = np.random.default_rng(8)
    m = np.linalg.inv(_spd(n, seed=7, cond=8)) + 0.3 * rng.standard_normal((n, n)) / np.sqrt(n)
    b = rng.standard_normal(n)

    def conjugacy(flexible):
        _, res, it = O.pcg_dense(a, m, b, rtol=1e-14, flexible=flexible,
                                 trace=6, maximum_iterations=6)
        xs = np.vstack([np.zeros(n), it])
        d = np.diff(xs, axis=0)
        worst = 0.0
        for k in range(len(d) - 1):
            num = abs(d[k + 1] @ a @ d[k])
            den = np.sqrt((d[k + 1] @ a @ d[k + 1]) * (d[k] @ a @ d[k]))
            worst = max(worst, num / den)
        return worst

    assert conjugacy(True) <= 1e-12
    assert conjugacy(False) >= 1e-4


def test_flexible_and_reference_beta_agree_on_symmetric_cycle():
    """The FP64 V(2,2) cycle with unit scaling is a fixed, nearly symmetric
    preconditioner: both beta forms take the same number of iterations and
    reach the same solution."""
    case = hf_cases.lognormal_sphere(n=(24, 24, 24), seed=23)
    (d, bx, by, bz), mg = O.from_case(case, O.policy())
    args = (case["n"], case["bc"], d, bx, by, bz, case["cell_active"], mg,
            case["rhs"])
    x0, r0 = O.pcg(*args, rtol=1e-10, flexible=False)
    x1, r1 = O.pcg(*args, rtol=1e-10, flexible=True)
    assert r0.status == 0 and r1.status == 0
    assert r0.iterations == r1.iterations
    act = case["cell_active"] != 0
    g = lambda v: np.where(act, v - v[act].mean(), 0.0)
    assert np.linalg.norm(g(x0) - g(x1)) <= 1e-8 * np.linalg.norm(g(x0))
