"""Pins the CPU oracle to the reference's own compiled hot path.

The fixtures under tests/golden/ were written by tests/golden/make_golden.py,
which runs the reference NativeCartesianMgPlan, solve_pcg and solve_fgmres
(compiled from the reference sources into oracle/_ref/).  The oracle restates
the same sequential FP64 arithmetic, so every comparison is bit-exact: the
level plan, the diagonal-shortcut choice, each level's diagonal, one
preconditioner application with every level's solution/rhs/residual slot
(red-black and Chebyshev-Jacobi smoothing, V and F cycles), and the Krylov
solution (PCG or FGMRES, plain or with volume-scaled rows) with its
iteration, operator and preconditioner counts.
"""
import glob
import os

import numpy as np
import pytest

import hf_oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "ref_*.npz")))


def load(path):
    with np.load(path, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def oracle_policy(g):
    p = g["policy"]
    return O.policy(anisotropy_threshold=float(p[0]), pre_sweeps=int(p[1]),
                    post_sweeps=int(p[2]), maximum_levels=int(p[3]),
                    coarse_sweeps=int(p[4]), minimum_coarse_extent=int(p[5]),
                    f_cycle=int(p[6]), null_space_constant=int(p[7]),
                    unit_linear=int(p[8]), point_smoother=int(p[9]),
                    chebyshev_lower_fraction=float(p[10]),
                    diagonal_shortcut_ratio=float(p[11]))


def oracle_mg(g):
    pol = oracle_policy(g)
    n = tuple(int(v) for v in g["n"])
    faces = [g["faces_x"], g["faces_y"], g["faces_z"]]
    return O.MG(n, faces, [int(v) for v in g["bc"]], pol, g["diag"], g["bx"],
                g["by"], g["bz"], g["cell_active"], g["x_active"],
                g["y_active"], g["z_active"])


def ids(paths):
    return [os.path.basename(p)[4:-4] for p in paths]


def test_fixtures_present():
    assert len(FIXTURES) >= 6


@pytest.mark.parametrize("path", FIXTURES, ids=ids(FIXTURES))
def test_level_plan_and_diagonals(path):
    g = load(path)
    mg = oracle_mg(g)
    plan = g["plan"]
    assert mg.level_count() == len(plan)
    assert mg.diagonal_shortcut() == bool(g["shortcut"])
    for level, row in enumerate(plan):
        info = mg.level_info(level)
        assert tuple(info["shape"]) == tuple(int(v) for v in row[:3])
        assert info["coarsen_mask"] == int(row[3])
        assert info["line_mask"] == int(row[4])
        assert np.array_equal(mg.level_array(level, 0), g[f"diag_{level}"])


@pytest.mark.parametrize("path", FIXTURES, ids=ids(FIXTURES))
def test_preconditioner_application_bitwise(path):
    g = load(path)
    mg = oracle_mg(g)
    z = mg.apply(g["r"])
    assert np.array_equal(z, g["z"]), np.max(np.abs(z - g["z"]))
    for level in range(len(g["plan"])):
        for which, slot in ((5, "x"), (6, "b"), (7, "r")):
            got = mg.level_array(level, which)
            want = g[f"slot_{slot}_{level}"]
            assert np.array_equal(got, want), (level, slot,
                                               np.max(np.abs(got - want)))


def test_fixture_coverage():
    """The fixture set covers both Krylov methods, both point smoothers,
    both cycles, the residual-minimising scale, the diagonal shortcut and
    the volume-scaled rows."""
    gs = [load(p) for p in FIXTURES]
    ctl = [g["krylov_control"] for g in gs]
    pol = [g["policy"] for g in gs]
    assert any(c[3] == 0 for c in ctl) and any(c[3] == 1 for c in ctl)
    assert any(p[9] == 1 for p in pol) and any(p[9] == 0 for p in pol)
    assert any(p[6] == 1 and p[9] == 1 for p in pol)
    assert any(c[3] == 1 and p[8] == 0 for c, p in zip(ctl, pol))
    assert any(int(g["shortcut"]) == 1 for g in gs)
    assert any(c[5] == 1 for c in ctl)
    assert any(c[3] == 1 and c[4] < 12 for c in ctl)


@pytest.mark.parametrize("path", FIXTURES, ids=ids(FIXTURES))
def test_krylov_solution_bitwise(path):
    g = load(path)
    mg = oracle_mg(g)
    n = tuple(int(v) for v in g["n"])
    bc = [int(v) for v in g["bc"]]
    rtol, maxit, interval, solver, restart, volume_rows = g["krylov_control"]
    if int(solver) == 0:
        x, res = O.pcg(n, bc, g["diag"], g["bx"], g["by"], g["bz"],
                       g["cell_active"], mg, g["b"], rtol=float(rtol),
                       maximum_iterations=int(maxit),
                       true_residual_interval=int(interval))
    else:
        faces = [g["faces_x"], g["faces_y"], g["faces_z"]]
        x, res = O.fgmres(n, faces, bc, g["diag"], g["bx"], g["by"], g["bz"],
                          g["cell_active"], mg, g["b"], rtol=float(rtol),
                          maximum_iterations=int(maxit),
                          true_residual_interval=int(interval),
                          restart=int(restart),
                          volume_rows=bool(volume_rows))
    assert res.status == int(g["krylov_status"]) == 0
    assert res.iterations == int(g["krylov_iterations"])
    ops, pcs, restarts = (int(v) for v in g["krylov_applies"])
    assert res.operator_applies == ops
    assert res.preconditioner_applies == pcs
    if int(solver) == 1:
        assert int(res.breakdown_value) == restarts
    assert res.final_true_residual == float(g["krylov_final"])
    assert np.array_equal(x, g["krylov_x"]), np.max(np.abs(x - g["krylov_x"]))
