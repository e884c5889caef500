"""Pins the CPU oracle to the reference's own compiled hot path.

The fixtures under tests/golden/ were written by tests/golden/make_golden.py,
which runs the reference NativeCartesianMgPlan and solve_pcg (compiled from
the reference sources into oracle/_ref/).  The oracle restates the same
sequential FP64 arithmetic, so every comparison is bit-exact: the level plan,
each level's diagonal, one preconditioner application with every level's
solution/rhs/residual slot, and the PCG solution with its iteration count.
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


def oracle_mg(g):
    p = g["policy"]
    pol = O.policy(anisotropy_threshold=float(p[0]), pre_sweeps=int(p[1]),
                   post_sweeps=int(p[2]), maximum_levels=int(p[3]),
                   coarse_sweeps=int(p[4]), minimum_coarse_extent=int(p[5]),
                   f_cycle=int(p[6]), null_space_constant=int(p[7]),
                   unit_linear=int(p[8]))
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


@pytest.mark.parametrize("path", FIXTURES, ids=ids(FIXTURES))
def test_pcg_solution_bitwise(path):
    g = load(path)
    mg = oracle_mg(g)
    n = tuple(int(v) for v in g["n"])
    rtol, maxit, interval = g["pcg_control"]
    x, res = O.pcg(n, [int(v) for v in g["bc"]], g["diag"], g["bx"], g["by"],
                   g["bz"], g["cell_active"], mg, g["b"], rtol=float(rtol),
                   maximum_iterations=int(maxit),
                   true_residual_interval=int(interval))
    assert res.status == int(g["pcg_status"]) == 0
    assert res.iterations == int(g["pcg_iterations"])
    assert res.final_true_residual == float(g["pcg_final"])
    assert np.array_equal(x, g["pcg_x"]), np.max(np.abs(x - g["pcg_x"]))
