"""Writes golden fixtures from the reference's own compiled hot path.

Runs oracle/_ref/hf_ref_driver (the reference NativeCartesianMgPlan,
solve_pcg and solve_fgmres, compiled from the reference sources by
oracle/ref.mk) on small seeded cases and stores inputs and outputs as .npz
fixtures next to this script.  Inputs: grid, boundary kinds, policy,
coefficients (D, beta_x, beta_y, beta_z from the face rule, or the
compressible pressure system's D), IBM activity, an MG input residual r and
a Krylov right-hand side b.  Outputs: the level plan, whether the plan chose
the diagonal shortcut, every level's diagonal, z = M r with the
solution/rhs/residual slot of every level after that application, and the
Krylov solution with its iteration and application counts.

Usage: python tests/golden/make_golden.py   (needs the reference tree)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "hundun-flow_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import hf_cases  # noqa: E402
import hf_oracle as O  # noqa: E402
import ref_case  # noqa: E402
from ref_case import run_driver  # noqa: E402



def with_bc(case, bc):
    n = case["n"]
    act = case["cell_active"].reshape(n[2], n[1], n[0])
    fx, fy, fz = hf_cases.face_activity(act, bc)
    case.update(bc=bc, x_active=fx, y_active=fy, z_active=fz)
    return case


def line_case(n, seed, stretch_axis, bc, ratio=8.0, depth=1.0):
    """Lognormal rho with a sphere mask on a grid tanh-stretched along one
    axis (line relaxation under the default anisotropy threshold); `depth`
    scales the z extent (a deep z makes x and y both strong)."""
    rng = np.random.Generator(np.random.Philox(seed))
    faces = []
    for a in range(3):
        f = (hf_cases.tanh_faces(n[a], ratio) if a == stretch_axis
             else hf_cases.uniform_faces(n[a]))
        if a == 2:
            f = hf_cases.real4(f * depth)
        faces.append(f)
    nx, ny, nz = n
    xi = rng.standard_normal((nz, ny, nx))
    rho = np.exp(hf_cases.box_filter(0.3 * xi, (False, False, False), 2))
    X, Y, Z = hf_cases._grid(faces)
    zc = 0.5 * depth
    solid = ((X - 0.45) ** 2 + (Y - 0.55) ** 2 + ((Z - zc) / depth) ** 2
             < 0.2 ** 2)
    rhs = rng.standard_normal((nz, ny, nx))
    return hf_cases._pack(n, faces, bc, rho, ~solid, rhs)


def cases():
    mixed = [0, 0, 1, 1, 2, 2]
    yield "mixed_sphere", with_bc(
        hf_cases.lognormal_sphere(n=(16, 12, 10), seed=21, radius=0.25),
        mixed), dict(null_space_constant=0)
    yield "odd_periodic", with_bc(
        hf_cases.lognormal_sphere(n=(20, 12, 18), seed=22, radius=0.2),
        mixed), dict(null_space_constant=0)
    yield "sphere_null_space", hf_cases.lognormal_sphere(
        n=(16, 16, 16), seed=23), dict()
    yield "step_fcycle_resmin", hf_cases.flame_step(
        n=(24, 12, 8), seed=24), dict(f_cycle=1, unit_linear=0)
    yield "stretched_point", hf_cases.flame_cylinder(
        n=(12, 16, 8), seed=25), dict(anisotropy_threshold=16.0)
    yield "neumann_all", with_bc(
        hf_cases.lognormal_sphere(n=(12, 10, 14), seed=26, radius=0.3),
        [1, 1, 1, 1, 1, 1]), dict(coarse_sweeps=6)
    # line relaxation (default anisotropy threshold 4)
    yield "line_y_cylinder", hf_cases.flame_cylinder(
        n=(12, 16, 8), seed=27), dict()
    yield "line_y_periodic", with_bc(hf_cases.flame_cylinder(
        n=(11, 16, 8), seed=28), [2, 2, 2, 2, 1, 1]), dict(coarse_sweeps=8)
    yield "line_x_dirichlet", line_case(
        (16, 10, 9), 29, 0, [0, 0, 1, 1, 2, 2]), dict(null_space_constant=0)
    yield "line_z_fcycle", line_case(
        (10, 9, 16), 30, 2, [1, 1, 0, 1, 1, 1]), dict(
            null_space_constant=0, f_cycle=1, unit_linear=0)
    yield "line_xy_deep", line_case(
        (12, 10, 6), 31, -1, [1, 1, 2, 2, 1, 1], depth=6.0), dict(
            coarse_sweeps=6)
    # the production pairing: Chebyshev-Jacobi F(1,2) cycles, diagonal
    # shortcut, FGMRES(12), the volume-scaled cold pairing
    yield "cheb_mixed", with_bc(
        hf_cases.lognormal_sphere(n=(16, 12, 10), seed=32, radius=0.25),
        mixed), dict(null_space_constant=0, point_smoother=1, pre_sweeps=1,
                     post_sweeps=2)
    yield "cheb_fcycle_null_space", hf_cases.lognormal_sphere(
        n=(16, 16, 12), seed=33), dict(point_smoother=1, f_cycle=1,
                                        chebyshev_lower_fraction=0.25)
    yield "fgmres_rb_sphere", with_bc(
        hf_cases.lognormal_sphere(n=(16, 12, 10), seed=34, radius=0.25),
        mixed), dict(null_space_constant=0, solver=1, restart=12)
    yield "fgmres_production", compressible(
        (20, 16, 12), 35, 6.0), dict(production=True, solver=1, restart=12)
    yield "fgmres_restart3_resmin", compressible(
        (16, 12, 12), 36, 12.0), dict(production=True, solver=1, restart=3,
                                      unit_linear=0)
    yield "fgmres_volume_rows", compressible(
        (16, 16, 10), 37, 6.0), dict(production=True, solver=1, restart=12,
                                     volume_rows=1)
    yield "shortcut_fgmres", compressible(
        (16, 12, 10), 38, 0.05), dict(production=True, solver=1, restart=12)


def compressible(n, seed, cfl):
    """The compressible pressure-correction system of
    hf_cases.compressible_flame (D = a0 V drho/dp + face sum) assembled by
    the oracle's restatement of src/solver_pressure.cpp:110-332."""
    case = hf_cases.compressible_flame(n=n, seed=seed, acoustic_cfl=cfl)
    d0, bx, by, bz = O.assemble(case["n"], case["faces"], case["bc"],
                                case["mobility"])
    st, D, b = O.assemble_pressure_system(
        case["n"], case["faces"], bx, by, bz, case["phi"], case["rho"],
        case["rho_n"], case["drho_dp"], case["bdf"], case["rho_nm1"],
        case["source"])
    assert st == 0
    case["coeffs"] = (D, bx, by, bz)
    case["rhs"] = np.where(case["cell_active"] != 0, b, 0.0)
    return case


def main():
    if not os.path.exists(ref_case.DRIVER):
        raise SystemExit("build the reference closure first: "
                         "make -C oracle -f ref.mk")
    for name, case, overrides in cases():
        overrides = dict(overrides)
        solver = overrides.pop("solver", 0)
        restart = overrides.pop("restart", 0)
        volume_rows = overrides.pop("volume_rows", 0)
        pol = (O.production_policy(**{k: v for k, v in overrides.items()
                                      if k != "production"})
               if overrides.get("production") else O.policy(**overrides))
        if "coeffs" not in case:
            case["coeffs"] = O.assemble(case["n"], case["faces"], case["bc"],
                                        case["mobility"])
        cells = case["n"][0] * case["n"][1] * case["n"][2]
        rng = np.random.Generator(np.random.Philox(1000 + len(name)))
        r = rng.standard_normal(cells)
        b = case["rhs"]
        if not pol.null_space_constant and "phi" not in case:
            b = np.where(case["cell_active"] != 0,
                         rng.standard_normal(cells), 0.0)
        rtol, maxit, interval = 1e-12, 400, 4
        out = run_driver(case, pol, r, b, rtol, maxit, interval, solver,
                         restart, volume_rows)
        fixture = {
            "n": np.array(case["n"], dtype=np.int32),
            "bc": np.array(case["bc"], dtype=np.int32),
            "policy": np.array([pol.anisotropy_threshold, pol.pre_sweeps,
                                pol.post_sweeps, pol.maximum_levels,
                                pol.coarse_sweeps, pol.minimum_coarse_extent,
                                pol.f_cycle, pol.null_space_constant,
                                pol.unit_linear, pol.point_smoother,
                                pol.chebyshev_lower_fraction,
                                pol.diagonal_shortcut_ratio],
                               dtype=np.float64),
            "faces_x": case["faces"][0], "faces_y": case["faces"][1],
            "faces_z": case["faces"][2],
            "diag": case["coeffs"][0], "bx": case["coeffs"][1],
            "by": case["coeffs"][2], "bz": case["coeffs"][3],
            "cell_active": case["cell_active"],
            "x_active": case["x_active"], "y_active": case["y_active"],
            "z_active": case["z_active"], "r": r, "b": b,
            "krylov_control": np.array([rtol, maxit, interval, solver,
                                        restart, volume_rows]),
        }
        out.pop("seconds")
        fixture.update(out)
        np.savez_compressed(os.path.join(HERE, f"ref_{name}.npz"), **fixture)
        print(f"{name}: levels={len(out['plan'])} "
              f"{'fgmres' if solver else 'pcg'} "
              f"it={int(out['krylov_iterations'])} "
              f"status={int(out['krylov_status'])} "
              f"shortcut={int(out['shortcut'])}")


if __name__ == "__main__":
    main()
