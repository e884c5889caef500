"""Writes golden fixtures from the reference's own compiled hot path.

Runs oracle/_ref/hf_ref_driver (the reference NativeCartesianMgPlan and
solve_pcg, compiled from the reference sources by oracle/ref.mk) on small
seeded cases and stores inputs and outputs as .npz fixtures next to this
script.  Inputs: grid, boundary kinds, policy, coefficients (D, beta_x,
beta_y, beta_z from the face rule), IBM activity, an MG input residual r and
a PCG right-hand side b.  Outputs: the level plan, every level's diagonal,
z = M r with the solution/rhs/residual slot of every level after that
application, and the PCG solution with its iteration count.

Usage: python tests/golden/make_golden.py   (needs the reference tree)
"""
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "hundun-flow_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import hf_cases  # noqa: E402
import hf_oracle as O  # noqa: E402

DRIVER = os.path.join(ROOT, "oracle", "_ref", "hf_ref_driver")


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


def write_case(path, case, pol, r, b, rtol, maxit, interval):
    n = case["n"]
    with open(path, "wb") as f:
        f.write(b"HFREF001")
        f.write(struct.pack("<3i", *n))
        f.write(struct.pack("<6i", *case["bc"]))
        f.write(struct.pack("<d", pol.anisotropy_threshold))
        f.write(struct.pack("<8i", pol.pre_sweeps, pol.post_sweeps,
                            pol.maximum_levels, pol.coarse_sweeps,
                            pol.minimum_coarse_extent, pol.f_cycle,
                            pol.null_space_constant, pol.unit_linear))
        f.write(struct.pack("<i", 1))
        f.write(struct.pack("<dii", rtol, maxit, interval))
        for a in case["faces"]:
            f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())
        for a in case["coeffs"]:
            f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())
        for k in ("cell_active", "x_active", "y_active", "z_active"):
            f.write(np.ascontiguousarray(case[k], dtype=np.uint8).tobytes())
        f.write(np.ascontiguousarray(r, dtype="<f8").tobytes())
        f.write(np.ascontiguousarray(b, dtype="<f8").tobytes())


def read_output(path):
    data = open(path, "rb").read()
    assert data[:8] == b"HFOUT001"
    off = 8

    def ints(k):
        nonlocal off
        v = struct.unpack_from(f"<{k}i", data, off)
        off += 4 * k
        return v

    def doubles(k):
        nonlocal off
        v = np.frombuffer(data, dtype="<f8", count=k, offset=off).copy()
        off += 8 * k
        return v

    (levels,) = ints(1)
    plan = np.array([ints(5) for _ in range(levels)], dtype=np.int32)
    cells = [int(p[0] * p[1] * p[2]) for p in plan]
    out = {"plan": plan}
    for l in range(levels):
        out[f"diag_{l}"] = doubles(cells[l])
    out["z"] = doubles(cells[0])
    for l in range(levels):
        for s in ("x", "b", "r"):
            out[f"slot_{s}_{l}"] = doubles(cells[l])
    status, termination, iterations = ints(3)
    out["pcg_status"] = np.int32(status)
    out["pcg_termination"] = np.int32(termination)
    out["pcg_iterations"] = np.int32(iterations)
    out["pcg_initial"], out["pcg_final"], out["pcg_rhs_norm"] = doubles(3)
    out["pcg_x"] = doubles(cells[0])
    assert off == len(data)
    return out


def main():
    if not os.path.exists(DRIVER):
        raise SystemExit("build the reference closure first: "
                         "make -C oracle -f ref.mk")
    for name, case, overrides in cases():
        pol = O.policy(**overrides)
        case["coeffs"] = O.assemble(case["n"], case["faces"], case["bc"],
                                    case["mobility"])
        cells = case["n"][0] * case["n"][1] * case["n"][2]
        rng = np.random.Generator(np.random.Philox(1000 + len(name)))
        r = rng.standard_normal(cells)
        b = case["rhs"]
        if not pol.null_space_constant:
            b = np.where(case["cell_active"] != 0,
                         rng.standard_normal(cells), 0.0)
        rtol, maxit, interval = 1e-12, 400, 4
        with tempfile.TemporaryDirectory() as tmp:
            cin = os.path.join(tmp, "case.bin")
            cout = os.path.join(tmp, "out.bin")
            write_case(cin, case, pol, r, b, rtol, maxit, interval)
            subprocess.run([DRIVER, cin, cout], check=True)
            out = read_output(cout)
        fixture = {
            "n": np.array(case["n"], dtype=np.int32),
            "bc": np.array(case["bc"], dtype=np.int32),
            "policy": np.array([pol.anisotropy_threshold, pol.pre_sweeps,
                                pol.post_sweeps, pol.maximum_levels,
                                pol.coarse_sweeps, pol.minimum_coarse_extent,
                                pol.f_cycle, pol.null_space_constant,
                                pol.unit_linear], dtype=np.float64),
            "faces_x": case["faces"][0], "faces_y": case["faces"][1],
            "faces_z": case["faces"][2],
            "diag": case["coeffs"][0], "bx": case["coeffs"][1],
            "by": case["coeffs"][2], "bz": case["coeffs"][3],
            "cell_active": case["cell_active"],
            "x_active": case["x_active"], "y_active": case["y_active"],
            "z_active": case["z_active"], "r": r, "b": b,
            "pcg_control": np.array([rtol, maxit, interval]),
        }
        fixture.update(out)
        np.savez_compressed(os.path.join(HERE, f"ref_{name}.npz"), **fixture)
        print(f"{name}: levels={len(out['plan'])} pcg_it="
              f"{int(out['pcg_iterations'])} status={int(out['pcg_status'])}")


if __name__ == "__main__":
    main()
