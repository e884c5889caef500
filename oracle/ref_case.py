"""Case files and outputs of the reference driver (TEST INFRASTRUCTURE).

oracle/_ref/hf_ref_driver is the reference's own hot path compiled from the
reference sources (oracle/ref.mk) around oracle/ref_driver.cpp.  This module
writes its binary case files and reads its outputs.  Only tests/golden/
make_golden.py (the fixtures) and the cpu_baseline leg of bench.py (timing
only) use it; the product never does.
"""
import os
import struct
import subprocess
import tempfile

import numpy as np

DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_ref",
                      "hf_ref_driver")


def write_case(path, case, pol, r, b, rtol, maxit, interval, solver=0,
               restart=0, volume_rows=0, repeats=0):
    n = case["n"]
    with open(path, "wb") as f:
        f.write(b"HFREF002")
        f.write(struct.pack("<3i", *n))
        f.write(struct.pack("<6i", *case["bc"]))
        f.write(struct.pack("<d", pol.anisotropy_threshold))
        f.write(struct.pack("<8i", pol.pre_sweeps, pol.post_sweeps,
                            pol.maximum_levels, pol.coarse_sweeps,
                            pol.minimum_coarse_extent, pol.f_cycle,
                            pol.null_space_constant, pol.unit_linear))
        f.write(struct.pack("<i", 1))
        f.write(struct.pack("<dii", rtol, maxit, interval))
        f.write(struct.pack("<i", pol.point_smoother))
        f.write(struct.pack("<dd", pol.chebyshev_lower_fraction,
                            pol.diagonal_shortcut_ratio))
        f.write(struct.pack("<4i", solver, restart, volume_rows, repeats))
        for a in case["faces"]:
            f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())
        for a in case["coeffs"]:
            f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())
        for k in ("cell_active", "x_active", "y_active", "z_active"):
            f.write(np.ascontiguousarray(case[k], dtype=np.uint8).tobytes())
        f.write(np.ascontiguousarray(r, dtype="<f8").tobytes())
        f.write(np.ascontiguousarray(b, dtype="<f8").tobytes())


def read_output(path, timing_only=False):
    data = open(path, "rb").read()
    assert data[:8] == b"HFOUT002"
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

    def int64s(k):
        nonlocal off
        v = struct.unpack_from(f"<{k}q", data, off)
        off += 8 * k
        return v

    (levels,) = ints(1)
    (shortcut,) = ints(1)
    plan = np.array([ints(5) for _ in range(levels)], dtype=np.int32)
    cells = [int(p[0] * p[1] * p[2]) for p in plan]
    out = {"plan": plan, "shortcut": np.int32(shortcut)}
    if not timing_only:
        for l in range(levels):
            out[f"diag_{l}"] = doubles(cells[l])
        out["z"] = doubles(cells[0])
        for l in range(levels):
            for s in ("x", "b", "r"):
                out[f"slot_{s}_{l}"] = doubles(cells[l])
    status, termination, iterations = ints(3)
    out["krylov_status"] = np.int32(status)
    out["krylov_termination"] = np.int32(termination)
    out["krylov_iterations"] = np.int32(iterations)
    (out["krylov_initial"], out["krylov_final"],
     out["krylov_rhs_norm"]) = doubles(3)
    applies = int64s(3)
    out["krylov_applies"] = np.array(applies, dtype=np.int64)
    if not timing_only:
        out["krylov_x"] = doubles(cells[0])
    (repeats,) = ints(1)
    out["seconds"] = doubles(repeats)
    if timing_only:
        (out["compile_seconds"],) = doubles(1)
    assert off == len(data)
    return out


def run_driver(case, pol, r, b, rtol=1e-12, maxit=400, interval=4, solver=0,
               restart=0, volume_rows=0, repeats=0, timing_only=False):
    """Runs the reference driver on one case; timing_only skips every array
    output (the CPU-baseline mode) and returns the timings."""
    env = dict(os.environ)
    if timing_only:
        env["HF_REF_TIMING_ONLY"] = "1"
    with tempfile.TemporaryDirectory() as tmp:
        cin = os.path.join(tmp, "case.bin")
        cout = os.path.join(tmp, "out.bin")
        write_case(cin, case, pol, r, b, rtol, maxit, interval, solver,
                   restart, volume_rows, repeats)
        subprocess.run([DRIVER, cin, cout], check=True, env=env)
        return read_output(cout, timing_only)
