"""hf_pcg_solve_static: the PCG entry point that allocates nothing.

Its reduction scratch lives in the library's static per-GPU array instead of
the plan's buffers; everything else is the hf_pcg_solve path, so both entry
points must give the same solution bit for bit, with the same termination,
iteration and application counts.  Two plans are solved back to back, the
second on a smaller grid, so the second call runs on an array that still
holds the first call's values.  A grid whose reductions do not fit the array
is refused with invalid_plan (detail 9110) before any device work.
"""
import ctypes
import sys

import numpy as np
import pytest

import hf_cases

pytestmark = pytest.mark.gpu

# (grid, precision, flexible beta): the second grid is the smaller one
CALLS = [((48, 40, 36), "fp32", True), ((20, 16, 12), "fp64", False)]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _solve(lib, plan, b, static, flexible):
    import hundun_flow as H
    import torch
    x = torch.zeros_like(b)
    ctl = plan.control(rtol=1e-10, maximum_iterations=200, flexible=flexible)
    res = H.hf_solve_result()
    fn = lib.hf_pcg_solve_static if static else lib.hf_pcg_solve
    st = fn(plan.handle, ctypes.c_void_p(b.data_ptr()),
            ctypes.c_void_p(x.data_ptr()), ctypes.byref(ctl),
            ctypes.byref(res))
    torch.cuda.synchronize()
    return st, x, H.SolveResult.from_c(res)


def test_static_solve_matches_pcg_solve_twice(torch_cuda):
    import hundun_flow as H
    torch = torch_cuda
    lib = H.load_library()
    for n, precision, flexible in CALLS:
        _say(f"[no-alloc] grid={n} precision={precision} flexible={flexible}")
        case = hf_cases.flame_step(n=n, seed=61)
        plan = H.plan_from_case(case, H.default_policy(
            precision=precision, coarse_face_scaling=1))
        b = torch.from_numpy(case["rhs"]).cuda()
        # static first: its array may hold a previous call's values
        st_s, x_s, r_s = _solve(lib, plan, b, True, flexible)
        st_p, x_p, r_p = _solve(lib, plan, b, False, flexible)
        assert st_s.code == 0 and st_p.code == 0, (st_s.code, st_p.code)
        assert r_s.termination == r_p.termination == "converged"
        assert r_s.iterations == r_p.iterations
        assert r_s.operator_applies == r_p.operator_applies
        assert r_s.preconditioner_applies == r_p.preconditioner_applies
        assert r_s.final_true_residual == r_p.final_true_residual
        assert r_s.rhs_norm == r_p.rhs_norm
        assert torch.equal(x_s, x_p)
        _say(f"  it={r_s.iterations} true_rel="
             f"{r_s.final_true_residual / r_s.rhs_norm:.3e} bitwise equal")
        plan.close()


def test_static_solve_refuses_grid_beyond_the_array(torch_cuda):
    """1024x512x256 (134M cells) needs about 12.7 MB of reduction scratch,
    more than the 8 MiB array: refused, the caller's solution untouched."""
    import hundun_flow as H
    torch = torch_cuda
    lib = H.load_library()
    n = (1024, 512, 256)
    _say(f"[no-alloc refuse] grid={n}")
    faces = [hf_cases.uniform_faces(v, 0.0, v / 1024.0) for v in n]
    plan = H.Plan(n, faces, [1, 1, 1, 1, 2, 2])
    cells = n[0] * n[1] * n[2]
    mob = torch.ones(cells, dtype=torch.float64, device="cuda")
    plan.set_coefficients(*plan.assemble(mob))
    del mob
    b = torch.zeros(cells, dtype=torch.float64, device="cuda")
    b[::7] = 1.0
    x = torch.full_like(b, 3.0)
    ctl = plan.control()
    res = H.hf_solve_result()
    st = lib.hf_pcg_solve_static(plan.handle, ctypes.c_void_p(b.data_ptr()),
                                 ctypes.c_void_p(x.data_ptr()),
                                 ctypes.byref(ctl), ctypes.byref(res))
    torch.cuda.synchronize()
    assert (st.code, st.detail) == (2, 9110)
    assert H.TERMINATIONS[res.termination] == "invalid_plan"
    assert bool((x == 3.0).all())
    plan.close()
