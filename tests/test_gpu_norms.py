"""Scaled 2-norm where a plain sum of squares overflows.

The reference takes every norm of its Krylov solves with global_norm
(src/solver_krylov.cpp:650-711: max |v| times the root of the compensated
sum of (v / max |v|)^2), which stays finite while the sum of squares of v
overflows.  The device takes that scaled norm wherever its one-pass sum of
squares is not finite.  The system of the `mixed` case is scaled so that
||b||^2 overflows (b ~ 1e160) while the inner products of the iteration
stay finite (operator x 2^66 on the fluid rows, solid rows kept as the
identity).  The scales are powers of two, so every product, quotient and
inner product of the iteration is the unscaled one's scaled exactly: both
solvers take the same iterations, report the unscaled norms times 2^532 (to
rounding: the scaled norm sums in another order) and return the unscaled
solution times 2^466 bit for bit.  A non-finite right-hand side still ends
as non-finite.
"""
import math
import sys

import numpy as np
import pytest

import hf_cases

pytestmark = pytest.mark.gpu

SCALE_B = 2.0 ** 532  # about 1.4e160
SCALE_A = 2.0 ** 66   # about 7.4e19


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _mixed(n=(24, 20, 16), seed=11):
    case = hf_cases.lognormal_sphere(n=n, seed=seed, radius=0.22)
    bc = [0, 0, 1, 1, 2, 2]
    act = case["cell_active"].reshape(n[2], n[1], n[0])
    fx, fy, fz = hf_cases.face_activity(act, bc)
    case.update(bc=bc, x_active=fx, y_active=fy, z_active=fz)
    return case


def _plans(torch):
    import hundun_flow as H
    case = _mixed()
    pol = H.default_policy(null_space_constant=0)
    base = H.plan_from_case(case, pol)
    d, bx, by, bz, *masks = base.coefficients
    scaled = H.Plan(case["n"], case["faces"], case["bc"], pol)
    fluid = masks[0] != 0
    scaled.set_coefficients(torch.where(fluid, d * SCALE_A, d), bx * SCALE_A,
                            by * SCALE_A, bz * SCALE_A, *masks)
    act = masks[0].to(torch.float64)
    b = torch.from_numpy(case["rhs"]).cuda() * act
    return base, scaled, b


@pytest.mark.parametrize("solver", ["pcg", "fgmres"])
def test_overflowing_norms_take_the_scaled_norm(torch_cuda, solver):
    torch = torch_cuda
    base, scaled, b = _plans(torch)
    bs = b * SCALE_B
    assert math.isinf(float(torch.dot(bs, bs)))
    _say(f"[scaled norm] solver={solver}")

    def solve(plan, rhs):
        if solver == "pcg":
            return plan.solve_pcg(rhs, rtol=1e-10)
        return plan.solve_fgmres(rhs, rtol=1e-10, restart=12)

    x0, r0 = solve(base, b)
    x1, r1 = solve(scaled, bs)
    x0 = x0.cpu().numpy()
    x1 = x1.cpu().numpy() / (SCALE_B / SCALE_A)
    err = np.linalg.norm(x1 - x0) / np.linalg.norm(x0)
    _say(f"  it {r0.iterations}/{r1.iterations} rhs norm {r0.rhs_norm:.6e} / "
         f"{r1.rhs_norm:.6e} true residual {r1.final_true_residual:.3e} "
         f"err={err:.2e}")
    assert r0.termination == "converged"
    assert r1.termination == "converged"
    assert abs(r1.rhs_norm / SCALE_B - r0.rhs_norm) <= 1e-12 * r0.rhs_norm
    assert r1.final_true_residual <= 1e-10 * r1.rhs_norm
    assert r1.iterations == r0.iterations
    assert np.array_equal(x1, x0)
    base.close()
    scaled.close()


def test_non_finite_rhs_still_fails(torch_cuda):
    torch = torch_cuda
    base, scaled, b = _plans(torch)
    bad = b.clone()
    bad[5] = float("inf")
    _, res = base.solve_pcg(bad, rtol=1e-10, check=False)
    assert res.termination == "non_finite"
    _, res = base.solve_fgmres(bad, rtol=1e-10, check=False)
    assert res.termination == "non_finite"
    base.close()
    scaled.close()
