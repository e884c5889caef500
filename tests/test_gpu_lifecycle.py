"""GPU plan lifecycle and Krylov breakdown gates through the C ABI.

* A coefficient refresh is transactional (src/solver_mg.cpp:4700-4770: the
  refresh builds into the inactive hierarchy and publishes it only on
  success): a rejected refresh leaves the previous hierarchy solving bit for
  bit as before, and an accepted one solves exactly as a fresh plan built
  from the same coefficients.
* A breakdown of r.z in the middle of the PCG loop (src/solver_krylov.cpp:
  1520-1544) ends the solve as breakdown with the oracle's iteration,
  operator and preconditioner counts.  The cases are weak cycles (one pre-
  and post-sweep, one or two coarse sweeps, residual-minimising scaling) on
  high-contrast coefficients, for which the oracle's r.z turns clearly
  negative (about -1e-2 of |b|^2).
"""
import sys

import numpy as np
import pytest

import hf_cases
import hf_oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _t(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _coefficients(torch, case):
    d, bx, by, bz = O.assemble(case["n"], case["faces"], case["bc"],
                               case["mobility"])
    masks = [_t(torch, case[k]) for k in ("cell_active", "x_active",
                                          "y_active", "z_active")]
    return [_t(torch, v) for v in (d, bx, by, bz)], masks


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_rejected_refresh_keeps_previous_hierarchy(torch_cuda, precision):
    import hundun_flow as H
    torch = torch_cuda
    _say(f"[refresh rejected] precision={precision}")
    case = hf_cases.lognormal_sphere(n=(32, 32, 32), seed=31)
    plan = H.Plan(case["n"], case["faces"], case["bc"],
                  H.default_policy(precision=precision))
    coeffs, masks = _coefficients(torch, case)
    plan.set_coefficients(*coeffs, *masks)
    b = _t(torch, case["rhs"])
    x1, r1 = plan.solve_pcg(b, rtol=1e-10, flexible=precision != "fp64")
    z1 = plan.mg_apply(b)
    # refreshes with a NaN face, an infinite face, and a face that makes its
    # row stop being diagonally dominant are rejected
    # (src/solver_mg.cpp:1965-2019)
    for poison in (float("nan"), float("inf"), 1e6):
        bad = [c.clone() for c in coeffs]
        bad[1][bad[1].numel() // 2 + 3] = poison
        with pytest.raises(H.HundunFlowError) as err:
            plan.set_coefficients(*bad, *masks)
        assert err.value.code == 5  # numerical_failure
        x2, r2 = plan.solve_pcg(b, rtol=1e-10, flexible=precision != "fp64")
        assert r2.termination == "converged"
        assert r2.iterations == r1.iterations
        assert torch.equal(x1, x2)
        assert torch.equal(z1, plan.mg_apply(b))
    plan.close()


@pytest.mark.parametrize("precision", ["fp64", "bf16"])
def test_accepted_refresh_equals_fresh_plan(torch_cuda, precision):
    import hundun_flow as H
    torch = torch_cuda
    _say(f"[refresh accepted] precision={precision}")
    first = hf_cases.lognormal_sphere(n=(32, 24, 16), seed=41)
    second = hf_cases.lognormal_sphere(n=(32, 24, 16), seed=42, radius=0.15)
    pol = H.default_policy(precision=precision)
    plan = H.Plan(first["n"], first["faces"], first["bc"], pol)
    c1, m1 = _coefficients(torch, first)
    plan.set_coefficients(*c1, *m1)
    c2, m2 = _coefficients(torch, second)
    plan.set_coefficients(*c2, *m2)  # into the spare hierarchy
    fresh = H.Plan(second["n"], second["faces"], second["bc"], pol)
    fresh.set_coefficients(*c2, *m2)
    for level in range(plan.level_count()):
        for which in range(5):
            assert torch.equal(plan.level_array(level, which),
                               fresh.level_array(level, which)), (level, which)
    b = _t(torch, second["rhs"])
    flexible = precision != "fp64"
    xa, ra = plan.solve_pcg(b, rtol=1e-10, flexible=flexible)
    xb, rb = fresh.solve_pcg(b, rtol=1e-10, flexible=flexible)
    assert ra.iterations == rb.iterations
    assert torch.equal(xa, xb)
    # and a third refresh returns to the first hierarchy (swapped back)
    plan.set_coefficients(*c1, *m1)
    x1, _ = plan.solve_pcg(_t(torch, first["rhs"]), rtol=1e-10,
                           flexible=flexible)
    ref = H.Plan(first["n"], first["faces"], first["bc"], pol)
    ref.set_coefficients(*c1, *m1)
    x1r, _ = ref.solve_pcg(_t(torch, first["rhs"]), rtol=1e-10,
                           flexible=flexible)
    assert torch.equal(x1, x1r)
    for p in (plan, fresh, ref):
        p.close()


def _breakdown_case(seed):
    """The weak-cycle, high-contrast case of `seed` (the oracle breaks down
    in the middle of the loop for seeds 11, 13 and 45)."""
    rng = np.random.default_rng(seed)
    n = (int(rng.integers(6, 14)), int(rng.integers(6, 14)),
         int(rng.integers(4, 10)))
    bc = ([0, 0, 1, 1, 2, 2] if seed % 3 == 0 else
          ([1, 1, 1, 1, 2, 2] if seed % 3 == 1 else [0, 1, 0, 1, 0, 1]))
    faces = [hf_cases.uniform_faces(v) for v in n]
    mob = np.exp(float(rng.uniform(1, 6)) *
                 rng.standard_normal(n[0] * n[1] * n[2]))
    null = 0 if 0 in bc else 1
    pol = dict(pre_sweeps=1, post_sweeps=1,
               coarse_sweeps=int(rng.integers(1, 3)), unit_linear=0,
               f_cycle=int(seed % 2), null_space_constant=null)
    b = rng.standard_normal(n[0] * n[1] * n[2])
    if null:
        b -= b.mean()
    interval = int(rng.integers(2, 6))
    return n, bc, faces, mob, pol, b, interval


@pytest.mark.parametrize("seed", [11, 13, 45])
def test_pcg_mid_loop_breakdown_matches_oracle(torch_cuda, seed):
    import hundun_flow as H
    torch = torch_cuda
    n, bc, faces, mob, pol, b, interval = _breakdown_case(seed)
    _say(f"[pcg breakdown] seed={seed} n={n} bc={bc} pol={pol} "
         f"interval={interval}")
    d, bx, by, bz = O.assemble(n, faces, bc, mob)
    mg = O.MG(n, faces, bc, O.policy(**pol), d, bx, by, bz)
    xo, ores = O.pcg(n, bc, d, bx, by, bz, None, mg, b, rtol=1e-12,
                     maximum_iterations=200, true_residual_interval=interval)
    assert ores.termination == 3 and ores.iterations > 0
    assert ores.breakdown_value < -1e-3 * float(b @ b)
    plan = H.Plan(n, faces, bc, H.default_policy(**pol))
    plan.set_coefficients(*[_t(torch, v) for v in (d, bx, by, bz)])
    x0 = torch.zeros(b.size, dtype=torch.float64, device="cuda")
    x, res = plan.solve_pcg(_t(torch, b), solution=x0, check=False,
                            rtol=1e-12, maximum_iterations=200,
                            true_residual_interval=interval)
    _say(f"  gpu {res.termination} it={res.iterations} "
         f"ops={res.operator_applies} pcs={res.preconditioner_applies} | "
         f"oracle it={ores.iterations} ops={ores.operator_applies} "
         f"pcs={ores.preconditioner_applies}")
    assert res.termination == "breakdown"
    assert res.status == 5
    assert res.iterations == ores.iterations
    assert res.operator_applies == ores.operator_applies
    assert res.preconditioner_applies == ores.preconditioner_applies
    # a failed solve publishes nothing
    assert torch.count_nonzero(x).item() == 0
    plan.close()
