"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bars: bit-exact where the arithmetic order is the reference's own (face
coefficients, hierarchy, masked operator, the FP64 V-cycle without a null
space); documented relative tolerances where a parallel reduction replaces a
sequential one (projection means, Krylov dots).
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


def mixed_case(n=(24, 20, 16), seed=11, periodic_z=True, dirichlet_x=True,
               radius=0.22):
    """Dirichlet x walls (no null space), Neumann y, periodic z, sphere
    mask, lognormal rho."""
    case = hf_cases.lognormal_sphere(n=n, seed=seed, radius=radius)
    bc = [0 if dirichlet_x else 1, 0 if dirichlet_x else 1, 1, 1,
          2 if periodic_z else 1, 2 if periodic_z else 1]
    act = case["cell_active"].reshape(n[2], n[1], n[0])
    fx, fy, fz = hf_cases.face_activity(act, bc)
    case.update(bc=bc, x_active=fx, y_active=fy, z_active=fz)
    return case


CASES = {
    "sphere32": lambda: hf_cases.lognormal_sphere(n=(32, 32, 32), seed=7),
    "stretched24": lambda: hf_cases.flame_cylinder(n=(24, 32, 16), seed=8),
    "step": lambda: hf_cases.flame_step(n=(48, 24, 16), seed=9),
    "mixed": lambda: mixed_case(),
    "odd_periodic": lambda: mixed_case(n=(20, 12, 18)),
    # fused-sweep levels (x >= 32, y >= 8) with partial 128 x 8 tiles
    "mixed_wide": lambda: mixed_case(n=(72, 40, 20), seed=12),
    # periodic x and y exercise the wrapped halo ring of the fused sweep
    "periodic_xy": lambda: periodic_xy_case(),
    # rows of 70 cells (nx % 4 == 2) fill part of one 128-wide tile of the
    # fused sweep and 36 rows leave a partial tile in y (spacing ratio 5.8:
    # planned with point smoothing under a threshold of 16)
    "half_quad": lambda: mixed_case(n=(70, 36, 12), seed=21),
}


def periodic_xy_case(n=(64, 48, 20), seed=13):
    case = hf_cases.lognormal_sphere(n=n, seed=seed, radius=0.2)
    bc = [2, 2, 2, 2, 0, 0]
    act = case["cell_active"].reshape(n[2], n[1], n[0])
    fx, fy, fz = hf_cases.face_activity(act, bc)
    case.update(bc=bc, x_active=fx, y_active=fy, z_active=fz)
    return case


def _policy(case_name, **kw):
    # stretched grids select line relaxation in the reference planner; a
    # threshold above the spacing ratio keeps point red-black + full
    # coarsening on both sides
    thr = 16.0 if case_name in ("stretched24", "half_quad") else 4.0
    null = 0 if case_name in ("mixed", "odd_periodic", "mixed_wide",
                              "periodic_xy", "half_quad") else 1
    base = dict(anisotropy_threshold=thr, null_space_constant=null)
    base.update(kw)
    return base


def _gpu_plan(torch, case, **policy):
    import hundun_flow as H
    pol = H.default_policy(**policy)
    plan = H.plan_from_case(case, pol)
    return plan


def _t(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("name", list(CASES))
def test_assemble_bitwise(torch_cuda, name):
    torch = torch_cuda
    _say(f"[assemble] case={name}")
    case = CASES[name]()
    import hundun_flow as H
    plan = H.Plan(case["n"], case["faces"], case["bc"],
                  H.default_policy(**_policy(name)))
    diag, bx, by, bz = plan.assemble(_t(torch, case["mobility"]))
    od, obx, oby, obz = O.assemble(case["n"], case["faces"], case["bc"],
                                   case["mobility"])
    for g, o in ((diag, od), (bx, obx), (by, oby), (bz, obz)):
        assert np.array_equal(g.cpu().numpy(), o)
    plan.close()


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("scaling", [0, 1])
def test_hierarchy_bitwise(torch_cuda, name, scaling):
    torch = torch_cuda
    _say(f"[hierarchy] case={name} coarse_face_scaling={scaling}")
    case = CASES[name]()
    pol = _policy(name, coarse_face_scaling=scaling)
    plan = _gpu_plan(torch, case, **pol)
    _, mg = O.from_case(case, O.policy(**pol))
    assert plan.level_count() == mg.level_count()
    for level in range(mg.level_count()):
        gi, oi = plan.level_info(level), mg.level_info(level)
        assert gi["shape"] == oi["shape"]
        assert gi["coarsen_mask"] == oi["coarsen_mask"]
        for which in range(5):
            g = plan.level_array(level, which).cpu().numpy()
            o = mg.level_array(level, which)
            assert np.array_equal(g, o), (level, which)
    plan.close()


@pytest.mark.parametrize("name", list(CASES))
def test_operator(torch_cuda, name):
    torch = torch_cuda
    _say(f"[operator] case={name}")
    case = CASES[name]()
    pol = _policy(name)
    plan = _gpu_plan(torch, case, **pol)
    (d, bx, by, bz), mg = O.from_case(case, O.policy(**pol))
    rng = np.random.default_rng(3)
    x = rng.standard_normal(plan.cells)
    y = plan.operator_apply(_t(torch, x)).cpu().numpy()
    # bitwise against the masked level-0 matrix of the oracle MG
    assert np.array_equal(y, mg.masked_apply(x))
    # and within rounding of the reference IBM-decorated operator
    ref = O.operator_apply(case["n"], case["bc"], d, bx, by, bz,
                           case["cell_active"], x)
    assert np.max(np.abs(y - ref)) <= 1e-13 * np.max(np.abs(ref))
    solid = case["cell_active"] == 0
    assert np.array_equal(y[solid], x[solid])  # identity rows
    plan.close()


@pytest.mark.parametrize("name", ["mixed", "odd_periodic", "mixed_wide",
                                  "periodic_xy"])
@pytest.mark.parametrize("f_cycle", [0, 1])
@pytest.mark.parametrize("tail", ["default", "0"])
@pytest.mark.parametrize("scaling", [0, 1])
@pytest.mark.parametrize("sweep", ["0", "1"])
def test_mg_apply_bitwise(torch_cuda, monkeypatch, name, f_cycle, tail,
                          scaling, sweep):
    """Without a null-space projection every V/F-cycle step is a fixed-order
    FP64 computation, so the GPU cycle equals the oracle bit for bit, both
    through the single-workgroup tail and through grid-level kernels."""
    torch = torch_cuda
    _say(f"[mg bitwise] case={name} f_cycle={f_cycle} tail={tail} "
         f"scaling={scaling} sweep={sweep}")
    if tail != "default":
        monkeypatch.setenv("HF_TAIL_CELLS", tail)
    monkeypatch.setenv("HF_FUSED_SWEEP", sweep)
    case = CASES[name]()
    pol = _policy(name, f_cycle=f_cycle, coarse_face_scaling=scaling)
    plan = _gpu_plan(torch, case, **pol)
    _, mg = O.from_case(case, O.policy(**pol))
    r = np.random.default_rng(5).standard_normal(plan.cells)
    z = plan.mg_apply(_t(torch, r)).cpu().numpy()
    zo = mg.apply(r)
    assert np.array_equal(z, zo), np.max(np.abs(z - zo))
    plan.close()


@pytest.mark.parametrize("name", ["sphere32", "step", "stretched24"])
@pytest.mark.parametrize("unit_linear", [1, 0])
@pytest.mark.parametrize("tail", ["default", "0"])
@pytest.mark.parametrize("sweep", ["1", "0"])
def test_mg_apply_projected(torch_cuda, monkeypatch, name, unit_linear, tail,
                            sweep):
    """With the constant null space the projection means come from a
    parallel reduction: agreement to 1e-12 relative."""
    torch = torch_cuda
    _say(f"[mg projected] case={name} unit_linear={unit_linear} tail={tail} "
         f"sweep={sweep}")
    if tail != "default":
        monkeypatch.setenv("HF_TAIL_CELLS", tail)
    monkeypatch.setenv("HF_FUSED_SWEEP", sweep)
    case = CASES[name]()
    pol = _policy(name, unit_linear=unit_linear)
    plan = _gpu_plan(torch, case, **pol)
    _, mg = O.from_case(case, O.policy(**pol))
    r = case["rhs"]
    z = plan.mg_apply(_t(torch, r)).cpu().numpy()
    zo = mg.apply(r)
    rel = np.linalg.norm(z - zo) / np.linalg.norm(zo)
    _say(f"  rel={rel:.3e}")
    assert rel <= 1e-12
    solid = case["cell_active"] == 0
    assert np.all(z[solid] == 0.0)
    plan.close()


_CHEB = dict(point_smoother=1, chebyshev_lower_fraction=0.3)


@pytest.mark.parametrize("name", ["mixed", "odd_periodic", "mixed_wide",
                                  "periodic_xy"])
@pytest.mark.parametrize("stages", [(1, 2), (3, 2)])
@pytest.mark.parametrize("f_cycle", [1, 0])
@pytest.mark.parametrize("tail", ["default", "0"])
def test_mg_apply_chebyshev_bitwise(torch_cuda, monkeypatch, name, stages,
                                    f_cycle, tail):
    """Chebyshev-Jacobi cycles without a null space are fixed-order FP64
    computations: the GPU cycle equals the oracle bit for bit, with the
    coarse Chebyshev levels inside the single-workgroup tail (default
    threshold) and as grid-level stage kernels (threshold 0)."""
    torch = torch_cuda
    _say(f"[mg chebyshev bitwise] case={name} stages={stages} "
         f"f_cycle={f_cycle} tail={tail}")
    if tail != "default":
        monkeypatch.setenv("HF_TAIL_CELLS", tail)
    case = CASES[name]()
    pol = _policy(name, f_cycle=f_cycle, pre_sweeps=stages[0],
                  post_sweeps=stages[1], **_CHEB)
    plan = _gpu_plan(torch, case, **pol)
    levels = plan.level_count()
    on_tail = [plan.level_info(l)["on_tail"] for l in range(levels)]
    # the default threshold takes Chebyshev levels (all but the coarsest)
    # into the tail; threshold 0 leaves it the coarsest alone
    assert sum(on_tail) >= (2 if tail == "default" else 1), on_tail
    assert tail == "default" or sum(on_tail) == 1, on_tail
    _, mg = O.from_case(case, O.policy(**pol))
    r = np.random.default_rng(7).standard_normal(plan.cells)
    z = plan.mg_apply(_t(torch, r)).cpu().numpy()
    zo = mg.apply(r)
    assert np.array_equal(z, zo), np.max(np.abs(z - zo))
    plan.close()


@pytest.mark.parametrize("name", ["sphere32", "step"])
@pytest.mark.parametrize("tail", ["default", "0"])
def test_mg_apply_chebyshev_projected(torch_cuda, monkeypatch, name, tail):
    """Chebyshev F-cycle with the constant null space (projection means from
    a parallel reduction): agreement to 1e-12 relative through the tail and
    through grid-level stage kernels."""
    torch = torch_cuda
    _say(f"[mg chebyshev projected] case={name} tail={tail}")
    if tail != "default":
        monkeypatch.setenv("HF_TAIL_CELLS", tail)
    case = CASES[name]()
    pol = _policy(name, f_cycle=1, pre_sweeps=1, post_sweeps=2, **_CHEB)
    plan = _gpu_plan(torch, case, **pol)
    _, mg = O.from_case(case, O.policy(**pol))
    r = case["rhs"]
    z = plan.mg_apply(_t(torch, r)).cpu().numpy()
    zo = mg.apply(r)
    rel = np.linalg.norm(z - zo) / np.linalg.norm(zo)
    _say(f"  rel={rel:.3e}")
    assert rel <= 1e-12
    plan.close()


def _gauge(x, act):
    fluid = act != 0
    y = x.copy()
    y[fluid] -= y[fluid].mean()
    y[~fluid] = 0.0
    return y


@pytest.mark.parametrize("name", ["sphere32", "step", "stretched24", "mixed"])
@pytest.mark.parametrize("flexible", [False, True])
@pytest.mark.parametrize("scaling", [0, 1])
def test_pcg_fp64_matches_oracle(torch_cuda, name, flexible, scaling):
    torch = torch_cuda
    _say(f"[pcg fp64] case={name} flexible={flexible} scaling={scaling}")
    case = CASES[name]()
    pol = _policy(name, coarse_face_scaling=scaling)
    plan = _gpu_plan(torch, case, **pol)
    (d, bx, by, bz), mg = O.from_case(case, O.policy(**pol))
    b = case["rhs"]
    xg, res = plan.solve_pcg(_t(torch, b), rtol=1e-12, flexible=flexible)
    xo, ores = O.pcg(case["n"], case["bc"], d, bx, by, bz,
                     case["cell_active"], mg, b, rtol=1e-12,
                     flexible=flexible)
    assert res.termination == "converged" and ores.status == 0
    xg = xg.cpu().numpy()
    act = case["cell_active"]
    # independent residual check with the reference-form operator
    Ax = O.operator_apply(case["n"], case["bc"], d, bx, by, bz, act, xg)
    true_rel = np.linalg.norm(b - Ax) / np.linalg.norm(b)
    ga, go = _gauge(xg, act), _gauge(xo, act)
    diff = np.linalg.norm(ga - go) / np.linalg.norm(go)
    _say(f"  it gpu={res.iterations} oracle={ores.iterations} "
         f"true_rel={true_rel:.2e} diff={diff:.2e}")
    assert true_rel <= 1.01e-12
    # the Krylov dots are summed in another order than the oracle's; the
    # iterates agree to about 1e-15, well inside the 1e-10 parity bar
    assert res.iterations == ores.iterations
    assert diff <= 1e-10
    assert np.all(xg[act == 0] == 0.0)
    plan.close()


# the oracle's termination codes, in the order of the C ABI's
_TERMS = ["converged", "zero_rhs", "maximum_iterations", "breakdown"]


@pytest.mark.parametrize("interval", [1, 3])
@pytest.mark.parametrize("maxit", [5, 1000])
@pytest.mark.parametrize("flexible", [False, True])
def test_pcg_gates_match_oracle(torch_cuda, interval, maxit, flexible):
    """The PCG gates (src/solver_krylov.cpp:1299-1555): the true-residual
    interval, the iteration limit and its failure, counted as the oracle
    counts them.  A solve stopped by the limit leaves the caller's solution
    untouched; iterations that verify on schedule update x alone, so the
    operator and preconditioner counts pin which iterations verified."""
    torch = torch_cuda
    _say(f"[pcg gates] interval={interval} maxit={maxit} flexible={flexible}")
    case = CASES["step"]()
    pol = _policy("step")
    plan = _gpu_plan(torch, case, **pol)
    (d, bx, by, bz), mg = O.from_case(case, O.policy(**pol))
    b = case["rhs"]
    cells = b.size
    x0 = np.random.default_rng(5).standard_normal(cells) * 1e-3
    x0[case["cell_active"] == 0] = 0.0
    xg_t = _t(torch, x0.copy())
    xg_t, res = plan.solve_pcg(_t(torch, b), solution=xg_t, check=False,
                               rtol=1e-12, maximum_iterations=maxit,
                               true_residual_interval=interval,
                               flexible=flexible)
    xo, ores = O.pcg(case["n"], case["bc"], d, bx, by, bz,
                     case["cell_active"], mg, b, x0=x0, rtol=1e-12,
                     maximum_iterations=maxit,
                     true_residual_interval=interval, flexible=flexible)
    want = _TERMS[ores.termination]
    _say(f"  gpu {res.termination} it={res.iterations} "
         f"ops={res.operator_applies} pcs={res.preconditioner_applies} | "
         f"oracle {want} it={ores.iterations} ops={ores.operator_applies} "
         f"pcs={ores.preconditioner_applies}")
    assert res.termination == want
    xg = xg_t.cpu().numpy()
    if want == "maximum_iterations":
        assert res.iterations == ores.iterations == maxit
        assert res.operator_applies == ores.operator_applies
        assert res.preconditioner_applies == ores.preconditioner_applies
        # same stopping point: the true residuals agree to rounding
        rel = abs(res.final_true_residual - ores.final_true_residual)
        assert rel <= 1e-8 * ores.final_true_residual
        # a failed solve publishes no solution
        assert np.array_equal(xg, x0)
    else:
        assert res.iterations == ores.iterations
        assert res.operator_applies == ores.operator_applies
        assert res.preconditioner_applies == ores.preconditioner_applies
        act = case["cell_active"]
        Ax = O.operator_apply(case["n"], case["bc"], d, bx, by, bz, act, xg)
        assert np.linalg.norm(b - Ax) <= 1.01e-12 * np.linalg.norm(b)
        ga, go = _gauge(xg, act), _gauge(xo, act)
        assert np.linalg.norm(ga - go) <= 1e-10 * np.linalg.norm(go)
    plan.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_pcg_start_gates(torch_cuda, precision):
    """The gates before the first iteration (src/solver_krylov.cpp:
    1321-1395): a zero right-hand side zeroes the solution; a non-finite
    right-hand side fails as non_finite and a non-finite initial guess as
    operator_failure, both leaving the caller's solution untouched; an
    initial guess that already meets the tolerance is returned as it is,
    after no iteration."""
    import hundun_flow as H
    torch = torch_cuda
    _say(f"[pcg start gates] precision={precision}")
    case = CASES["step"]()
    plan = _gpu_plan(torch, case, **_policy("step", precision=precision))
    act = case["cell_active"]
    b = case["rhs"]
    fluid = np.flatnonzero(act)
    x0 = np.random.default_rng(9).standard_normal(b.size)
    x0[act == 0] = 0.0
    numerical = H.STATUS_NAMES.index("numerical_failure")

    x, res = plan.solve_pcg(_t(torch, np.zeros_like(b)),
                            solution=_t(torch, x0), rtol=1e-10)
    assert res.termination == "zero_rhs" and res.status == 0
    assert res.iterations == 0 and res.rhs_norm == 0.0
    assert not x.cpu().numpy().any()

    bad = b.copy()
    bad[fluid[len(fluid) // 2]] = np.nan
    x, res = plan.solve_pcg(_t(torch, bad), solution=_t(torch, x0),
                            check=False, rtol=1e-10)
    assert res.termination == "non_finite" and res.status == numerical
    assert np.array_equal(x.cpu().numpy(), x0)

    xinf = x0.copy()
    xinf[fluid[0]] = np.inf
    x, res = plan.solve_pcg(_t(torch, b), solution=_t(torch, xinf),
                            check=False, rtol=1e-10)
    assert res.termination == "operator_failure" and res.status == numerical
    assert res.iterations == 0
    assert np.array_equal(x.cpu().numpy(), xinf)

    xs, res = plan.solve_pcg(_t(torch, b), rtol=1e-10, flexible=True)
    assert res.termination == "converged" and res.iterations > 0
    want = xs.cpu().numpy().copy()
    x, res = plan.solve_pcg(_t(torch, b), solution=xs, rtol=1e-8,
                            flexible=True)
    assert res.termination == "converged" and res.iterations == 0
    assert res.operator_applies == 1 and res.preconditioner_applies == 0
    assert np.array_equal(x.cpu().numpy(), want)
    # the norms the gates use: ||b|| and the initial true residual
    assert abs(res.rhs_norm - np.linalg.norm(b)) <= 1e-12 * np.linalg.norm(b)
    (d, bx, by, bz), _ = O.from_case(case, O.policy(**_policy("step")))
    r0 = np.linalg.norm(b - O.operator_apply(case["n"], case["bc"], d, bx, by,
                                             bz, act, want))
    assert abs(res.initial_true_residual - r0) <= 1e-6 * r0
    plan.close()


_SPECTRUM = {}


def _spectrum(name, case, coef, mg, null):
    """Extreme eigenvalues of the FP64 operator on the fluid cells (with the
    constant null space, on its complement): the largest by Lanczos on A,
    the smallest by Lanczos on A^-1 applied with the oracle's PCG."""
    if name in _SPECTRUM:
        return _SPECTRUM[name]
    from scipy.sparse.linalg import LinearOperator, eigsh
    d, bx, by, bz = coef
    n, bc = case["n"], case["bc"]
    act = case["cell_active"] != 0
    cells = int(np.prod(n))

    def proj(v):
        v = np.where(act, v, 0.0)
        if null:
            v[act] -= v[act].mean()
        return v

    def inv(v):
        y, _ = O.pcg(n, bc, d, bx, by, bz, case["cell_active"], mg, proj(v),
                     rtol=1e-12, maximum_iterations=400)
        return proj(y)

    def fwd(v):
        return proj(O.operator_apply(n, bc, d, bx, by, bz,
                                     case["cell_active"], proj(v)))

    v0 = proj(np.random.default_rng(1).standard_normal(cells))
    top = eigsh(LinearOperator((cells, cells), matvec=fwd), k=1, which="LA",
                v0=v0, tol=1e-8, return_eigenvectors=False)[0]
    low = 1.0 / eigsh(LinearOperator((cells, cells), matvec=inv), k=1,
                      which="LA", v0=v0, tol=1e-8,
                      return_eigenvectors=False)[0]
    _SPECTRUM[name] = (low, top)
    return low, top


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
@pytest.mark.parametrize("name", ["sphere32", "step"])
@pytest.mark.parametrize("scaling", [0, 1])
def test_mixed_precision_reaches_tolerance(torch_cuda, monkeypatch, precision,
                                           name, scaling):
    """FP32 / BF16-storage V-cycle inside the FP64 CG: the FP64 true residual
    reaches 1e-10 and the solution agrees with the FP64 oracle to the
    conditioning bound.  A zero tail threshold keeps every level but the
    coarsest in the storage precision.

    The bound: A e = -(r - r_o) for e = x - x_o, so on the complement of the
    null space |e| <= (|r| + |r_o|) / lambda_min, and |e| / |x_o| is at most
    that over |x_o| (printed with kappa times the relative residual, the
    looser textbook form).  lambda_min and lambda_max are Lanczos estimates
    on the oracle's operator."""
    torch = torch_cuda
    _say(f"[mixed precision] case={name} precision={precision} "
         f"scaling={scaling}")
    monkeypatch.setenv("HF_TAIL_CELLS", "0")
    case = CASES[name]()
    pol = _policy(name, precision=precision, coarse_face_scaling=scaling)
    plan = _gpu_plan(torch, case, **pol)
    (d, bx, by, bz), mg = O.from_case(case, O.policy(**_policy(name)))
    b = case["rhs"]
    xg, res = plan.solve_pcg(_t(torch, b), rtol=1e-10, flexible=True)
    xo, ores = O.pcg(case["n"], case["bc"], d, bx, by, bz,
                     case["cell_active"], mg, b, rtol=1e-12)
    xg = xg.cpu().numpy()
    act = case["cell_active"]
    Ax = O.operator_apply(case["n"], case["bc"], d, bx, by, bz, act, xg)
    Axo = O.operator_apply(case["n"], case["bc"], d, bx, by, bz, act, xo)
    true_rel = np.linalg.norm(b - Ax) / np.linalg.norm(b)
    diff = (np.linalg.norm(_gauge(xg, act) - _gauge(xo, act))
            / np.linalg.norm(_gauge(xo, act)))
    null = pol.get("null_space_constant", 1)
    low, top = _spectrum(name, case, (d, bx, by, bz), mg, null)
    bound = ((np.linalg.norm(b - Ax) + np.linalg.norm(b - Axo))
             / (low * np.linalg.norm(_gauge(xo, act))))
    _say(f"  it={res.iterations} (fp64 oracle {ores.iterations}) "
         f"true_rel={true_rel:.2e} err_vs_fp64={diff:.2e} bound "
         f"|r|/(lambda_min |x|)={bound:.2e} kappa={top / low:.0f} "
         f"kappa*true_rel={top / low * true_rel:.2e}")
    assert res.termination == "converged"
    assert true_rel <= 1.01e-10
    assert diff <= 1.05 * bound
    assert np.all(xg[act == 0] == 0.0)
    plan.close()


@pytest.mark.parametrize("name", ["mixed_wide", "periodic_xy", "sphere32",
                                  "half_quad"])
@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_fused_sweep_matches_two_pass(torch_cuda, monkeypatch, name,
                                      precision):
    """The fused finest-level sweep and the two colour passes perform the
    same updates in the same arithmetic order: one cycle on each route gives
    the same correction bit for bit."""
    torch = torch_cuda
    _say(f"[sweep routes] case={name} precision={precision}")
    case = CASES[name]()
    r = _t(torch, np.random.default_rng(8).standard_normal(
        int(np.prod(case["n"]))))
    out = {}
    for sweep in ("1", "0"):
        monkeypatch.setenv("HF_FUSED_SWEEP", sweep)
        plan = _gpu_plan(torch, case, **_policy(name, precision=precision))
        out[sweep] = plan.mg_apply(r).cpu().numpy()
        plan.close()
    assert np.array_equal(out["1"], out["0"]), \
        np.max(np.abs(out["1"] - out["0"]))


def _walls(n, bc, seed):
    case = hf_cases.lognormal_sphere(n=n, seed=seed, radius=0.2)
    act = case["cell_active"].reshape(n[2], n[1], n[0])
    fx, fy, fz = hf_cases.face_activity(act, bc)
    case.update(bc=bc, x_active=fx, y_active=fy, z_active=fz)
    return case


# Finest levels run by the fused sweep with every boundary kind in its halo:
# Neumann and Dirichlet mixes per axis, periodic faces on tiles cut short by
# the domain (the wrapped ring cell then sits inside the tile's last column
# or row), and an odd x extent (planned on the two-pass route instead).
SWEEP_EDGES = {
    "neumann_all": lambda: _walls((160, 20, 12), [1, 1, 1, 1, 1, 1], 75),
    "mixed_dn": lambda: _walls((136, 18, 10), [0, 1, 1, 0, 0, 1], 76),
    "perz_neumann_xy": lambda: _walls((128, 16, 8), [1, 1, 0, 1, 2, 2], 77),
    "periodic_x_partial": lambda: _walls((160, 16, 8), [2, 2, 1, 1, 0, 0], 79),
    "periodic_xy_partial": lambda: _walls((200, 12, 6), [2, 2, 2, 2, 1, 1],
                                          80),
    "odd_x": lambda: _walls((97, 13, 9), [0, 0, 0, 0, 0, 0], 78),
}


# "all": every grid level large enough runs the fused sweep (coarse levels
# too, as the large grids do), not only the finest
@pytest.mark.parametrize("levels", ["finest", "all"])
@pytest.mark.parametrize("name", list(SWEEP_EDGES))
@pytest.mark.parametrize("f_cycle", [0, 1])
def test_fused_sweep_edges_fp64_bitwise(torch_cuda, monkeypatch, name,
                                        f_cycle, levels):
    """FP64 cycle through the fused sweep equals the oracle bit for bit."""
    torch = torch_cuda
    _say(f"[sweep edges fp64] case={name} f_cycle={f_cycle} levels={levels}")
    if levels == "all":
        monkeypatch.setenv("HF_SWEEP_MIN_CELLS", "0")
    case = SWEEP_EDGES[name]()
    pol = dict(null_space_constant=0, f_cycle=f_cycle,
               anisotropy_threshold=1e9)
    plan = _gpu_plan(torch, case, **pol)
    _, mg = O.from_case(case, O.policy(**pol))
    r = np.random.default_rng(5).standard_normal(plan.cells)
    z = plan.mg_apply(_t(torch, r)).cpu().numpy()
    zo = mg.apply(r)
    assert np.array_equal(z, zo), np.max(np.abs(z - zo))
    plan.close()


@pytest.mark.parametrize("levels", ["finest", "all"])
@pytest.mark.parametrize("name", list(SWEEP_EDGES))
@pytest.mark.parametrize("precision", ["fp32", "bf16"])
@pytest.mark.parametrize("null", [0, 1])
def test_fused_sweep_edges_match_two_pass(torch_cuda, monkeypatch, name,
                                          precision, null, levels):
    """32- and 16-bit storage: the fused sweep and the two colour passes
    give the same cycle bit for bit (with the constant null space too: the
    first sweep then applies the projection of b on the fly)."""
    torch = torch_cuda
    _say(f"[sweep edges] case={name} precision={precision} null={null} "
         f"levels={levels}")
    if levels == "all":
        monkeypatch.setenv("HF_SWEEP_MIN_CELLS", "0")
    case = SWEEP_EDGES[name]()
    pol = dict(null_space_constant=null, anisotropy_threshold=1e9,
               precision=precision)
    r = _t(torch, np.random.default_rng(8).standard_normal(
        int(np.prod(case["n"]))))
    out = {}
    for sweep in ("1", "0"):
        monkeypatch.setenv("HF_FUSED_SWEEP", sweep)
        plan = _gpu_plan(torch, case, **pol)
        out[sweep] = plan.mg_apply(r).cpu().numpy()
        plan.close()
    # the projection means come from the same reductions on either route
    assert np.array_equal(out["1"], out["0"]), \
        np.max(np.abs(out["1"] - out["0"]))


# Levels where the retained defect and its restriction run as one kernel:
# fused-sweep levels with even extents and full z-marching tiles.
DEFECT_RESTRICT = {
    "periodic_xy": CASES["periodic_xy"],
    "sphere32": CASES["sphere32"],
    "perz_neumann_xy": SWEEP_EDGES["perz_neumann_xy"],
    "two_tiles_periodic": lambda: _walls((256, 24, 8), [2, 2, 2, 2, 0, 0], 72),
}


@pytest.mark.parametrize("name", list(DEFECT_RESTRICT))
@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16"])
@pytest.mark.parametrize("null", [0, 1])
@pytest.mark.parametrize("smoother", ["red_black", "chebyshev"])
def test_defect_restrict_fused_matches_two_kernels(torch_cuda, monkeypatch,
                                                    name, precision, null,
                                                    smoother):
    """The fused defect + restriction stores the coarse b of the two
    kernels bit for bit, so the cycle is unchanged without the null space;
    with it, the coarse mean's partial sums come in another block order.
    Fused-sweep levels and Chebyshev levels (the production F-cycle) alike."""
    torch = torch_cuda
    _say(f"[defect restrict] case={name} precision={precision} null={null} "
         f"smoother={smoother}")
    case = DEFECT_RESTRICT[name]()
    pol = dict(null_space_constant=null, anisotropy_threshold=1e9,
               precision=precision)
    if smoother == "chebyshev":
        pol.update(point_smoother=1, f_cycle=1, pre_sweeps=1, post_sweeps=2)
    r = _t(torch, np.random.default_rng(6).standard_normal(
        int(np.prod(case["n"]))))
    out = {}
    for fused in ("1", "0"):
        monkeypatch.setenv("HF_DEFECT_RESTRICT", fused)
        plan = _gpu_plan(torch, case, **pol)
        out[fused] = plan.mg_apply(r).cpu().numpy()
        plan.close()
    if null == 0:
        assert np.array_equal(out["1"], out["0"]), \
            np.max(np.abs(out["1"] - out["0"]))
    else:
        rel = np.linalg.norm(out["1"] - out["0"]) / np.linalg.norm(out["0"])
        bound = 1e-12 if precision == "fp64" else 1e-4
        assert rel <= bound, rel


# Even and odd row lengths on every level (72 -> 36 -> 18 -> 9; 20 -> 10 ->
# 5), the periodic wrap along x and Dirichlet/Neumann walls.
ROW_PAIRS = {
    "mixed_wide": CASES["mixed_wide"],
    "odd_periodic": CASES["odd_periodic"],
    "periodic_xy": CASES["periodic_xy"],
    # periodic in every direction, levels far narrower than their tiles
    "periodic_all": lambda: _walls((36, 10, 14), [2, 2, 2, 2, 2, 2], 74),
}


@pytest.mark.parametrize("name", list(ROW_PAIRS))
@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16"])
@pytest.mark.parametrize("null", [0, 1])
@pytest.mark.parametrize("smoother", ["red_black", "chebyshev"])
def test_row_pair_kernels_match_per_cell(torch_cuda, monkeypatch, name,
                                         precision, null, smoother):
    """The prolongation and Chebyshev stages over pairs of cells of a row,
    and the first Chebyshev stage of a from-zero visit formed from b and D
    alone, store what the per-cell kernels over an explicit zero x store,
    bit for bit, in every storage precision (the projection sums come from
    the same kernels on both routes)."""
    torch = torch_cuda
    _say(f"[row pairs] case={name} precision={precision} null={null} "
         f"smoother={smoother}")
    case = ROW_PAIRS[name]()
    pol = _policy(name, null_space_constant=null, precision=precision)
    if smoother == "chebyshev":
        pol.update(point_smoother=1, f_cycle=1, pre_sweeps=1, post_sweeps=2)
    r = _t(torch, np.random.default_rng(8).standard_normal(
        int(np.prod(case["n"]))))
    out = {}
    for on in ("1", "0"):
        monkeypatch.setenv("HF_ROW_PAIRS", on)
        monkeypatch.setenv("HF_CHEB_FIRST", on)
        # grid levels down to the coarsest
        monkeypatch.setenv("HF_TAIL_CELLS", "0")
        plan = _gpu_plan(torch, case, **pol)
        out[on] = plan.mg_apply(r).cpu().numpy()
        plan.close()
    assert np.array_equal(out["1"], out["0"]), \
        np.max(np.abs(out["1"] - out["0"]))


@pytest.mark.parametrize("name", ["step", "mixed_wide", "periodic_xy"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_pcg_row_pair_operator(torch_cuda, monkeypatch, name, precision):
    """The PCG operator over pairs of cells of a row (q = A p bit for bit;
    p.q summed in another order) against the one-cell-per-lane operator:
    the same iteration count and the solution to 1e-12."""
    torch = torch_cuda
    _say(f"[pcg pairs] case={name} precision={precision}")
    case = CASES[name]()
    pol = _policy(name, precision=precision)
    b = _t(torch, case["rhs"])  # consistent with a constant null space
    out = {}
    for on in ("1", "0"):
        monkeypatch.setenv("HF_KRYLOV_PAIRS", on)
        plan = _gpu_plan(torch, case, **pol)
        x, res = plan.solve_pcg(b, rtol=1e-10,
                                flexible=precision != "fp64")
        out[on] = (x.cpu().numpy(), res.iterations, res.termination)
        plan.close()
    rel = (np.linalg.norm(out["1"][0] - out["0"][0])
           / np.linalg.norm(out["0"][0]))
    _say(f"  it={out['1'][1]}/{out['0'][1]} rel={rel:.2e}")
    assert out["1"][2] == "converged" and out["0"][2] == "converged"
    assert out["1"][1] == out["0"][1]
    assert rel <= 1e-12, rel
