"""Full-size workloads on the GPU, checked by size-independent properties.

At these sizes the CPU oracle is too slow for a full solve, so each GPU
solution is verified with the oracle's own FP64 IBM operator (an independent
code path): the true relative residual ||b - A p|| / ||b|| must meet the
solve tolerance, solid cells must be exactly zero, and the mixed-precision
solutions must agree with the FP64 solution to the conditioning bound.  The
64^3 case is also compared with a full oracle solve.
"""
import functools
import sys
import time

import numpy as np
import pytest

import hf_cases
import hf_oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@functools.lru_cache(maxsize=2)
def cached_case(name, shape, seed):
    return hf_cases.make(name, n=shape, seed=seed)


def _solve(case, **policy):
    import torch

    import hundun_flow as H
    pol = H.default_policy(**policy)
    plan = H.plan_from_case(case, pol)
    b = torch.from_numpy(case["rhs"]).cuda()
    flexible = policy.get("precision", "fp64") != "fp64" or \
        policy.get("coarse_face_scaling", 0) == 1
    t0 = time.perf_counter()
    x, res = plan.solve_pcg(b, rtol=1e-10, maximum_iterations=1000,
                            flexible=flexible)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3
    xh = x.cpu().numpy()
    plan.close()
    return xh, res, ms


def _oracle_residual(case, x):
    d, bx, by, bz = O.assemble(case["n"], case["faces"], case["bc"],
                               case["mobility"])
    ax = O.operator_apply(case["n"], case["bc"], d, bx, by, bz,
                          case["cell_active"], x)
    return np.linalg.norm(case["rhs"] - ax) / np.linalg.norm(case["rhs"])


def _gauge(x, act):
    fluid = act != 0
    y = x.copy()
    y[fluid] -= y[fluid].mean()
    y[~fluid] = 0.0
    return y


def test_lognormal_sphere_64_matches_oracle_solve():
    """64^3 lognormal rho, sphere mask, Neumann walls, periodic z: the GPU
    FP64 solve with the reference hierarchy takes the oracle's iterations and
    reaches the same solution."""
    case = hf_cases.lognormal_sphere(n=(64, 64, 64), seed=1)
    x, res, ms = _solve(case)
    (d, bx, by, bz), mg = O.from_case(case, O.policy())
    xo, ores = O.pcg(case["n"], case["bc"], d, bx, by, bz,
                     case["cell_active"], mg, case["rhs"], rtol=1e-10)
    act = case["cell_active"]
    rel = _oracle_residual(case, x)
    diff = np.linalg.norm(_gauge(x, act) - _gauge(xo, act)) / \
        np.linalg.norm(_gauge(xo, act))
    _say(f"[64^3] gpu it={res.iterations} ({ms:.1f} ms) oracle "
         f"it={ores.iterations} rel={rel:.2e} diff={diff:.2e}")
    assert res.iterations == ores.iterations
    assert rel <= 1.01e-10
    assert diff <= 1e-10
    assert np.all(x[act == 0] == 0.0)


@pytest.mark.parametrize("scaling", [0, 1])
def test_flame_cylinder_256_stretched(scaling):
    """256^3, tanh-stretched y (spacing ratio 8), flame sheet 1:7, cylinder
    mask.  Point red-black with full coarsening on both sides (anisotropy
    threshold 16 keeps the reference planner off line relaxation)."""
    case = cached_case("flame_cylinder", (256, 256, 256), 2)
    x, res, ms = _solve(case, anisotropy_threshold=16.0,
                        coarse_face_scaling=scaling)
    rel = _oracle_residual(case, x)
    _say(f"[256^3 stretched] scaling={scaling} it={res.iterations} "
         f"{ms:.1f} ms rel={rel:.2e}")
    assert res.termination == "converged"
    assert rel <= 1.01e-10
    assert np.all(x[case["cell_active"] == 0] == 0.0)


def test_flame_step_512_precisions():
    """512x256x256 flame sheet + 5% noise, backward-facing step: FP64, FP32
    and BF16-storage V-cycles all reach 1e-10 with the FP64 operator; the
    mixed-precision solutions agree with the FP64 one."""
    case = hf_cases.flame_step(n=(512, 256, 256), seed=3)
    act = case["cell_active"]
    sols = {}
    for prec in ("fp64", "fp32", "bf16"):
        x, res, ms = _solve(case, precision=prec, coarse_face_scaling=1)
        rel = _oracle_residual(case, x)
        sols[prec] = _gauge(x, act)
        err = (np.linalg.norm(sols[prec] - sols["fp64"]) /
               np.linalg.norm(sols["fp64"]))
        _say(f"[512x256x256] {prec}: it={res.iterations} {ms:.1f} ms "
             f"rel={rel:.2e} err_vs_fp64={err:.2e}")
        assert res.termination == "converged"
        assert rel <= 1.01e-10
        assert np.all(x[act == 0] == 0.0)
        assert err <= 1e-6


def test_flame_probe_256_adjoint_64_directions():
    """Differentiation workload at 256^3: gradients of the probe-box pressure
    sum w.r.t. log rho along 64 seeded directions against central finite
    differences (h = 1e-6), and w.r.t. b against the linear response.

    Each central difference L(rho e^{+h d}) - L(rho e^{-h d}) is evaluated
    against the base solution p0: A+- (p+- - p0) = b - A+- p0 has an O(h)
    right-hand side, so the relative tolerance bounds the error of each
    correction, and p+ - p- = (p+ - p0) - (p- - p0) holds exactly whatever
    the error of p0 -- the difference carries no 1/h amplification of the
    solve tolerance."""
    import torch

    import hundun_flow as H
    case = hf_cases.flame_probe(n=(256, 256, 256), seed=4)
    dev = torch.device("cuda")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    log_rho = t(np.log(case["rho"]))
    act = case["cell_active"] != 0
    actf = t(act.astype(np.float64))
    masks = [t(case[k]) for k in ("cell_active", "x_active", "y_active",
                                  "z_active")]
    probe = t(case["probe"])
    b0 = t(case["rhs"])
    pol = H.default_policy(coarse_face_scaling=1)
    ctl = dict(rtol=1e-12, maximum_iterations=2000, flexible=True)

    def build(lr):
        plan = H.Plan(case["n"], case["faces"], case["bc"], pol)
        mob = torch.exp(-lr)
        diag, bx, by, bz = plan.assemble(mob)
        plan.set_coefficients(diag, bx, by, bz, *masks)
        plan.mobility = mob
        return plan

    def solve(plan, rhs, **kw):
        rhs = rhs * actf
        plan.project_mean(rhs)
        x, res = plan.solve_pcg(rhs, **kw)
        plan.project_mean(x)
        return x, res

    plan = build(log_rho)
    x, res = solve(plan, b0, **ctl)
    gb, gr, ares = plan.adjoint(plan.mobility, x, probe, **ctl)
    gr = gr.cpu().numpy()
    gb = gb.cpu().numpy()
    plan.close()
    rng = np.random.default_rng(64)
    h = 1e-6
    h5 = 1e-5
    adjs, fds, fds5, scales = [], [], [], []
    gnorm = float(np.linalg.norm(gr))
    t0 = time.perf_counter()
    rhs0 = b0 * actf
    pbase = build(log_rho)
    pbase.project_mean(rhs0)
    pbase.close()
    for i in range(64):
        d = rng.standard_normal(len(act)) * act
        dt = t(d)
        for step, out in ((h, fds), (h5, fds5)):
            corr = []
            for sgn in (1.0, -1.0):
                pl = build(log_rho + sgn * step * dt)
                delta, _ = solve(pl, rhs0 - pl.operator_apply(x), rtol=1e-10,
                                 maximum_iterations=2000, flexible=True)
                pl.close()
                corr.append(delta)
            out.append(float((corr[0] - corr[1]) @ probe) / (2 * step))
        adjs.append(float(gr @ d))
        scales.append(gnorm * float(np.linalg.norm(d)))
    d = rng.standard_normal(len(act)) * act
    dt = t(d)
    pl = build(log_rho)
    xp, _ = solve(pl, b0 + dt, **ctl)
    xm, _ = solve(pl, b0 - dt, **ctl)
    pl.close()
    fd_b = float((xp - xm) @ probe) / 2.0
    adj_b = float(gb @ d)
    adjs, fds, scales = np.array(adjs), np.array(fds), np.array(scales)
    fds5 = np.array(fds5)
    # the central difference at h = 1e-6 carries ~1e-13 of rounding in L,
    # i.e. ~1e-7 relative noise where g.d is of typical size and more where
    # g.d happens to be small; the 64-vector of directional derivatives and
    # the error scaled by |g| |d| are the noise-robust measures
    vec_rel = float(np.linalg.norm(adjs - fds) / np.linalg.norm(fds))
    scaled = float(np.max(np.abs(adjs - fds) / scales))
    worst = float(np.max(np.abs(adjs - fds) / np.abs(fds)))
    # per direction at h = 1e-5 (the rounding of L 10x smaller): the
    # relative error of one direction is large only where g.d itself is a
    # small fraction of |g| |d| (cancellation), which the ratio printed for
    # the worst direction shows; scaled by |g| |d| every direction agrees
    rel5 = np.abs(adjs - fds5) / np.abs(fds5)
    worst5 = float(np.max(rel5))
    i5 = int(np.argmax(rel5))
    scaled5 = float(np.max(np.abs(adjs - fds5) / scales))
    _say(f"[256^3 adjoint] forward it={res.iterations} adjoint "
         f"it={ares.iterations} 64-direction rel={vec_rel:.2e} "
         f"max scaled={scaled:.2e} worst per-direction rel h=1e-6 "
         f"{worst:.2e}, h=1e-5 {worst5:.2e} (median "
         f"{float(np.median(rel5)):.2e}; worst direction has |g.d| = "
         f"{abs(adjs[i5]) / scales[i5]:.1e} |g||d|, scaled error "
         f"{scaled5:.1e}) grad_b rel={abs(adj_b - fd_b) / abs(fd_b):.2e} "
         f"({time.perf_counter() - t0:.1f} s for 258 solves)")
    assert vec_rel <= 1e-6
    assert scaled <= 1e-8
    assert float(np.median(rel5)) <= 1e-7
    assert scaled5 <= 1e-9
    assert abs(adj_b - fd_b) <= 1e-8 * abs(fd_b)


def test_random_spheres_1024_forward_adjoint():
    """Combined large case (1024x512x512, 268M cells): flame sheet 1:8 with
    200 seeded spheres masked.  The FP32-storage V-cycle inside the FP64 CG
    reaches 1e-10, checked with the oracle's own FP64 operator, with every
    solid cell exactly 0.  Then dL/dlog rho for L = sum of p over a 64^3
    probe box, with the hierarchy rebuilt for the adjoint after the forward
    plan is freed; the directional derivative along the gradient itself is
    checked by a central difference (h = 1e-5, each side solved as a
    correction of the base solution so the difference carries no 1/h
    amplification of the tolerance).  Peak device memory is sampled with
    hipMemGetInfo after every phase and must stay far below the 288 GB."""
    import torch

    import hundun_flow as H
    case = hf_cases.random_spheres(n=(1024, 512, 512), seed=5)
    n = case["n"]
    nx, ny, nz = n
    dev = torch.device("cuda")
    free0, total = torch.cuda.mem_get_info(dev)
    low = [free0]

    def sample():
        torch.cuda.synchronize()
        low[0] = min(low[0], torch.cuda.mem_get_info(dev)[0])

    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    act = case["cell_active"] != 0
    masks = [t(case[k]) for k in ("cell_active", "x_active", "y_active",
                                  "z_active")]
    actf = masks[0].to(torch.float64)
    log_rho = t(np.log(case["rho"]))
    b = t(case["rhs"])
    pol = H.default_policy(precision="fp32", coarse_face_scaling=1)

    def build(lr):
        plan = H.Plan(n, case["faces"], case["bc"], pol)
        mob = torch.exp(-lr)
        diag, bx, by, bz = plan.assemble(mob)
        plan.set_coefficients(diag, bx, by, bz, *masks)
        plan.mobility = mob
        sample()
        return plan

    _say("[1024x512x512] forward")
    plan = build(log_rho)
    x, res = plan.solve_pcg(b, rtol=1e-10, maximum_iterations=1000,
                            flexible=True)
    torch.cuda.synchronize()
    sample()
    xh = x.cpu().numpy()
    rel = _oracle_residual(case, xh)
    _say(f"[1024x512x512] forward it={res.iterations} "
         f"{res.solve_milliseconds:.1f} ms oracle true_rel={rel:.2e}")
    assert res.termination == "converged"
    assert rel <= 1.01e-10
    assert np.all(xh[~act] == 0.0)
    del xh
    plan.project_mean(x)
    plan.close()
    probe = torch.zeros(nz, ny, nx, dtype=torch.float64, device=dev)
    probe[nz // 2 - 32:nz // 2 + 32, ny // 2 - 32:ny // 2 + 32,
          nx // 2 - 32:nx // 2 + 32] = 1.0
    probe = probe.reshape(-1) * actf
    plan = build(log_rho)
    # 1e-11: the FP64 rounding floor of the true residual of a 64^3 probe
    # source on 268M cells lies near 1e-12 of its norm
    ctl = dict(rtol=1e-11, maximum_iterations=400, flexible=True)
    _, gr, ares = plan.adjoint(plan.mobility, x, probe, want_b=False,
                               check=False, **ctl)
    torch.cuda.synchronize()
    sample()
    _say(f"[1024x512x512] adjoint it={ares.iterations} "
         f"{ares.solve_milliseconds:.1f} ms (hierarchy rebuilt) "
         f"termination={ares.termination} true residual "
         f"{ares.initial_true_residual:.3e} -> {ares.final_true_residual:.3e}")
    assert ares.termination == "converged"
    plan.close()
    # central difference along d = g / |g|
    g_norm = float(torch.linalg.norm(gr))
    d = gr / g_norm
    del gr
    rhs0 = b * actf
    h = 1e-5
    corr = []
    for sgn in (1.0, -1.0):
        pl = build(log_rho + sgn * h * d)
        r = rhs0 - pl.operator_apply(x)
        r = r * actf
        pl.project_mean(r)
        delta, dres = pl.solve_pcg(r, rtol=1e-10, maximum_iterations=1000,
                                   flexible=True)
        assert dres.termination == "converged"
        pl.project_mean(delta)
        pl.close()
        corr.append(delta)
    fd = float((corr[0] - corr[1]) @ probe) / (2 * h)
    err = abs(fd - g_norm) / g_norm
    peak = (total - low[0]) / 2**30
    _say(f"[1024x512x512] g.d = {g_norm:.6e} fd = {fd:.6e} rel={err:.2e} "
         f"peak device memory {peak:.1f} GiB of {total / 2**30:.0f}")
    assert err <= 1e-6
    assert peak <= 96.0
