"""GPU adjoint (differentiation form) checks through the C ABI.

L(b, log rho) = probe . p with p the zero-fluid-mean solution of
A(rho) p = P b.  The adjoint solve (A is symmetric, so the transpose solve is
the same solve) gives dL/db and dL/dlog(rho).  Pins:
  * a dense reference on a tiny grid: the oracle operator assembled column by
    column and solved exactly (pseudo-inverse on the fluid block), with
    central finite differences of the exact solves for every log-rho entry;
  * dL/db against an independent forward solve with the probe as rhs;
  * central finite differences (h = 1e-6) of tight GPU solves along seeded
    random directions;
  * the torch.autograd.Function binding against the same finite differences.
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


def _t(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def tiny_case(n=(6, 5, 4), seed=31):
    rng = np.random.default_rng(seed)
    nx, ny, nz = n
    faces = [hf_cases.uniform_faces(nx), hf_cases.uniform_faces(ny),
             hf_cases.uniform_faces(nz)]
    bc = [1, 1, 1, 1, 2, 2]
    act = np.ones((nz, ny, nx), dtype=np.uint8)
    act[1, 2, 3] = 0
    act[2, 2, 3] = 0
    fx, fy, fz = hf_cases.face_activity(act, bc)
    log_rho = 0.4 * rng.standard_normal(nx * ny * nz)
    b = np.where(act.ravel() != 0, rng.standard_normal(nx * ny * nz), 0.0)
    probe = np.zeros(nx * ny * nz)
    probe[[7, 8, 30, 31, 65]] = 1.0
    probe *= act.ravel()
    return {"n": n, "faces": faces, "bc": bc, "log_rho": log_rho,
            "mobility": np.exp(-log_rho), "cell_active": act.ravel(),
            "x_active": fx, "y_active": fy, "z_active": fz, "b": b,
            "probe": probe}


def dense_solution(case, log_rho, b):
    n = case["n"]
    N = n[0] * n[1] * n[2]
    d, bx, by, bz = O.assemble(n, case["faces"], case["bc"], np.exp(-log_rho))
    act = case["cell_active"] != 0
    A = np.zeros((N, N))
    for c in range(N):
        e = np.zeros(N)
        e[c] = 1.0
        A[:, c] = O.operator_apply(n, case["bc"], d, bx, by, bz,
                                   case["cell_active"], e)
    F = A[np.ix_(act, act)]
    bf = b[act] - b[act].mean()
    p = np.zeros(N)
    p[act] = np.linalg.pinv(F) @ bf
    p[act] -= p[act].mean()
    return p


def plan_for(case, precision="fp64", scaling=1):
    import hundun_flow as H
    pol = H.default_policy(precision=precision, coarse_face_scaling=scaling)
    c = dict(case)
    c["mobility"] = np.exp(-case["log_rho"])
    return H.plan_from_case(c, pol)


def gpu_forward(plan, b, rtol):
    rhs = _t(np.where(plan.masks[0].cpu().numpy() != 0, b, 0.0))
    plan.project_mean(rhs)
    x, res = plan.solve_pcg(rhs, rtol=rtol, maximum_iterations=1000,
                            flexible=True)
    plan.project_mean(x)
    return x, res


def test_adjoint_matches_dense_reference():
    case = tiny_case()
    _say("[adjoint dense] tiny 6x5x4")
    plan = plan_for(case)
    x, res = gpu_forward(plan, case["b"], 1e-14)
    p_dense = dense_solution(case, case["log_rho"], case["b"])
    xg = x.cpu().numpy()
    assert np.linalg.norm(xg - p_dense) <= 1e-11 * np.linalg.norm(p_dense)
    gb, gr, _ = plan.adjoint(plan.mobility, x, _t(case["probe"]), rtol=1e-14,
                             maximum_iterations=1000, flexible=True)
    gb, gr = gb.cpu().numpy(), gr.cpu().numpy()
    # dense reference gradients: exact solves, central differences in log rho
    h = 1e-5
    N = len(case["b"])
    ref_gr = np.zeros(N)
    for c in range(N):
        lp = case["log_rho"].copy()
        lm = case["log_rho"].copy()
        lp[c] += h
        lm[c] -= h
        ref_gr[c] = (case["probe"] @ dense_solution(case, lp, case["b"])
                     - case["probe"] @ dense_solution(case, lm, case["b"])) / (2 * h)
    act = case["cell_active"] != 0
    ref_gb = np.zeros(N)
    pr = case["probe"][act] - case["probe"][act].mean()
    d, bx, by, bz = O.assemble(case["n"], case["faces"], case["bc"],
                               case["mobility"])
    lam = dense_solution(case, case["log_rho"], case["probe"])
    ref_gb[act] = lam[act]
    rel_r = np.linalg.norm(gr - ref_gr) / np.linalg.norm(ref_gr)
    rel_b = np.linalg.norm(gb - ref_gb) / np.linalg.norm(ref_gb)
    _say(f"  grad_logrho rel={rel_r:.2e} grad_b rel={rel_b:.2e}")
    assert rel_r <= 1e-6
    assert rel_b <= 1e-10
    assert np.all(gr[~act] == 0.0)
    plan.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_adjoint_grad_b_is_the_transpose_solve(precision):
    case = hf_cases.flame_probe(n=(32, 32, 32), seed=41)
    case["log_rho"] = np.log(case["rho"])
    case["b"] = case["rhs"]
    _say(f"[adjoint transpose] 32^3 precision={precision}")
    plan = plan_for(case, precision)
    x, _ = gpu_forward(plan, case["b"], 1e-12)
    gb, _, res = plan.adjoint(plan.mobility, x, _t(case["probe"]), rtol=1e-12,
                              maximum_iterations=1000, flexible=True)
    lam, _ = gpu_forward(plan, case["probe"], 1e-12)
    rel = float((gb - lam).norm() / lam.norm())
    _say(f"  rel={rel:.2e} iterations={res.iterations}")
    assert rel <= 1e-9
    # symmetry of the operator (the transpose solve is the same solve)
    rng = np.random.default_rng(2)
    u = _t(rng.standard_normal(plan.cells))
    v = _t(rng.standard_normal(plan.cells))
    uav = float(u @ plan.operator_apply(v))
    vau = float(v @ plan.operator_apply(u))
    assert abs(uav - vau) <= 1e-12 * abs(uav)
    plan.close()


def fd_directions(case, plan_factory, directions, h, rtol, seed=7):
    """Central FD of L along random log-rho and b directions with tight
    solves; returns (adjoint . d, FD) pairs."""
    import torch
    rng = np.random.default_rng(seed)
    plan = plan_factory(case["log_rho"])
    x, _ = gpu_forward(plan, case["b"], rtol)
    gb, gr, _ = plan.adjoint(plan.mobility, x, _t(case["probe"]), rtol=rtol,
                             maximum_iterations=2000, flexible=True)
    gb, gr = gb.cpu().numpy(), gr.cpu().numpy()
    plan.close()
    probe = case["probe"]
    out = []
    for i in range(directions):
        d = rng.standard_normal(len(case["b"])) * (case["cell_active"] != 0)
        vals = []
        for sgn in (1.0, -1.0):
            pl = plan_factory(case["log_rho"] + sgn * h * d)
            xs, _ = gpu_forward(pl, case["b"], rtol)
            vals.append(float(xs.cpu().numpy() @ probe))
            pl.close()
        fd = (vals[0] - vals[1]) / (2 * h)
        out.append((float(gr @ d), fd))
    # grad_b: L is linear in b, so the FD is exact up to the solve
    d = rng.standard_normal(len(case["b"])) * (case["cell_active"] != 0)
    pl = plan_factory(case["log_rho"])
    xp, _ = gpu_forward(pl, case["b"] + d, rtol)
    xm, _ = gpu_forward(pl, case["b"] - d, rtol)
    pl.close()
    fd_b = float((xp - xm).cpu().numpy() @ probe) / 2.0
    return out, (float(gb @ d), fd_b)


def test_adjoint_finite_differences_random_directions():
    case = hf_cases.flame_probe(n=(24, 24, 24), seed=43)
    case["log_rho"] = np.log(case["rho"])
    case["b"] = case["rhs"]
    _say("[adjoint fd] 24^3, 8 directions, h=1e-6")

    def factory(log_rho):
        c = dict(case)
        c["log_rho"] = log_rho
        return plan_for(c)
    pairs, (ab, fb) = fd_directions(case, factory, 8, 1e-6, 1e-12)
    worst = 0.0
    for adj, fd in pairs:
        rel = abs(adj - fd) / max(abs(fd), 1e-300)
        worst = max(worst, rel)
        _say(f"  adj={adj:+.12e} fd={fd:+.12e} rel={rel:.2e}")
    _say(f"  grad_b adj={ab:+.12e} fd={fb:+.12e}")
    assert worst <= 1e-6
    assert abs(ab - fb) <= 1e-9 * abs(fb)


def test_autograd_binding():
    import torch

    import hundun_flow as H
    case = hf_cases.flame_probe(n=(16, 16, 16), seed=47)
    log_rho = torch.from_numpy(np.log(case["rho"])).cuda().requires_grad_(True)
    b = torch.from_numpy(case["rhs"]).cuda().requires_grad_(True)
    probe = torch.from_numpy(case["probe"]).cuda()
    fn = H.PressureSolve.build()
    policy = H.default_policy(coarse_face_scaling=1)
    control = dict(rtol=1e-12, maximum_iterations=1000, flexible=True)
    p = fn.apply(b, log_rho, case, policy, control)
    loss = (p * probe).sum()
    loss.backward()
    _say("[autograd] 16^3")
    rng = np.random.default_rng(3)
    d = torch.from_numpy(rng.standard_normal(len(case["rhs"])) *
                         (case["cell_active"] != 0)).cuda()
    h = 1e-6
    with torch.no_grad():
        lp = (fn.apply(b, log_rho + h * d, case, policy, control) * probe).sum()
        lm = (fn.apply(b, log_rho - h * d, case, policy, control) * probe).sum()
        fd = float((lp - lm) / (2 * h))
        adj = float(log_rho.grad @ d)
        lpb = (fn.apply(b + d, log_rho, case, policy, control) * probe).sum()
        lmb = (fn.apply(b - d, log_rho, case, policy, control) * probe).sum()
        fdb = float((lpb - lmb) / 2.0)
        adjb = float(b.grad @ d)
    _say(f"  logrho adj={adj:+.12e} fd={fd:+.12e}; b adj={adjb:+.12e} "
         f"fd={fdb:+.12e}")
    assert abs(adj - fd) <= 1e-6 * abs(fd)
    assert abs(adjb - fdb) <= 1e-9 * abs(fdb)
