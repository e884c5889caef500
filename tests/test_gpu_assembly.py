"""On-device assembly and the reference's production pressure pairing.

* hf_assemble_pressure_system (src/solver_pressure.cpp:110-332) and
  hf_cold_coefficients (src/core_product_freeze.cpp:14279-14317) against
  the oracle, bit for bit: both are elementwise passes in the reference's
  operation order.  A refused input publishes nothing.
* The production pairing end to end on the compressible low-Mach system:
  face rule, system assembly and hierarchy on the device, then FGMRES(12)
  with the Chebyshev-Jacobi F(1,2) cycle and the diagonal shortcut policy
  (src/solver_piso_detail.hpp:10-19), against the oracle's FGMRES on the
  oracle-assembled system: same iteration and application counts, solution
  to 1e-10 relative.  A small acoustic CFL number selects the shortcut.
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
    return None if a is None else torch.from_numpy(
        np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("bdf2,source", [(True, True), (False, False),
                                         (True, False)])
def test_pressure_system_bitwise(torch_cuda, bdf2, source):
    import hundun_flow as H
    torch = torch_cuda
    case = hf_cases.compressible_flame(n=(24, 20, 16), seed=71, bdf2=bdf2,
                                       source=source)
    _say(f"[pressure system] bdf2={bdf2} source={source}")
    n, faces = case["n"], case["faces"]
    _, bx, by, bz = O.assemble(n, faces, case["bc"], case["mobility"])
    st, d_o, b_o = O.assemble_pressure_system(
        n, faces, bx, by, bz, case["phi"], case["rho"], case["rho_n"],
        case["drho_dp"], case["bdf"], rho_nm1=case["rho_nm1"],
        source=case["source"])
    assert st == 0
    plan = H.Plan(n, faces, case["bc"])
    d, b = plan.assemble_pressure_system(
        _t(torch, bx), _t(torch, by), _t(torch, bz),
        [_t(torch, p) for p in case["phi"]], _t(torch, case["rho"]),
        _t(torch, case["rho_n"]), _t(torch, case["drho_dp"]), case["bdf"],
        rho_nm1=_t(torch, case["rho_nm1"]), source=_t(torch, case["source"]))
    assert np.array_equal(d.cpu().numpy(), d_o)
    assert np.array_equal(b.cpu().numpy(), b_o)
    plan.close()


@pytest.mark.parametrize("poison", ["rho", "face", "flux"])
def test_pressure_system_refusal_publishes_nothing(torch_cuda, poison):
    import hundun_flow as H
    torch = torch_cuda
    case = hf_cases.compressible_flame(n=(16, 12, 10), seed=72)
    n, faces = case["n"], case["faces"]
    _, bx, by, bz = O.assemble(n, faces, case["bc"], case["mobility"])
    rho = case["rho"].copy()
    phi = [p.copy() for p in case["phi"]]
    if poison == "rho":
        rho[5] = -1.0
    elif poison == "face":
        by = by.copy()
        by[7] = -0.5
    else:
        phi[2][3] = np.inf
    plan = H.Plan(n, faces, case["bc"])
    diag = torch.full((plan.cells,), 7.0, dtype=torch.float64, device="cuda")
    rhs = torch.full_like(diag, 7.0)
    st = plan.lib.hf_assemble_pressure_system(
        plan.handle, *[H._ptr(_t(torch, v)) for v in (bx, by, bz)],
        *[H._ptr(_t(torch, p)) for p in phi], H._ptr(_t(torch, rho)),
        H._ptr(_t(torch, case["rho_n"])), H._ptr(_t(torch, case["rho_nm1"])),
        H._ptr(_t(torch, case["drho_dp"])), H._ptr(_t(torch, case["source"])),
        *[float(a) for a in case["bdf"]], H._ptr(diag), H._ptr(rhs))
    torch.cuda.synchronize()
    assert (st.code, st.detail) == (5, 9111)
    assert bool((diag == 7.0).all()) and bool((rhs == 7.0).all())
    plan.close()


def _cold_inputs(n, seed):
    rng = np.random.default_rng(seed)
    fc = O.face_counts(n)
    proj = [0.2 + rng.random(c) for c in fc]
    lim = [rng.random(c) for c in fc]
    cells = n[0] * n[1] * n[2]
    return proj, lim, 0.5 + rng.random(cells), 0.1 * rng.standard_normal(cells)


@pytest.mark.parametrize("kinds,fixed", [([0] * 6, False),
                                         ([2, 0, 1, 1, 0, 2], False),
                                         ([1, 1, 1, 1, 1, 1], True)])
def test_cold_coefficients_bitwise(torch_cuda, kinds, fixed):
    import hundun_flow as H
    torch = torch_cuda
    case = hf_cases.lognormal_sphere(n=(20, 18, 14), seed=73, radius=0.25)
    n, faces = case["n"], case["faces"]
    proj, lim, rho, pres = _cold_inputs(n, 74)
    fluid = case["cell_active"]
    _say(f"[cold coefficients] kinds={kinds} fixed_eos={fixed}")
    st, d_o, cx_o, cy_o, cz_o = O.cold_coefficients(
        n, faces, proj, lim, fluid, rho, pres, 1.5, 0.02, kinds,
        fixed_eos=fixed)
    assert st == 0
    plan = H.Plan(n, faces, case["bc"])
    d, cx, cy, cz = plan.cold_coefficients(
        [_t(torch, p) for p in proj], [_t(torch, v) for v in lim],
        _t(torch, fluid), _t(torch, rho), _t(torch, pres), 1.5, 0.02, kinds,
        fixed_eos=fixed)
    for got, want in ((d, d_o), (cx, cx_o), (cy, cy_o), (cz, cz_o)):
        assert np.array_equal(got.cpu().numpy(), want)
    plan.close()


# (an odd cell count takes the one-element-per-thread basis combinations and
# odd rows the per-cell Krylov and cycle kernels)
@pytest.mark.parametrize("cfl,shortcut,shape", [
    (4.0, False, (32, 32, 24)), (0.05, True, (32, 32, 24)),
    (4.0, False, (21, 17, 13))])
def test_production_pairing_matches_oracle(torch_cuda, cfl, shortcut, shape):
    import hundun_flow as H
    torch = torch_cuda
    case = hf_cases.compressible_flame(n=shape, seed=75, acoustic_cfl=cfl)
    n, faces, bc = case["n"], case["faces"], case["bc"]
    _say(f"[production pairing] cfl={cfl} shape={shape}")
    # oracle: face rule, system, production policy, FGMRES(12)
    _, bx, by, bz = O.assemble(n, faces, bc, case["mobility"])
    st, d_o, b_o = O.assemble_pressure_system(
        n, faces, bx, by, bz, case["phi"], case["rho"], case["rho_n"],
        case["drho_dp"], case["bdf"], rho_nm1=case["rho_nm1"],
        source=case["source"])
    assert st == 0
    act = case["cell_active"]
    b_o = np.where(act != 0, b_o, 0.0)
    mg = O.MG(n, faces, bc, O.production_policy(), d_o, bx, by, bz, act,
              case["x_active"], case["y_active"], case["z_active"])
    assert mg.diagonal_shortcut() == shortcut
    xo, ores = O.fgmres(n, faces, bc, d_o, bx, by, bz, act, mg, b_o,
                        rtol=1e-10, restart=12)
    assert ores.termination == 0
    # device: the same chain on the GPU
    pol = H.default_policy(pre_sweeps=1, post_sweeps=2, point_smoother=1,
                           f_cycle=1, chebyshev_lower_fraction=0.3,
                           diagonal_shortcut_ratio=0.1, null_space_constant=0)
    plan = H.Plan(n, faces, bc, pol)
    _, gbx, gby, gbz = plan.assemble(_t(torch, case["mobility"]))
    d, b = plan.assemble_pressure_system(
        gbx, gby, gbz, [_t(torch, p) for p in case["phi"]],
        _t(torch, case["rho"]), _t(torch, case["rho_n"]),
        _t(torch, case["drho_dp"]), case["bdf"],
        rho_nm1=_t(torch, case["rho_nm1"]), source=_t(torch, case["source"]))
    masks = [_t(torch, case[k]) for k in ("cell_active", "x_active",
                                          "y_active", "z_active")]
    plan.set_coefficients(d, gbx, gby, gbz, *masks)
    b = b * masks[0].to(b.dtype)
    assert np.array_equal(b.cpu().numpy(), b_o)
    x, res = plan.solve_fgmres(b, rtol=1e-10, restart=12)
    err = np.linalg.norm(x.cpu().numpy() - xo) / np.linalg.norm(xo)
    _say(f"  it gpu={res.iterations} oracle={ores.iterations} "
         f"ops={res.operator_applies}/{ores.operator_applies} "
         f"pcs={res.preconditioner_applies}/{ores.preconditioner_applies} "
         f"err={err:.3e}")
    assert res.termination == "converged"
    assert res.iterations == ores.iterations
    assert res.operator_applies == ores.operator_applies
    assert res.preconditioner_applies == ores.preconditioner_applies
    assert err <= 1e-10, err
    # the FP32 cycle of the same pairing reaches the same tolerance, checked
    # with the oracle's FP64 operator
    plan32 = H.Plan(n, faces, bc, H.default_policy(
        pre_sweeps=1, post_sweeps=2, point_smoother=1, f_cycle=1,
        diagonal_shortcut_ratio=0.1, null_space_constant=0, precision="fp32"))
    plan32.set_coefficients(d, gbx, gby, gbz, *masks)
    x32, r32 = plan32.solve_fgmres(b, rtol=1e-10, restart=12)
    xs = x32.cpu().numpy()
    ax = O.operator_apply(n, bc, d_o, bx, by, bz, act, xs)
    rel = np.linalg.norm(b_o - ax) / np.linalg.norm(b_o)
    e32 = np.linalg.norm(xs - xo) / np.linalg.norm(xo)
    _say(f"  fp32 it={r32.iterations} true_rel={rel:.3e} err={e32:.3e}")
    assert r32.termination == "converged" and rel <= 1.01e-10
    plan.close()
    plan32.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16"])
def test_production_solve_same_on_per_cell_route(torch_cuda, monkeypatch,
                                                 precision):
    """The production solve (FGMRES(12), Chebyshev F(1,2)) through the pair
    kernels, the b-and-D-only first stage and the combine that stores the
    multigrid entry gives the iterates of the per-cell route over a stored
    zero x with the separate entry: the same iteration count and the same
    solution bit for bit."""
    torch = torch_cuda
    import hundun_flow as H
    case = hf_cases.compressible_flame(n=(40, 24, 20), seed=76,
                                       acoustic_cfl=4.0)
    masks = [_t(torch, case[k]) for k in ("cell_active", "x_active",
                                          "y_active", "z_active")]
    out = {}
    for on in ("1", "0"):
        monkeypatch.setenv("HF_ROW_PAIRS", on)
        monkeypatch.setenv("HF_CHEB_FIRST", on)
        plan = H.Plan(case["n"], case["faces"], case["bc"], H.default_policy(
            pre_sweeps=1, post_sweeps=2, point_smoother=1, f_cycle=1,
            diagonal_shortcut_ratio=0.0, null_space_constant=0,
            precision=precision))  # (ratio 0: the cycle always runs)
        _, bx, by, bz = plan.assemble(_t(torch, case["mobility"]))
        d, b = plan.assemble_pressure_system(
            bx, by, bz, [_t(torch, p) for p in case["phi"]],
            _t(torch, case["rho"]), _t(torch, case["rho_n"]),
            _t(torch, case["drho_dp"]), case["bdf"],
            rho_nm1=_t(torch, case["rho_nm1"]),
            source=_t(torch, case["source"]))
        b = b * masks[0].to(b.dtype)
        plan.set_coefficients(d, bx, by, bz, *masks)
        x, res = plan.solve_fgmres(b, rtol=1e-10, restart=12)
        out[on] = (x.cpu().numpy(), res.iterations, res.termination)
        plan.close()
    _say(f"[production routes] {precision} it={out['1'][1]}/{out['0'][1]}")
    assert out["1"][2] == "converged"
    assert out["1"][1] == out["0"][1]
    assert np.array_equal(out["1"][0], out["0"][0]), \
        np.max(np.abs(out["1"][0] - out["0"][0]))


@pytest.mark.parametrize("volume_rows", [False, True])
def test_fgmres_row_pair_kernels(torch_cuda, monkeypatch, volume_rows):
    """FGMRES with its level-0 operator, residual and dot kernels over pairs
    of cells of a row against the one-cell-per-lane kernels: the pair forms
    add each lane's two products in cell order, so the inner products round
    differently; the same iteration count and the solution to 1e-12."""
    torch = torch_cuda
    import hundun_flow as H
    case = hf_cases.compressible_flame(n=(40, 24, 20), seed=77,
                                       acoustic_cfl=4.0)
    masks = [_t(torch, case[k]) for k in ("cell_active", "x_active",
                                          "y_active", "z_active")]
    out = {}
    for on in ("1", "0"):
        monkeypatch.setenv("HF_KRYLOV_PAIRS", on)
        plan = H.Plan(case["n"], case["faces"], case["bc"], H.default_policy(
            pre_sweeps=1, post_sweeps=2, point_smoother=1, f_cycle=1,
            diagonal_shortcut_ratio=0.0, null_space_constant=0))
        _, bx, by, bz = plan.assemble(_t(torch, case["mobility"]))
        d, b = plan.assemble_pressure_system(
            bx, by, bz, [_t(torch, p) for p in case["phi"]],
            _t(torch, case["rho"]), _t(torch, case["rho_n"]),
            _t(torch, case["drho_dp"]), case["bdf"],
            rho_nm1=_t(torch, case["rho_nm1"]),
            source=_t(torch, case["source"]))
        b = b * masks[0].to(b.dtype)
        plan.set_coefficients(d, bx, by, bz, *masks)
        x, res = plan.solve_fgmres(b, rtol=1e-10, restart=12,
                                   volume_rows=volume_rows)
        out[on] = (x.cpu().numpy(), res.iterations, res.termination)
        plan.close()
    rel = (np.linalg.norm(out["1"][0] - out["0"][0])
           / np.linalg.norm(out["0"][0]))
    _say(f"[fgmres pairs] volume_rows={volume_rows} "
         f"it={out['1'][1]}/{out['0'][1]} rel={rel:.2e}")
    assert out["1"][2] == "converged" and out["0"][2] == "converged"
    assert out["1"][1] == out["0"][1]
    assert rel <= 1e-12, rel
