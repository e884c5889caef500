"""The C++ drop-in layer (include/hundun_flow_linear.hpp).

HipCartesianMgPlan / HipPressureOperator / solve_pcg mirror the reference's
NativeCartesianMgPlan / LinearOperator / solve_pcg
(versions/v0.4/include/hundun/v04_linear.hpp:396-452, :952-1013, :1070-1075).
A small test shim (tests/cpp/linear_api_check.cpp) drives that layer; the GPU
test checks it against the C ABI, bit for bit, on the same device buffers, and
the CPU test covers its host-only error paths.
"""
import ctypes
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SHIM = os.path.join(HERE, "cpp", "libhf_linear_check.so")
INVALID_PLAN = 2


def shim():
    if not os.path.exists(SHIM):
        pytest.skip("tests/cpp/libhf_linear_check.so not built (build())")
    import hundun_flow
    hundun_flow.load_library()  # one HIP runtime: torch's, loaded first
    lib = ctypes.CDLL(SHIM)
    P = ctypes.c_void_p
    lib.hf_linear_check.argtypes = [P] * 20
    lib.hf_linear_check.restype = ctypes.c_int32
    lib.hf_linear_check_errors.argtypes = [P]
    lib.hf_linear_check_errors.restype = None
    lib.hf_linear_check_surface.argtypes = [P] * 19
    lib.hf_linear_check_surface.restype = ctypes.c_int32
    return lib


def _policy_arrays(pol):
    pf = np.array([pol.anisotropy_threshold, pol.chebyshev_lower_fraction,
                   pol.diagonal_shortcut_ratio], dtype=np.float64)
    pi = np.array([pol.pre_sweeps, pol.post_sweeps, pol.maximum_levels,
                   pol.minimum_coarse_extent, pol.coarse_sweeps, pol.f_cycle,
                   pol.null_space_constant, pol.unit_linear, pol.precision,
                   pol.coarse_face_scaling, pol.line_order,
                   pol.point_smoother], dtype=np.int32)
    return pf, pi


def test_error_paths_without_device_work():
    lib = shim()
    out = np.zeros(2, dtype=np.int32)
    lib.hf_linear_check_errors(out.ctypes.data)
    # compile without face coordinates, solve_pcg over an empty plan
    assert out.tolist() == [INVALID_PLAN, INVALID_PLAN]


@pytest.mark.gpu
@pytest.mark.parametrize("precision,scaling", [(0, 0), (1, 1)])
def test_cpp_layer_matches_c_abi(precision, scaling):
    import torch

    import hf_cases
    import hundun_flow as H
    lib = shim()
    case = hf_cases.lognormal_sphere(n=(24, 20, 16), seed=5)
    print(f"[linear api] 24x20x16 precision={precision} scaling={scaling}",
          flush=True)
    pol = H.default_policy(precision=precision, coarse_face_scaling=scaling)
    plan = H.plan_from_case(case, pol)
    diag, bx, by, bz = plan.coefficients[:4]
    rhs = torch.from_numpy(case["rhs"]).cuda()
    flexible = 1 if precision != 0 or scaling else 0
    y_ref = plan.operator_apply(rhs)
    z_ref = plan.mg_apply(rhs)
    x_ref, r_ref = plan.solve_pcg(rhs, rtol=1e-10, maximum_iterations=400,
                                  flexible=bool(flexible))
    torch.cuda.synchronize()

    n = np.array(case["n"], dtype=np.int32)
    faces = [np.ascontiguousarray(f, dtype=np.float64) for f in case["faces"]]
    bc = np.array(case["bc"], dtype=np.int32)
    act = (ctypes.c_void_p * 4)(*[m.data_ptr() for m in plan.masks])
    pf, pi = _policy_arrays(pol)
    y = torch.empty_like(rhs)
    z = torch.empty_like(rhs)
    x = torch.zeros_like(rhs)
    cf = np.array([0.0, 1e-10], dtype=np.float64)
    ci = np.array([400, 4, flexible], dtype=np.int32)
    oi = np.zeros(6, dtype=np.int32)
    of = np.zeros(2, dtype=np.float64)
    torch.cuda.synchronize()
    code = lib.hf_linear_check(
        n.ctypes.data, faces[0].ctypes.data, faces[1].ctypes.data,
        faces[2].ctypes.data, bc.ctypes.data, ctypes.addressof(act),
        pf.ctypes.data, pi.ctypes.data, diag.data_ptr(), bx.data_ptr(),
        by.data_ptr(), bz.data_ptr(), rhs.data_ptr(), y.data_ptr(),
        z.data_ptr(), x.data_ptr(), cf.ctypes.data, ci.ctypes.data,
        oi.ctypes.data, of.ctypes.data)
    torch.cuda.synchronize()
    assert code == 0, oi
    assert oi[2] == 0  # converged
    assert oi[3] == r_ref.iterations
    assert oi[4] == plan.level_count()
    # unit-linear FP64 cycles certify fixed_general, otherwise flexible
    assert oi[5] == (1 if precision == 0 and pol.unit_linear else 2)
    assert torch.equal(y, y_ref)
    assert torch.equal(z, z_ref)
    assert torch.equal(x, x_ref)
    assert of[0] == r_ref.final_true_residual
    plan.close()


@pytest.mark.gpu
def test_cpp_layer_drop_in_surface():
    """solve_pcg over a caller's own operator (the host-driven path over the
    workspace's vector engine) against the device path; the prepared-batch
    lifecycle (apply_prepared equals apply bit for bit; a ticket from before
    a coefficient refresh, or of another owner, is refused with 603); the
    failure provenance of a refused operator apply; solve_fgmres and the plan
    counters through the layer."""
    import torch

    import hf_cases
    import hundun_flow as H
    lib = shim()
    case = hf_cases.lognormal_sphere(n=(24, 20, 16), seed=6)
    print("[linear api surface] 24x20x16 fp64", flush=True)
    pol = H.default_policy()
    plan = H.plan_from_case(case, pol)
    diag, bx, by, bz = plan.coefficients[:4]
    rhs = torch.from_numpy(case["rhs"]).cuda()
    x_ref, r_ref = plan.solve_pcg(rhs, rtol=1e-10, maximum_iterations=400)
    z_ref = plan.mg_apply(rhs)
    torch.cuda.synchronize()
    n = np.array(case["n"], dtype=np.int32)
    faces = [np.ascontiguousarray(f, dtype=np.float64) for f in case["faces"]]
    bc = np.array(case["bc"], dtype=np.int32)
    act = (ctypes.c_void_p * 4)(*[m.data_ptr() for m in plan.masks])
    pf, pi = _policy_arrays(pol)
    xg, xd, z, zp, xf = [torch.zeros_like(rhs) for _ in range(5)]
    oi = np.zeros(12, dtype=np.int32)
    torch.cuda.synchronize()
    code = lib.hf_linear_check_surface(
        n.ctypes.data, faces[0].ctypes.data, faces[1].ctypes.data,
        faces[2].ctypes.data, bc.ctypes.data, ctypes.addressof(act),
        pf.ctypes.data, pi.ctypes.data, diag.data_ptr(), bx.data_ptr(),
        by.data_ptr(), bz.data_ptr(), rhs.data_ptr(), xg.data_ptr(),
        xd.data_ptr(), z.data_ptr(), zp.data_ptr(), xf.data_ptr(),
        oi.ctypes.data)
    torch.cuda.synchronize()
    assert code == 0, (code, oi)
    err = float((xg - xd).norm() / xd.norm())
    print(f"  generic it={oi[1]} device it={oi[2]} rel diff={err:.3e}",
          flush=True)
    assert oi[0] == 0 and oi[1] == oi[2] == r_ref.iterations
    assert torch.equal(xd, x_ref)
    # the host-driven path sums in another order than the device solve
    assert err <= 1e-10
    assert oi[3] == 0 and torch.equal(z, z_ref) and torch.equal(zp, z)
    assert oi[4] == INVALID_PLAN and oi[5] == INVALID_PLAN
    assert oi[6] == INVALID_PLAN
    x_fg, r_fg = plan.solve_fgmres(rhs, rtol=1e-10, restart=12)
    assert oi[7] == 0 and oi[8] == r_fg.iterations
    assert torch.equal(xf, x_fg)
    assert oi[9] >= 1 + oi[1] and oi[10] == 2 and oi[11] > 0
    plan.close()


REFERENCE_INCLUDE = "/root/reference/versions/v0.4/include"


def test_reference_adapter_compiles_against_reference_headers():
    """tests/cpp/reference_adapter.cpp -- the reference-side binding of
    INTEGRATION.md (NativeCartesianMgSpec, MgCoefficientViews and
    LinearSolveInvocation of the reference translated to this library's
    C++ layer) -- compiles against the reference's own headers.  Compiled
    only; skipped where the reference tree or its MPI headers are absent."""
    import shutil
    import subprocess
    mpi = "/opt/conda/include"
    if not os.path.isdir(REFERENCE_INCLUDE) or not os.path.exists(
            os.path.join(mpi, "mpi.h")) or shutil.which("g++") is None:
        pytest.skip("reference headers or MPI headers not present")
    root = os.path.dirname(HERE)
    cmd = ["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra",
           "-Werror", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           "-I" + REFERENCE_INCLUDE, "-I" + mpi,
           "-I" + os.path.join(root, "include"),
           os.path.join(HERE, "cpp", "reference_adapter.cpp")]
    out = subprocess.run(cmd, capture_output=True, text=True)
    assert out.returncode == 0, out.stderr[-2000:]
