"""Python binding of libhundun_flow.so (the gfx950 pressure solver C ABI).

The binding mirrors the reference's solver_* interface for this path:
`Plan` plays NativeCartesianMgPlan (compile = create + set_coefficients,
apply = mg_apply), `Plan.operator_apply` plays the IBM pressure
LinearOperator, and `Plan.solve_pcg` plays solve_pcg.  Arrays are torch CUDA
tensors (device memory); the library runs every kernel on the torch stream
current at plan creation.  There is no CPU fallback: without the built HIP
library every entry point raises.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY_PATH = os.path.join(_HERE, "libhundun_flow.so")

# StatusCode (include/hundun/v04_status.hpp:13-22 of the reference)
STATUS_NAMES = ["ok", "invalid_case", "invalid_plan", "allocation_failure",
                "mpi_failure", "numerical_failure", "rejected_step",
                "io_failure"]
TERMINATIONS = ["converged", "zero_rhs", "maximum_iterations", "breakdown",
                "non_finite", "operator_failure", "preconditioner_failure",
                "collective_failure", "invalid_plan"]
PRECISIONS = {"fp64": 0, "fp32": 1, "bf16": 2}
# kernel classes of hf_plan_profile, in order
PROFILE_CLASSES = ["rb_finest", "rb_coarse", "defect", "restrict", "prolong",
                   "tail", "operator", "update_xr", "true_residual",
                   "mg_entry_exit", "project", "update_p", "finalize",
                   "other", "sweep_finest", "sweep_coarse", "line",
                   "rb_first", "sweep_mean"]

ABI_FUNCTIONS = [
    "hf_policy_default", "hf_plan_create", "hf_plan_destroy",
    "hf_assemble_coefficients", "hf_assemble_pressure_system",
    "hf_cold_coefficients", "hf_plan_set_coefficients",
    "hf_operator_apply", "hf_mg_apply", "hf_pcg_solve",
    "hf_pcg_solve_static", "hf_fgmres_solve", "hf_adjoint_gradients",
    "hf_project_mean", "hf_plan_level_count", "hf_plan_level_info",
    "hf_plan_level_array", "hf_plan_device_bytes", "hf_plan_set_profiling",
    "hf_plan_profile", "hf_status_message", "hf_vector_engine_create",
    "hf_vector_engine_destroy", "hf_vector_engine_slot", "hf_vector_dot",
    "hf_vector_axpby", "hf_vector_copy",
]


class hf_status(ctypes.Structure):
    _fields_ = [("code", ctypes.c_int32), ("detail", ctypes.c_uint32)]


class hf_grid_desc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
                ("nz", ctypes.c_int32),
                ("faces_x", ctypes.POINTER(ctypes.c_double)),
                ("faces_y", ctypes.POINTER(ctypes.c_double)),
                ("faces_z", ctypes.POINTER(ctypes.c_double)),
                ("boundary", ctypes.c_int32 * 6)]


class hf_policy(ctypes.Structure):
    _fields_ = [("anisotropy_threshold", ctypes.c_double),
                ("pre_sweeps", ctypes.c_uint8),
                ("post_sweeps", ctypes.c_uint8),
                ("maximum_levels", ctypes.c_uint8),
                ("minimum_coarse_extent", ctypes.c_uint8),
                ("coarse_sweeps", ctypes.c_uint16),
                ("f_cycle", ctypes.c_uint8),
                ("null_space_constant", ctypes.c_uint8),
                ("unit_linear", ctypes.c_uint8),
                ("precision", ctypes.c_uint8),
                ("coarse_face_scaling", ctypes.c_uint8),
                ("line_order", ctypes.c_uint8),
                ("point_smoother", ctypes.c_uint8),
                ("chebyshev_lower_fraction", ctypes.c_double),
                ("diagonal_shortcut_ratio", ctypes.c_double)]


class hf_solve_control(ctypes.Structure):
    _fields_ = [("absolute_tolerance", ctypes.c_double),
                ("relative_tolerance", ctypes.c_double),
                ("maximum_iterations", ctypes.c_uint32),
                ("true_residual_interval", ctypes.c_uint32),
                ("flexible", ctypes.c_uint8)]


class hf_fgmres_control(ctypes.Structure):
    _fields_ = [("absolute_tolerance", ctypes.c_double),
                ("relative_tolerance", ctypes.c_double),
                ("maximum_iterations", ctypes.c_uint32),
                ("true_residual_interval", ctypes.c_uint32),
                ("restart", ctypes.c_uint32),
                ("volume_rows", ctypes.c_uint8)]


class hf_solve_result(ctypes.Structure):
    _fields_ = [("status", hf_status),
                ("termination", ctypes.c_int32),
                ("iterations", ctypes.c_uint32),
                ("initial_true_residual", ctypes.c_double),
                ("final_true_residual", ctypes.c_double),
                ("recursive_residual", ctypes.c_double),
                ("rhs_norm", ctypes.c_double),
                ("true_residual_limit", ctypes.c_double),
                ("operator_applies", ctypes.c_uint64),
                ("preconditioner_applies", ctypes.c_uint64),
                ("reduction_calls", ctypes.c_uint64),
                ("solve_milliseconds", ctypes.c_double),
                ("norm_breakdown_restarts", ctypes.c_uint64)]


class hf_level_info(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
                ("nz", ctypes.c_int32), ("coarsen_mask", ctypes.c_uint8),
                ("line_mask", ctypes.c_uint8), ("precision", ctypes.c_uint8),
                ("on_tail", ctypes.c_uint8)]


class HundunFlowError(RuntimeError):
    def __init__(self, what: str, status: hf_status):
        name = STATUS_NAMES[status.code] if 0 <= status.code < 8 else "?"
        super().__init__(f"{what}: {name} (detail {status.detail})")
        self.code = status.code
        self.detail = status.detail


_lib = None


def load_library(path: str = LIBRARY_PATH):
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"libhundun_flow.so not found at {path}: build it first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    # torch ships its own HIP and HSA runtimes.  Loading torch first lets this
    # library bind to those by soname, so the process holds one HIP runtime;
    # loaded the other way round, two runtimes compete for the device and the
    # second one finds none.
    import torch  # noqa: F401
    lib = ctypes.CDLL(path)
    P = ctypes.c_void_p
    D = ctypes.c_void_p  # device pointers
    sigs = {
        "hf_policy_default": [ctypes.POINTER(hf_policy)],
        "hf_plan_create": [ctypes.POINTER(hf_grid_desc),
                           ctypes.POINTER(hf_policy), P,
                           ctypes.POINTER(ctypes.c_void_p)],
        "hf_plan_destroy": [P],
        "hf_assemble_coefficients": [P, D, D, D, D, D],
        "hf_plan_set_coefficients": [P, D, D, D, D, D, D, D, D],
        "hf_assemble_pressure_system": [P, D, D, D, D, D, D, D, D, D, D, D,
                                        ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, D, D],
        "hf_cold_coefficients": [P, D, D, D, D, D, D, D, D, D,
                                 ctypes.c_double, ctypes.c_double,
                                 ctypes.c_int32,
                                 ctypes.POINTER(ctypes.c_int32), D, D, D, D],
        "hf_operator_apply": [P, D, D],
        "hf_mg_apply": [P, D, D],
        "hf_pcg_solve": [P, D, D, ctypes.POINTER(hf_solve_control),
                         ctypes.POINTER(hf_solve_result)],
        "hf_pcg_solve_static": [P, D, D, ctypes.POINTER(hf_solve_control),
                                ctypes.POINTER(hf_solve_result)],
        "hf_fgmres_solve": [P, D, D, ctypes.POINTER(hf_fgmres_control),
                            ctypes.POINTER(hf_solve_result)],
        "hf_adjoint_gradients": [P, D, D, D, ctypes.POINTER(hf_solve_control),
                                 D, D, ctypes.POINTER(hf_solve_result)],
        "hf_project_mean": [P, D],
        "hf_vector_engine_create": [ctypes.c_int64, ctypes.c_int32, P,
                                    ctypes.POINTER(ctypes.c_void_p)],
        "hf_vector_engine_destroy": [P],
        "hf_vector_engine_slot": [P, ctypes.c_int32,
                                  ctypes.POINTER(ctypes.c_void_p)],
        "hf_vector_dot": [P, ctypes.c_int64, D, D,
                          ctypes.POINTER(ctypes.c_double)],
        "hf_vector_axpby": [P, ctypes.c_int64, ctypes.c_double, D,
                            ctypes.c_double, D, D],
        "hf_vector_copy": [P, ctypes.c_int64, D, D],
        "hf_plan_level_count": [P, ctypes.POINTER(ctypes.c_int32)],
        "hf_plan_level_info": [P, ctypes.c_int32,
                               ctypes.POINTER(hf_level_info)],
        "hf_plan_level_array": [P, ctypes.c_int32, ctypes.c_int32, D],
        "hf_plan_device_bytes": [P, ctypes.POINTER(ctypes.c_uint64)],
        "hf_plan_set_profiling": [P, ctypes.c_int32],
        "hf_plan_profile": [P, ctypes.POINTER(ctypes.c_double),
                            ctypes.POINTER(ctypes.c_uint64), ctypes.c_int32,
                            ctypes.POINTER(ctypes.c_int32)],
    }
    for name, args in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = hf_status
    lib.hf_status_message.argtypes = [hf_status]
    lib.hf_status_message.restype = ctypes.c_char_p
    _lib = lib
    return lib


def _check(what: str, status: hf_status):
    if status.code != 0:
        raise HundunFlowError(what, status)


def default_policy(**overrides) -> hf_policy:
    lib = load_library()
    p = hf_policy()
    _check("hf_policy_default", lib.hf_policy_default(ctypes.byref(p)))
    for k, v in overrides.items():
        if k == "precision" and isinstance(v, str):
            v = PRECISIONS[v]
        setattr(p, k, v)
    return p


@dataclass
class SolveResult:
    status: int
    detail: int
    termination: str
    iterations: int
    initial_true_residual: float
    final_true_residual: float
    recursive_residual: float
    rhs_norm: float
    true_residual_limit: float
    operator_applies: int
    preconditioner_applies: int
    reduction_calls: int
    solve_milliseconds: float
    norm_breakdown_restarts: int = 0

    @classmethod
    def from_c(cls, r: hf_solve_result) -> "SolveResult":
        t = r.termination
        return cls(r.status.code, r.status.detail,
                   TERMINATIONS[t] if 0 <= t < len(TERMINATIONS) else str(t),
                   r.iterations, r.initial_true_residual,
                   r.final_true_residual, r.recursive_residual, r.rhs_norm,
                   r.true_residual_limit, r.operator_applies,
                   r.preconditioner_applies, r.reduction_calls,
                   r.solve_milliseconds, r.norm_breakdown_restarts)


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


class Plan:
    """A compiled hierarchy on one GPU (NativeCartesianMgPlan counterpart)."""

    def __init__(self, n, faces, bc, policy: hf_policy | None = None,
                 stream=None):
        import torch
        self._torch = torch
        self.lib = load_library()
        self.n = tuple(int(v) for v in n)
        self.cells = self.n[0] * self.n[1] * self.n[2]
        self.faces = [np.ascontiguousarray(f, dtype=np.float64) for f in faces]
        self.bc = list(bc)
        self.policy = policy if policy is not None else default_policy()
        if stream is None:
            stream = torch.cuda.current_stream()
        self.stream = stream
        g = hf_grid_desc()
        g.nx, g.ny, g.nz = self.n
        g.faces_x = self.faces[0].ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        g.faces_y = self.faces[1].ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        g.faces_z = self.faces[2].ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        for i, v in enumerate(self.bc):
            g.boundary[i] = int(v)
        handle = ctypes.c_void_p()
        _check("hf_plan_create",
               self.lib.hf_plan_create(ctypes.byref(g), ctypes.byref(self.policy),
                                       ctypes.c_void_p(stream.cuda_stream),
                                       ctypes.byref(handle)))
        self.handle = handle
        self.device = torch.device("cuda", torch.cuda.current_device())
        self.coefficients = None

    def close(self):
        if getattr(self, "handle", None):
            self.lib.hf_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- helpers -------------------------------------------------------
    def _f64(self, count):
        return self._torch.empty(count, dtype=self._torch.float64,
                                 device=self.device)

    def face_counts(self):
        nx, ny, nz = self.n
        return ((nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1))

    # -- setup ---------------------------------------------------------
    def assemble(self, mobility):
        fx, fy, fz = self.face_counts()
        diag = self._f64(self.cells)
        bx, by, bz = self._f64(fx), self._f64(fy), self._f64(fz)
        _check("hf_assemble_coefficients",
               self.lib.hf_assemble_coefficients(self.handle, _ptr(mobility),
                                                 _ptr(diag), _ptr(bx), _ptr(by),
                                                 _ptr(bz)))
        return diag, bx, by, bz

    def assemble_pressure_system(self, bx, by, bz, phi, rho, rho_n, drho_dp,
                                 a, rho_nm1=None, source=None):
        """D and b of the pressure-correction system
        (assemble_pressure_system_impl); phi = (phi_x, phi_y, phi_z) face
        fluxes, a = (a0, a1, a2) BDF coefficients."""
        diag = self._f64(self.cells)
        rhs = self._f64(self.cells)
        _check("hf_assemble_pressure_system",
               self.lib.hf_assemble_pressure_system(
                   self.handle, _ptr(bx), _ptr(by), _ptr(bz), _ptr(phi[0]),
                   _ptr(phi[1]), _ptr(phi[2]), _ptr(rho), _ptr(rho_n),
                   _ptr(rho_nm1), _ptr(drho_dp), _ptr(source), float(a[0]),
                   float(a[1]), float(a[2]), _ptr(diag), _ptr(rhs)))
        return diag, rhs

    def cold_coefficients(self, proj, lim, fluid, rho, pressure,
                          pressure_reference, dt, kinds, fixed_eos=False):
        """(D, cx, cy, cz) of the cold-path pressure coefficient loop;
        proj, lim = per-axis face arrays, kinds = six side kinds (0 other,
        1 periodic, 2 pressure outlet)."""
        fx, fy, fz = self.face_counts()
        diag = self._f64(self.cells)
        c = [self._f64(fx), self._f64(fy), self._f64(fz)]
        k = (ctypes.c_int32 * 6)(*[int(v) for v in kinds])
        _check("hf_cold_coefficients",
               self.lib.hf_cold_coefficients(
                   self.handle, _ptr(proj[0]), _ptr(proj[1]), _ptr(proj[2]),
                   _ptr(lim[0]), _ptr(lim[1]), _ptr(lim[2]), _ptr(fluid),
                   _ptr(rho), _ptr(pressure), float(pressure_reference),
                   float(dt), 1 if fixed_eos else 0, k, _ptr(diag),
                   _ptr(c[0]), _ptr(c[1]), _ptr(c[2])))
        return diag, c[0], c[1], c[2]

    def set_coefficients(self, diag, bx, by, bz, cell_active=None,
                         x_active=None, y_active=None, z_active=None):
        _check("hf_plan_set_coefficients",
               self.lib.hf_plan_set_coefficients(
                   self.handle, _ptr(diag), _ptr(bx), _ptr(by), _ptr(bz),
                   _ptr(cell_active), _ptr(x_active), _ptr(y_active),
                   _ptr(z_active)))
        self.coefficients = (diag, bx, by, bz, cell_active, x_active,
                             y_active, z_active)

    # -- compute -------------------------------------------------------
    def operator_apply(self, x, out=None):
        y = self._f64(self.cells) if out is None else out
        _check("hf_operator_apply",
               self.lib.hf_operator_apply(self.handle, _ptr(x), _ptr(y)))
        return y

    def mg_apply(self, r, out=None):
        z = self._f64(self.cells) if out is None else out
        _check("hf_mg_apply", self.lib.hf_mg_apply(self.handle, _ptr(r),
                                                   _ptr(z)))
        return z

    @staticmethod
    def control(rtol=1e-10, atol=0.0, maximum_iterations=400,
                true_residual_interval=4, flexible=False) -> hf_solve_control:
        c = hf_solve_control()
        c.absolute_tolerance = atol
        c.relative_tolerance = rtol
        c.maximum_iterations = maximum_iterations
        c.true_residual_interval = true_residual_interval
        c.flexible = 1 if flexible else 0
        return c

    def solve_pcg(self, rhs, solution=None, check=True, static=False,
                  **control):
        """PCG solve; static=True calls hf_pcg_solve_static, whose reduction
        scratch is the library's static per-GPU array (nothing allocated)."""
        x = self._torch.zeros_like(rhs) if solution is None else solution
        ctl = self.control(**control)
        res = hf_solve_result()
        name = "hf_pcg_solve_static" if static else "hf_pcg_solve"
        status = getattr(self.lib, name)(self.handle, _ptr(rhs), _ptr(x),
                                         ctypes.byref(ctl), ctypes.byref(res))
        result = SolveResult.from_c(res)
        if check:
            _check(name, status)
        return x, result

    def solve_fgmres(self, rhs, solution=None, check=True, rtol=1e-10,
                     atol=0.0, maximum_iterations=400,
                     true_residual_interval=4, restart=12,
                     volume_rows=False):
        """Restarted single-reduction FGMRES with the plan's multigrid as the
        flexible preconditioner (solve_fgmres); volume_rows selects the
        per-volume rows of the cold pressure system with the volume-scaled
        preconditioner."""
        x = self._torch.zeros_like(rhs) if solution is None else solution
        c = hf_fgmres_control()
        c.absolute_tolerance = atol
        c.relative_tolerance = rtol
        c.maximum_iterations = maximum_iterations
        c.true_residual_interval = true_residual_interval
        c.restart = restart
        c.volume_rows = 1 if volume_rows else 0
        res = hf_solve_result()
        status = self.lib.hf_fgmres_solve(self.handle, _ptr(rhs), _ptr(x),
                                          ctypes.byref(c), ctypes.byref(res))
        result = SolveResult.from_c(res)
        if check:
            _check("hf_fgmres_solve", status)
        return x, result

    def project_mean(self, field):
        _check("hf_project_mean",
               self.lib.hf_project_mean(self.handle, _ptr(field)))
        return field

    def adjoint(self, mobility, pressure, grad_p, want_b=True,
                want_logrho=True, check=True, **control):
        gb = self._f64(self.cells) if want_b else None
        gr = self._f64(self.cells) if want_logrho else None
        ctl = self.control(**control)
        res = hf_solve_result()
        status = self.lib.hf_adjoint_gradients(self.handle, _ptr(mobility),
                                               _ptr(pressure), _ptr(grad_p),
                                               ctypes.byref(ctl), _ptr(gb),
                                               _ptr(gr), ctypes.byref(res))
        if check:
            _check("hf_adjoint_gradients", status)
        return gb, gr, SolveResult.from_c(res)

    # -- inspection ----------------------------------------------------
    def level_count(self) -> int:
        v = ctypes.c_int32()
        _check("hf_plan_level_count",
               self.lib.hf_plan_level_count(self.handle, ctypes.byref(v)))
        return v.value

    def level_info(self, level: int) -> dict:
        info = hf_level_info()
        _check("hf_plan_level_info",
               self.lib.hf_plan_level_info(self.handle, level,
                                           ctypes.byref(info)))
        return {"shape": (info.nx, info.ny, info.nz),
                "coarsen_mask": info.coarsen_mask,
                "line_mask": info.line_mask, "precision": info.precision,
                "on_tail": info.on_tail}

    def level_array(self, level: int, which: int):
        s = self.level_info(level)["shape"]
        nx, ny, nz = s
        counts = [nx * ny * nz, (nx + 1) * ny * nz, nx * (ny + 1) * nz,
                  nx * ny * (nz + 1), nx * ny * nz]
        out = self._f64(counts[which])
        _check("hf_plan_level_array",
               self.lib.hf_plan_level_array(self.handle, level, which,
                                            _ptr(out)))
        return out

    def set_profiling(self, enabled: bool):
        _check("hf_plan_set_profiling",
               self.lib.hf_plan_set_profiling(self.handle, 1 if enabled else 0))

    def profile(self) -> dict:
        cap = len(PROFILE_CLASSES)
        ms = (ctypes.c_double * cap)()
        launches = (ctypes.c_uint64 * cap)()
        count = ctypes.c_int32()
        _check("hf_plan_profile",
               self.lib.hf_plan_profile(self.handle, ms, launches, cap,
                                        ctypes.byref(count)))
        return {PROFILE_CLASSES[i]: (ms[i], launches[i])
                for i in range(min(cap, count.value))}

    def device_bytes(self) -> int:
        v = ctypes.c_uint64()
        _check("hf_plan_device_bytes",
               self.lib.hf_plan_device_bytes(self.handle, ctypes.byref(v)))
        return v.value


def plan_from_case(case: dict, policy: hf_policy | None = None, device=None):
    """Create a plan, assemble coefficients on the GPU from the case's
    mobility and install them with the IBM activity masks."""
    import torch
    dev = device or torch.device("cuda", torch.cuda.current_device())
    plan = Plan(case["n"], case["faces"], case["bc"], policy)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    mob = t(case["mobility"])
    diag, bx, by, bz = plan.assemble(mob)
    masks = [t(case[k]) for k in ("cell_active", "x_active", "y_active",
                                  "z_active")]
    plan.set_coefficients(diag, bx, by, bz, *masks)
    plan.mobility = mob
    plan.masks = masks
    return plan


class PressureSolve:
    """torch.autograd.Function test binding: p = solve(b, log_rho).

    forward assembles beta(rho) on the GPU, builds the hierarchy and solves
    A(rho) p = b to the requested tolerance, returning the gauge-fixed p
    (zero fluid mean).  backward runs the adjoint solve (A is symmetric) with
    the hierarchy recomputed from rho instead of stored."""

    @staticmethod
    def build():
        import torch

        class _Fn(torch.autograd.Function):
            @staticmethod
            def forward(ctx, b, log_rho, case, policy, control):
                plan = Plan(case["n"], case["faces"], case["bc"], policy)
                mob = torch.exp(-log_rho).contiguous()
                diag, bx, by, bz = plan.assemble(mob)
                masks = [torch.from_numpy(np.ascontiguousarray(case[k])).to(b.device)
                         for k in ("cell_active", "x_active", "y_active",
                                   "z_active")]
                plan.set_coefficients(diag, bx, by, bz, *masks)
                rhs = b.contiguous().clone()
                rhs = rhs * masks[0].to(rhs.dtype)
                plan.project_mean(rhs)
                p, res = plan.solve_pcg(rhs, **control)
                plan.project_mean(p)
                plan.close()
                ctx.save_for_backward(log_rho, p)
                ctx.case = case
                ctx.policy = policy
                ctx.control = control
                return p

            @staticmethod
            def backward(ctx, grad_p):
                log_rho, p = ctx.saved_tensors
                case = ctx.case
                plan = Plan(case["n"], case["faces"], case["bc"], ctx.policy)
                mob = torch.exp(-log_rho).contiguous()
                diag, bx, by, bz = plan.assemble(mob)
                masks = [torch.from_numpy(np.ascontiguousarray(case[k])).to(p.device)
                         for k in ("cell_active", "x_active", "y_active",
                                   "z_active")]
                plan.set_coefficients(diag, bx, by, bz, *masks)
                gb, gr, _ = plan.adjoint(mob, p, grad_p.contiguous(),
                                         **ctx.control)
                gb = gb * masks[0].to(gb.dtype)
                # the forward projects b to zero fluid mean before solving
                fluid = masks[0].to(gb.dtype)
                gb = gb - fluid * (gb * fluid).sum() / fluid.sum()
                plan.close()
                return gb, gr, None, None, None

        return _Fn
