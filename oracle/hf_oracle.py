"""ctypes binding of the CPU oracle (liboracle_hf.so) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  It is the checker, never part of the product path.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY_PATH = os.path.join(_HERE, "liboracle_hf.so")


class or_policy(ctypes.Structure):
    _fields_ = [("anisotropy_threshold", ctypes.c_double),
                ("pre_sweeps", ctypes.c_int), ("post_sweeps", ctypes.c_int),
                ("maximum_levels", ctypes.c_int),
                ("coarse_sweeps", ctypes.c_int),
                ("minimum_coarse_extent", ctypes.c_int),
                ("f_cycle", ctypes.c_int),
                ("null_space_constant", ctypes.c_int),
                ("unit_linear", ctypes.c_int),
                ("coarse_face_scaling", ctypes.c_int),
                ("line_order", ctypes.c_int),
                ("point_smoother", ctypes.c_int),
                ("chebyshev_lower_fraction", ctypes.c_double),
                ("diagonal_shortcut_ratio", ctypes.c_double)]


class or_fgmres_control(ctypes.Structure):
    _fields_ = [("absolute_tolerance", ctypes.c_double),
                ("relative_tolerance", ctypes.c_double),
                ("maximum_iterations", ctypes.c_uint32),
                ("true_residual_interval", ctypes.c_uint32),
                ("restart", ctypes.c_uint32),
                ("volume_rows", ctypes.c_int)]


class or_control(ctypes.Structure):
    _fields_ = [("absolute_tolerance", ctypes.c_double),
                ("relative_tolerance", ctypes.c_double),
                ("maximum_iterations", ctypes.c_uint32),
                ("true_residual_interval", ctypes.c_uint32),
                ("flexible", ctypes.c_int)]


class or_result(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("termination", ctypes.c_int),
                ("iterations", ctypes.c_uint32),
                ("initial_true_residual", ctypes.c_double),
                ("final_true_residual", ctypes.c_double),
                ("rhs_norm", ctypes.c_double),
                ("operator_applies", ctypes.c_uint64),
                ("preconditioner_applies", ctypes.c_uint64),
                ("breakdown_value", ctypes.c_double)]


_lib = None
_D = ctypes.POINTER(ctypes.c_double)
_U8 = ctypes.POINTER(ctypes.c_uint8)
_I = ctypes.POINTER(ctypes.c_int)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBRARY_PATH):
            raise RuntimeError("oracle not built: make -C oracle")
        L = ctypes.CDLL(LIBRARY_PATH)
        L.or_assemble.argtypes = [_I, _D, _D, _D, _I, _D, _D, _D, _D, _D]
        L.or_assemble.restype = ctypes.c_int
        L.or_operator_apply.argtypes = [_I, _I, _D, _D, _D, _D, _U8, _D, _D]
        L.or_operator_apply.restype = None
        L.or_masked_apply.argtypes = [ctypes.c_void_p, _D, _D]
        L.or_masked_apply.restype = None
        L.or_mg_create.argtypes = [_I, _D, _D, _D, _I,
                                   ctypes.POINTER(or_policy), _D, _D, _D, _D,
                                   _U8, _U8, _U8, _U8, _I]
        L.or_mg_create.restype = ctypes.c_void_p
        L.or_mg_destroy.argtypes = [ctypes.c_void_p]
        L.or_mg_destroy.restype = None
        L.or_mg_apply.argtypes = [ctypes.c_void_p, _D, _D]
        L.or_mg_apply.restype = ctypes.c_int
        L.or_mg_level_count.argtypes = [ctypes.c_void_p]
        L.or_mg_level_count.restype = ctypes.c_int
        L.or_mg_level_info.argtypes = [ctypes.c_void_p, ctypes.c_int, _I, _I, _I]
        L.or_mg_level_info.restype = ctypes.c_int
        L.or_mg_level_array.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.or_mg_level_array.restype = _D
        L.or_mg_level_slot.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.or_mg_level_slot.restype = _D
        L.or_mg_smooth.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int]
        L.or_mg_smooth.restype = ctypes.c_int
        L.or_mg_restrict.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_mg_restrict.restype = ctypes.c_int
        L.or_mg_prolongate.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_mg_prolongate.restype = ctypes.c_int
        L.or_pcg_solve.argtypes = [_I, _I, _D, _D, _D, _D, _U8, ctypes.c_void_p,
                                   _D, _D, ctypes.POINTER(or_control),
                                   ctypes.POINTER(or_result)]
        L.or_pcg_solve.restype = ctypes.c_int
        L.or_pcg_dense.argtypes = [ctypes.c_int, _D, _D, _D, _D,
                                   ctypes.POINTER(or_control),
                                   ctypes.POINTER(or_result), _D,
                                   ctypes.c_uint32]
        L.or_pcg_dense.restype = ctypes.c_int
        L.or_fgmres_solve.argtypes = [_I, _D, _D, _D, _I, _D, _D, _D, _D,
                                      _U8, ctypes.c_void_p, _D, _D,
                                      ctypes.POINTER(or_fgmres_control),
                                      ctypes.POINTER(or_result)]
        L.or_fgmres_solve.restype = ctypes.c_int
        L.or_fgmres_dense.argtypes = [ctypes.c_int, _D, _D, _D, _D,
                                      ctypes.POINTER(or_fgmres_control),
                                      ctypes.POINTER(or_result), _D,
                                      ctypes.c_uint32]
        L.or_fgmres_dense.restype = ctypes.c_int
        L.or_assemble_pressure_system.argtypes = (
            [_I] + [_D] * 14 + [ctypes.c_double] * 3 + [_D, _D])
        L.or_assemble_pressure_system.restype = ctypes.c_int
        L.or_cold_coefficients.argtypes = (
            [_I, _D, _D, _D, _D, _D, _D, _D, _D, _D, _U8, _D, _D,
             ctypes.c_double, ctypes.c_double, ctypes.c_int, _I, _D, _D, _D,
             _D])
        L.or_cold_coefficients.restype = ctypes.c_int
        L.or_mg_diagonal_shortcut.argtypes = [ctypes.c_void_p]
        L.or_mg_diagonal_shortcut.restype = ctypes.c_int
        L.or_dot.argtypes = [_D, _D, ctypes.c_size_t]
        L.or_dot.restype = ctypes.c_double
        L.or_norm2.argtypes = [_D, ctypes.c_size_t]
        L.or_norm2.restype = ctypes.c_double
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_D)


def _u8(a):
    return None if a is None else a.ctypes.data_as(_U8)


def _i(seq):
    arr = (ctypes.c_int * len(seq))(*[int(v) for v in seq])
    return ctypes.cast(arr, _I), arr


def face_counts(n):
    nx, ny, nz = n
    return (nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1)


def assemble(n, faces, bc, mobility):
    fx, fy, fz = face_counts(n)
    cells = n[0] * n[1] * n[2]
    bx, by, bz = np.empty(fx), np.empty(fy), np.empty(fz)
    diag = np.empty(cells)
    pn, _k1 = _i(n)
    pb, _k2 = _i(bc)
    f = [np.ascontiguousarray(v, dtype=np.float64) for v in faces]
    m = np.ascontiguousarray(mobility, dtype=np.float64)
    st = lib().or_assemble(pn, _d(f[0]), _d(f[1]), _d(f[2]), pb, _d(m),
                           _d(bx), _d(by), _d(bz), _d(diag))
    if st != 0:
        raise RuntimeError(f"or_assemble status {st}")
    return diag, bx, by, bz


def operator_apply(n, bc, diag, bx, by, bz, cell_active, x):
    y = np.empty(n[0] * n[1] * n[2])
    pn, _k1 = _i(n)
    pb, _k2 = _i(bc)
    x = np.ascontiguousarray(x, dtype=np.float64)
    lib().or_operator_apply(pn, pb, _d(diag), _d(bx), _d(by), _d(bz),
                            _u8(cell_active), _d(x), _d(y))
    return y


def policy(**kw) -> or_policy:
    p = or_policy(anisotropy_threshold=4.0, pre_sweeps=2, post_sweeps=2,
                  maximum_levels=32, coarse_sweeps=24,
                  minimum_coarse_extent=3, f_cycle=0, null_space_constant=1,
                  unit_linear=1, coarse_face_scaling=0, line_order=0,
                  point_smoother=0, chebyshev_lower_fraction=0.3,
                  diagonal_shortcut_ratio=0.0)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class MG:
    """Oracle multigrid plan (NativeCartesianMgPlan restatement)."""

    def __init__(self, n, faces, bc, pol: or_policy, diag, bx, by, bz,
                 cell_active=None, x_active=None, y_active=None,
                 z_active=None):
        self.n = tuple(n)
        self.cells = n[0] * n[1] * n[2]
        w = [np.ascontiguousarray(np.diff(f), dtype=np.float64) for f in faces]
        self._keep = [w, diag, bx, by, bz, cell_active, x_active, y_active,
                      z_active]
        pn, self._k1 = _i(n)
        pb, self._k2 = _i(bc)
        st = ctypes.c_int()
        self.pol = pol
        self.h = lib().or_mg_create(pn, _d(w[0]), _d(w[1]), _d(w[2]), pb,
                                    ctypes.byref(pol), _d(diag), _d(bx),
                                    _d(by), _d(bz), _u8(cell_active),
                                    _u8(x_active), _u8(y_active),
                                    _u8(z_active), ctypes.byref(st))
        if not self.h:
            raise RuntimeError(f"or_mg_create status {st.value}")

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_mg_destroy(self.h)
            self.h = None

    def apply(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.empty(self.cells)
        st = lib().or_mg_apply(self.h, _d(r), _d(z))
        if st != 0:
            raise RuntimeError(f"or_mg_apply status {st}")
        return z

    def masked_apply(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.cells)
        lib().or_masked_apply(self.h, _d(x), _d(y))
        return y

    def level_count(self):
        return lib().or_mg_level_count(self.h)

    def level_info(self, level):
        shape = (ctypes.c_int * 3)()
        cm = ctypes.c_int()
        lm = ctypes.c_int()
        lib().or_mg_level_info(self.h, level, ctypes.cast(shape, _I),
                               ctypes.byref(cm), ctypes.byref(lm))
        return {"shape": tuple(shape), "coarsen_mask": cm.value,
                "line_mask": lm.value}

    def level_array(self, level, which):
        nx, ny, nz = self.level_info(level)["shape"]
        counts = [nx * ny * nz, (nx + 1) * ny * nz, nx * (ny + 1) * nz,
                  nx * ny * (nz + 1), nx * ny * nz, nx * ny * nz,
                  nx * ny * nz, nx * ny * nz]
        ptr = lib().or_mg_level_array(self.h, level, which)
        return np.ctypeslib.as_array(ptr, shape=(counts[which],)).copy()

    def diagonal_shortcut(self):
        return bool(lib().or_mg_diagonal_shortcut(self.h))

    def smooth(self, level, sweeps, reverse=0, retain=0):
        st = lib().or_mg_smooth(self.h, level, sweeps, reverse, retain)
        if st != 0:
            raise RuntimeError(f"or_mg_smooth status {st}")

    def slot(self, level, which):
        nx, ny, nz = self.level_info(level)["shape"]
        ptr = lib().or_mg_level_slot(self.h, level, which)
        return np.ctypeslib.as_array(ptr, shape=(nx * ny * nz,))


def pcg(n, bc, diag, bx, by, bz, cell_active, mg, rhs, x0=None, rtol=1e-10,
        atol=0.0, maximum_iterations=400, true_residual_interval=4,
        flexible=False):
    cells = n[0] * n[1] * n[2]
    x = np.zeros(cells) if x0 is None else np.array(x0, dtype=np.float64)
    ctl = or_control(atol, rtol, maximum_iterations, true_residual_interval,
                     1 if flexible else 0)
    res = or_result()
    pn, _k1 = _i(n)
    pb, _k2 = _i(bc)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    lib().or_pcg_solve(pn, pb, _d(diag), _d(bx), _d(by), _d(bz),
                       _u8(cell_active), None if mg is None else mg.h,
                       _d(rhs), _d(x), ctypes.byref(ctl), ctypes.byref(res))
    return x, res


def pcg_dense(a, m, rhs, x0=None, rtol=1e-12, atol=0.0,
              maximum_iterations=200, true_residual_interval=1000,
              flexible=False, trace=0):
    """The oracle PCG recurrence on a dense operator `a` with dense
    preconditioner `m` (None: identity).  Returns (x, result, iterates)."""
    n = a.shape[0]
    a = np.ascontiguousarray(a, dtype=np.float64)
    mm = None if m is None else np.ascontiguousarray(m, dtype=np.float64)
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    ctl = or_control(atol, rtol, maximum_iterations, true_residual_interval,
                     1 if flexible else 0)
    res = or_result()
    tr = np.zeros((max(trace, 1), n))
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    lib().or_pcg_dense(n, _d(a), None if mm is None else _d(mm), _d(rhs),
                       _d(x), ctypes.byref(ctl), ctypes.byref(res), _d(tr),
                       trace)
    return x, res, tr[:min(trace, res.iterations)]


def assemble_pressure_system(n, faces, bx, by, bz, phi, rho, rho_n,
                             drho_dp, a, rho_nm1=None, source=None):
    """(status, D, b) of the pressure-correction system; phi = (phi_x,
    phi_y, phi_z) face fluxes, a = (a0, a1, a2) BDF coefficients."""
    cells = n[0] * n[1] * n[2]
    w = [np.ascontiguousarray(np.diff(f), dtype=np.float64) for f in faces]
    diag = np.full(cells, np.nan)
    rhs = np.full(cells, np.nan)
    c = lambda v: None if v is None else _d(np.ascontiguousarray(
        v, dtype=np.float64))
    keep = [np.ascontiguousarray(v, dtype=np.float64) for v in
            (bx, by, bz, phi[0], phi[1], phi[2], rho, rho_n, drho_dp)]
    nm1 = None if rho_nm1 is None else np.ascontiguousarray(rho_nm1,
                                                            dtype=np.float64)
    src = None if source is None else np.ascontiguousarray(source,
                                                           dtype=np.float64)
    pn, _k1 = _i(n)
    st = lib().or_assemble_pressure_system(
        pn, _d(w[0]), _d(w[1]), _d(w[2]), *[_d(k) for k in keep[:8]],
        c(nm1), _d(keep[8]), c(src), float(a[0]), float(a[1]), float(a[2]),
        _d(diag), _d(rhs))
    return st, diag, rhs


def cold_coefficients(n, faces, proj, lim, fluid, rho, pressure,
                      pressure_reference, dt, kinds, fixed_eos=False):
    """(status, D, cx, cy, cz) of the cold-path pressure coefficient loop;
    proj, lim = per-axis face arrays, kinds = six side kinds (0 other,
    1 periodic, 2 pressure outlet)."""
    cells = n[0] * n[1] * n[2]
    w = [np.ascontiguousarray(np.diff(f), dtype=np.float64) for f in faces]
    fc = face_counts(n)
    diag = np.full(cells, np.nan)
    out = [np.full(c, np.nan) for c in fc]
    keep = [np.ascontiguousarray(v, dtype=np.float64)
            for v in (*proj, *lim, rho, pressure)]
    fl = None if fluid is None else np.ascontiguousarray(fluid,
                                                         dtype=np.uint8)
    pn, _k1 = _i(n)
    pk, _k2 = _i(kinds)
    st = lib().or_cold_coefficients(
        pn, _d(w[0]), _d(w[1]), _d(w[2]), *[_d(k) for k in keep[:6]],
        _u8(fl), _d(keep[6]), _d(keep[7]), float(pressure_reference),
        float(dt), 1 if fixed_eos else 0, pk, _d(diag), _d(out[0]),
        _d(out[1]), _d(out[2]))
    return st, diag, out[0], out[1], out[2]


def production_policy(**kw) -> or_policy:
    """The reference's production pressure policy
    (src/solver_piso_detail.hpp:10-19): Chebyshev-Jacobi F(1,2) cycle,
    lower spectrum fraction 0.3, diagonal shortcut at ratio 0.1; no null
    space (the compressible pressure rows carry a diagonal reaction)."""
    base = dict(pre_sweeps=1, post_sweeps=2, point_smoother=1, f_cycle=1,
                chebyshev_lower_fraction=0.3, diagonal_shortcut_ratio=0.1,
                null_space_constant=0)
    base.update(kw)
    return policy(**base)


def fgmres(n, faces, bc, diag, bx, by, bz, cell_active, mg, rhs, x0=None,
           rtol=1e-10, atol=0.0, maximum_iterations=400,
           true_residual_interval=4, restart=12, volume_rows=False):
    cells = n[0] * n[1] * n[2]
    x = np.zeros(cells) if x0 is None else np.array(x0, dtype=np.float64)
    ctl = or_fgmres_control(atol, rtol, maximum_iterations,
                            true_residual_interval, restart,
                            1 if volume_rows else 0)
    res = or_result()
    w = [np.ascontiguousarray(np.diff(f), dtype=np.float64) for f in faces]
    pn, _k1 = _i(n)
    pb, _k2 = _i(bc)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    lib().or_fgmres_solve(pn, _d(w[0]), _d(w[1]), _d(w[2]), pb, _d(diag),
                          _d(bx), _d(by), _d(bz), _u8(cell_active),
                          None if mg is None else mg.h, _d(rhs), _d(x),
                          ctypes.byref(ctl), ctypes.byref(res))
    return x, res


def fgmres_dense(a, m, rhs, x0=None, rtol=1e-12, atol=0.0,
                 maximum_iterations=200, true_residual_interval=1000,
                 restart=12, trace=0):
    n = a.shape[0]
    a = np.ascontiguousarray(a, dtype=np.float64)
    mm = None if m is None else np.ascontiguousarray(m, dtype=np.float64)
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    ctl = or_fgmres_control(atol, rtol, maximum_iterations,
                            true_residual_interval, restart, 0)
    res = or_result()
    tr = np.zeros((max(trace, 1), n))
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    lib().or_fgmres_dense(n, _d(a), None if mm is None else _d(mm), _d(rhs),
                          _d(x), ctypes.byref(ctl), ctypes.byref(res),
                          _d(tr), trace)
    return x, res, tr[:min(trace, res.iterations)]


def dot(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return lib().or_dot(_d(a), _d(b), a.size)


def norm2(v):
    v = np.ascontiguousarray(v, dtype=np.float64)
    return lib().or_norm2(_d(v), v.size)


def from_case(case, pol: or_policy | None = None):
    """Assemble coefficients and build the oracle MG for a case dict."""
    diag, bx, by, bz = assemble(case["n"], case["faces"], case["bc"],
                                case["mobility"])
    mg = MG(case["n"], case["faces"], case["bc"], pol or policy(), diag, bx,
            by, bz, case["cell_active"], case["x_active"], case["y_active"],
            case["z_active"])
    return (diag, bx, by, bz), mg
