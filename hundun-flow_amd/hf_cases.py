"""Synthetic pressure-solve workloads (numpy, seeded, deterministic).

Each case is a dict with
  n        (nx, ny, nz)
  faces    [faces_x, faces_y, faces_z]  face coordinates (n+1 each)
  bc       six boundary kinds (x_min, x_max, y_min, y_max, z_min, z_max);
           0 Dirichlet, 1 Neumann, 2 periodic
  mobility per-cell m = 1/rho (float64, x-fastest)
  cell_active, x_active, y_active, z_active  uint8 activity (IBM masks)
  rhs      right-hand side (float64): zero on solids, zero Euclidean mean over
           fluid cells (the Neumann/periodic operator is singular)

Arrays are flat, x-fastest: cell (i,j,k) -> i + nx*(j + ny*k).  Face arrays:
x-faces (nx+1)*ny*nz, y-faces nx*(ny+1)*nz, z-faces nx*ny*(nz+1).

The workloads follow the reference's low-Mach pressure step: coefficient
1/rho on faces by the harmonic rule, Neumann walls, periodic span, immersed
solids as masked cells.
"""
from __future__ import annotations

import math

import numpy as np

NEUMANN = 1
PERIODIC = 2
DIRICHLET = 0


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(seed))


def real4(faces: np.ndarray) -> np.ndarray:
    """Face coordinates as REAL4 values: the reference's runtime-axes
    geometry stores face coordinates in single precision."""
    return faces.astype(np.float32).astype(np.float64)


def uniform_faces(n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
    return real4(lo + (hi - lo) * np.arange(n + 1, dtype=np.float64) / n)


def tanh_faces(n: int, ratio: float) -> np.ndarray:
    """tanh-stretched faces on [0,1] clustered at both ends; the ratio of the
    largest to the smallest width is `ratio` in the continuum limit
    (cosh^2(s) = ratio)."""
    s = math.acosh(math.sqrt(ratio))
    eta = np.arange(n + 1, dtype=np.float64) / n
    f = 0.5 + 0.5 * np.tanh(s * (2.0 * eta - 1.0)) / math.tanh(s)
    f[0] = 0.0
    f[-1] = 1.0
    return real4(f)


def centres(faces: np.ndarray) -> np.ndarray:
    w = np.diff(faces)
    return faces[:-1] + 0.5 * w


def box_filter(field: np.ndarray, periodic: tuple[bool, bool, bool],
               passes: int) -> np.ndarray:
    """3x3x3 box filter; field shape (nz, ny, nx); periodic per (x, y, z)."""
    f = field.astype(np.float64, copy=True)
    axes = {0: 2, 1: 1, 2: 0}  # x -> numpy axis 2, ...
    for _ in range(passes):
        for a in range(3):
            ax = axes[a]
            if periodic[a]:
                f = (np.roll(f, 1, axis=ax) + f + np.roll(f, -1, axis=ax)) / 3.0
            else:
                pad = [(0, 0)] * 3
                pad[ax] = (1, 1)
                g = np.pad(f, pad, mode="edge")
                sl = [slice(None)] * 3
                lo = sl.copy()
                lo[ax] = slice(0, -2)
                mid = sl.copy()
                mid[ax] = slice(1, -1)
                hi = sl.copy()
                hi[ax] = slice(2, None)
                f = (g[tuple(lo)] + g[tuple(mid)] + g[tuple(hi)]) / 3.0
    return f


def face_activity(active: np.ndarray, bc) -> tuple[np.ndarray, np.ndarray,
                                                   np.ndarray]:
    """Face is active iff both adjacent cells are fluid; a physical boundary
    face follows its inside cell, a periodic one the two wrapped cells."""
    a = active.astype(bool)
    nz, ny, nx = a.shape
    out = []
    for axis, npax in ((0, 2), (1, 1), (2, 0)):
        lo_cells = np.concatenate(
            [np.take(a, [a.shape[npax] - 1], axis=npax), a], axis=npax)
        hi_cells = np.concatenate([a, np.take(a, [0], axis=npax)], axis=npax)
        f = lo_cells & hi_cells
        if bc[2 * axis] != PERIODIC:
            idx0 = [slice(None)] * 3
            idx0[npax] = 0
            f[tuple(idx0)] = np.take(a, 0, axis=npax)
            idxn = [slice(None)] * 3
            idxn[npax] = a.shape[npax]
            f[tuple(idxn)] = np.take(a, a.shape[npax] - 1, axis=npax)
        out.append(f.astype(np.uint8).ravel())
    return out[0], out[1], out[2]


def finish_rhs(rhs: np.ndarray, active: np.ndarray) -> np.ndarray:
    b = np.where(active.ravel() != 0, rhs.ravel(), 0.0).astype(np.float64)
    fluid = active.ravel() != 0
    b[fluid] -= b[fluid].sum() / fluid.sum()
    b[~fluid] = 0.0
    return b


def _grid(faces):
    xc, yc, zc = (centres(f) for f in faces)
    Z, Y, X = np.meshgrid(zc, yc, xc, indexing="ij")
    return X, Y, Z


def _pack(n, faces, bc, rho, active, rhs):
    active = active.astype(np.uint8)
    fx, fy, fz = face_activity(active, bc)
    return {
        "n": tuple(int(v) for v in n),
        "faces": [np.ascontiguousarray(f, dtype=np.float64) for f in faces],
        "bc": list(bc),
        "mobility": np.ascontiguousarray((1.0 / rho).ravel(), dtype=np.float64),
        "rho": np.ascontiguousarray(rho.ravel(), dtype=np.float64),
        "cell_active": np.ascontiguousarray(active.ravel()),
        "x_active": fx,
        "y_active": fy,
        "z_active": fz,
        "rhs": finish_rhs(rhs, active),
    }


def flame_sheet(Y: np.ndarray, ratio: float = 7.0) -> np.ndarray:
    return 1.0 + (ratio - 1.0) * (1.0 + np.tanh((Y - 0.5) / 0.02)) / 2.0


def lognormal_sphere(n=(64, 64, 64), seed: int = 1, radius: float = 0.2,
                     sigma: float = 0.3, passes: int = 3) -> dict:
    """64^3 unit cube: rho = exp(sigma*xi) with xi ~ N(0,1) box-filtered
    (3 passes, in log space), sphere of radius 0.2 at the centre masked,
    Neumann walls in x/y, periodic z, rhs ~ N(0,1) projected."""
    nx, ny, nz = n
    rng = _rng(seed)
    faces = [uniform_faces(nx), uniform_faces(ny), uniform_faces(nz)]
    bc = [NEUMANN, NEUMANN, NEUMANN, NEUMANN, PERIODIC, PERIODIC]
    xi = rng.standard_normal((nz, ny, nx))
    log_rho = box_filter(sigma * xi, (False, False, True), passes)
    rho = np.exp(log_rho)
    X, Y, Z = _grid(faces)
    solid = (X - 0.5) ** 2 + (Y - 0.5) ** 2 + (Z - 0.5) ** 2 < radius * radius
    rhs = rng.standard_normal((nz, ny, nx))
    return _pack(n, faces, bc, rho, ~solid, rhs)


def divergence_rhs(faces, bc, rng, noise: float = 0.1) -> np.ndarray:
    """Discrete divergence of a random solenoidal-plus-noise face velocity
    (face_velocity), integrated over each cell."""
    u, v, w = face_velocity(faces, bc, rng, noise)
    dx, dy, dz = (np.diff(f) for f in faces)
    ax = dy[None, :, None] * dz[:, None, None]
    ay = dx[None, None, :] * dz[:, None, None]
    az = dx[None, None, :] * dy[None, :, None]
    div = ((u[:, :, 1:] - u[:, :, :-1]) * ax
           + (v[:, 1:, :] - v[:, :-1, :]) * ay
           + (w[1:, :, :] - w[:-1, :, :]) * az)
    return div


def face_velocity(faces, bc, rng, noise: float = 0.1):
    """Random solenoidal-plus-noise face velocity: u = curl(psi) sampled on
    faces (psi random, box-filtered) plus white noise; wall-normal
    velocities are zero, periodic faces wrap.  Arrays (nz, ny, nx+1),
    (nz, ny+1, nx), (nz+1, ny, nx)."""
    nx, ny, nz = (len(f) - 1 for f in faces)
    psi = [box_filter(rng.standard_normal((nz + 1, ny + 1, nx + 1)),
                      (False, False, False), 2) for _ in range(3)]
    dx, dy, dz = (np.diff(f) for f in faces)
    # face-normal velocities from edge potentials (discrete curl)
    u = ((psi[2][:-1, 1:, :] - psi[2][:-1, :-1, :]) / dy[None, :, None]
         - (psi[1][1:, :-1, :] - psi[1][:-1, :-1, :]) / dz[:, None, None])
    v = ((psi[0][1:, :, :-1] - psi[0][:-1, :, :-1]) / dz[:, None, None]
         - (psi[2][:-1, :, 1:] - psi[2][:-1, :, :-1]) / dx[None, None, :])
    w = ((psi[1][:, :-1, 1:] - psi[1][:, :-1, :-1]) / dx[None, None, :]
         - (psi[0][:, 1:, :-1] - psi[0][:, :-1, :-1]) / dy[None, :, None])
    u = u + noise * rng.standard_normal(u.shape)
    v = v + noise * rng.standard_normal(v.shape)
    w = w + noise * rng.standard_normal(w.shape)
    for arr, axis, npax in ((u, 0, 2), (v, 1, 1), (w, 2, 0)):
        if bc[2 * axis] == PERIODIC:
            first = [slice(None)] * 3
            first[npax] = 0
            last = [slice(None)] * 3
            last[npax] = -1
            arr[tuple(last)] = arr[tuple(first)]
        else:
            first = [slice(None)] * 3
            first[npax] = 0
            last = [slice(None)] * 3
            last[npax] = -1
            arr[tuple(first)] = 0.0
            arr[tuple(last)] = 0.0
    return u, v, w


def flame_cylinder(n=(256, 256, 256), seed: int = 2, ratio: float = 8.0,
                   density_ratio: float = 7.0, stretched: bool = True) -> dict:
    """256^3, y tanh-stretched (max/min spacing 8), flame-sheet rho (1:7),
    cylinder r=0.1 along z through (0.5, 0.5) masked, rhs = divergence of a
    random solenoidal-plus-noise velocity field.  Neumann x/y, periodic z."""
    nx, ny, nz = n
    rng = _rng(seed)
    faces = [uniform_faces(nx),
             tanh_faces(ny, ratio) if stretched else uniform_faces(ny),
             uniform_faces(nz)]
    bc = [NEUMANN, NEUMANN, NEUMANN, NEUMANN, PERIODIC, PERIODIC]
    X, Y, Z = _grid(faces)
    rho = flame_sheet(Y, density_ratio)
    solid = (X - 0.5) ** 2 + (Y - 0.5) ** 2 < 0.1 * 0.1
    rhs = divergence_rhs(faces, bc, rng)
    return _pack(n, faces, bc, rho, ~solid, rhs)


def flame_step(n=(512, 256, 256), seed: int = 3, noise: float = 0.05,
               density_ratio: float = 7.0) -> dict:
    """512x256x256 on [0,2]x[0,1]x[0,1]: flame-sheet rho (1:7) times 5%
    lognormal noise; backward-facing step (i < nx/8 and y < 0.25) masked;
    rhs ~ N(0,1) projected.  Neumann x/y, periodic z."""
    nx, ny, nz = n
    rng = _rng(seed)
    faces = [uniform_faces(nx, 0.0, 2.0), uniform_faces(ny), uniform_faces(nz)]
    bc = [NEUMANN, NEUMANN, NEUMANN, NEUMANN, PERIODIC, PERIODIC]
    X, Y, Z = _grid(faces)
    rho = flame_sheet(Y, density_ratio) * np.exp(
        noise * rng.standard_normal((nz, ny, nx)))
    ii = np.arange(nx)[None, None, :]
    solid = (ii < max(nx // 8, 1)) & (Y < 0.25)
    rhs = rng.standard_normal((nz, ny, nx))
    return _pack(n, faces, bc, rho, ~solid, rhs)


def flame_probe(n=(256, 256, 256), seed: int = 4) -> dict:
    """Differentiation workload: uniform 256^3, flame sheet 1:7, cylinder
    mask; loss = sum of pressure over a 32^3 probe box centred at
    (0.75, 0.5, 0.5) (probe mask returned as `probe`)."""
    case = flame_cylinder(n, seed=seed, stretched=False)
    nx, ny, nz = n
    faces = case["faces"]
    X, Y, Z = _grid(faces)
    # 32 cells across on large grids, never more than a quarter of the
    # domain (a box covering every fluid cell would sum a zero-mean field)
    half = [min(16 / nx, 0.125), min(16 / ny, 0.125), min(16 / nz, 0.125)]
    probe = ((np.abs(X - 0.75) < half[0]) & (np.abs(Y - 0.5) < half[1])
             & (np.abs(Z - 0.5) < half[2]))
    case["probe"] = np.ascontiguousarray(
        (probe.ravel() & (case["cell_active"] != 0)).astype(np.float64))
    return case


def random_spheres(n=(1024, 512, 512), seed: int = 5, count: int = 200,
                   density_ratio: float = 8.0) -> dict:
    """1024x512x512 on [0,2]x[0,1]x[0,1], flame sheet 1:8, 200 seeded
    spheres of radius 0.01-0.05 masked, rhs ~ N(0,1) projected."""
    nx, ny, nz = n
    rng = _rng(seed)
    faces = [uniform_faces(nx, 0.0, 2.0), uniform_faces(ny), uniform_faces(nz)]
    bc = [NEUMANN, NEUMANN, NEUMANN, NEUMANN, PERIODIC, PERIODIC]
    xc, yc, zc = (centres(f) for f in faces)
    solid = np.zeros((nz, ny, nx), dtype=bool)
    centre = rng.uniform([0.1, 0.1, 0.0], [1.9, 0.9, 1.0], size=(count, 3))
    radius = rng.uniform(0.01, 0.05, size=count)
    for (cx, cy, cz), r in zip(centre, radius):
        i0, i1 = np.searchsorted(xc, [cx - r, cx + r])
        j0, j1 = np.searchsorted(yc, [cy - r, cy + r])
        kz = np.arange(nz)
        dzw = np.minimum(np.abs(zc - cz), 1.0 - np.abs(zc - cz))
        ks = kz[dzw < r]
        if i1 <= i0 or j1 <= j0 or ks.size == 0:
            continue
        Xs = xc[i0:i1][None, None, :]
        Ys = yc[j0:j1][None, :, None]
        Zs = dzw[ks][:, None, None]
        inside = (Xs - cx) ** 2 + (Ys - cy) ** 2 + Zs ** 2 < r * r
        sub = solid[ks, j0:j1, i0:i1]
        solid[ks, j0:j1, i0:i1] = sub | inside
    Y = yc[None, :, None] * np.ones((nz, 1, nx))
    rho = flame_sheet(Y, density_ratio)
    rhs = rng.standard_normal((nz, ny, nx))
    return _pack(n, faces, bc, rho, ~solid, rhs)


def compressible_flame(n=(64, 64, 64), seed: int = 6,
                       acoustic_cfl: float = 4.0, bdf2: bool = True,
                       solid: bool = True, source: bool = True,
                       density_ratio: float = 7.0,
                       gamma: float = 1.4) -> dict:
    """Inputs of the low-Mach pressure-correction system
    (src/solver_pressure.cpp:110-332) on a flame-sheet density field: trial
    density rho, accepted rho_n (and rho_nm1 for BDF2) perturbed by 1%,
    drho/dp = rho / (gamma p) at p = 1, BDF coefficients for the step dt
    chosen so that c_max dt / h_min = acoustic_cfl, face coefficients by the
    face rule from the mobility dt / rho, face fluxes phi = u . A of a
    random solenoidal-plus-noise velocity, an optional small mass source,
    an optional cylinder mask.  Neumann x/y walls, periodic z.  The
    diagonal carries a0 V drho/dp > 0, so the system has no null space;
    small acoustic CFL numbers make it strongly diagonally dominant."""
    nx, ny, nz = n
    rng = _rng(seed)
    faces = [uniform_faces(nx), uniform_faces(ny), uniform_faces(nz)]
    bc = [NEUMANN, NEUMANN, NEUMANN, NEUMANN, PERIODIC, PERIODIC]
    X, Y, Z = _grid(faces)
    rho = flame_sheet(Y, density_ratio) * np.exp(
        0.02 * rng.standard_normal((nz, ny, nx)))
    h = min(float(np.min(np.diff(f))) for f in faces)
    c_max = math.sqrt(gamma / float(rho.min()))
    dt = acoustic_cfl * h / c_max
    if bdf2:
        a = (1.5 / dt, -2.0 / dt, 0.5 / dt)
    else:
        a = (1.0 / dt, -1.0 / dt, 0.0)
    rho_n = rho * (1.0 + 0.01 * rng.standard_normal(rho.shape))
    rho_nm1 = rho * (1.0 + 0.01 * rng.standard_normal(rho.shape))
    u, v, w = face_velocity(faces, bc, rng, noise=0.1)
    dx, dy, dz = (np.diff(f) for f in faces)
    phi = [u * (dy[None, :, None] * dz[:, None, None]),
           v * (dx[None, None, :] * dz[:, None, None]),
           w * (dx[None, None, :] * dy[None, :, None])]
    active = np.ones((nz, ny, nx), dtype=bool)
    if solid:
        active = ~((X - 0.5) ** 2 + (Y - 0.5) ** 2 < 0.1 * 0.1)
    act = active.astype(np.uint8)
    fx, fy, fz = face_activity(act, bc)
    flat = lambda arr: np.ascontiguousarray(arr.ravel(), dtype=np.float64)
    case = {
        "n": tuple(int(v) for v in n),
        "faces": [np.ascontiguousarray(f, dtype=np.float64) for f in faces],
        "bc": list(bc),
        "mobility": flat(dt / rho),
        "rho": flat(rho),
        "rho_n": flat(rho_n),
        "rho_nm1": flat(rho_nm1) if bdf2 else None,
        "drho_dp": flat(rho / gamma),
        "source": flat(0.01 * rng.standard_normal(rho.shape)) if source
        else None,
        "bdf": a,
        "phi": [flat(phi[0]), flat(phi[1]), flat(phi[2])],
        "cell_active": np.ascontiguousarray(act.ravel()),
        "x_active": fx,
        "y_active": fy,
        "z_active": fz,
        "dt": dt,
    }
    return case


CASES = {
    "lognormal_sphere": lognormal_sphere,
    "flame_cylinder": flame_cylinder,
    "flame_step": flame_step,
    "flame_probe": flame_probe,
    "random_spheres": random_spheres,
    "compressible_flame": compressible_flame,
}


def make(name: str, **kw) -> dict:
    return CASES[name](**kw)
