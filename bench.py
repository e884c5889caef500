"""Pressure-solve benchmark on one MI355X.

Workload (default): the 512x256x256 precision case -- flame-sheet rho (1:7)
with 5% lognormal noise, a backward-facing step masked, Neumann walls and a
periodic span; rhs ~ N(0,1) projected to zero fluid mean.  One step = one
complete pressure solve from x0 = 0 to a relative FP64 true residual of 1e-10,
with every input already resident in HBM.  Timed: the FP32-storage V(2,2)
multigrid inside the FP64 CG (the value), next to the plain FP64 solve.

Prints one JSON line.  `value` is ms per solve (lower is better);
`roofline` is the dominant kernel's algorithmic bytes per launch over its
average launch time (HIP events on the plan stream, measured in this run);
`cpu_baseline` times the CPU oracle on a bounded sample of the same grid.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for sub in ("hundun-flow_amd", "oracle"):
    sys.path.insert(0, os.path.join(ROOT, sub))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def storage_bytes(precision: str) -> int:
    return {"fp64": 8, "fp32": 4, "bf16": 2}[precision]


# Algorithmic bytes per cell of one launch, per kernel class (each array
# streamed once at its storage width s; neighbour reuse assumed perfect).
def class_bytes_per_cell(cls: str, s: int):
    return {
        # half-sweep: read x (all), b and D (active half), 3 face arrays,
        # write x (active half)
        "rb_finest": 5.5 * s,
        # first half-sweep from x = 0: read b, D; write b (projected), x
        "rb_first": 4 * s,
        # fused two-colour sweep: read x, b, D, 3 face arrays, write x (the
        # first sweep from zero reads no x and writes the projected b)
        "sweep_finest": 7 * s,
        # the cycle's last fused sweep, which also sums the level-0 mean:
        # the sweep's arrays and the fluid flags
        "sweep_mean": 7 * s + 1,
        # retained defect r = b - A x: x, b, D, 3 faces, r
        "defect": 7 * s,
        # Krylov q = A p + dot: p, D, 3 faces, q (FP64)
        "operator": 6 * 8,
        # x += a p, r -= a q: read x p r q, write x r
        "update_xr": 6 * 8,
        # r = b - A x: x, b, D, 3 faces, r
        "true_residual": 7 * 8,
        # p = z + beta p, z formed from the level-0 candidate: read x (s),
        # the fluid flags and p, write p
        "update_p": s + 1 + 2 * 8,
    }.get(cls)


# kernel class -> symbol of its kernel for the rocprof traffic lookup
CLASS_KERNELS = {
    "rb_finest": "void hf::k_rb<{t}>(hf::Tiles, hf::Lv<{t}>, hf::Bc, int)",
    "sweep_finest": "void hf::k_rb_sweep<{t}, false, false>(hf::SweepArgs<{t}>)",
    "operator": "void hf::k_stencil_z<double, 0>(hf::Tiles, hf::Lv<double>, hf::Bc, double const*, double const*, double*, double*, int)",
}


def measured_traffic(workload_tag, cls, precision, grid):
    """HBM bytes per launch of the class's kernel from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (copy-calibrated), or None."""
    path = os.path.join(ROOT, "profiles", f"{workload_tag}_hbm_traffic.json")
    if not os.path.exists(path) or cls not in CLASS_KERNELS:
        return None
    t = {"fp32": "float", "bf16": "hf::bf16", "fp64": "double"}[precision]
    name = CLASS_KERNELS[cls].format(t=t)
    with open(path) as f:
        data = json.load(f)
    entry = data["kernels"].get(f"{name}|{grid}")
    if entry is None and cls == "sweep_finest":
        # the fused sweep also runs on large coarse levels: the finest
        # level's launches are the ones moving the most bytes
        hits = [v for k, v in data["kernels"].items()
                if k.split("|")[0] == name]
        entry = max(hits, key=lambda v: v["fetch_bytes"] + v["write_bytes"],
                    default=None)
    if entry is None:
        return None
    return int(entry["fetch_bytes"] + entry["write_bytes"])


def make_case(name, nx, ny, nz):
    import hf_cases
    if name == "flame_step":
        return hf_cases.flame_step(n=(nx, ny, nz))
    if name == "flame_cylinder":
        return hf_cases.flame_cylinder(n=(nx, ny, nz))
    if name == "lognormal_sphere":
        return hf_cases.lognormal_sphere(n=(nx, ny, nz))
    if name == "random_spheres":
        return hf_cases.random_spheres(n=(nx, ny, nz))
    raise SystemExit(f"unknown workload {name}")


def gpu_solves(plan, b, steps, warmup, flexible, torch):
    x = torch.zeros_like(b)
    result = None
    for _ in range(warmup):
        x.zero_()
        x, result = plan.solve_pcg(b, x, rtol=1e-10, maximum_iterations=1000,
                                   flexible=flexible)
    torch.cuda.synchronize()
    times = []
    for _ in range(steps):
        x.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x, result = plan.solve_pcg(b, x, rtol=1e-10, maximum_iterations=1000,
                                   flexible=flexible)
        torch.cuda.synchronize()
        times.append((time.perf_counter() - t0) * 1e3)
    return x, result, times


def production_pairing(shape, steps, warmup, torch, H, dev):
    """The reference's production pressure path at full size: the low-Mach
    pressure-correction system assembled on the device from the compressible
    flame case (acoustic CFL 4, BDF2, cylinder mask), then FGMRES(12) with the
    Chebyshev-Jacobi F(1,2) cycle and the diagonal-shortcut policy
    (src/solver_piso_detail.hpp:10-19), FP64 and FP32 cycle storage, to 1e-10.
    Times are per solve (assembly and hierarchy build excluded, timed apart)."""
    import numpy as np

    import hf_cases
    nx, ny, nz = shape
    log(f"[bench] production pairing: compressible_flame {nx}x{ny}x{nz}")
    case = hf_cases.compressible_flame(n=(nx, ny, nz), seed=6,
                                       acoustic_cfl=4.0)
    tt = lambda a: None if a is None else torch.from_numpy(
        np.ascontiguousarray(a)).to(dev)
    masks = [tt(case[k]) for k in ("cell_active", "x_active", "y_active",
                                   "z_active")]
    inputs = dict(phi=[tt(v) for v in case["phi"]], rho=tt(case["rho"]),
                  rho_n=tt(case["rho_n"]), drho_dp=tt(case["drho_dp"]),
                  rho_nm1=tt(case["rho_nm1"]), source=tt(case["source"]))
    mobility = tt(case["mobility"])
    out = {"workload": f"compressible_flame {nx}x{ny}x{nz}, acoustic CFL 4",
           "krylov": "FGMRES(12), single reduction",
           "cycle": "F(1,2) Chebyshev-Jacobi, diagonal shortcut policy"}
    for prec in ("fp64", "fp32"):
        pol = H.default_policy(pre_sweeps=1, post_sweeps=2, point_smoother=1,
                               f_cycle=1, chebyshev_lower_fraction=0.3,
                               diagonal_shortcut_ratio=0.1,
                               null_space_constant=0, precision=prec)
        plan = H.Plan(case["n"], case["faces"], case["bc"], pol)
        # assembly and coefficient refresh as in a time step: the second of
        # two rounds is timed (the first also loads the kernels)
        for _ in range(2):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            _, bx, by, bz = plan.assemble(mobility)
            d, b = plan.assemble_pressure_system(
                bx, by, bz, inputs["phi"], inputs["rho"], inputs["rho_n"],
                inputs["drho_dp"], case["bdf"], rho_nm1=inputs["rho_nm1"],
                source=inputs["source"])
            b = b * masks[0].to(b.dtype)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            plan.set_coefficients(d, bx, by, bz, *masks)
            torch.cuda.synchronize()
            t2 = time.perf_counter()
        res = None
        times = []
        for i in range(warmup + steps):
            torch.cuda.synchronize()
            t = time.perf_counter()
            x, res = plan.solve_fgmres(b, rtol=1e-10, restart=12,
                                       maximum_iterations=400)
            torch.cuda.synchronize()
            if i >= warmup:
                times.append((time.perf_counter() - t) * 1e3)
        ax = plan.operator_apply(x)
        true_rel = float(torch.linalg.norm(b - ax) / torch.linalg.norm(b))
        log(f"[bench] production {prec}: iterations={res.iterations} "
            f"ms={['%.2f' % v for v in times]} true_rel={true_rel:.3e}")
        assert res.termination == "converged", res
        assert true_rel <= 1.01e-10, true_rel
        out[prec] = {"ms_per_solve": round(sum(times) / len(times), 3),
                     "iterations": res.iterations,
                     "true_relative_residual": true_rel,
                     "assembly_ms": round((t1 - t0) * 1e3, 3),
                     "hierarchy_build_ms": round((t2 - t1) * 1e3, 3)}
        plan.close()
    out["speedup_fp32_vs_fp64"] = round(out["fp64"]["ms_per_solve"] /
                                        out["fp32"]["ms_per_solve"], 3)
    return out


def cpu_port_sample(case):
    """The oracle (single-thread C restatement) on the full grid: one V(2,2)
    application + one operator application + three dot passes, i.e. one
    PCG iteration (seconds)."""
    import hf_oracle as O
    (d, bx, by, bz), mg = O.from_case(case, O.policy())
    r = case["rhs"]
    t0 = time.perf_counter()
    z = mg.apply(r)
    O.operator_apply(case["n"], case["bc"], d, bx, by, bz,
                     case["cell_active"], z)
    for _ in range(3):
        O.dot(r, z)
    per_iteration = time.perf_counter() - t0
    del mg
    return per_iteration


def cpu_reference_sample(case, short=1, long=3):
    """The reference's own CPU path (oracle/_ref/hf_ref_driver: the
    reference NativeCartesianMgPlan in the reference solve_pcg, compiled from
    the reference sources, one MPI rank) on the full grid, timed for `short`
    and `long` PCG iterations: the solve costs start + iterations x
    per_iteration.  Returns (start, per_iteration, compile) in seconds, or
    None when the driver was not built."""
    import hf_oracle as O
    import ref_case
    if not os.path.exists(ref_case.DRIVER):
        return None
    pol = O.policy()  # V(2,2) red-black, 24 coarse sweeps, constant null space
    run = dict(case)
    run["coeffs"] = O.assemble(case["n"], case["faces"], case["bc"],
                               case["mobility"])
    b = case["rhs"]
    seconds = {}
    compile_s = 0.0
    for its in (short, long):
        out = ref_case.run_driver(run, pol, b, b, rtol=1e-10, maxit=its,
                                  interval=4, repeats=1, timing_only=True)
        if int(out["krylov_iterations"]) != its:
            return None
        seconds[its] = float(out["seconds"][0])
        compile_s = float(out["compile_seconds"])
    per_iteration = (seconds[long] - seconds[short]) / (long - short)
    start = max(0.0, seconds[short] - short * per_iteration)
    return start, per_iteration, compile_s


def cpu_baseline(case, iterations):
    """CPU baseline on the host cores: the reference's CPU path when its
    build is present (kind "reference"), else the C oracle (kind "port"),
    on a bounded sample of the same grid, scaled to one solve of
    `iterations` FP64 PCG iterations (the reference hierarchy's count)."""
    ref = cpu_reference_sample(case)
    port = cpu_port_sample(case)
    if ref is not None:
        start, per_it, compile_s = ref
        return {
            "value": round((start + iterations * per_it) * 1e3, 1),
            "unit": "ms",
            "cores": 1,
            "kind": "reference",
            "sample": (f"reference solve_pcg + NativeCartesianMgPlan (one MPI "
                       f"rank) on the full grid, timed at 1 and 3 iterations: "
                       f"{per_it:.2f} s per iteration + {start:.2f} s start "
                       f"(plan compile {compile_s:.1f} s untimed), scaled to "
                       f"{iterations} FP64 iterations"),
            "port_seconds_per_iteration": round(port, 3),
            "port_over_reference": round(port / per_it, 3),
        }
    return {
        "value": round(port * iterations * 1e3, 1),
        "unit": "ms",
        "cores": 1,
        "kind": "port",
        "sample": (f"1 PCG iteration of the C oracle on the full grid "
                   f"({port:.2f} s) x {iterations} FP64 iterations"),
    }


def large_case(args, torch, H):
    """Forward + adjoint on the combined large case: FP32-storage V-cycle in
    FP64 CG to 1e-10, then dL/dlog(rho) for L = sum of p over the fluid
    cells of a 64^3 probe box, with the hierarchy rebuilt for the adjoint
    (the forward plan is freed first).  Peak device memory comes from
    hipMemGetInfo sampled after every phase."""
    import numpy as np
    nx, ny, nz = (int(v) for v in args.shape.split(","))
    log(f"[large] generating {args.workload} {nx}x{ny}x{nz}")
    case = make_case(args.workload, nx, ny, nz)
    dev = torch.device("cuda", 0)
    free0, total = torch.cuda.mem_get_info(dev)
    low = [free0]

    def sample():
        torch.cuda.synchronize()
        low[0] = min(low[0], torch.cuda.mem_get_info(dev)[0])
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    b = t(case["rhs"])
    masks = [t(case[k]) for k in ("cell_active", "x_active", "y_active",
                                  "z_active")]
    log_rho = t(np.log(case["rho"]))
    del case["rho"], case["mobility"]
    pol = H.default_policy(precision=args.precision, coarse_face_scaling=1)
    ctl = dict(rtol=1e-10, maximum_iterations=1000, flexible=True)

    def build(lr):
        plan = H.Plan(case["n"], case["faces"], case["bc"], pol)
        mob = torch.exp(-lr)
        diag, bx, by, bz = plan.assemble(mob)
        plan.set_coefficients(diag, bx, by, bz, *masks)
        sample()
        return plan, mob, (diag, bx, by, bz)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan, mob, coeffs = build(log_rho)
    setup_ms = (time.perf_counter() - t0) * 1e3
    x = torch.zeros_like(b)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x, res = plan.solve_pcg(b, x, **ctl)
    torch.cuda.synchronize()
    fwd_ms = (time.perf_counter() - t0) * 1e3
    sample()
    true_rel = float(torch.linalg.norm(b - plan.operator_apply(x)) /
                     torch.linalg.norm(b))
    log(f"[large] forward it={res.iterations} {fwd_ms:.1f} ms "
        f"true_rel={true_rel:.2e}")
    plan.project_mean(x)
    plan.close()
    del coeffs, mob
    # loss: sum of p over a 64^3 probe box at the domain centre
    probe = torch.zeros(nz, ny, nx, dtype=torch.float64, device=dev)
    probe[nz // 2 - 32:nz // 2 + 32, ny // 2 - 32:ny // 2 + 32,
          nx // 2 - 32:nx // 2 + 32] = 1.0
    probe = probe.reshape(-1) * masks[0].to(torch.float64)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan, mob, coeffs = build(log_rho)
    gb, gr, ares = plan.adjoint(mob, x, probe, want_b=False, **ctl)
    torch.cuda.synchronize()
    adj_ms = (time.perf_counter() - t0) * 1e3
    sample()
    log(f"[large] adjoint it={ares.iterations} {adj_ms:.1f} ms "
        f"(hierarchy rebuilt)")
    plan.close()
    peak = total - low[0]
    cells = nx * ny * nz
    print(json.dumps({
        "metric": "combined large case: forward + adjoint, peak HBM",
        "value": round(fwd_ms + adj_ms, 1), "unit": "ms",
        "n_gpus": 1, "higher_is_better": False,
        "dtype": f"{args.precision} V-cycle storage / f64 Krylov",
        "data": "synthetic (seeded numpy generator)",
        "config": {"workload": f"{args.workload} {nx}x{ny}x{nz}",
                   "cells": cells, "tolerance": 1e-10,
                   "hierarchy": "distance-corrected face aggregation"},
        "setup_ms": round(setup_ms, 1),
        "forward_ms": round(fwd_ms, 1), "forward_iterations": res.iterations,
        "forward_true_relative_residual": true_rel,
        "adjoint_ms_including_rebuild": round(adj_ms, 1),
        "adjoint_iterations": ares.iterations,
        "peak_device_bytes": int(peak),
        "peak_device_gib": round(peak / 2**30, 2),
        "device_total_gib": round(total / 2**30, 1),
        "grad_logrho_norm": float(torch.linalg.norm(gr)),
    }), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="flame_step")
    ap.add_argument("--shape", default="512,256,256")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "bf16"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--cycle", default="v", choices=["v", "f"])
    ap.add_argument("--no-reference-hierarchy", dest="reference_hierarchy",
                    action="store_false",
                    help="skip the solves with the reference's face "
                         "aggregation")
    ap.add_argument("--large", action="store_true",
                    help="forward + adjoint on the combined large case "
                         "(random_spheres 1024x512x512 unless --workload/"
                         "--shape say otherwise) and peak device memory")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device copy of this many MiB (x10) as "
                         "the known-size reference for HBM counters")
    ap.add_argument("--no-production", dest="production",
                    action="store_false",
                    help="skip the production pairing (FGMRES + Chebyshev "
                         "F-cycle on the assembled compressible system)")
    ap.add_argument("--production-only", action="store_true",
                    help="only time the production pairing and print its "
                         "record (for profiling it alone)")
    ap.add_argument("--cpu-only", type=int, default=0, metavar="ITERATIONS",
                    help="only time the CPU baseline (reference CPU path, or "
                         "the C oracle without its build) for a solve of this "
                         "many FP64 iterations; needs no GPU")
    args = ap.parse_args()
    if args.cpu_only:
        nx, ny, nz = (int(v) for v in args.shape.split(","))
        log(f"[bench] generating {args.workload} {nx}x{ny}x{nz}")
        case = make_case(args.workload, nx, ny, nz)
        print(json.dumps({"workload": f"{args.workload} {nx}x{ny}x{nz}",
                          "cpu_baseline": cpu_baseline(case, args.cpu_only)}),
              flush=True)
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build measures the "
              "single-GPU share of the pressure solve (multi-GPU sharding is "
              "a design note, not a measured path).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import hundun_flow as H

    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(10):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": 8 * n, "launches": 10}))
        return

    if args.large:
        if args.workload == "flame_step":
            args.workload = "random_spheres"
        if args.shape == "512,256,256":
            args.shape = "1024,512,512"
        large_case(args, torch, H)
        return

    nx, ny, nz = (int(v) for v in args.shape.split(","))
    if args.production_only:
        dev = torch.device("cuda", 0)
        torch.cuda.set_device(dev)
        print(json.dumps({"production_pairing": production_pairing(
            (nx, ny, nz), args.steps, args.warmup, torch, H, dev)}),
            flush=True)
        return
    log(f"[bench] generating {args.workload} {nx}x{ny}x{nz}")
    case = make_case(args.workload, nx, ny, nz)
    cells = nx * ny * nz
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    tt = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    b = tt(case["rhs"])
    act = case["cell_active"]

    results = {}
    plans = {}
    # (storage precision, coarse-face scaling, flexible beta).  The
    # reduced-precision cycle runs under the flexible (Polak-Ribiere) beta;
    # FP64 runs twice: the plain solve of the reference algorithm (FP64
    # cycle, beta = rho'/rho) and the same flexible Krylov as the timed run.
    runs = [(args.precision, 1, True)]
    if not args.no_fp64:
        runs += [("fp64", 1, False), ("fp64", 1, True)]
    if args.reference_hierarchy:
        runs += [(args.precision, 0, True)] + (
            [] if args.no_fp64 else [("fp64", 0, False)])
    for prec, scaling, flexible in runs:
        key = (prec, scaling, flexible)
        log(f"[bench] plan {prec} coarse_face_scaling={scaling} "
            f"flexible={flexible}")
        pol = H.default_policy(precision=prec, coarse_face_scaling=scaling,
                               f_cycle=1 if args.cycle == "f" else 0)
        plan = H.plan_from_case(case, pol, dev)
        plans[key] = plan
        x, res, times = gpu_solves(plan, b, args.steps, args.warmup, flexible,
                                   torch)
        # independent FP64 check of the published solution
        ax = plan.operator_apply(x)
        true_rel = float(torch.linalg.norm(b - ax) / torch.linalg.norm(b))
        solid_max = float(x[torch.from_numpy(act == 0).to(dev)].abs().max()) \
            if (act == 0).any() else 0.0
        results[key] = dict(x=x, res=res, times=times, true_rel=true_rel,
                            solid_max=solid_max)
        log(f"[bench] {prec} scaling={scaling} flexible={flexible}: "
            f"iterations={res.iterations} "
            f"ms={['%.2f' % t for t in times]} true_rel={true_rel:.3e} "
            f"solid_max={solid_max}")
        assert res.termination == "converged", res
        assert true_rel <= 1.01e-10, true_rel
        assert solid_max == 0.0
        if key != (args.precision, 1, True):
            plan.close()
            plans.pop(key)

    main_prec = args.precision
    main = results[(main_prec, 1, True)]
    ms = sum(main["times"]) / len(main["times"])

    # dominant kernel: one more profiled solve of the timed configuration
    plan = plans[(main_prec, 1, True)]
    plan.set_profiling(True)
    x = torch.zeros_like(b)
    plan.solve_pcg(b, x, rtol=1e-10, maximum_iterations=1000, flexible=True)
    prof = plan.profile()
    plan.set_profiling(False)
    s = storage_bytes(main_prec)
    ranked = sorted(prof.items(), key=lambda kv: -kv[1][0])
    log("[bench] kernel classes (ms, launches): " +
        ", ".join(f"{k}={v[0]:.2f}/{v[1]}" for k, v in ranked if v[1]))
    dominant, (dom_ms, dom_n) = ranked[0]
    per_cell = class_bytes_per_cell(dominant, s)
    roofline = None
    if per_cell is not None and dom_n:
        avg_ms = dom_ms / dom_n
        bytes_per_launch = per_cell * cells
        achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
        if dominant == "rb_finest":
            work = ((nx + 1) // 2) * ny * nz
            grid = -(-work // 256) * 256
        elif dominant == "sweep_finest":
            grid = None  # looked up by kernel name (one launch geometry)
        else:
            grid = -(-cells // 256) * 256
        tag = f"{args.workload}_{nx}x{ny}x{nz}_{main_prec}"
        traffic = measured_traffic(tag, dominant, main_prec, grid)
        roofline = {"bound": "hbm", "achieved": round(achieved, 1),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": traffic,
                    "traffic_source": (f"profiles/{tag}_hbm_traffic.json"
                                       if traffic else None),
                    "kernel": dominant,
                    "avg_launch_ms": round(avg_ms, 4),
                    "algorithmic_bytes_per_launch": int(bytes_per_launch)}

    def mean_ms(r):
        return sum(r["times"]) / len(r["times"])

    extra = {}
    if ("fp64", 1, False) in results:
        r64 = results[("fp64", 1, False)]
        ms64 = mean_ms(r64)
        fluid = torch.from_numpy(act != 0).to(dev)

        def gauge(v):
            v = v.clone()
            v[fluid] -= v[fluid].mean()
            v[~fluid] = 0.0
            return v
        err = float(torch.linalg.norm(gauge(main["x"]) - gauge(r64["x"]))
                    / torch.linalg.norm(gauge(r64["x"])))
        r64f = results[("fp64", 1, True)]
        extra = {"fp64_ms_per_solve": round(ms64, 3),
                 "fp64_iterations": r64["res"].iterations,
                 "speedup_vs_fp64": round(ms64 / ms, 3),
                 "fp64_krylov": "PCG, beta = rho'/rho (the reference solve)",
                 "fp64_flexible_ms_per_solve": round(mean_ms(r64f), 3),
                 "fp64_flexible_iterations": r64f["res"].iterations,
                 "speedup_vs_fp64_flexible": round(mean_ms(r64f) / ms, 3),
                 "error_vs_fp64_solution": err}

    reference = None
    if (main_prec, 0, True) in results:
        r32 = results[(main_prec, 0, True)]
        reference = {"hierarchy": "reference face aggregation",
                     "ms_per_solve": round(mean_ms(r32), 3),
                     "iterations": r32["res"].iterations}
        if ("fp64", 0, False) in results:
            r64 = results[("fp64", 0, False)]
            reference.update({"fp64_ms_per_solve": round(mean_ms(r64), 3),
                              "fp64_iterations": r64["res"].iterations,
                              "speedup_vs_fp64": round(mean_ms(r64) /
                                                       mean_ms(r32), 3)})

    production = None
    if args.production and args.workload == "flame_step":
        for key in list(plans):
            plans.pop(key).close()
        production = production_pairing((nx, ny, nz), args.steps, args.warmup,
                                         torch, H, dev)

    cpu = None
    if not args.no_cpu_baseline:
        # the reference CPU algorithm: reference aggregation, FP64 PCG
        if ("fp64", 0, False) in results:
            its = results[("fp64", 0, False)]["res"].iterations
        elif ("fp64", 1, False) in results:
            its = results[("fp64", 1, False)]["res"].iterations
        else:
            its = main["res"].iterations
        log("[bench] cpu baseline (reference CPU path, bounded sample)")
        cpu = cpu_baseline(case, its)

    line = {
        "metric": ("Pressure solve ms to 1e-10 rel residual, FP32-MG vs "
                   "FP64-MG; % HBM roofline"),
        "value": round(ms, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"{main_prec} V-cycle storage / f64 Krylov",
        "data": "synthetic (seeded numpy generator)",
        "config": {"workload": f"{args.workload} {nx}x{ny}x{nz}",
                   "cells": cells, "cycle": f"{args.cycle.upper()}(2,2) "
                   "red-black, 24 coarse sweeps, constant null space",
                   "hierarchy": "distance-corrected face aggregation "
                   "(reference aggregation timed in reference_hierarchy)", "krylov": "PCG, flexible "
                   "(Polak-Ribiere) beta", "tolerance": 1e-10},
        "iterations": main["res"].iterations,
        "mcell_solves_per_s": round(cells / (ms * 1e-3) / 1e6, 2),
        "true_relative_residual": main["true_rel"],
        **extra,
        "reference_hierarchy": reference,
        "production_pairing": production,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
