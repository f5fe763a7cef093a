#!/usr/bin/env python3
"""Benchmark of the pressure Poisson solve on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--line-path]

A step is one solve of the 512^3 stretched-z channel (periodic x and y,
Neumann z walls, Lx = 2 pi, Ly = pi, tanh stretching 1.5) on a seeded N(0,1)
right-hand side already resident in HBM.  The line reports ns per grid
point for the plain FP64 solve (the `value`), the FP32 + FP64-refinement
solve to a 1e-10 relative residual, per-kernel HBM roofline figures from
HIP events on the library stream, and the FP64 CPU oracle timed on the
same workload on the host cores.

Only one GPU is used: --gpus above 1 exits with a message.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("ns/gridpoint pressure solve @512³: FP64 vs FP32+refine @1e-10; "
          "% HBM peak")
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# Algorithmic HBM bytes per grid point of each pipeline stage: one read and
# one write of the field at FP64 width.  The whole solve is priced on the
# five-sweep pencil model (x-fwd, y-fwd, z-solve, y-bwd, x-bwd): 80 B/pt.
STAGE_BYTES_FP64 = 16
SOLVE_BYTES_FP64 = 80
# stage name -> kernel (as rocprofv3 names it) that the stage launches
STAGE_KERNEL = {"P1_x_fwd": "k_x_fwd<double", "P2_yz_fwd": "k_yz_fwd<double",
                "P3_zy_bwd": "k_zy_bwd<double", "P4_x_bwd": "k_x_bwd<double"}
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "pmc_traffic_fused_fp64_512_channel.txt")
# Plain copy used to calibrate the rocprofv3 byte counters: the interior of a
# 512^3 halo-1 field (510^3 FP64 points) read once and written once.
COPY_N = 510
COPY_BYTES = COPY_N ** 3 * 8


def time_steps(fn, steps, warmup, api):
    """Mean, median and min ms per step from HIP events on the library
    stream around each step, and the mean wall-clock ms per step."""
    for _ in range(warmup):
        fn()
    api.synchronize()
    ev = [api.Event() for _ in range(steps + 1)]
    t0 = time.perf_counter()
    ev[0].record()
    for s in range(steps):
        fn()
        ev[s + 1].record()
    api.synchronize()
    wall = (time.perf_counter() - t0) * 1e3
    per = [ev[s].elapsed_ms(ev[s + 1]) for s in range(steps)]
    return {"mean": ev[0].elapsed_ms(ev[steps]) / steps, "median": float(np.median(per)),
            "min": min(per), "wall": wall / steps}


def cpu_baseline(case, reps=5):
    """The C FP64 oracle (solver_cpu restated, own FFT, OpenMP) on the same
    512^3 workload, timed over `reps` solves; ns per grid point (median).
    Also returns the oracle's solution (interior), the reference the GPU
    results are compared with."""
    from oracle import pyoracle as po
    from flutas_amd import workloads as wl
    nx, ny, nz = case.ng
    dzc, dzf, zc, zf = po.initgrid(nz, case.alpha, case.lengths[2], 1)
    s = po.Solver(case.ng, case.dli, 1.0 / dzc, 1.0 / dzf, case.cbc)
    p0 = wl.rhs_normal(case, dzf)
    times = []
    for _ in range(reps):
        p = p0.copy(order="F")
        t0 = time.perf_counter()
        s.solve(p)
        times.append(time.perf_counter() - t0)
    npts = nx * ny * nz
    med = float(np.median(times))
    return ({"value": med * 1e9 / npts, "unit": "ns/gridpoint",
             "cores": int(po.lib().or_max_threads()), "kind": "port",
             "best": min(times) * 1e9 / npts,
             "sample": f"C FP64 oracle (solver_cpu restated, own FFT, OpenMP) on the same "
                       f"{nx}x{ny}x{nz} {case.cbc} stretch-{case.alpha} solve; median of {reps}"},
            p[1:-1, 1:-1, 1:-1])


def relmax(a, b, gauge_free=False):
    """max|a - b| / max|b|; gauge_free removes the mean of both first (the
    additive constant of a singular problem)."""
    if gauge_free:
        a, b = a - a.mean(), b - b.mean()
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


# Algorithmic bytes per grid point of the field passes around the solve in
# one projection step (main__single_phase.f90:692-747, constant density):
# reference order pold = p (16), fillps (32), correc (56), p = pold + p (24);
# fused order fillps_rhs (32), correc_update (72: phi, pold, u, v, w read,
# u, v, w, pnew written).  Face-only passes (updt_rhs_b, boundp) are not
# counted.
STEP_SIDE_BYTES = {"reference_order": 128, "fused": 104}


def step_bench(api, case, dzc, dzf, dzci, dzfi, f_host, steps, warmup):
    """One projection step around the FP64 solve at the bench size, in the
    reference's call order (main__single_phase.f90:705-747, bounduvw after
    correc) and in the fused order with swapped pressure buffers (fillps_rhs,
    solve, correc_update_bc, bounduvw), eagerly and as one replayed graph;
    ms per step from HIP events.  Velocity BCs follow the pressure pattern:
    periodic stays periodic, a pressure N wall is a no-slip (D, 0) wall."""
    nx, ny, nz = case.ng
    D = api.DeviceArray.from_host
    u, v, w = D(f_host), D(f_host), D(f_host)
    p, pold, scratch = D(f_host), api.DeviceArray(f_host.shape), api.DeviceArray(f_host.shape)
    dzci_d, dzfi_d = D(dzci), D(dzfi)
    dli = case.dli
    dl = tuple(1.0 / d for d in dli)
    plan = api.PoissonPlan(case.ng, dli, dzci, dzfi, case.cbc)
    rx, ry, rz = plan.rhsb_device()
    zero_bc = (0.0,) * 6
    dt = 1e-3
    vel = {"P": "P", "N": "D", "D": "N"}
    cbcvel = "".join(vel[c] for _ in range(3) for c in case.cbc)

    def bounduvw():
        api.bounduvw(cbcvel, case.ng, (0.0,) * 18, 1, 1, [0] * 6, dl, dzc, dzf, u, v, w)

    def reference_order():
        api.copy_interior(nx, ny, nz, p, pold)
        api.fillps(nx, ny, nz, 1, 1, dli[0], dli[1], dli[2], dzfi_d, 1 / dt, 1.0, u, v, w, p)
        api.updt_rhs_b(nx, ny, nz, "ccc", case.cbc, rx, ry, rz, p)
        plan.solve_async(p)
        api.boundp(case.cbc, case.ng, zero_bc, dl, dzc[0], dzc[nz], p)
        api.correc(nx, ny, nz, 1, 1, dli[0], dli[1], dli[2], dzci_d, dt, 1.0, p, u, v, w)
        bounduvw()
        api.add_interior(nx, ny, nz, pold, p)
        api.boundp(case.cbc, case.ng, zero_bc, dl, dzc[0], dzc[nz], p)

    def fused():
        # boundp(phi), correc, p = pold + phi and boundp(p) in one pass
        api.fillps_rhs(nx, ny, nz, 1, 1, dli[0], dli[1], dli[2], dzfi_d, 1 / dt, 1.0,
                       u, v, w, "ccc", case.cbc, rx, ry, rz, scratch)
        plan.solve_async(scratch)
        api.correc_update_bc(nx, ny, nz, 1, 1, dli[0], dli[1], dli[2], dzci_d, dt, 1.0,
                             scratch, p, p, u, v, w, case.cbc, zero_bc, dl, dzc[0], dzc[nz])
        bounduvw()

    t_ref = time_steps(reference_order, steps, warmup, api)
    t_fused = time_steps(fused, steps, warmup, api)
    t_solve = time_steps(lambda: plan.solve_async(scratch), steps, warmup, api)
    # the fused step recorded once (flutas_graph_*) and replayed
    with api.Graph() as graph:
        fused()
    t_graph = time_steps(graph.launch, steps, warmup, api)
    graph.destroy()
    plan.destroy()
    npts = case.npoints
    out = {"reference_order_ms": t_ref["mean"], "fused_ms": t_fused["mean"],
           "fused_graph_ms": t_graph["mean"],
           "solve_ms": t_solve["mean"], "side_bytes_per_point": STEP_SIDE_BYTES}
    for key, t in (("reference_order", t_ref), ("fused", t_fused)):
        side = t["mean"] - t_solve["mean"]
        out[key + "_side_ms"] = side
        out[key + "_side_GBps"] = STEP_SIDE_BYTES[key] * npts / (side * 1e6)
    return out


def batch_bench(api, wl, reps=5):
    """The uneven batch (64 problems, nx, ny in {64..256}, nz in {64, 128,
    256}, three BC patterns, half stretched): aggregate grid points per
    second of the persistent batched launch, plain FP64 and FP32 + FP64
    refinement to 1e-10 (before and after autotuning), against the same 64
    problems solved one after another with single plain FP64 plans; wall
    clock per round."""
    cases = wl.batch_problems()
    problems, fs, ps, singles = [], [], [], []
    npts = 0
    for case in cases:
        dzc, dzf, dzci, dzfi = wl.z_grid(api, case)
        f = api.DeviceArray.from_host(wl.rhs_normal(case, dzf))
        p = api.DeviceArray(f.shape)
        p.zero()
        problems.append((case.ng, case.dli, dzci, dzfi, case.cbc))
        fs.append(f)
        ps.append(p)
        npts += case.npoints

    def rate(fn):
        fn()
        api.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        api.synchronize()
        dt = (time.perf_counter() - t0) / reps
        return dt * 1e3, npts / dt

    out = {"problems": len(cases), "grid_points": npts}
    for key, prec in (("fp64", api.FP64), ("refined", api.FP32_REFINE)):
        batch = api.BatchPlan(problems, precision=prec, tol=1e-10)
        batch.set_fields(fs, ps)
        ms_b, pps_b = rate(batch.solve)
        t0 = time.perf_counter()
        batch.tune(reps=2)
        tune_s = time.perf_counter() - t0
        ms_t, pps_t = rate(batch.solve)
        th, gh = batch.tuning()
        wg, items, per_cu, lds = batch.launch_info()
        shape, kxp = batch.launch_shape()
        shape = ("launch per pass and line length", "persistent launch per pass")[shape]
        res = {"untuned_ms": ms_b, "untuned_points_per_s": pps_b, "launch_shape": shape,
               "x_item_planes": kxp,
               "tuned_ms": ms_t, "tuned_points_per_s": pps_t, "tune_seconds": tune_s,
               "tile_variants": {str(v): th.count(v) for v in sorted(set(th))},
               "group_variants": {str(v): gh.count(v) for v in sorted(set(gh))},
               "workgroups": wg, "queue_items": items, "workgroups_per_cu": per_cu,
               "lds_bytes": lds}
        if prec == api.FP32_REFINE:
            iters, rel = batch.stats()
            res["rounds"] = {str(v): iters.count(v) for v in sorted(set(iters))}
            res["worst_relres"] = max(rel)
        batch.destroy()
        out[key] = res
    for (ng, dli, dzci, dzfi, cbc) in problems:
        singles.append(api.PoissonPlan(ng, dli, dzci, dzfi, cbc))

    def sequential():
        for plan, f, p in zip(singles, fs, ps):
            plan.solve_into(f, p)

    ms_s, pps_s = rate(sequential)
    for plan in singles:
        plan.destroy()
    out.update({"sequential_fp64_ms": ms_s, "sequential_fp64_points_per_s": pps_s,
                "fp64_speedup_vs_sequential": ms_s / out["fp64"]["tuned_ms"],
                "refined_speedup_vs_sequential": ms_s / out["refined"]["tuned_ms"]})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="channel_512_stretched")
    ap.add_argument("--profile-reps", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-refine", action="store_true")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the uneven 64-problem batch")
    ap.add_argument("--no-step", action="store_true",
                    help="skip the projection step around the solve")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run the plain copy the byte-counter summaries calibrate on")
    ap.add_argument("--line-path", action="store_true",
                    help="run the five-pass line pipeline (separate x, y, z, y, x launches; "
                         "FLUTAS_UNFUSED=1) instead of the four fused passes")
    args = ap.parse_args()
    if args.line_path:
        # read by the library when a plan is created; the batched solves
        # run on the fused passes only
        os.environ["FLUTAS_UNFUSED"] = "1"
        args.no_batch = True
    if args.gpus != 1:
        print("bench.py measures one GPU only; run it with --gpus 1 "
              "(the multi-GPU slab exchange is a design, not built).", file=sys.stderr)
        sys.exit(2)

    from flutas_amd import api, workloads as wl
    case = wl.WORKLOADS[args.workload]
    nx, ny, nz = case.ng
    npts = case.npoints
    log(f"bench: {case.name} {case.ng} {case.cbc} stretch={case.alpha}")
    dzc, dzf, dzci, dzfi = wl.z_grid(api, case)
    f_host = wl.rhs_normal(case, dzf)
    f = api.DeviceArray.from_host(f_host)
    p = api.DeviceArray(f.shape)
    p.zero()

    # ---- FP64 ---------------------------------------------------------
    plan = api.PoissonPlan(case.ng, case.dli, dzci, dzfi, case.cbc)
    z_chunks = plan.zchunks
    t64 = time_steps(lambda: plan.solve_into(f, p), args.steps, args.warmup, api)
    ms = t64["mean"]
    relres64 = plan.residual(f, p)
    sol64 = p.to_host()[1:-1, 1:-1, 1:-1]
    log(f"bench: fp64 {ms:.3f} ms/solve (median {t64['median']:.3f}, min {t64['min']:.3f}, "
        f"wall {t64['wall']:.3f}), relres {relres64:.2e}")
    # per-stage kernel times: each stage's kernel launched back to back
    # between two HIP events on the library stream (its own average
    # duration, as rocprofv3 reports it); the events-between-stages profile
    # of whole solves is kept beside it
    stages = plan.profile(f, p, args.profile_reps)
    stage_rows = []
    for s_idx, (name, t_between) in enumerate(stages):
        # the line pipeline's stages are timed by the events between them
        t = t_between if args.line_path else plan.time_stage(f, p, s_idx, reps=8)
        b = STAGE_BYTES_FP64 * npts
        stage_rows.append({"stage": name, "ms": t, "GBps": b / (t * 1e6),
                           "ms_events_between_stages": t_between})
    dom = max(stage_rows, key=lambda r: r["ms"])
    plan.solve_into(f, p)
    traffic = None
    if os.path.exists(TRAFFIC_SUMMARY) and dom["stage"] in STAGE_KERNEL:
        from flutas_amd.profsum import read_traffic
        key = STAGE_KERNEL[dom["stage"]]
        for kname, rec in read_traffic(TRAFFIC_SUMMARY).items():
            if kname.startswith(key):
                traffic = rec["total"]
    log("bench: stages " + ", ".join(f"{r['stage']} {r['ms']:.3f} ms ({r['GBps']:.0f} GB/s)"
                                     for r in stage_rows))
    solve_gbps = SOLVE_BYTES_FP64 * npts / (ms * 1e6)
    plan.destroy()

    # ---- FP32 + FP64 refinement ---------------------------------------
    refine = None
    if not args.no_refine:
        plan32 = api.PoissonPlan(case.ng, case.dli, dzci, dzfi, case.cbc,
                                 precision=api.FP32_REFINE, tol=1e-10)
        t32 = time_steps(lambda: plan32.solve_into(f, p), max(3, args.steps // 2),
                         args.warmup, api)
        ms32 = t32["mean"]
        iters, rel = plan32.stats()
        check = plan32.residual(f, p)
        stages32 = plan32.profile(f, p, max(1, args.profile_reps // 2))
        log("bench: refine stages " + ", ".join(f"{s} {t:.3f}" for s, t in stages32))
        refine = {"ns_per_gridpoint": ms32 * 1e6 / npts, "ms_per_solve": ms32,
                  "iterations": iters, "relres": rel, "relres_check": check,
                  "throughput_vs_fp64": ms / ms32,
                  "stages_ms": [[s, t] for s, t in stages32],
                  "algorithmic_bytes_per_point": 48 + 96 * iters + 24}
        log(f"bench: fp32+refine {ms32:.3f} ms/solve, {iters} iters, relres {rel:.2e}")
        plan32.solve_into(f, p)
        sol32 = p.to_host()[1:-1, 1:-1, 1:-1]
        plan32.destroy()

    result = {
        "metric": METRIC,
        "value": ms * 1e6 / npts,
        "unit": "ns/gridpoint",
        "higher_is_better": False,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "ms_median": t64["median"],
        "ms_min": t64["min"],
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{case.name}: {nx}x{ny}x{nz} pressure Poisson solve, "
                               f"BC {case.cbc}, Lx,Ly,Lz={tuple(round(v, 6) for v in case.lengths)}, "
                               f"tanh z stretching {case.alpha}, seeded N(0,1) RHS "
                               "(dz-weighted mean removed), inputs resident in HBM",
                   "ng": list(case.ng), "bc": case.cbc,
                   "pipeline": "five-pass line path" if args.line_path else "four fused passes",
                   "z_chunks": z_chunks},
        "roofline": {"bound": "hbm", "kernel": dom["stage"],
                     "time_source": ("HIP events between the stages of whole solves "
                                     "(flutas_psolve_profile)" if args.line_path else
                                     "HIP events on the library stream around 8 back-to-back "
                                     "launches of the kernel after a solve "
                                     "(flutas_psolve_time_stage)"),
                     "achieved": dom["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": dom["GBps"] / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes": STAGE_BYTES_FP64 * npts,
                     "traffic_source": os.path.relpath(TRAFFIC_SUMMARY, ROOT) if traffic else None},
        "solve_roofline": {"bytes_per_point": SOLVE_BYTES_FP64, "achieved": solve_gbps,
                           "frac": solve_gbps / HBM_PEAK_GBS},
        "stages": stage_rows,
        "fp64_relres": relres64,
        "fp32_refine": refine,
    }
    if not args.no_step:
        log("bench: projection step (reference order vs fused) ...")
        result["projection_step"] = step_bench(api, case, dzc, dzf, dzci, dzfi, f_host,
                                               max(3, args.steps // 2), args.warmup)
        st = result["projection_step"]
        log(f"bench: step {st['reference_order_ms']:.3f} ms in the reference order, "
            f"{st['fused_ms']:.3f} ms fused, {st['fused_graph_ms']:.3f} ms as one graph "
            f"(solve {st['solve_ms']:.3f} ms; side passes "
            f"{st['reference_order_side_GBps']:.0f} / {st['fused_side_GBps']:.0f} GB/s)")
    if not args.no_batch:
        log("bench: uneven batch of 64 problems ...")
        result["batch"] = batch_bench(api, wl)
        bb = result["batch"]
        log(f"bench: batch {bb['grid_points'] / 1e6:.1f} M pts: FP64 {bb['fp64']['tuned_ms']:.2f} ms "
            f"(untuned {bb['fp64']['untuned_ms']:.2f}), refined {bb['refined']['tuned_ms']:.2f} ms "
            f"(untuned {bb['refined']['untuned_ms']:.2f}), sequential FP64 "
            f"{bb['sequential_fp64_ms']:.2f} ms")
    if args.calibrate_copy:
        del f, p
        a = api.DeviceArray((COPY_N + 2,) * 3)
        b = api.DeviceArray((COPY_N + 2,) * 3)
        a.zero()
        for _ in range(2):
            api.copy_interior(COPY_N, COPY_N, COPY_N, a, b)
        api.synchronize()
        log(f"bench: calibration copy of {COPY_BYTES} B (interior of {COPY_N + 2}^3)")
    if not args.no_cpu:
        log("bench: cpu baseline (oracle) ...")
        result["cpu_baseline"], ref = cpu_baseline(case)
        # the north star's "FP64-oracle error per result": relative max-norm
        # differences from the oracle's solution (same pinned gauge for
        # FP64; gauge-free for the refined solve)
        result["fp64_oracle_err"] = relmax(sol64, ref)
        if refine is not None:
            refine["oracle_err"] = relmax(sol32, ref, gauge_free=True)
            result["refine_oracle_err"] = refine["oracle_err"]
        log(f"bench: oracle error fp64 {result['fp64_oracle_err']:.2e}" +
            (f", refined {refine['oracle_err']:.2e}" if refine is not None else ""))
        del ref
        log(f"bench: cpu {result['cpu_baseline']['value']:.2f} ns/pt on "
            f"{result['cpu_baseline']['cores']} threads")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
