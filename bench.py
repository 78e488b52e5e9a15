"""Benchmark of the RKDG hot path on one MI355X.

Workload (the headline configuration): 2D compressible Euler, gamma 1.4,
isentropic vortex (strength 5, centre (5,5), background rho=u=v=p=1) on
periodic [0,10]^2, 2x128x128 = 32768 jittered triangles (+/-15% h, seed
0x5EED), modal Dubiner P2 (6 modes x 4 variables), HLLC, no limiting, CFL
0.15, FP64. A step is one full SSP-RK3 step of the whole mesh: three fused
stage kernels with the CFL dt computed on the device.

value = DOF-updates/s = E * Np * V * steps / wall time of the stepping loop
(state resident in HBM; setup, projection and upload excluded).
"""

from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Euler DG DOF-updates/s (P2, 32k tris) + p50/p99 solve latency, mixed vs plain"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6

# HBM bytes per launch of the stage kernel, measured with rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE (one pass each) and scaled by what the same
# counters read for a 256 MiB device copy (FETCH x 2.0, WRITE x 1.0); keyed by
# (kernel, order, triangles, precision). The commands and counter values are in
# profiles/euler_stage_p2_32k_fp64.md, profiles/euler_stage_mx_p3_524k.md,
# profiles/advect_stage_524k.md and profiles/euler_line_stage_1m.md.
# Other workloads report null.
MEASURED_TRAFFIC = {
    ("whip::euler_stage<2,4>", 2, 32768, "fp64"): (7895.22 * 2.0 + 6149.39) * 1024.0,
    ("whip::euler_stage<2,2>", 2, 32768, "fp64"): (7965.57 * 2.0 + 6146.72) * 1024.0,
    ("whip::euler_stage_mx<3>", 3, 524288, "mixed"): (226335.0 * 2.0 + 262171.0) * 1024.0,
    ("whip::advect_stage<1>", 1, 524288, "fp64"): (26979.5 * 2.0 + 12288.1) * 1024.0,
    ("whip::advect_stage<3>", 3, 524288, "fp64"): (51864.7 * 2.0 + 40960.1) * 1024.0,
    ("whip::euler_line_stage<2>", 2, 1048576, "fp64"): (96774.8 * 2.0 + 98346.8) * 1024.0,
}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def algorithmic_bytes_per_element_step(Np: int, V: int) -> float:
    """Each array touched once per fused stage: stage 1 reads U^n and writes U1,
    stages 2 and 3 read their input and U^n and write (8 S per step, S = Np V 8 B),
    plus geometry 3 x (40 + 1.5 x 36) B and the CFL speed 32 B."""
    S = Np * V * 8
    return 8 * S + 3 * (40 + 1.5 * 36) + 32


def flops_per_element_stage(p: int) -> float:
    """FP64 operations of one stage per element (counted from the kernel):
    modal->quadrature 2 Nq Np V, flux contraction 4 Nq Np V, pointwise flux
    ~30 Nq, edges 3 (p+1) (2 traces 2 Np V + HLLC ~110 + lift 2 Np V) with
    each face evaluated from both sides, plus 3 Np V for the RK combination."""
    Np = (p + 1) * (p + 2) // 2
    Nq = (p + 3) ** 2
    V = 4
    vol = 2 * Nq * Np * V + 4 * Nq * Np * V + 30 * Nq
    face = 3 * (p + 1) * (2 * 2 * Np * V + 110 + 2 * Np * V)
    return vol + face + 3 * Np * V


def cpu_baseline(mesh, order, cfl, budget_s: float):
    """The reference's own CPU path (oracle/_ref, compiled from the reference
    sources) or, where that was not built, the project's C++ restatement,
    timed single-threaded on a bounded number of steps of the same case."""
    from oracle import oracle as O

    kind = "reference" if O.available("reference") else "port"
    c = O.tri_euler(mesh.node_tags, mesh.node_xy, mesh.triangles, order, kind=kind)
    U = c.project(O.IC_VORTEX, (5.0, 5.0, 5.0))
    steps = 0
    t0 = time.perf_counter()
    while True:
        dt = c.stable_dt(U, cfl)
        U = c.step(U, dt)
        steps += 1
        el = time.perf_counter() - t0
        if el > budget_s or steps >= 200:
            break
    dofs = c.E * c.Np * c.V
    return {
        "value": dofs * steps / el,
        "unit": "DOF-updates/s",
        "cores": 1,
        "kind": kind,
        "sample": f"{steps} SSP-RK3 steps of the same P2 32768-triangle vortex case from t=0 "
                  f"({el:.1f} s, single thread, {cpu_model()}, nproc {os.cpu_count()})",
    }


def precision_error(mesh, p: int, precision: int, steps: int) -> dict:
    """The reduced-precision stage against the FP64 engine (the plain path,
    parity-tested against the CPU oracle) on the bench workload: both start
    from the same projected vortex and take `steps` steps of the same fixed
    dt (the FP64 CFL step of the initial state). Density is compared as cell
    averages (L1 weighted by element area) and as all modal coefficients
    (L-inf), relative to the initial density's maximum cell average."""
    import wasl_amd

    runs = {}
    for name, prec in (("fp64", wasl_amd.FP64), ("reduced", precision)):
        g = wasl_amd.Engine(mesh, wasl_amd.EULER, p, precision=prec)
        g.project(wasl_amd.IC_VORTEX, (5.0, 5.0, 5.0))
        if name == "fp64":
            dt = g.stable_time_step(0.15)
            rho0 = g.state.reshape(4, -1, g.E)[0, 0] / np.sqrt(2.0)
        for _ in range(steps):
            g.step(dt)
        runs[name] = g.state.reshape(4, -1, g.E)[0]
        E = g.E
        del g
    d = runs["reduced"] - runs["fp64"]
    avg = d[0] / np.sqrt(2.0)
    scale = float(np.max(np.abs(rho0)))
    detj = mesh.prepared()["element_metrics"][:E, 4]   # proportional to the element's area
    return {
        "steps": steps, "dt": dt,
        "density_l1": float(np.sum(np.abs(avg) * detj) / np.sum(detj)),
        "density_linf": float(np.max(np.abs(d))),
        "density_linf_relative": float(np.max(np.abs(d))) / scale,
        "reference": "FP64 engine, same initial state and fixed dt",
    }


def rarefaction_run(cells: int, end_time: float) -> dict:
    """The radial double rarefaction (rho = 1, p = 0.4, velocity 2 r_hat) on
    [-1,1]^2 with transmissive boundaries, 2 cells^2 jittered triangles, P2,
    characteristic WENO and positivity limiting (floors 1e-13), CFL 0.1, run
    to end_time through the FP64 engine. Reports the run's DOF-updates/s
    (setup excluded) and the smallest cell-average density and pressure at
    the end, or the reference's stop message if a cell average turns
    unphysical."""
    import torch

    import wasl_amd
    from wasl_amd import meshgen as MG

    mesh = wasl_amd.Mesh.from_input(MG.jittered_triangles(cells, (-1.0, -1.0), (1.0, 1.0), 0.15))
    g = wasl_amd.Engine(mesh, wasl_amd.EULER, 2, boundary=(wasl_amd.OUTFLOW, wasl_amd.OUTFLOW),
                        limiting=True, density_floor=1e-13, pressure_floor=1e-13)
    g.project(wasl_amd.IC_RAREFACTION, (1.0, 0.4, 2.0))
    g.limit()   # the reference limits the projected initial state
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = {"workload": f"radial double rarefaction, P2, 2x{cells}x{cells}={g.E} jittered "
                       f"triangles on [-1,1]^2, transmissive, char-WENO + positivity (1e-13), "
                       f"CFL 0.1, T = {end_time}"}
    try:
        t, steps, _ = g.advance(0.1, end_time)
        out["status"] = "completed"
    except wasl_amd.SimulationError as exc:
        t, steps = float("nan"), -1
        out["status"] = f"stopped: {exc}"
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    out.update({"elements": g.E, "steps": steps, "time": t, "wall_s": wall})
    if steps > 0:
        out["dof_updates_per_s"] = g.E * g.Np * 4 * steps / wall
        U = g.state.reshape(4, g.Np, g.E)[:, 0] / np.sqrt(2.0)
        rho, mx, my, en = U
        p = 0.4 * (en - 0.5 * (mx * mx + my * my) / rho)
        out["min_cell_density"] = float(rho.min())
        out["min_cell_pressure"] = float(p.min())
    return out


def advect_bytes_per_element_stage(Np: int) -> float:
    """The 2D scalar advection stage reads its input and writes its output
    (8 Np B each), reads U^n in two stages of three (8 Np B x 2/3), and reads
    the contravariant velocity (16 B), the edge flux factors (24 B) and the
    edge records (neighbour and meta, 24 B). The neighbours' coefficients are
    the same array as the input and count once."""
    return 8 * Np * (2 + 2 / 3) + 64


def advect2d_run(args) -> int:
    """2D scalar advection u_t + div(v u) = 0, v = (0.7, -0.4), plane wave
    sin(2 pi (x + 2 y) / 10) on periodic [0,10]^2, 2 cells^2 jittered
    triangles, P = --order, CFL 0.3, FP64 (scalar_advection_triangle.cpp).
    value = DOF-updates/s of the stepping loop."""
    import torch

    import wasl_amd
    from wasl_amd import meshgen as MG

    p, cfl, vel, wave = args.order, 0.3, (0.7, -0.4), (10.0, 1.0, 2.0)
    mesh_in = MG.jittered_triangles(args.cells, (0.0, 0.0), (10.0, 10.0), 0.15)
    g = wasl_amd.Engine(wasl_amd.Mesh.from_input(mesh_in), wasl_amd.SCALAR_ADVECTION, p,
                        wavespeed=vel[0], velocity_y=vel[1])
    g.project(wasl_amd.IC_SCALAR_WAVE, wave)
    E, Np = g.E, g.Np
    t, s, _ = g.advance(cfl, 1e9, max_steps=args.warmup, stop_at_max_steps=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    t, s, _ = g.advance(cfl, 1e9, max_steps=s + args.steps, time=t, step=s,
                        stop_at_max_steps=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    value = E * Np * args.steps / wall
    g.set_timing(True)
    t, s, _ = g.advance(cfl, 1e9, max_steps=s + args.steps, time=t, step=s,
                        stop_at_max_steps=True)
    g.set_timing(False)
    _, stage_total, launches = g.last_timing()
    stage_ms = stage_total / max(launches, 1)
    bytes_stage = E * advect_bytes_per_element_stage(Np)
    achieved = bytes_stage / (stage_ms * 1e-3) / 1e9
    log(f"advect2d P{p}: {E} triangles, {wall * 1e3 / args.steps:.4f} ms/step, "
        f"{value:.4e} DOF-updates/s, stage {stage_ms * 1e3:.2f} us, {achieved:.0f} GB/s")
    traffic = MEASURED_TRAFFIC.get((f"whip::advect_stage<{p}>", p, E, "fp64"))
    result = {
        "metric": "2D scalar advection DOF-updates/s",
        "value": value, "unit": "DOF-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"2D scalar advection, v = {vel}, plane wave, P{p} modal Dubiner, "
                        f"2x{args.cells}x{args.cells}={E} jittered periodic triangles on "
                        f"[0,10]^2, upwind, SSP-RK3, CFL {cfl}, FP64",
            "elements": E, "modes": Np, "dofs": E * Np,
            "state_finite": bool(np.isfinite(g.state).all()),
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "traffic_frac_of_algorithmic": None if traffic is None else traffic / bytes_stage,
            "kernel": f"whip::advect_stage<{p}>", "stage_kernel_us": stage_ms * 1e3,
            "algorithmic_bytes_per_launch": bytes_stage,
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        from oracle import oracle as O

        kind = "reference" if O.available("reference") else "port"
        c = O.tri_scalar(mesh_in.node_tags, mesh_in.node_xy, mesh_in.triangles, p, *vel,
                         kind=kind)
        U = c.project(O.IC_SCALAR_WAVE, wave)
        dt = c.stable_dt(U, cfl)
        steps = 0
        t0 = time.perf_counter()
        while True:
            U = c.step(U, dt)
            steps += 1
            el = time.perf_counter() - t0
            if el > args.cpu_seconds or steps >= 200:
                break
        result["cpu_baseline"] = {
            "value": c.E * c.Np * steps / el, "unit": "DOF-updates/s", "cores": 1, "kind": kind,
            "sample": f"{steps} SSP-RK3 steps of the same P{p} {c.E}-triangle advection case "
                      f"({el:.1f} s, single thread, {cpu_model()})",
        }
    print(json.dumps(result), flush=True)
    return 0


def euler1d_bytes_per_element_stage(Np: int) -> float:
    """The 1D Euler stage reads its input and writes its output (4 Np 8 B
    each), reads U^n in two stages of three, 1/J (8 B), and the neighbours'
    coefficients, which are the same array as the input and count once."""
    return 32 * Np * (2 + 2 / 3) + 8


def sod_run(args) -> int:
    """Sod shock tube (cases/tutorials/euler/sod-shock-tube): P = --order on
    --cells uniform cells of [0,1], outflow ends, HLLC, characteristic WENO +
    positivity limiting (threshold 1e-3), CFL 0.2, FP64, a fixed number of
    steps from t = 0. value = DOF-updates/s of the stepping loop."""
    import torch

    import wasl_amd

    p, N, cfl = args.order, args.cells, 0.2
    tags = np.arange(1, N + 2)
    x = np.arange(N + 1) / N
    segs = np.stack([tags[:-1], tags[1:]], 1)
    g = wasl_amd.Engine(wasl_amd.Mesh.lines(tags, x, segs), wasl_amd.EULER, p,
                        boundary=(wasl_amd.OUTFLOW, wasl_amd.OUTFLOW), limiting=True)
    sod = (0, 0, 0, 0, 0, 0, 1.0, 0.0, 0.0, 1.0, 1, 0.5, 1.0, 0, 0, 0, 0.125, 0.0, 0.0, 0.1)
    g.project(wasl_amd.IC_EULER_REGIONS, sod)
    g.limit()
    E, Np = g.E, g.Np
    t, s, _ = g.advance(cfl, 1e9, max_steps=args.warmup, stop_at_max_steps=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    t, s, _ = g.advance(cfl, 1e9, max_steps=s + args.steps, time=t, step=s,
                        stop_at_max_steps=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    value = E * Np * 4 * args.steps / wall
    g.set_timing(True)
    t, s, _ = g.advance(cfl, 1e9, max_steps=s + args.steps, time=t, step=s,
                        stop_at_max_steps=True)
    g.set_timing(False)
    _, stage_total, launches = g.last_timing()
    stage_ms = stage_total / max(launches, 1)
    bytes_stage = E * euler1d_bytes_per_element_stage(Np)
    achieved = bytes_stage / (stage_ms * 1e-3) / 1e9
    log(f"sod P{p}: {E} cells, {wall * 1e3 / args.steps:.4f} ms/step, {value:.4e} DOF-updates/s, "
        f"stage {stage_ms * 1e3:.2f} us, {achieved:.0f} GB/s, t = {t:.5f}")
    result = {
        "metric": "1D Euler (limited Sod) DOF-updates/s",
        "value": value, "unit": "DOF-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"1D Euler Sod shock tube, P{p} modal Legendre, {E} uniform cells on "
                        f"[0,1], outflow ends, HLLC, char-WENO + positivity, CFL {cfl}, FP64",
            "elements": E, "modes": Np, "variables": 4, "dofs": E * Np * 4,
            "state_finite": bool(np.isfinite(g.state).all()),
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": MEASURED_TRAFFIC.get((f"whip::euler_line_stage<{p}>", p, E, "fp64")),
            "kernel": f"whip::euler_line_stage<{p}>", "stage_kernel_us": stage_ms * 1e3,
            "algorithmic_bytes_per_launch": bytes_stage,
        },
        "cpu_baseline": None,
    }
    if result["roofline"]["traffic"] is not None:
        result["roofline"]["traffic_frac_of_algorithmic"] = (
            result["roofline"]["traffic"] / bytes_stage)
    if not args.no_cpu:
        from oracle import oracle as O

        kind = "reference" if O.available("reference") else "port"
        c = O.line_euler(tags, x, segs, p, left=O.OUTFLOW, right=O.OUTFLOW, kind=kind)
        U, _ = c.limit(c.project(O.IC_EULER_REGIONS, sod), 1e-3)
        steps = 0
        t0 = time.perf_counter()
        while True:
            U = c.step(U, c.stable_dt(U, cfl), True, 1e-3)
            steps += 1
            el = time.perf_counter() - t0
            if el > args.cpu_seconds or steps >= 200:
                break
        result["cpu_baseline"] = {
            "value": c.E * c.Np * 4 * steps / el, "unit": "DOF-updates/s", "cores": 1,
            "kind": kind,
            "sample": f"{steps} limited SSP-RK3 steps of the same P{p} {c.E}-cell Sod case "
                      f"({el:.1f} s, single thread, {cpu_model()})",
        }
    print(json.dumps(result), flush=True)
    return 0


def serving_measurement(count: int, seed: int, steps_per_batch: int = 16,
                        setup_threads: int = 4, fp64_tiles=None,
                        fp64_max_order: int = 1) -> dict:
    """The serving mix run plain (FP64, one case at a time) and served
    (mixed precision, continuous batching): aggregate DOF-updates/s of each,
    and time to solution under Poisson arrivals at 80 % of the served
    capacity measured closed-loop."""
    import numpy as np

    from wasl_amd import workload as WL
    from wasl_amd.serve import default_fp64_tiles

    mix = WL.serving_mix(count, seed)
    meshes = WL.MeshCache()
    for r in mix:
        meshes.get(r.mesh_key)
    log(f"serving mix: {count} requests ({sum(r.kind != 'advection' for r in mix)} Euler)")
    WL.run_plain(mix, meshes)
    failed = [r.index for r in mix if r.plain_error]
    for r in mix:
        if r.plain_error:
            log(f"  plain stop: request {r.index} {r.kind} P{r.order} {r.size} elements, "
                f"{r.steps} steps: {r.plain_error}")
    mix = [r for r in mix if not r.plain_error]
    count = len(mix)
    plain_total = sum(r.plain_seconds for r in mix)
    updates = sum(r.dofs * r.plain_steps for r in mix)
    plain_rate = updates / plain_total
    log(f"plain, one case at a time: {plain_total:.2f} s, {plain_rate:.4e} DOF-updates/s"
        f" ({len(failed)} requests stopped by the reference's physicality checks, left out)")

    done, makespan, stats = WL.run_served(mix, meshes, steps_per_batch=steps_per_batch,
                                          setup_threads=setup_threads, fp64_tiles=fp64_tiles,
                                          fp64_max_order=fp64_max_order)
    served_updates = sum(mix[k].dofs * v[2] for k, v in done.items())
    served_rate = served_updates / makespan
    capacity = count / makespan
    log(f"served closed-loop: {makespan:.2f} s, {served_rate:.4e} DOF-updates/s, "
        f"x{served_rate / plain_rate:.2f} plain; {stats}")

    rng = np.random.default_rng(seed + 1)
    rate = 0.8 * capacity
    arrivals = np.cumsum(rng.exponential(1.0 / rate, count)).tolist()
    done_o, makespan_o, _ = WL.run_served(mix, meshes, arrivals=arrivals,
                                          steps_per_batch=steps_per_batch,
                                          setup_threads=setup_threads, fp64_tiles=fp64_tiles,
                                          fp64_max_order=fp64_max_order)
    tts = [done_o[k][1] - done_o[k][0] for k in range(count)]
    plain_tts = WL.fifo_times(arrivals, [r.plain_seconds for r in mix])

    worst = {}
    served_failed = sum(1 for v in done.values() if v[3] is None)
    for k, r in enumerate(mix):
        if done[k][3] is None:
            log(f"  served stop: request {r.index} {r.kind} P{r.order} {r.size} elements")
            continue
        if r.kind == "advection":
            continue
        Np = (r.order + 1) * (r.order + 2) // 2
        E = r.size
        a = done[k][3].reshape(4, Np, E)[0]
        b = r.plain_state.reshape(4, Np, E)[0]
        worst[r.kind] = max(worst.get(r.kind, 0.0), float(np.max(np.abs(a - b))))
    out = {
        "requests": count,
        "requests_failed_plain": len(failed),
        "requests_failed_served": served_failed,
        "mix": "70% 2D Euler (vortex | radial rarefaction, limited | smooth Fourier), "
               "30% 1D advection; tris log-U[512,131072], cells log-U[256,65536], P U{1,2,3}, "
               "200-2000 steps",
        "plain_dof_updates_per_s": plain_rate,
        "served_dof_updates_per_s": served_rate,
        "served_vs_plain": served_rate / plain_rate,
        "served_precision": "unlimited 2D Euler P>=2: mixed (FP32 matrix-core contractions, "
                            "FP64 state and RK); P<=1 and limited cases: FP64 stage; limiter, "
                            "CFL and 1D advection FP64",
        # unlimited FP64 classes; limited cases keep each case's autotuned tile
        "served_fp64_tiles": {f"P{o}": ({"stage_kernel": k, "lanes": n} if k else "autotuned")
                              for o, (k, n) in sorted({**default_fp64_tiles(),
                                                       **(fp64_tiles or {})}.items())},
        "capacity_requests_per_s": capacity,
        "arrival_rate_per_s": rate,
        "tts_p50_s": WL.percentile(tts, 50),
        "tts_p99_s": WL.percentile(tts, 99),
        "plain_tts_p50_s": WL.percentile(plain_tts, 50),
        "plain_tts_p99_s": WL.percentile(plain_tts, 99),
        "open_loop_makespan_s": makespan_o,
        "closed_loop_server": stats,
        "density_linf_served_vs_fp64": worst,
        "density_linf_note": "max over requests of each kind of |rho coefficients, served "
                             "(mixed) - plain (FP64)|; limited rarefactions can flip a "
                             "troubled-cell decision between the two",
    }
    log(f"served open-loop at {rate:.2f} req/s: TTS p50 {out['tts_p50_s']:.3f} s "
        f"p99 {out['tts_p99_s']:.3f} s (plain p50 {out['plain_tts_p50_s']:.3f} s "
        f"p99 {out['plain_tts_p99_s']:.3f} s); density L-inf vs FP64 {worst}")
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    # 64 warmup steps are one full captured batch: the step graph is recorded
    # and instantiated before the timed region.
    ap.add_argument("--warmup", type=int, default=64)
    ap.add_argument("--order", type=int, default=2)
    ap.add_argument("--cells", type=int, default=128, help="quads per axis (2 n^2 triangles)")
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--stage-kernel", type=int, default=0,
                    help="0 autotuned, 1 point split, 2 variable split")
    ap.add_argument("--precision", choices=["fp64", "mixed", "bf16"], default="fp64",
                    help="fp64: the plain drop-in; mixed: FP32 matrix-core contractions; "
                         "bf16: the BF16 ablation")
    ap.add_argument("--serving-requests", type=int, default=64,
                    help="requests of the serving mix (0: skip the serving measurement)")
    ap.add_argument("--serving-seed", type=int, default=0x5EED)
    ap.add_argument("--workload", choices=["vortex", "rarefaction", "advect2d", "sod"],
                    default="vortex",
                    help="vortex: the headline (config 2 sizes by default); rarefaction: the "
                         "limited radial rarefaction run to --end-time; advect2d: 2D scalar "
                         "advection at --order and --cells; sod: 1D Euler Sod tube, --cells "
                         "cells; each other workload prints one JSON line of its own")
    ap.add_argument("--end-time", type=float, default=0.3)
    ap.add_argument("--error-steps", type=int, default=20,
                    help="steps of the reduced-precision error check (--precision mixed/bf16)")
    ap.add_argument("--serving-steps", type=int, default=16,
                    help="SSP-RK3 steps per scheduling round of the server (even)")
    ap.add_argument("--serving-threads", type=int, default=4,
                    help="host threads that prepare served cases")
    ap.add_argument("--serving-fp64-max-order", type=int, default=1,
                    help="served unlimited 2D Euler cases up to this order run the FP64 stage")
    ap.add_argument("--serving-fp64-tiles", default="",
                    help="tiles of the served FP64 2D Euler classes, order:kernel:lanes,... "
                         "(kernel 1 point split, 2 variable split); default: each case's "
                         "autotuned tile")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device to device (a plain copy kernel of known "
                         "size, to scale hardware byte counters)")
    args = ap.parse_args(argv)

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this engine runs one case per GPU and the "
              "multi-GPU path is not part of this build; run with --gpus 1", file=sys.stderr)
        return 2

    import torch

    import wasl_amd
    from wasl_amd import meshgen as MG

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 2

    if args.calibrate_copy:
        n = args.calibrate_copy * 2 ** 20 // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.calibrate_copy} MiB read + {args.calibrate_copy} MiB written")
        del src, dst

    if args.workload == "advect2d":
        return advect2d_run(args)
    if args.workload == "sod":
        return sod_run(args)

    if args.workload == "rarefaction":
        res = rarefaction_run(args.cells, args.end_time)
        log(json.dumps(res))
        print(json.dumps({"metric": "limited rarefaction DOF-updates/s (P2)",
                          "value": res.get("dof_updates_per_s"), "unit": "DOF-updates/s",
                          "n_gpus": 1, "higher_is_better": True, "dtype": "f64",
                          "data": "synthetic", "config": res}), flush=True)
        return 0

    p = args.order
    mesh_in = MG.jittered_triangles(args.cells, (0.0, 0.0), (10.0, 10.0), 0.15)
    mesh = wasl_amd.Mesh.from_input(mesh_in)
    precision = {"fp64": wasl_amd.FP64, "mixed": wasl_amd.MIXED, "bf16": wasl_amd.BF16}[
        args.precision]
    eng = wasl_amd.Engine(mesh, wasl_amd.EULER, p, lanes=args.lanes,
                          stage_kernel=args.stage_kernel, precision=precision)
    eng.project(wasl_amd.IC_VORTEX, (5.0, 5.0, 5.0))
    E, Np, V = eng.E, eng.Np, eng.V
    log(f"engine: {E} triangles, P{p}, stage kernel {eng.stage_kernel}, "
        f"lanes/element {eng.lanes}, "
        f"{eng.device_bytes / 2**20:.1f} MiB on device")

    cfl = 0.15
    t, s, _ = eng.advance(cfl, 1e9, max_steps=args.warmup, stop_at_max_steps=True)
    torch.cuda.synchronize()
    log(f"warmup {s} steps done (t = {t:.4f})")

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    t, s, _ = eng.advance(cfl, 1e9, max_steps=s + args.steps, time=t, step=s,
                          stop_at_max_steps=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    value = E * Np * V * args.steps / wall
    log(f"timed {args.steps} steps: {ms_per_step:.4f} ms/step, {value:.4e} DOF-updates/s")

    # Stage-kernel duration: HIP events around every stage kernel on the
    # engine's stream over a second pass of the same length.
    eng.set_timing(True)
    t, s, _ = eng.advance(cfl, 1e9, max_steps=s + args.steps, time=t, step=s,
                          stop_at_max_steps=True)
    eng.set_timing(False)
    _, stage_ms_total, launches = eng.last_timing()
    stage_ms = stage_ms_total / max(launches, 1)
    bytes_stage = E * algorithmic_bytes_per_element_step(Np, V) / 3.0
    achieved = bytes_stage / (stage_ms * 1e-3) / 1e9
    flops_stage = E * flops_per_element_stage(p)
    log(f"stage kernel: {stage_ms * 1e3:.2f} us average over {launches} launches, "
        f"{achieved:.1f} GB/s algorithmic, {flops_stage / (stage_ms * 1e-3) / 1e12:.2f} TFLOP/s")

    U = eng.state
    finite = bool(np.isfinite(U).all())
    kernel = ("whip::euler_stage_mx" if args.precision != "fp64" else
              ["whip::euler_stage", "whip::euler_stage_vs"][eng.stage_kernel - 1])
    tile = (f"<{p}>" if args.precision != "fp64" else
            f"<{p},{eng.lanes}>" if eng.stage_kernel == 1 else f"<{p},{eng.lanes // 4}>")
    traffic = MEASURED_TRAFFIC.get((kernel + tile, p, E, args.precision))

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64" if args.precision == "mixed"
                 else "bf16+f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"2D Euler isentropic vortex, P{p} modal Dubiner, "
                        f"2x{args.cells}x{args.cells}={E} jittered periodic triangles on [0,10]^2, "
                        f"HLLC, SSP-RK3, CFL {cfl}, device dt, " +
                        {"fp64": "plain FP64",
                         "mixed": "mixed precision (FP32 matrix-core contractions, FP64 state)",
                         "bf16": "BF16 volume contractions (ablation)"}[args.precision],
            "elements": E, "modes": Np, "variables": V,
            "dofs": E * Np * V, "lanes_per_element": eng.lanes,
            "stage_kernel": ("matrix-core mixed" if args.precision != "fp64" else
                             ["point split", "variable split"][eng.stage_kernel - 1]),
            "stage_rate": 3 * args.steps / wall, "state_finite": finite,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_frac_of_algorithmic": None if traffic is None else traffic / bytes_stage,
            "kernel": kernel + tile,
            "stage_kernel_us": stage_ms * 1e3,
            "algorithmic_bytes_per_launch": bytes_stage,
            "fp64_tflops": flops_stage / (stage_ms * 1e-3) / 1e12,
            "fp64_frac": flops_stage / (stage_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
        },
        "cpu_baseline": None,
    }
    if args.precision != "fp64":
        result["precision_error"] = precision_error(mesh, p, precision, args.error_steps)
        log(f"{args.precision} vs FP64 after {args.error_steps} steps: density L1 "
            f"{result['precision_error']['density_l1']:.3e}, L-inf "
            f"{result['precision_error']['density_linf']:.3e}")
    if args.serving_requests > 0:
        del eng
        tiles = {int(o): (int(k), int(n)) for o, k, n in
                 (t.split(":") for t in args.serving_fp64_tiles.split(",") if t)}
        result["serving"] = serving_measurement(args.serving_requests, args.serving_seed,
                                                args.serving_steps, args.serving_threads,
                                                tiles or None, args.serving_fp64_max_order)
    if not args.no_cpu:
        log(f"cpu baseline: timing up to {args.cpu_seconds:.0f} s of the CPU path")
        result["cpu_baseline"] = cpu_baseline(mesh_in, p, cfl, args.cpu_seconds)
        log(f"cpu baseline: {result['cpu_baseline']['value']:.4e} DOF-updates/s")
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
