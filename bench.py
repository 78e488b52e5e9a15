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
    ("whip::euler_stage<2,4>", 2, 32768, "fp64"): (7827.3 * 2.0 + 6144.3) * 1024.0,
    ("whip::euler_stage<2,2>", 2, 32768, "fp64"): (7965.57 * 2.0 + 6146.72) * 1024.0,
    ("whip::euler_stage_mx<3>", 3, 524288, "mixed"): (224747.1 * 2.0 + 262158.2) * 1024.0,
    ("whip::advect_stage<1>", 1, 524288, "fp64"): (26869.3 * 2.0 + 12288.0) * 1024.0,
    ("whip::advect_stage<3>", 3, 524288, "fp64"): (51451.3 * 2.0 + 40960.0) * 1024.0,
    ("whip::euler_line_stage<2>", 2, 1048576, "fp64"): (94410.7 * 2.0 + 98319.2) * 1024.0,
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


def _cpu_kind() -> str:
    """The reference build the CPU baseline times: its -O3 x86-64-v3 build
    where the host has AVX2 and FMA, else the parity build (-O2), else the
    project's restatement."""
    from oracle import oracle as O

    flags = ""
    try:
        with open("/proc/cpuinfo") as fh:
            flags = fh.read()
    except OSError:
        pass
    if O.available("reference_fast") and " avx2" in flags and " fma" in flags:
        return "reference_fast"
    return "reference" if O.available("reference") else "port"


def _cpu_worker(args):
    """One independent vortex case on one core: (DOF-updates, seconds, steps)."""
    node_tags, node_xy, triangles, order, cfl, budget_s, kind = args
    from oracle import oracle as O

    c = O.tri_euler(node_tags, node_xy, triangles, order, kind=kind)
    U = c.project(O.IC_VORTEX, (5.0, 5.0, 5.0))
    steps = 0
    t0 = time.perf_counter()
    while True:
        U = c.step(U, c.stable_dt(U, cfl))
        steps += 1
        el = time.perf_counter() - t0
        if el > budget_s or steps >= 200:
            break
    return c.E * c.Np * c.V * steps, el, steps


def cpu_baseline(mesh, order, cfl, budget_s: float, all_cores_s: float = 10.0):
    """The reference's own CPU path (oracle/_ref, compiled from the reference
    sources; its -O3 AVX2/FMA build where the host allows) timed on a bounded
    number of steps of the same case: single-threaded, as the reference runs,
    and, for the aggregate, one independent case per core (BASELINE's CPU
    plan, item 4)."""
    import multiprocessing as mp

    kind = _cpu_kind()
    build = {"reference_fast": "reference sources, g++ -O3 -march=x86-64-v3 (AVX2, FMA)",
             "reference": "reference sources, g++ -O2 -ffp-contract=off (parity build)",
             "port": "project restatement, g++ -O2"}[kind]
    args = (mesh.node_tags, mesh.node_xy, mesh.triangles, order, cfl, budget_s, kind)
    updates, el, steps = _cpu_worker(args)
    out = {
        "value": updates / el,
        "unit": "DOF-updates/s",
        "cores": 1,
        "kind": "port" if kind == "port" else "reference",
        "build": build,
        "sample": f"{steps} SSP-RK3 steps of the same P{order} {mesh.triangles.shape[0]}-triangle vortex "
                  f"case from t=0 ({el:.1f} s, single thread, {cpu_model()}, nproc "
                  f"{os.cpu_count()})",
    }
    if all_cores_s > 0:
        n = max(1, min(os.cpu_count() or 1, 16))
        ctx = mp.get_context("spawn")
        t0 = time.perf_counter()
        with ctx.Pool(n) as pool:
            res = pool.map(_cpu_worker, [args[:5] + (all_cores_s, kind)] * n)
        wall = time.perf_counter() - t0
        out["all_cores"] = {
            "value": sum(u / e for u, e, _ in res),
            "unit": "DOF-updates/s",
            "cores": n,
            "sample": f"{n} independent copies of the case, one per process, each timed over "
                      f"up to {all_cores_s:.0f} s or 200 steps; value = sum of their rates "
                      f"({wall:.1f} s wall including setup)",
        }
    return out


def precision_error(mesh, p: int, precision: int, steps: int, end_time: float = 0.0) -> dict:
    """The reduced-precision stage against the FP64 engine (the plain path,
    parity-tested against the CPU oracle) on the bench workload, both from the
    same projected vortex: `steps` steps of the same fixed dt (the FP64 CFL
    step of the initial state), or, with end_time > 0, the device loop of
    each engine to end_time, each choosing its own CFL dt and landing on
    end_time exactly. Density is compared as cell averages (L1 weighted by
    element area), as all modal coefficients (L-inf, also relative to the
    initial density's maximum cell average) and as the density field sampled
    at lattice points of every element (L-inf)."""
    import wasl_amd

    runs = {}
    for name, prec in (("fp64", wasl_amd.FP64), ("reduced", precision)):
        g = wasl_amd.Engine(mesh, wasl_amd.EULER, p, precision=prec)
        g.project(wasl_amd.IC_VORTEX, (5.0, 5.0, 5.0))
        if name == "fp64":
            dt = g.stable_time_step(0.15)
            rho0 = g.state.reshape(4, -1, g.E)[0, 0] / np.sqrt(2.0)
        t0 = time.perf_counter()
        if end_time > 0:
            t, n, _ = g.advance(0.15, end_time)
        else:
            for _ in range(steps):
                g.step(dt)
            t, n = steps * dt, steps
        runs[name] = (g.state.reshape(4, -1, g.E)[0], t, n, time.perf_counter() - t0, g.state)
        E = g.E
        del g
    from wasl_amd import workload as WL

    d = runs["reduced"][0] - runs["fp64"][0]
    avg = d[0] / np.sqrt(2.0)
    scale = float(np.max(np.abs(rho0)))
    detj = mesh.prepared()["element_metrics"][:E, 4]   # proportional to the element's area
    out = {
        "steps": {k: v[2] for k, v in runs.items()},
        "time": {k: v[1] for k, v in runs.items()},
        "wall_s": {k: v[3] for k, v in runs.items()},
        "density_l1": float(np.sum(np.abs(avg) * detj) / np.sum(detj)),
        "density_linf": float(np.max(np.abs(d))),
        "density_linf_relative": float(np.max(np.abs(d))) / scale,
        "density_field_linf": WL.density_field_linf(runs["reduced"][4], runs["fp64"][4], p, E),
    }
    if end_time > 0:
        out["reference"] = (f"FP64 engine, same initial state, each engine's own device dt to "
                            f"T = {end_time}")
    else:
        out.update(dt=dt, reference="FP64 engine, same initial state and fixed dt")
    return out


def rarefaction_run(cells: int, end_time: float, steps: int = 20, reps: int = 10) -> dict:
    """The radial double rarefaction (rho = 1, p = 0.4, velocity 2 r_hat) on
    [-1,1]^2 with transmissive boundaries, 2 cells^2 jittered triangles, P2,
    characteristic WENO and positivity limiting (floors 1e-13), CFL 0.1,
    through the FP64 engine, from the limited projection as the reference
    composes it (project, limit, advance).

    With these settings the reference itself loses a physical cell average
    in its 22nd step at 2x256x256 (NaN face fluxes next to the origin), and
    so does the engine. The run to end_time is therefore reported as its
    status, and the rate is measured over `reps` runs of the first `steps`
    limited steps from the same initial state (setup excluded)."""
    import torch

    import wasl_amd
    from wasl_amd import meshgen as MG

    mesh = wasl_amd.Mesh.from_input(MG.jittered_triangles(cells, (-1.0, -1.0), (1.0, 1.0), 0.15))
    g = wasl_amd.Engine(mesh, wasl_amd.EULER, 2, boundary=(wasl_amd.OUTFLOW, wasl_amd.OUTFLOW),
                        limiting=True, density_floor=1e-13, pressure_floor=1e-13)
    g.project(wasl_amd.IC_RAREFACTION, (1.0, 0.4, 2.0))
    g.limit()   # the reference limits the projected initial state
    U0 = g.state
    g.capture_graphs()
    out = {"workload": f"radial double rarefaction, P2, 2x{cells}x{cells}={g.E} jittered "
                       f"triangles on [-1,1]^2, transmissive, char-WENO + positivity (1e-13), "
                       f"CFL 0.1, from the limited projection", "elements": g.E}
    try:
        t, n, _ = g.advance(0.1, end_time)
        out["run_to_end_time"] = {"end_time": end_time, "status": "completed", "steps": n}
    except wasl_amd.SimulationError as exc:
        out["run_to_end_time"] = {"end_time": end_time, "status": f"stopped: {exc}"}
    done = 0
    wall = 0.0
    try:
        for _ in range(reps):
            g.state = U0
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            t, n, _ = g.advance(0.1, 1e9, max_steps=steps, stop_at_max_steps=True)
            torch.cuda.synchronize()
            wall += time.perf_counter() - t0
            done += n
        out["status"] = f"{reps} runs of {steps} limited steps from the limited projection"
    except wasl_amd.SimulationError as exc:
        out["status"] = f"stopped: {exc}"
    out.update({"steps_timed": done, "wall_s": wall})
    if done > 0 and wall > 0.0:
        out["dof_updates_per_s"] = g.E * g.Np * 4 * done / wall
        U = g.state.reshape(4, g.Np, g.E)[:, 0] / np.sqrt(2.0)
        rho, mx, my, en = U
        p = 0.4 * (en - 0.5 * (mx * mx + my * my) / rho)
        out["time_after_steps"] = t
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
    g.capture_graphs()
    t, s, _ = g.advance(cfl, 1e9, max_steps=args.warmup, stop_at_max_steps=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    t, s, _ = g.advance(cfl, 1e9, max_steps=s + args.steps, time=t, step=s,
                        stop_at_max_steps=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    value = E * Np * args.steps / wall
    finite = bool(np.isfinite(g.state).all())
    stage_ms = float(np.mean(g.time_stages(args.stage_reps, g.stable_time_step(cfl)))) * 1e-3
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
            "state_finite": finite,
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
    g.capture_graphs()
    t, s, _ = g.advance(cfl, 1e9, max_steps=args.warmup, stop_at_max_steps=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    t, s, _ = g.advance(cfl, 1e9, max_steps=s + args.steps, time=t, step=s,
                        stop_at_max_steps=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    value = E * Np * 4 * args.steps / wall
    finite = bool(np.isfinite(g.state).all())
    stage_ms = float(np.mean(g.time_stages(args.stage_reps, g.stable_time_step(cfl)))) * 1e-3
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
            "state_finite": finite,
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
    """The serving mix served (mixed precision, continuous batching) and run
    plain (FP64, one case at a time): aggregate DOF-updates/s of each, and
    time to solution of both under the same Poisson arrivals at 80 % of the
    served capacity measured closed-loop. Every tile-autotuning key of the
    mix is tuned before either leg is timed. Requests that the reference's
    own physicality checks stop (on the plain path) are listed and left out
    of both rates."""
    import numpy as np

    from wasl_amd import workload as WL
    from wasl_amd.serve import default_fp64_tiles

    mix = WL.serving_mix(count, seed)
    meshes = WL.MeshCache()
    t0 = time.perf_counter()
    for r in mix:
        meshes.get(r.mesh_key)
    keys = WL.warm_autotune(mix, meshes)
    log(f"serving mix: {count} requests ({sum(r.kind != 'advection' for r in mix)} Euler), "
        f"{len(meshes.meshes)} meshes and {keys} tile keys prepared in "
        f"{time.perf_counter() - t0:.1f} s")
    kw = dict(steps_per_batch=steps_per_batch, setup_threads=setup_threads,
              fp64_tiles=fp64_tiles, fp64_max_order=fp64_max_order)

    done, makespan, stats = WL.run_served(mix, meshes, **kw)
    capacity = count / makespan
    log(f"served closed-loop: {makespan:.2f} s for {count} requests; {stats}")
    rng = np.random.default_rng(seed + 1)
    rate = 0.8 * capacity
    arrivals = np.cumsum(rng.exponential(1.0 / rate, count)).tolist()
    done_o, makespan_o, _ = WL.run_served(mix, meshes, arrivals=arrivals, **kw)
    log(f"served open-loop at {rate:.2f} req/s: {makespan_o:.2f} s")

    t0 = time.perf_counter()
    WL.run_plain(mix, meshes, arrivals=arrivals)
    plain_wall = time.perf_counter() - t0
    failed = [k for k, r in enumerate(mix) if r.plain_error]
    for k in failed:
        r = mix[k]
        log(f"  plain stop: request {r.index} {r.kind} P{r.order} {r.size} elements, "
            f"{r.steps} steps: {r.plain_error}")
    ok = [k for k in range(count) if not mix[k].plain_error]
    served_failed = [k for k in ok if done[k][3] is None]
    for k in served_failed:
        r = mix[k]
        log(f"  served stop: request {r.index} {r.kind} P{r.order} {r.size} elements")
    both = [k for k in ok if done[k][3] is not None]
    plain_total = sum(mix[k].plain_seconds for k in both)
    updates = sum(mix[k].dofs * mix[k].plain_steps for k in both)
    plain_rate = updates / plain_total
    served_updates = sum(mix[k].dofs * done[k][2] for k in both)
    served_rate = served_updates / makespan
    end_mismatch = sum(1 for k in both if done[k][4] != mix[k].end_time)
    log(f"plain: {plain_total:.2f} s of service ({plain_wall:.2f} s open-loop wall), "
        f"{plain_rate:.4e} DOF-updates/s; served {served_rate:.4e} DOF-updates/s, "
        f"x{served_rate / plain_rate:.2f} plain ({len(failed)} requests stopped by the "
        f"reference's physicality checks, left out)")

    tts = [done_o[k][1] - done_o[k][0] for k in both]
    plain_tts = [mix[k].plain_done - arrivals[k] for k in both]
    worst, worst_coef = {}, {}
    for k in both:
        r = mix[k]
        if r.kind == "advection":
            continue
        Np = (r.order + 1) * (r.order + 2) // 2
        a = done[k][3].reshape(4, Np, r.size)[0]
        b = r.plain_state.reshape(4, Np, r.size)[0]
        worst_coef[r.kind] = max(worst_coef.get(r.kind, 0.0), float(np.max(np.abs(a - b))))
        worst[r.kind] = max(worst.get(r.kind, 0.0),
                            WL.density_field_linf(done[k][3], r.plain_state, r.order, r.size))
    out = {
        "requests": count,
        "requests_measured": len(both),
        "requests_failed_plain": len(failed),
        "requests_failed_served_only": len(served_failed),
        "served_end_time_mismatches": end_mismatch,
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
        "plain_tts_note": "measured: the plain path serves the same arrivals one case at a time "
                          "in arrival order",
        "open_loop_makespan_s": makespan_o,
        "plain_open_loop_wall_s": plain_wall,
        "closed_loop_server": stats,
        "density_linf_served_vs_fp64": worst,
        "density_linf_note": "max over requests of each kind of |rho_served(x) - rho_plain(x)|, "
                             "the density fields (mixed against FP64) sampled at the "
                             "(2p+3)(2p+4)/2 lattice points of every element; limited cases run "
                             "the FP64 stage",
        "density_coefficient_linf_served_vs_fp64": worst_coef,
    }
    log(f"TTS p50 {out['tts_p50_s']:.3f} s p99 {out['tts_p99_s']:.3f} s (plain p50 "
        f"{out['plain_tts_p50_s']:.3f} s p99 {out['plain_tts_p99_s']:.3f} s); density L-inf "
        f"vs FP64 {worst}")
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
    ap.add_argument("--serving-requests", type=int, default=1024,
                    help="requests of the serving mix (0: skip the serving measurement)")
    ap.add_argument("--serving-seed", type=int, default=0x5EED)
    ap.add_argument("--workload", choices=["vortex", "rarefaction", "advect2d", "sod"],
                    default="vortex",
                    help="vortex: the headline (config 2 sizes by default); rarefaction: the "
                         "limited radial rarefaction (its run to --end-time reported, 10 x 20 "
                         "limited steps timed); advect2d: 2D scalar "
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
    ap.add_argument("--stage-reps", type=int, default=100,
                    help="back-to-back launches of each stage kernel for its duration")
    ap.add_argument("--error-end-time", type=float, default=0.0,
                    help="with --precision mixed/bf16: compare with the FP64 engine at this end "
                         "time (each engine choosing its own dt) instead of --error-steps steps")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-all-cores-seconds", type=float, default=10.0,
                    help="per-process budget of the all-cores CPU figure (0: skip it)")
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
    # setup: every step graph of the device loop is captured now, so none
    # is captured inside the timed region whatever its length
    eng.capture_graphs()
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

    U = eng.state
    finite = bool(np.isfinite(U).all())
    # Stage-kernel duration: each stage launched back to back from a captured
    # graph, HIP events on the engine's stream around the replay (the state
    # is not meaningful afterwards, so it was read above).
    stage_each = eng.time_stages(args.stage_reps, eng.stable_time_step(cfl))
    stage_ms = float(np.mean(stage_each)) * 1e-3
    assert 3 * stage_ms <= 1.02 * ms_per_step, (
        f"stage kernel {stage_ms * 1e3:.2f} us x 3 exceeds the step time {ms_per_step * 1e3:.2f} us")
    launches = 3 * args.stage_reps
    bytes_stage = E * algorithmic_bytes_per_element_step(Np, V) / 3.0
    achieved = bytes_stage / (stage_ms * 1e-3) / 1e9
    flops_stage = E * flops_per_element_stage(p)
    log(f"stage kernel: {stage_ms * 1e3:.2f} us average over {launches} back-to-back launches "
        f"(stages {', '.join(f'{x:.2f}' for x in stage_each)} us), {achieved:.1f} GB/s "
        f"algorithmic, {flops_stage / (stage_ms * 1e-3) / 1e12:.2f} TFLOP/s")

    kernel = ("whip::euler_stage_mx" if args.precision != "fp64" else
              ["whip::euler_stage", "whip::euler_stage_vs"][eng.stage_kernel - 1])
    tile = (f"<{p}>" if args.precision != "fp64" else
            f"<{p},{eng.lanes}>" if eng.stage_kernel == 1 else f"<{p},{eng.lanes // 4}>")
    traffic = MEASURED_TRAFFIC.get((kernel + tile, p, E, args.precision))
    ceilings = wasl_amd.fp64_ceilings()
    log(f"FP64 ceilings on this device: vector ALU {ceilings[0]:.1f} TFLOP/s, matrix core "
        f"{ceilings[1]:.1f} TFLOP/s")

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
            "stage_kernel_us_by_stage": [float(x) for x in stage_each],
            "stage_timing": f"{args.stage_reps} back-to-back launches of each stage replayed "
                            f"from a captured graph, HIP events on the engine's stream",
            "algorithmic_bytes_per_launch": bytes_stage,
            "fp64_tflops": flops_stage / (stage_ms * 1e-3) / 1e12,
            "fp64_frac": flops_stage / (stage_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
            # what this device sustains on plain FMA chains: the FP64 matrix
            # core (16x16x4) runs below the vector ALU, so the FP64 stage
            # keeps its contractions on the vector ALU
            "fp64_valu_ceiling_tflops": ceilings[0],
            "fp64_mfma_ceiling_tflops": ceilings[1],
        },
        "cpu_baseline": None,
    }
    if args.precision != "fp64":
        result["precision_error"] = precision_error(mesh, p, precision, args.error_steps,
                                                    args.error_end_time)
        log(f"{args.precision} vs FP64 ({result['precision_error']['steps']} steps): density L1 "
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
        result["cpu_baseline"] = cpu_baseline(mesh_in, p, cfl, args.cpu_seconds,
                                              args.cpu_all_cores_seconds)
        log(f"cpu baseline: {result['cpu_baseline']['value']:.4e} DOF-updates/s")
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
