"""The serving mix and its two executions, plain and served.

The mix: seeded requests, 70 % 2D Euler and 30 % 1D advection; triangle
counts log-uniform in [512, 131072] (quad-split n x n grids, n = round(sqrt(N/2)))
and line cells log-uniform in [256, 65536]; order uniform in P1-P3; Euler
initial states drawn from {isentropic vortex, radial double rarefaction,
smooth Fourier field}; each run lasts a uniform 200-2000 steps of its initial
CFL step. The rarefaction runs on [-1,1]^2 with transmissive boundaries,
characteristic WENO and positivity limiting (floors 1e-13, CFL 0.1) and ends
at t = 0.3 at the latest (the near-vacuum time of the rarefaction case); vortex
and smooth fields on periodic [0,10]^2 at CFL 0.15; advection of the pulse on
periodic [0,1] at CFL 0.1.

plain:  one case at a time through the FP64 engine (create, project, limit
        when limiting, advance), in arrival order as requests arrive
served: all cases through a Server (mixed precision, continuous batching)
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import engine as W
from . import meshgen as MG


@dataclass
class Request:
    index: int
    kind: str            # "vortex", "smooth", "rarefaction", "advection"
    size: int            # triangles or cells
    order: int
    steps: int
    mesh_key: tuple
    equation: int
    ic: int
    params: tuple
    cfl: float
    boundary: tuple = (W.PERIODIC, W.PERIODIC)
    limiting: bool = False
    floors: float = 1e-10
    # filled by the plain run
    end_time: float = 0.0
    dofs: int = 0
    plain_seconds: float = 0.0
    plain_steps: int = 0
    plain_state: np.ndarray | None = field(default=None, repr=False)
    plain_error: str = ""
    plain_done: float = 0.0
    end_cap: float = float("inf")


def serving_mix(count: int, seed: int = 0x5EED) -> list[Request]:
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        order = int(rng.integers(1, 4))
        steps = int(rng.integers(200, 2001))
        if rng.random() < 0.7:
            tris = math.exp(rng.uniform(math.log(512), math.log(131072)))
            n = max(2, int(round(math.sqrt(tris / 2))))
            kind = ("vortex", "rarefaction", "smooth")[int(rng.integers(0, 3))]
            if kind == "vortex":
                out.append(Request(i, kind, 2 * n * n, order, steps, ("box10", n), W.EULER,
                                   W.IC_VORTEX, (5.0, 5.0, 5.0), 0.15))
            elif kind == "smooth":
                prm = [10.0]
                for _ in range(4):
                    prm += [float(rng.uniform(-1, 1)), int(rng.integers(1, 4)),
                            int(rng.integers(1, 4)), float(rng.uniform(0, 2 * math.pi))]
                out.append(Request(i, kind, 2 * n * n, order, steps, ("box10", n), W.EULER,
                                   W.IC_SMOOTH, tuple(prm), 0.15))
            else:
                out.append(Request(i, kind, 2 * n * n, order, steps, ("box1", n), W.EULER,
                                   W.IC_RAREFACTION, (1.0, 0.4, 2.0), 0.1,
                                   (W.OUTFLOW, W.OUTFLOW), True, 1e-13, end_cap=0.3))
        else:
            cells = int(round(math.exp(rng.uniform(math.log(256), math.log(65536)))))
            out.append(Request(i, "advection", cells, order, steps, ("line", cells),
                               W.SCALAR_ADVECTION, W.IC_SCALAR_PULSE, (), 0.1))
    return out


class MeshCache:
    def __init__(self):
        self.meshes = {}

    def get(self, key):
        if key not in self.meshes:
            kind, n = key
            if kind == "box10":
                m = MG.jittered_triangles(n, (0.0, 0.0), (10.0, 10.0), 0.15)
            elif kind == "box1":
                m = MG.jittered_triangles(n, (-1.0, -1.0), (1.0, 1.0), 0.15)
            else:
                m = MG.jittered_line(n)
            self.meshes[key] = W.Mesh.from_input(m)
        return self.meshes[key]


def _engine(r: Request, mesh, precision=W.FP64):
    return W.Engine(mesh, r.equation, r.order, boundary=r.boundary, limiting=r.limiting,
                    density_floor=r.floors, pressure_floor=r.floors, precision=precision,
                    wavespeed=1.0)


def run_plain(reqs: list[Request], meshes: MeshCache, keep_states: bool = True,
              arrivals=None) -> float:
    """One case at a time, FP64: engine setup, projection (and the initial
    limit of limited cases) and the stepping loop are the service time. With
    arrivals (offsets in seconds from the start) the requests are served in
    arrival order as they arrive, a single FIFO server that idles until the
    next arrival when it has caught up; each request's completion offset goes
    to plain_done. Sets each request's end time (its steps times the initial
    CFL step) and DOF count; returns the total service seconds."""
    import torch

    order = list(range(len(reqs))) if arrivals is None else \
        sorted(range(len(reqs)), key=lambda k: arrivals[k])
    total = 0.0
    start = time.perf_counter()
    for k in order:
        r = reqs[k]
        if arrivals is not None:
            wait = arrivals[k] - (time.perf_counter() - start)
            if wait > 0:
                time.sleep(wait)
        mesh = meshes.get(r.mesh_key)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g = _engine(r, mesh)
        g.project(r.ic, r.params)
        if r.limiting:
            g.limit()   # the reference limits the projected state (euler_triangle.cpp:37-41)
        r.end_time = min(r.steps * g.stable_time_step(r.cfl), r.end_cap)
        r.dofs = g.E * g.Np * g.V
        try:
            t, s, _ = g.advance(r.cfl, r.end_time)
            r.plain_steps = s
            if keep_states:
                r.plain_state = g.state
        except W.SimulationError as exc:
            # the reference stops the same way (cfl_timestep.cpp:39-40,
            # euler_positivity.hpp:74-75); the request is left out of the rates
            r.plain_error = str(exc)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        r.plain_seconds = t1 - t0
        r.plain_done = t1 - start
        total += r.plain_seconds
        del g
    return total


def warm_autotune(reqs: list[Request], meshes: MeshCache) -> int:
    """Creates one FP64 engine per tile-autotuning key of the mix (2D Euler:
    order, limiting and the engine's mesh-size class), so neither timed leg
    pays the process-wide tile search. Returns the number of keys."""
    keys = {}
    for r in reqs:
        if r.equation != W.EULER:
            continue
        size_class = 0 if r.size < 8192 else 1 if r.size < 131072 else 2
        keys.setdefault((r.order, r.limiting, size_class), r)
    for r in keys.values():
        g = _engine(r, meshes.get(r.mesh_key))
        del g
    return len(keys)


def run_served(reqs: list[Request], meshes: MeshCache, arrivals=None, precision=W.MIXED,
               steps_per_batch: int = 16, max_resident: int = 256, setup_threads: int = 4,
               fp64_tiles=None, fp64_max_order: int = 1):
    """Submit each request at its arrival offset (all at once when arrivals is
    None) and pump the server until every request is finished. Returns
    (per-request (arrival, completion, steps, state) dict, makespan seconds,
    server counters: rounds, graphs, resident-set changes, and the seconds
    spent rebuilding launch tables and graphs, waiting for rounds and
    preparing cases summed over the setup threads)."""
    from .serve import Server

    s = Server(precision=precision, steps_per_batch=steps_per_batch, max_resident=max_resident,
               setup_threads=setup_threads, fp64_tiles=fp64_tiles,
               fp64_max_order=fp64_max_order)
    order = sorted(range(len(reqs)), key=lambda k: 0.0 if arrivals is None else arrivals[k])
    pending = list(order)
    ids = {}
    done = {}
    t0 = time.perf_counter()
    while pending or len(done) < len(reqs):
        now = time.perf_counter() - t0
        while pending and (arrivals is None or arrivals[pending[0]] <= now):
            k = pending.pop(0)
            r = reqs[k]
            # the server fixes the end time as the plain run does: the
            # request's steps times its initial CFL step, capped at end_cap
            rid = s.submit(meshes.get(r.mesh_key), r.equation, r.order, r.ic, r.params, cfl=r.cfl,
                           end_steps=r.steps,
                           end_time=r.end_cap if math.isfinite(r.end_cap) else 0.0,
                           boundary=r.boundary, limiting=r.limiting,
                           density_floor=r.floors, pressure_floor=r.floors, wavespeed=1.0)
            ids[rid] = (k, now if arrivals is not None else 0.0)
        if s.idle and not s.count(3):
            if pending:
                time.sleep(max(0.0, min(0.01, arrivals[pending[0]] - now)))
            continue
        for rid in s.pump():
            k, arrived = ids[rid]
            finish = time.perf_counter() - t0
            try:
                state, t, steps = s.result(rid)
            except W.WaslHipError:
                state, t, steps = None, float("nan"), 0
            done[k] = (arrived, finish, steps, state, t)
    makespan = time.perf_counter() - t0
    stats = {"rounds": s.count(6), "graphs": s.count(5), "resident_changes": s.count(7),
             "rebuild_s": s.count(8) * 1e-6, "round_wait_s": s.count(9) * 1e-6,
             "prepare_s": s.count(10) * 1e-6, "admit_s": s.count(11) * 1e-6,
             "retire_s": s.count(12) * 1e-6, "idle_s": s.count(13) * 1e-6,
             "submit_s": s.count(14) * 1e-6, "release_s": s.count(15) * 1e-6}
    s.close()
    return done, makespan, stats


def percentile(xs, q):
    return float(np.percentile(np.asarray(xs, np.float64), q))


def _jacobi(n: int, alpha: float, x: np.ndarray) -> np.ndarray:
    """P_n^(alpha, 0)(x) by the three-term recurrence."""
    p0 = np.ones_like(x)
    if n == 0:
        return p0
    p1 = 0.5 * ((alpha + 2.0) * x + alpha)
    for k in range(2, n + 1):
        s = 2 * k + alpha
        a1 = 2.0 * k * (k + alpha) * (s - 2.0)
        a2 = (s - 1.0) * alpha * alpha
        a3 = (s - 2.0) * (s - 1.0) * s
        a4 = 2.0 * (k + alpha - 1.0) * (k - 1.0) * s
        p0, p1 = p1, ((a2 + a3 * x) * p1 - a4 * p0) / a1
    return p1


def dubiner_lattice(order: int, n: int = 0) -> np.ndarray:
    """Orthonormal Dubiner basis [points][modes] (modes by degree, then i) at
    the lattice points (xi, eta) = (-1 + 2a/n, -1 + 2b/n), a + b <= n, of the
    reference triangle (n = 2 order + 2 by default): vertices, edges and
    interior, so a maximum over these points samples a field's L-inf norm."""
    n = n or 2 * order + 2
    pts = [(-1.0 + 2.0 * a / n, -1.0 + 2.0 * b / n) for b in range(n + 1) for a in range(n + 1 - b)]
    xi = np.array([p[0] for p in pts])
    eta = np.array([p[1] for p in pts])
    top = eta >= 1.0
    a = np.where(top, -1.0, 2.0 * (1.0 + xi) / np.where(top, 1.0, 1.0 - eta) - 1.0)
    cols = []
    for deg in range(order + 1):
        for i in range(deg + 1):
            j = deg - i
            norm = math.sqrt((2 * i + 1) * (deg + 1) / 2.0)
            cols.append(norm * _jacobi(i, 0.0, a) * (0.5 * (1.0 - eta)) ** i *
                        _jacobi(j, 2.0 * i + 1.0, eta))
    return np.stack(cols, axis=1)


def density_field_linf(a: np.ndarray, b: np.ndarray, order: int, elements: int) -> float:
    """max |rho_a(x) - rho_b(x)| over the dubiner_lattice points of every
    element, for two 2D Euler states in the U[variable][mode][element] layout."""
    Np = (order + 1) * (order + 2) // 2
    d = a.reshape(4, Np, elements)[0] - b.reshape(4, Np, elements)[0]
    return float(np.max(np.abs(dubiner_lattice(order) @ d)))
