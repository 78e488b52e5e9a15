"""1D Euler on lines through the HIP engine (dg/euler/line_residual.hpp:23-86,
src/boundaries/line_euler_exterior.cpp:13-43, the line limiter stencil and
src/case/euler_line.cpp:14-48) against the oracle, the vectors the
reference's own code wrote (tests/golden/reference_vectors.npz) and the Sod
tutorial's published accuracy gate.

Tolerances: residual, limiter and single steps relative L-inf <= 1e-12 per
variable (max-norm scaled); troubled flags bit-exact; end of run <= 1e-10."""

import math
import os

import numpy as np
import pytest

import cases as K
import refmeshes as RM
from conftest import GOLDEN
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _engine(gpu, mesh, p, bc=O.PERIODIC, **kw):
    return gpu.Engine(gpu.Mesh.lines(*mesh), gpu.EULER, p, boundary=(bc, bc), **kw)


def _uniform_line(N):
    tags = np.arange(1, N + 2)
    return tags, np.arange(N + 1) / N, np.stack([tags[:-1], tags[1:]], 1)


@pytest.mark.parametrize("p", [0, 1, 2, 3])
@pytest.mark.parametrize("bc", [O.PERIODIC, O.OUTFLOW])
def test_reference_vectors(gpu, p, bc):
    f = np.load(os.path.join(GOLDEN, "reference_vectors.npz"))
    tag = "periodic" if bc == O.PERIODIC else "outflow"
    key = f"euler1d_p{p}_{tag}"
    mesh = RM.irregular_line(12)
    g = _engine(gpu, mesh, p, bc, limiting=True)
    assert (g.V, g.Np) == (4, p + 1)
    U = g.project(O.IC_EULER_REGIONS, K.SHOCK_TUBE)
    assert K.rel_inf(U, f[key + "_state"]) <= 1e-15
    err = K.rel_inf_per_variable(g.residual(f[key + "_state"]), f[key + "_residual"], 4)
    print(f"P{p} {tag}: residual rel {err:.2e}", flush=True)
    assert err <= K.STEP_TOL
    g.state = f[key + "_state"]
    flags = g.limit()
    assert np.array_equal(flags, f[key + "_flags"])
    assert K.rel_inf_per_variable(g.state, f[key + "_limited"], 4) <= K.STEP_TOL
    assert abs(g.stable_time_step(0.2) / f[key + "_dt"][0] - 1.0) <= 1e-15
    traj = f[key + "_steps"]
    g.state = traj[0]
    for k in range(1, traj.shape[0]):
        g.step(g.stable_time_step(0.2))
        err = K.rel_inf_per_variable(g.state, traj[k], 4)
        assert err <= K.STEP_TOL * k, (k, err)


@pytest.mark.parametrize("p", [1, 2])
def test_limited_run_matches_oracle(gpu, p):
    """The device-driven limited loop against the oracle's runner, 60 cells,
    periodic ends, the case-regression shock tube."""
    mesh = RM.irregular_line(60)
    o = O.line_euler(*mesh, p)
    U0, _ = o.limit(o.project(O.IC_EULER_REGIONS, K.SHOCK_TUBE), 1e-3)
    Uo, so, to, _ = o.run(U0, 0.2, 0.05, limiting=True, threshold=1e-3)
    g = _engine(gpu, mesh, p, limiting=True)
    g.state = U0
    t, s, _ = g.advance(0.2, 0.05)
    assert s == so and abs(t - to) <= 1e-14
    err = K.rel_inf_per_variable(g.state, Uo, 4)
    print(f"P{p}: {s} steps, end rel {err:.2e}", flush=True)
    assert err <= K.RUN_TOL


def test_sod_tutorial_accuracy_gate(gpu):
    """cases/tutorials/euler/sod-shock-tube/README.md on the device: density L1
    errors 1.083e-2, 5.673e-3, 2.992e-3 at 50/100/200 cells to the digits
    quoted, ratios >= 1.6, positivity of every cell average."""
    from exact_riemann import RiemannProblem

    rp = RiemannProblem(1.4, (1.0, 0.0, 1.0), (0.125, 0.0, 0.1))
    quoted = {50: 1.083e-2, 100: 5.673e-3, 200: 2.992e-3}
    errs = []
    for N, ref in quoted.items():
        tags, x, segs = _uniform_line(N)
        g = gpu.Engine(gpu.Mesh.lines(tags, x, segs), gpu.EULER, 2,
                       boundary=(O.OUTFLOW, O.OUTFLOW), limiting=True)
        g.project(O.IC_EULER_REGIONS, K.SOD)
        g.limit()
        t, _, _ = g.advance(0.2, 0.2)
        avg = g.cell_averages()
        rho, mx, en = avg[:, 0], avg[:, 1], avg[:, 3]
        assert rho.min() > 0 and (0.4 * (en - 0.5 * mx * mx / rho)).min() > 0
        err = sum((x[i + 1] - x[i]) * abs(rho[i] - rp.cell_average_density(x[i], x[i + 1], 0.5, t))
                  for i in range(N))
        print(f"Sod {N} cells: density L1 {err:.4e} (quoted {ref:.3e})", flush=True)
        assert abs(err - ref) <= 0.5e-3 * ref * 1.01
        errs.append(err)
    assert errs[0] / errs[1] >= 1.6 and errs[1] / errs[2] >= 1.6 and errs[2] <= 4e-3


@pytest.mark.parametrize("cells", [2, 61, 62, 63, 124, 125, 187])
@pytest.mark.parametrize("bc", [O.PERIODIC, O.OUTFLOW])
def test_wave_ends(gpu, cells, bc):
    """Cell counts either side of the 62 cells a wave of the stage kernel
    finalises (its first and last lanes hold the cells beyond them, wrapped
    at periodic ends and clamped at outflow ends): residual, CFL dt and two
    SSP-RK3 steps against the oracle at P1 and P3, on the irregular mesh
    whose elements are listed in reverse with alternating orientation."""
    mesh = RM.irregular_line(cells)
    for p in (1, 3):
        g = _engine(gpu, mesh, p, bc)
        o = O.line_euler(*mesh, p, left=bc, right=bc)
        U = o.project(O.IC_EULER_REGIONS, K.SHOCK_TUBE)
        err = K.rel_inf_per_variable(g.residual(U), o.residual(U), 4)
        print(f"{cells} cells P{p} bc {bc}: residual rel {err:.2e}", flush=True)
        assert err <= K.STEP_TOL
        g.state = U
        ref = U
        for k in range(1, 3):
            dt = o.stable_dt(ref, 0.2)
            assert abs(g.stable_time_step(0.2) / dt - 1.0) <= 1e-15
            g.step(dt)
            ref = o.step(ref, dt)
            assert K.rel_inf_per_variable(g.state, ref, 4) <= K.STEP_TOL * k


def test_edge_cases(gpu):
    mesh = RM.irregular_line(12)
    with pytest.raises(gpu.WaslHipError, match="inflow is not implemented"):
        _engine(gpu, mesh, 1, O.INFLOW)
    with pytest.raises(gpu.WaslHipError, match="paired"):
        gpu.Engine(gpu.Mesh.lines(*mesh), gpu.EULER, 1, boundary=(O.PERIODIC, O.OUTFLOW))
    with pytest.raises(gpu.WaslHipError):
        _engine(gpu, mesh, 1, precision=gpu.MIXED)
    # a single element with periodic ends is its own neighbour on both sides
    g = gpu.Engine(gpu.Mesh.lines([1, 2], [0.0, 1.0], [[1, 2]]), gpu.EULER, 2)
    o = O.line_euler([1, 2], [0.0, 1.0], [[1, 2]], 2)
    U = o.project(O.IC_EULER_REGIONS, K.SHOCK_TUBE)
    assert K.rel_inf_per_variable(g.residual(U), o.residual(U), 4) <= K.STEP_TOL


def test_full_size_properties(gpu):
    """65536 periodic cells, P2, limited shock tube: 200 steps conserve every
    variable's integral (density through the engine's mass) and keep every
    cell average physical."""
    mesh = _uniform_line(65536)
    g = gpu.Engine(gpu.Mesh.lines(*mesh), gpu.EULER, 2, limiting=True)
    g.project(O.IC_EULER_REGIONS, K.SHOCK_TUBE)
    g.limit()
    m0 = g.mass()
    t, s, _ = g.advance(0.2, 1e9, max_steps=200, stop_at_max_steps=True)
    assert s == 200
    assert abs(g.mass() - m0) <= 1e-12 * abs(m0)
    avg = g.cell_averages()
    assert avg[:, 0].min() > 0
    assert (0.4 * (avg[:, 3] - 0.5 * avg[:, 1] ** 2 / avg[:, 0])).min() > 0
    assert math.isfinite(t)
