"""End-to-end runs through the HIP engine against the reference's own stored
results: all 28 case-regression rows (relative 1e-10, the reference's
tolerance, case_regression_test.cpp:40-44), vectors written by the reference's
code (tests/golden/reference_vectors.npz) and the vortex convergence order."""

import math
import os

import numpy as np
import pytest

import cases as K
import refmeshes as RM
from conftest import GOLDEN
from test_oracle_pins import _vortex_error

pytestmark = pytest.mark.gpu


class _GpuCase:
    """Adapts the engine to the regression driver's oracle-shaped calls; a
    limited and an unlimited engine share the mesh."""

    def __init__(self, wasl, kind, mesh, p):
        if kind == "line":
            self.plain = wasl.Engine(wasl.Mesh.lines(*mesh), wasl.SCALAR_ADVECTION, p,
                                     wavespeed=0.7)
            self.limited = None
        elif kind == "line_euler":
            m = wasl.Mesh.lines(*mesh)
            self.plain = wasl.Engine(m, wasl.EULER, p)
            self.limited = wasl.Engine(m, wasl.EULER, p, limiting=True, sensor_threshold=1e-4)
        elif kind == "tri_scalar":
            self.plain = wasl.Engine(wasl.Mesh.triangles(*mesh), wasl.SCALAR_ADVECTION, p,
                                     wavespeed=0.7, velocity_y=-0.4)
            self.limited = None
        else:
            m = wasl.Mesh.triangles(*mesh)
            self.plain = wasl.Engine(m, wasl.EULER, p)
            self.limited = wasl.Engine(m, wasl.EULER, p, limiting=True, sensor_threshold=1e-4)
        self.E, self.Np = self.plain.E, self.plain.Np

    def project(self, ic, prm):
        return self.plain.project(ic, prm)

    def limit(self, U, threshold):
        self.limited.state = U
        flags = self.limited.limit()
        return self.limited.state, flags

    def run(self, U, cfl, end, max_steps, hist, limiting=False, threshold=1e-3, start_time=0.0,
            start_step=0):
        eng = self.limited if limiting else self.plain
        eng.state = U
        t, s, h = eng.advance(cfl, end, max_steps=max_steps, history_interval=hist,
                              time=start_time, step=start_step, history_capacity=512)
        return eng.state, s, t, h


def test_case_regression_rows_on_gpu(gpu):
    from test_oracle_pins import run_regression

    def factory(kind, mesh, p):
        return _GpuCase(gpu, kind, mesh, p)

    got = run_regression(factory)
    ref = RM.read_case_regression(os.path.join(GOLDEN, "case_regression_rows.txt"))
    assert set(got) == set(ref)
    for name, exp in sorted(ref.items()):
        act = got[name]
        rel = [abs(a - e) / max(1.0, abs(e)) for a, e in zip(act[1:], exp[1:])]
        print(f"{name}: steps {act[0]}/{exp[0]} max rel {max(rel):.2e}")
        assert act[0] == exp[0], name
        assert max(rel) <= 1e-10, (name, act, exp)


def test_reference_vectors_on_gpu(gpu):
    f = np.load(os.path.join(GOLDEN, "reference_vectors.npz"))
    m = K.vortex_mesh(8)
    mesh = gpu.Mesh.from_input(m)
    for p in range(4):
        g = gpu.Engine(mesh, gpu.EULER, p)
        U = f[f"vortex_p{p}_state"]
        assert K.rel_inf_per_variable(g.residual(U), f[f"vortex_p{p}_residual"], 4) <= 1e-12
        g.state = U
        assert abs(g.stable_time_step(0.15) / f[f"vortex_p{p}_dt"][0] - 1.0) <= 1e-15
    g = gpu.Engine(mesh, gpu.EULER, 2)
    traj = f["vortex_p2_steps"]
    g.state = traj[0]
    for k in range(1, traj.shape[0]):
        g.step(g.stable_time_step(0.15))
        assert K.rel_inf_per_variable(g.state, traj[k], 4) <= 1e-12
    r = gpu.Mesh.from_input(K.rarefaction_mesh(8))
    for p in (1, 2, 3):
        g = gpu.Engine(r, gpu.EULER, p, boundary=(gpu.OUTFLOW, gpu.OUTFLOW), limiting=True,
                       density_floor=1e-13, pressure_floor=1e-13)
        g.state = f[f"jump_p{p}_state"]
        flags = g.limit()
        assert np.array_equal(flags, f[f"jump_p{p}_flags"])
        assert K.rel_inf_per_variable(g.state, f[f"jump_p{p}_limited"], 4) <= 1e-12
    g = gpu.Engine(r, gpu.EULER, 2, boundary=(gpu.OUTFLOW, gpu.OUTFLOW), limiting=True,
                   density_floor=1e-13, pressure_floor=1e-13)
    traj = f["rarefaction_p2_limited_steps"]
    g.state = traj[0]
    for k in range(1, traj.shape[0]):
        g.step(g.stable_time_step(0.1))
        err = K.rel_inf_per_variable(g.state, traj[k], 4)
        print(f"rarefaction limited step {k}: rel L-inf {err:.3e}")
        assert err <= 1e-12
    line = gpu.Mesh.from_input(K.MG.jittered_line(200))
    g = gpu.Engine(line, gpu.SCALAR_ADVECTION, 2, wavespeed=1.0)
    traj = f["pulse_p2_steps"]
    g.state = traj[0]
    for k in range(1, traj.shape[0]):
        g.step(f["pulse_p2_dt"][0])
    assert K.rel_inf(g.state, traj[-1]) <= 1e-12


@pytest.mark.parametrize("p", [1, 2, 3])
def test_vortex_convergence_on_gpu(gpu, p):
    """Density L2 error of the translating vortex, jittered periodic meshes
    n = 8..32 (..64 for P1), rate >= p + 0.6 at the finest pair."""
    from oracle import oracle as O

    levels = (16, 32, 64) if p == 1 else (8, 16, 32)
    errs = []
    for n in levels:
        m = K.vortex_mesh(n)
        geometry = K.oracle_tri(m, p)  # element maps for the error integral
        g = gpu.Engine(gpu.Mesh.from_input(m), gpu.EULER, p)
        g.project(gpu.IC_VORTEX, K.VORTEX)
        t, _, _ = g.advance(0.15, 0.5)
        errs.append(_vortex_error(geometry, g.state, t, p))
    rates = [math.log2(errs[i] / errs[i + 1]) for i in range(len(errs) - 1)]
    print(f"P{p} errors {['%.3e' % e for e in errs]} rates {['%.2f' % r for r in rates]}")
    assert rates[-1] >= p + 0.6
