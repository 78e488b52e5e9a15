"""2D scalar advection on periodic triangles through the HIP engine
(dg/scalar_advection/triangle_residual.hpp:25-105,
src/case/scalar_advection_triangle.cpp:15-46) against the oracle and the
vectors the reference's own code wrote (tests/golden/reference_vectors.npz).

Tolerances: residual and single step relative L-inf <= 1e-12 (max-norm
scaled); end of run <= 1e-10; the CFL step to the last bit but one."""

import math
import os

import numpy as np
import pytest

import cases as K
from conftest import GOLDEN
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _engine(gpu, m, p, velocity=K.ADVECTION):
    return gpu.Engine(gpu.Mesh.from_input(m), gpu.SCALAR_ADVECTION, p, wavespeed=velocity[0],
                      velocity_y=velocity[1])


@pytest.mark.parametrize("p", [0, 1, 2, 3])
def test_residual_matches_oracle(gpu, p):
    m = K.vortex_mesh(9)
    g, o = _engine(gpu, m, p), K.oracle_tri_scalar(m, p)
    assert (g.V, g.Np, g.E) == (1, (p + 1) * (p + 2) // 2, o.E)
    rng = np.random.default_rng(40 + p)
    for ic, prm in ((O.IC_SCALAR_WAVE, K.SCALAR_WAVE), (O.IC_SCALAR_CIRCLES, K.SCALAR_DISC)):
        U = o.project(ic, prm) + 1e-3 * rng.standard_normal(o.state_size)
        err = K.rel_inf(g.residual(U), o.residual(U))
        print(f"P{p} ic {ic}: residual rel {err:.2e}", flush=True)
        assert err <= K.STEP_TOL


@pytest.mark.parametrize("p", [0, 1, 2, 3])
def test_residual_periodic_donors(gpu, p):
    """Per-point periodic donors on the x boundaries (shifted x = 10 nodes)."""
    m = K.nonconforming_vortex_mesh(9)
    g, o = _engine(gpu, m, p), K.oracle_tri_scalar(m, p)
    assert g.boundary_faces == 18
    rng = np.random.default_rng(50 + p)
    U = o.project(O.IC_SCALAR_WAVE, K.SCALAR_WAVE) + 1e-3 * rng.standard_normal(o.state_size)
    err = K.rel_inf(g.residual(U), o.residual(U))
    print(f"P{p} donor-path residual rel {err:.2e}", flush=True)
    assert err <= K.STEP_TOL


@pytest.mark.parametrize("p", [0, 1, 2, 3])
def test_reference_vectors(gpu, p):
    f = np.load(os.path.join(GOLDEN, "reference_vectors.npz"))
    g = _engine(gpu, K.vortex_mesh(8), p)
    U = f[f"advect2d_p{p}_state"]
    assert K.rel_inf(g.residual(U), f[f"advect2d_p{p}_residual"]) <= K.STEP_TOL
    g.state = U
    assert abs(g.stable_time_step(0.3) / f[f"advect2d_p{p}_dt"][0] - 1.0) <= 1e-15


def test_reference_trajectory(gpu):
    f = np.load(os.path.join(GOLDEN, "reference_vectors.npz"))
    traj = f["advect2d_p2_steps"]
    g = _engine(gpu, K.vortex_mesh(8), 2)
    assert np.array_equal(g.project(O.IC_SCALAR_CIRCLES, K.SCALAR_DISC), traj[0])
    dt = g.stable_time_step(0.3)
    for k in range(1, traj.shape[0]):
        g.step(dt)
        err = K.rel_inf(g.state, traj[k])
        print(f"step {k}: rel {err:.2e}", flush=True)
        assert err <= K.STEP_TOL * k


@pytest.mark.parametrize("p", [1, 3])
def test_advance_matches_oracle_run(gpu, p):
    """The device-driven loop (constant CFL step, last step clipped to the end
    time) against the oracle's runner loop."""
    m = K.vortex_mesh(10)
    g, o = _engine(gpu, m, p), K.oracle_tri_scalar(m, p)
    U0 = o.project(O.IC_SCALAR_WAVE, K.SCALAR_WAVE)
    Uo, so, to, ho = o.run(U0, 0.3, 0.9, history_interval=5)
    g.state = U0
    t, s, h = g.advance(0.3, 0.9, history_interval=5, history_capacity=64)
    assert s == so and abs(t - to) <= 1e-14 * to
    assert np.array_equal(h[:, 0], ho[:, 0])
    assert np.abs(h[:, 1:3] - ho[:, 1:3]).max() <= 1e-14
    # the wave's mass is zero: compare on the scale of area x amplitude (100)
    assert np.abs(h[:, 3] - ho[:, 3]).max() <= 1e-12 * 100.0
    err = K.rel_inf(g.state, Uo)
    print(f"P{p}: {s} steps, end rel {err:.2e}", flush=True)
    assert err <= K.RUN_TOL


def test_edge_cases(gpu):
    """Zero velocity (residual 0, dt 0 like the reference), a velocity along
    an axis (normal speed 0 on half the faces: the upwind branch of v.n = 0),
    and an odd element count from the irregular box."""
    import refmeshes as RM

    m = K.vortex_mesh(6)
    g0 = _engine(gpu, m, 2, (0.0, 0.0))
    U = g0.project(O.IC_SCALAR_WAVE, K.SCALAR_WAVE)
    assert np.abs(g0.residual(U)).max() == 0.0
    assert g0.stable_time_step(0.3) == 0.0
    g1, o1 = _engine(gpu, m, 2, (1.0, 0.0)), O.tri_scalar(m.node_tags, m.node_xy, m.triangles,
                                                           2, 1.0, 0.0)
    assert K.rel_inf(g1.residual(U), o1.residual(U)) <= K.STEP_TOL
    bt, bxy, btri = RM.irregular_box(5)
    gb = gpu.Engine(gpu.Mesh.triangles(bt, bxy, btri), gpu.SCALAR_ADVECTION, 3, wavespeed=0.7,
                    velocity_y=-0.4)
    ob = O.tri_scalar(bt, bxy, btri, 3, 0.7, -0.4)
    Ub = ob.project(O.IC_SCALAR_CIRCLES, [0.25, 0.5, 0.4, 0.25, 2.0])
    assert K.rel_inf(gb.residual(Ub), ob.residual(Ub)) <= K.STEP_TOL
    with pytest.raises(gpu.WaslHipError):
        gpu.Engine(gpu.Mesh.from_input(m), gpu.SCALAR_ADVECTION, 1, boundary=(2, 2))
    with pytest.raises(gpu.WaslHipError):
        gpu.Engine(gpu.Mesh.from_input(m), gpu.SCALAR_ADVECTION, 1, limiting=True)
    with pytest.raises(gpu.WaslHipError):
        g1.project(O.IC_VORTEX, K.VORTEX)


def test_full_size_properties(gpu):
    """2 x 256 x 256 triangles, P2: 100 steps keep the mass to round-off and
    a constant state stays constant."""
    from wasl_amd import meshgen as MG

    m = MG.jittered_triangles(256, (0.0, 0.0), (10.0, 10.0), 0.15, MG.DEFAULT_SEED)
    g = _engine(gpu, m, 2)
    U = g.project(O.IC_SCALAR_CIRCLES, K.SCALAR_DISC)
    m0 = g.mass()
    t, s, _ = g.advance(0.3, 1e9, max_steps=100, stop_at_max_steps=True)
    assert s == 100
    assert abs(g.mass() - m0) <= 1e-12 * abs(m0)
    c = np.zeros_like(U)
    c[::6] = 1.7 * math.sqrt(2.0)
    g.state = c
    g.advance(0.3, 1e9, max_steps=20, stop_at_max_steps=True)
    assert np.abs(g.state - c).max() <= 1e-12
