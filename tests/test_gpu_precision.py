"""The precision form on the device: the mixed kernel (FP32 contractions on
the matrix cores, FP32 accumulation, FP64 state, RK combination, sensor, CFL
speed and limiter) and its BF16 ablation, against the FP64 oracle.

Tolerances (written here, they are the contract of the precision form):
  residual, MIXED   max-norm-scaled L-inf per variable <= 2e-5
  residual, BF16    max-norm-scaled L-inf per variable <= 0.25 (an ablation:
                    8-bit mantissas in the volume products; the value is printed)
  runs, MIXED       density L-inf <= 1e-5 (absolute; densities are O(1))
The L1 and L-inf errors of every run are printed.
"""

import numpy as np
import pytest

import cases as K
from oracle import oracle as O

pytestmark = pytest.mark.gpu

MIXED_RESIDUAL_TOL = 2e-5
BF16_RESIDUAL_TOL = 0.25
DENSITY_TOL = 1e-5


def _errors(a, b, E, Np):
    """L-inf and element-mean L1 of the density coefficients."""
    d = np.abs(np.asarray(a).reshape(4, Np, E)[0] - np.asarray(b).reshape(4, Np, E)[0])
    return float(d.max()), float(d.sum() / E)


@pytest.mark.parametrize("order", [0, 1, 2, 3])
def test_mixed_residual_periodic(gpu, order):
    m = K.vortex_mesh(10)
    o = K.oracle_tri(m, order)
    U = o.project(O.IC_VORTEX, K.VORTEX)
    rng = np.random.default_rng(order)
    U = U * (1.0 + 1e-3 * rng.standard_normal(U.size))
    ref = o.residual(U)
    for prec, tol in ((gpu.MIXED, MIXED_RESIDUAL_TOL), (gpu.BF16, BF16_RESIDUAL_TOL)):
        g = K.engine_tri(gpu, m, order, precision=prec)
        err = K.rel_inf_per_variable(g.residual(U), ref, 4)
        print(f"P{order} precision {prec} residual rel L-inf {err:.3e}")
        assert err <= tol


@pytest.mark.parametrize("order", [1, 2, 3])
def test_mixed_residual_periodic_donors(gpu, order):
    """The matrix-core stage on a mesh whose x boundaries keep per-point
    periodic donors (the y boundaries are paired as interior faces)."""
    m = K.nonconforming_vortex_mesh(10)
    o = K.oracle_tri(m, order)
    U = o.project(O.IC_VORTEX, K.VORTEX)
    g = K.engine_tri(gpu, m, order, precision=gpu.MIXED)
    assert g.boundary_faces == 20
    err = K.rel_inf_per_variable(g.residual(U), o.residual(U), 4)
    print(f"P{order} donor-path mixed residual rel L-inf {err:.3e}")
    assert err <= MIXED_RESIDUAL_TOL


@pytest.mark.parametrize("order", [2, 3])
def test_mixed_residual_transmissive(gpu, order):
    # (the unlimited P1 projection of the rarefaction is unphysical at volume
    # points, in the oracle as well; P1 is covered by the periodic test)
    m = K.rarefaction_mesh(10)
    o = K.oracle_tri(m, order, boundary=O.OUTFLOW)
    # the radial rarefaction: outflow through every boundary face, O(1) residuals
    U = o.project(O.IC_RAREFACTION, K.RAREFACTION)
    g = K.engine_tri(gpu, m, order, boundary=gpu.OUTFLOW, precision=gpu.MIXED)
    err = K.rel_inf_per_variable(g.residual(U), o.residual(U), 4)
    print(f"P{order} transmissive mixed residual rel L-inf {err:.3e}")
    assert err <= MIXED_RESIDUAL_TOL


@pytest.mark.parametrize("order", [1, 2, 3])
def test_mixed_vortex_run(gpu, order):
    """Isentropic vortex to t = 0.5 on a 2x16x16 jittered periodic mesh."""
    m = K.vortex_mesh(16)
    o = K.oracle_tri(m, order)
    U0 = o.project(O.IC_VORTEX, K.VORTEX)
    Uo, so, to, _ = o.run(U0, 0.15, 0.5, 100000, 100000)
    g = K.engine_tri(gpu, m, order, precision=gpu.MIXED)
    g.state = U0
    tg, sg, h = g.advance(0.15, 0.5, max_steps=100000, history_interval=100000,
                          history_capacity=4)
    linf, l1 = _errors(g.state, Uo, g.E, g.Np)
    drift = abs(h[-1, 3] - h[0, 3]) / abs(h[0, 3])
    print(f"P{order} mixed vortex: steps {sg}/{so} density L-inf {linf:.3e} L1 {l1:.3e} "
          f"mass drift {drift:.2e}")
    assert tg == to and abs(int(sg) - int(so)) <= 1
    assert linf <= DENSITY_TOL
    assert drift <= 1e-6


def test_mixed_limited_rarefaction(gpu):
    """Radial double rarefaction (transmissive; characteristic WENO and
    positivity limiting in FP64 on the mixed state) on 2x12x12, P2: ten steps
    with the oracle's dt sequence."""
    m = K.rarefaction_mesh(12)
    o = K.oracle_tri(m, 2, boundary=O.OUTFLOW)
    U = o.project(O.IC_RAREFACTION, K.RAREFACTION)
    U, _ = o.limit(U, 1e-3, 1e-13, 1e-13)
    g = K.engine_tri(gpu, m, 2, boundary=gpu.OUTFLOW, limiting=True, threshold=1e-3,
                     floors=1e-13, precision=gpu.MIXED)
    g.state = U
    for k in range(10):
        dt = o.stable_dt(U, 0.1)
        U = o.step(U, dt, True, 1e-3, 1e-13, 1e-13)
        g.step(dt)
    Ug = g.state
    linf, l1 = _errors(Ug, U, g.E, g.Np)
    print(f"mixed rarefaction, 10 steps: density L-inf {linf:.3e} L1 {l1:.3e}")
    assert np.isfinite(Ug).all()
    assert linf <= DENSITY_TOL
