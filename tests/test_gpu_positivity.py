"""GPU parity of the limiter where the Zhang-Shu positivity limiter acts.

The inputs are tests/cases.py positivity_state; the expected outputs are the
reference's own (oracle/_ref, committed in reference_vectors.npz), and the
oracle's recomposition of each call from the reference's single routines
confirms that density passes and pressure bisections (theta < 1) occur on
the very inputs the device limits:
  troubled flags   bit-exact
  limited state    relative L-inf <= 1e-12 per variable
  limited steps    relative L-inf <= 1e-12 per step, or 4x the oracle's own
                   change under one-ulp input noise where the positivity-limited
                   state makes the step ill conditioned (see the test)
  floor error      the reference's message, and its partial state: elements
                   before the first failing one limited, the rest unchanged
"""

import os

import numpy as np
import pytest

import cases as K
from conftest import GOLDEN
from oracle import oracle as O

pytestmark = pytest.mark.gpu

F = K.POSITIVITY_FLOORS
COMBOS = [(tag, bnd, p) for tag, bnd in (("transmissive", O.OUTFLOW), ("periodic", O.PERIODIC))
          for p in (1, 2, 3)]


def _mesh(tag):
    return K.rarefaction_mesh(8) if tag == "transmissive" else K.vortex_mesh(8)


@pytest.fixture(scope="module")
def vectors():
    return np.load(os.path.join(GOLDEN, "reference_vectors.npz"))


def _engine(gpu, tag, p, threshold):
    bnd = gpu.OUTFLOW if tag == "transmissive" else gpu.PERIODIC
    return K.engine_tri(gpu, _mesh(tag), p, boundary=bnd, limiting=True, threshold=threshold,
                        floors=F)


@pytest.mark.parametrize("tag,bnd,p", COMBOS)
def test_positivity_limiter_parity(gpu, vectors, tag, bnd, p):
    key = f"positivity_{tag}_p{p}"
    U = vectors[key + "_state"]
    c = K.oracle_tri(_mesh(tag), p, boundary=bnd)
    _, dens, scale, _ = K.positivity_scales(c, U, 1e-3, F, F, O.Routines("port"))
    print(f"{key}: oracle scaled {(scale < 1).sum()} cells ({dens.sum()} by the density pass)",
          flush=True)
    assert dens.sum() >= 10 and ((scale < 1) & ~dens).sum() >= 10
    g = _engine(gpu, tag, p, 1e-3)
    g.state = U
    flags = g.limit()
    assert np.array_equal(flags, vectors[key + "_flags"])
    err = K.rel_inf_per_variable(g.state, vectors[key + "_limited"], 4)
    print(f"{key}: limited state rel L-inf {err:.3e}")
    assert err <= K.STEP_TOL


def _ulp_sensitivity(c, U, dt, seeds=4):
    """How far the oracle's own limited step moves when its input is
    perturbed by about one ulp (relative 2.2e-16 noise): the conditioning of
    the step at U."""
    F = K.POSITIVITY_FLOORS
    base = c.step(U, dt, True, 1e-10, F, F)
    worst = 0.0
    for seed in range(seeds):
        rng = np.random.default_rng(seed)
        Up = U * (1.0 + 2.2e-16 * rng.standard_normal(U.size))
        worst = max(worst, K.rel_inf_per_variable(c.step(Up, dt, True, 1e-10, F, F), base, 4))
    return worst


@pytest.mark.parametrize("tag,bnd,p", COMBOS)
def test_positivity_limited_steps(gpu, vectors, tag, bnd, p):
    """Limited steps with every cell troubled, from the positivity-limited
    state. Its check points sit on the 1e-13 floors, and the face points are
    check points, so HLLC sees sound speeds near 1e-6: the step is ill
    conditioned, and the oracle itself moves by up to ~2e-11 when its input
    is perturbed by one ulp. The bar is the 1e-12 per-step tolerance or four
    times that measured one-ulp sensitivity, whichever is larger."""
    key = f"positivity_{tag}_p{p}"
    traj = vectors[key + "_steps"]
    c = K.oracle_tri(_mesh(tag), p, boundary=bnd)
    g = _engine(gpu, tag, p, 1e-10)
    g.state = vectors[key + "_state"]
    g.limit()
    err = K.rel_inf_per_variable(g.state, traj[0], 4)
    assert err <= K.STEP_TOL
    g.state = traj[0]
    for k in range(1, traj.shape[0]):
        dt = c.stable_dt(traj[k - 1], 0.1)
        g.step(dt)
        err = K.rel_inf_per_variable(g.state, traj[k], 4)
        sens = _ulp_sensitivity(c, traj[k - 1], dt)
        bound = max(K.STEP_TOL, 4.0 * sens)
        print(f"{key}: limited step {k} rel L-inf {err:.3e} (oracle one-ulp sensitivity "
              f"{sens:.3e}, bound {bound:.3e})", flush=True)
        assert err <= bound
        g.state = traj[k]   # each step from the reference's own state


def test_positivity_floor_error(gpu, vectors):
    from golden.make_fixtures import FLOOR_ERROR_ELEMENTS

    g = _engine(gpu, "transmissive", 2, 1e-3)
    g.state = vectors["floor_error_p2_state"]
    with pytest.raises(gpu.SimulationError,
                       match="^Euler positivity limiter: cell average is below a floor$"):
        g.limit()
    expected = vectors["floor_error_p2_partial"]
    err = K.rel_inf_per_variable(g.state, expected, 4)
    print(f"floor error: partial state rel L-inf {err:.3e}")
    assert err <= K.STEP_TOL
    first = FLOOR_ERROR_ELEMENTS[0]
    S = g.state.reshape(4, 6, -1)
    assert np.array_equal(S[:, :, first:],
                          vectors["floor_error_p2_state"].reshape(4, 6, -1)[:, :, first:])
    # the engine is usable afterwards: the error word was cleared
    g.state = vectors["positivity_transmissive_p2_state"]
    g.limit()
    assert K.rel_inf_per_variable(g.state, vectors["positivity_transmissive_p2_limited"],
                                  4) <= K.STEP_TOL

