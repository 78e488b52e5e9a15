"""Pins of the Zhang-Shu positivity limiter on inputs where it acts.

The inputs (tests/cases.py positivity_state) make most troubled cells reach
limit_euler_positivity (euler_positivity.hpp:63-102) with a check-point
density or pressure below the floor. The reference's own outputs on them
(oracle/_ref through tests/golden/make_fixtures.py) are committed in
reference_vectors.npz. These CPU tests check:
  - the restatement reproduces the reference's vectors bit for bit, and the
    per-element recomposition from the reference's single routines equals
    the case-level limiter (flags, state, every element);
  - the limiter actually acted: density passes, pressure bisections and
    untouched troubled cells all occur;
  - what the mathematics fixes, independent of the oracle's formulas: mode 0
    is never changed, untroubled cells are never changed, every check point
    respects both floors after limiting, and where the limiter scaled, the
    binding constraint sits on its floor (the scale is the largest admissible
    one, so a wrong theta formula fails either way);
  - the floor error: the reference throws at the first element (ascending)
    whose cell average is below a floor, leaving that element and every later
    one as they were.
"""

import os

import numpy as np
import pytest

import cases as K
from conftest import GOLDEN
from oracle import oracle as O

F = K.POSITIVITY_FLOORS
COMBOS = [(tag, bnd, p) for tag, bnd in (("transmissive", O.OUTFLOW), ("periodic", O.PERIODIC))
          for p in (1, 2, 3)]


def _mesh(tag):
    return K.rarefaction_mesh(8) if tag == "transmissive" else K.vortex_mesh(8)


@pytest.fixture(scope="module")
def vectors():
    return np.load(os.path.join(GOLDEN, "reference_vectors.npz"))


def _pressure(s, g=K.GAMMA):
    return (g - 1.0) * (s[3] - 0.5 * (s[1] * s[1] + s[2] * s[2]) / s[0])


@pytest.mark.parametrize("tag,bnd,p", COMBOS)
def test_positivity_limiter_acts_and_matches_reference(vectors, tag, bnd, p):
    c = K.oracle_tri(_mesh(tag), p, boundary=bnd)
    key = f"positivity_{tag}_p{p}"
    U = K.positivity_state(c, p, 40 + p)
    assert np.array_equal(U, vectors[key + "_state"])
    flags, dens, scale, Ul = K.positivity_scales(c, U, 1e-3, F, F, O.Routines("port"))
    assert np.array_equal(flags, vectors[key + "_flags"])
    assert np.array_equal(Ul, vectors[key + "_limited"])
    troubled = flags.astype(bool)
    scaled = scale < 1.0
    print(f"{key}: {troubled.sum()} troubled of {c.E}, {scaled.sum()} scaled, "
          f"{dens.sum()} with the density pass, min scale {scale.min():.3e}")
    # the limiter acted, through both passes, and left some troubled cells alone
    assert dens.sum() >= 10
    assert (scaled & ~dens).sum() >= 10
    assert troubled.sum() > scaled.sum()

    Np, E = c.Np, c.E
    cb = c.stencil()[3]
    U3, L3 = U.reshape(4, Np, E), Ul.reshape(4, Np, E)
    assert np.array_equal(L3[:, 0, :], U3[:, 0, :])           # means kept exactly
    assert np.array_equal(L3[:, :, ~troubled], U3[:, :, ~troubled])
    for e in np.nonzero(troubled)[0]:
        pts = np.einsum("qm,vm->vq", cb, L3[:, :, e])
        rho_min, p_min = pts[0].min(), _pressure(pts).min()
        assert rho_min >= F * (1 - 1e-9) - 1e-16, (e, rho_min)
        assert p_min >= F - 1e-15, (e, p_min)
        if scaled[e]:
            # the largest admissible scale puts the binding check point on its
            # floor: density when only the density pass scaled, else pressure
            binding = min(rho_min - F, p_min - F)
            assert abs(binding) <= 1e-15, (e, rho_min, p_min)


@pytest.mark.parametrize("tag,bnd,p", COMBOS)
def test_positivity_limited_steps_match_reference(vectors, tag, bnd, p):
    """Two limited SSP-RK3 steps from the limited state with every cell
    troubled (threshold 1e-10): the positivity limiter acts inside the
    stages, and the restatement follows the reference bit for bit."""
    c = K.oracle_tri(_mesh(tag), p, boundary=bnd)
    traj = vectors[f"positivity_{tag}_p{p}_steps"]
    U0, _ = c.limit(vectors[f"positivity_{tag}_p{p}_state"], 1e-10, F, F)
    assert np.array_equal(U0, traj[0])
    U1 = U0 + c.stable_dt(U0, 0.1) * c.residual(U0)
    _, _, scale, _ = K.positivity_scales(c, U1, 1e-10, F, F, O.Routines("port"))
    assert (scale < 1.0).sum() >= 1, "the positivity limiter does not act in stage 1"
    U = U0
    for k in range(1, traj.shape[0]):
        U = c.step(U, c.stable_dt(U, 0.1), True, 1e-10, F, F)
        assert np.array_equal(U, traj[k])


def test_floor_error_partial_state(vectors):
    from golden.make_fixtures import FLOOR_ERROR_ELEMENTS, floor_error_state

    c = K.oracle_tri(_mesh("transmissive"), 2, boundary=O.OUTFLOW)
    U = floor_error_state(c)
    assert np.array_equal(U, vectors["floor_error_p2_state"])
    part, _, msg = c.limit(U, 1e-3, F, F, partial=True)
    assert msg == "Euler positivity limiter: cell average is below a floor"
    assert np.array_equal(part, vectors["floor_error_p2_partial"])
    first = FLOOR_ERROR_ELEMENTS[0]
    full, _ = c.limit(K.positivity_state(c, 2, 42), 1e-3, F, F)
    P3, U3, L3 = part.reshape(4, 6, -1), U.reshape(4, 6, -1), full.reshape(4, 6, -1)
    assert np.array_equal(P3[:, :, first:], U3[:, :, first:])
    # before the failing element, the limited values (candidates read the
    # unlimited input, so elements away from the two changed ones match)
    nb = c.stencil()[0]
    changed = set(FLOOR_ERROR_ELEMENTS)
    for e in range(first):
        if changed.isdisjoint(nb[e].tolist()):
            assert np.array_equal(P3[:, :, e], L3[:, :, e]), e


@pytest.mark.parametrize("p", [1, 2, 3])
def test_positivity_routine_bitwise_equals_reference_build(reference_oracle, p):
    """limit_euler_positivity alone on engaged inputs: the restatement and
    the reference build agree bit for bit, scale included."""
    a, b = O.Routines("port"), O.Routines("reference")
    c = K.oracle_tri(_mesh("periodic"), p)
    cb = c.stencil()[3]
    U = K.positivity_state(c, p, 70 + p).reshape(4, c.Np, c.E)
    acted = 0
    for e in range(c.E):
        poly = U[:, :, e]
        oa, sa = a.positivity(K.GAMMA, poly, c.Np, cb, F, F)
        ob, sb = b.positivity(K.GAMMA, poly, c.Np, cb, F, F)
        assert np.array_equal(oa, ob, equal_nan=True) and (sa == sb or (sa != sa and sb != sb))
        acted += sa < 1.0
    assert acted >= 20


def test_floor_error_bitwise_equals_reference_build(reference_oracle, vectors):
    from golden.make_fixtures import floor_error_state

    mesh = _mesh("transmissive")
    a = K.oracle_tri(mesh, 2, boundary=O.OUTFLOW)
    b = K.oracle_tri(mesh, 2, boundary=O.OUTFLOW, kind="reference")
    U = floor_error_state(a)
    pa, _, ma = a.limit(U, 1e-3, F, F, partial=True)
    pb, _, mb = b.limit(U, 1e-3, F, F, partial=True)
    assert ma == mb and np.array_equal(pa, pb)
