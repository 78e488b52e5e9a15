"""Serving on the device: continuous batching of uneven cases.

  alone vs in the mix     bit-identical final states (a tile of the batched
                          kernels computes exactly the single-case block)
  server vs engine        bit-identical to Engine(precision=MIXED).advance
  few slots               every request completes with max_resident = 2
  failures stay local     a failing case does not disturb the others
"""

import numpy as np
import pytest

import cases as K
from wasl_amd import meshgen as MG

pytestmark = pytest.mark.gpu


def _requests(w):
    """A small mix: Euler P1-P3 of uneven sizes (vortex, smooth field), a
    limited radial rarefaction and 1D advection."""
    vm = {n: w.Mesh.from_input(K.vortex_mesh(n)) for n in (6, 9, 14)}
    rm = w.Mesh.from_input(K.rarefaction_mesh(10))
    lm = w.Mesh.from_input(MG.jittered_line(300))
    return [
        dict(mesh=vm[6], equation=w.EULER, order=1, ic=w.IC_VORTEX, params=K.VORTEX, end_steps=40),
        dict(mesh=vm[9], equation=w.EULER, order=2, ic=w.IC_VORTEX, params=K.VORTEX, end_steps=30),
        dict(mesh=vm[14], equation=w.EULER, order=3, ic=w.IC_VORTEX, params=K.VORTEX,
             end_steps=20),
        dict(mesh=vm[9], equation=w.EULER, order=2, ic=w.IC_SMOOTH, params=K.SMOOTH,
             end_steps=50),
        dict(mesh=rm, equation=w.EULER, order=2, ic=w.IC_RAREFACTION, params=K.RAREFACTION,
             end_steps=12, boundary=(w.OUTFLOW, w.OUTFLOW), limiting=True, cfl=0.1,
             density_floor=1e-13, pressure_floor=1e-13),
        dict(mesh=lm, equation=w.SCALAR_ADVECTION, order=2, ic=w.IC_SCALAR_PULSE, params=(),
             end_steps=60, cfl=0.1),
    ]


def _run(w, reqs, **server_args):
    from wasl_amd.serve import Server

    s = Server(**server_args)
    ids = [s.submit(**r) for r in reqs]
    out = s.drain()
    s.close()
    return [out[i] for i in ids]


def test_alone_equals_mix(gpu):
    reqs = _requests(gpu)
    mix = _run(gpu, reqs)
    for k, r in enumerate(reqs):
        alone = _run(gpu, [r])[0]
        for x in (alone, mix[k]):
            assert not isinstance(x, Exception), x
        assert alone[2] == mix[k][2] and alone[1] == mix[k][1]
        assert np.array_equal(alone[0], mix[k][0]), f"request {k} differs alone vs in the mix"
        assert np.isfinite(alone[0]).all()
        print(f"request {k}: {alone[2]} steps to t = {alone[1]:.6g}, identical alone and mixed")


def test_server_matches_single_engine(gpu):
    m = gpu.Mesh.from_input(K.vortex_mesh(12))
    g = gpu.Engine(m, gpu.EULER, 2, precision=gpu.MIXED)
    g.project(gpu.IC_VORTEX, K.VORTEX)
    t, s, _ = g.advance(0.15, 0.2)
    (state, ts, ss), = _run(gpu, [dict(mesh=m, equation=gpu.EULER, order=2, ic=gpu.IC_VORTEX,
                                       params=K.VORTEX, end_time=0.2)])
    assert (ts, ss) == (t, s)
    assert np.array_equal(state, g.state)


def test_few_slots_admit_continuously(gpu):
    reqs = _requests(gpu)
    full = _run(gpu, reqs)
    tight = _run(gpu, reqs, max_resident=2, steps_per_batch=4)
    for a, b in zip(full, tight):
        assert np.array_equal(a[0], b[0]) and a[1:] == b[1:]


def test_failure_stays_local(gpu):
    """An unlimited P1 rarefaction on transmissive boundaries loses a physical
    cell average (the FP64 reference fails the same way); the other cases of
    its batch finish unchanged."""
    reqs = _requests(gpu)
    bad = dict(reqs[4])
    bad.update(order=1, limiting=False, end_steps=400)
    out = _run(gpu, reqs + [bad])
    assert isinstance(out[-1], gpu.SimulationError)
    assert "not physical" in str(out[-1])
    ref = _run(gpu, reqs)
    for a, b in zip(out[:-1], ref):
        assert np.array_equal(a[0], b[0])
