"""Serving on the device: continuous batching of uneven cases.

  alone vs in the mix     bit-identical final states (a tile of the batched
                          kernels computes exactly the single-case block)
  server vs engine        bit-identical to Engine(precision=MIXED).advance, and
                          for the FP64 classes (P1, limited cases, or every
                          order with fp64_max_order=3) to the FP64 engine
                          with its autotuned tile
  few slots               every request completes with max_resident = 2
  failures stay local     a failing case does not disturb the others
"""

import numpy as np
import pytest

import cases as K
from wasl_amd import meshgen as MG

pytestmark = pytest.mark.gpu


def _requests(w):
    """A small mix: Euler P1-P3 of uneven sizes (vortex, smooth field),
    a limited radial rarefaction, limited low-density discs and 1D advection. Every launch class holds
    at least two cases, so each case's first tile (not the launch's) must
    lead its step control and clear its troubled-element count."""
    vm = {n: w.Mesh.from_input(K.vortex_mesh(n)) for n in (6, 9, 14)}
    rm = w.Mesh.from_input(K.rarefaction_mesh(10))
    rm8 = w.Mesh.from_input(K.rarefaction_mesh(8))
    lm = w.Mesh.from_input(MG.jittered_line(300))
    lm2 = w.Mesh.from_input(MG.jittered_line(170))
    # limited discs of low density and pressure (troubled from the start; a
    # P1 rarefaction on these coarse meshes loses a physical cell average in
    # its first step, in the reference as well)
    lim = dict(equation=w.EULER, ic=w.IC_EULER_REGIONS, params=K.JUMP,
               boundary=(w.OUTFLOW, w.OUTFLOW), limiting=True, cfl=0.1, density_floor=1e-13,
               pressure_floor=1e-13)
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
        dict(mesh=vm[9], equation=w.EULER, order=1, ic=w.IC_VORTEX, params=K.VORTEX,
             end_steps=25),
        dict(mesh=vm[6], equation=w.EULER, order=3, ic=w.IC_SMOOTH, params=K.SMOOTH,
             end_steps=15),
        dict(mesh=rm8, order=2, end_steps=9, **lim),
        dict(mesh=rm, order=1, end_steps=14, **lim),
        dict(mesh=rm8, order=1, end_steps=10, **lim),
        dict(mesh=lm2, equation=w.SCALAR_ADVECTION, order=2, ic=w.IC_SCALAR_PULSE, params=(),
             end_steps=35, cfl=0.1),
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


@pytest.mark.parametrize("order,limited,fp64_max_order", [
    (1, False, 1), (2, True, 1), (3, True, 1), (2, False, 3), (3, False, 3)])
def test_fp64_classes_match_single_engine(gpu, order, limited, fp64_max_order):
    if limited:
        m = gpu.Mesh.from_input(K.rarefaction_mesh(10))
        kw = dict(boundary=(gpu.OUTFLOW, gpu.OUTFLOW), limiting=True, density_floor=1e-13,
                  pressure_floor=1e-13)
        # a few steps: at P3 this coarse rarefaction loses a physical cell
        # average by t = 0.02 in FP64 as well
        ic, prm, cfl, end = gpu.IC_RAREFACTION, K.RAREFACTION, 0.1, 0.004
    else:
        m = gpu.Mesh.from_input(K.vortex_mesh(12))
        kw = {}
        ic, prm, cfl, end = gpu.IC_VORTEX, K.VORTEX, 0.15, 0.2
    # unlimited FP64 classes run the server's tile for the order, limited
    # ones the engine's autotuned tile
    from wasl_amd.serve import default_fp64_tiles
    kernel, lanes = (0, 0) if limited else default_fp64_tiles()[order]
    g = gpu.Engine(m, gpu.EULER, order, precision=gpu.FP64, wavespeed=1.0, stage_kernel=kernel,
                   lanes=lanes, **kw)
    g.project(ic, prm)
    if limited:
        g.limit()
    t, s, _ = g.advance(cfl, end)
    (state, ts, ss), = _run(gpu, [dict(mesh=m, equation=gpu.EULER, order=order, ic=ic,
                                       params=prm, end_time=end, cfl=cfl, **kw)],
                            fp64_max_order=fp64_max_order)
    assert (ts, ss) == (t, s)
    assert np.array_equal(state, g.state)


@pytest.mark.parametrize("order", [1, 2, 3])
def test_served_limited_case_is_the_case_composition(gpu, order):
    """A served limited case follows the reference's case driver: project,
    limit the projected state, then step (euler_triangle.cpp:23-47). On the
    disc of low density and pressure (cases.JUMP) the initial limit decides
    the run: the served result equals the engine's project -> limit ->
    advance bit for bit, while project -> advance loses a physical cell
    average after its first step, on the device as in the oracle (the
    projection's negative face pressures give NaN fluxes)."""
    from oracle import oracle as O

    mesh_in = K.rarefaction_mesh(10)
    m = gpu.Mesh.from_input(mesh_in)
    kw = dict(boundary=(gpu.OUTFLOW, gpu.OUTFLOW), limiting=True, density_floor=1e-13,
              pressure_floor=1e-13)
    cfl, end = 0.1, 0.01
    g = gpu.Engine(m, gpu.EULER, order, **kw)
    g.project(gpu.IC_EULER_REGIONS, K.JUMP)
    assert g.limit().sum() > 0
    t, s, _ = g.advance(cfl, end)
    (state, ts, ss), = _run(gpu, [dict(mesh=m, equation=gpu.EULER, order=order,
                                       ic=gpu.IC_EULER_REGIONS, params=K.JUMP, end_time=end,
                                       cfl=cfl, **kw)])
    assert (ts, ss) == (t, s)
    assert np.array_equal(state, g.state)

    o = K.oracle_tri(mesh_in, order, boundary=O.OUTFLOW)
    with pytest.raises(O.OracleError, match="Euler cell average is not physical"):
        o.run(o.project(O.IC_EULER_REGIONS, K.JUMP), cfl, end, limiting=True,
              density_floor=1e-13, pressure_floor=1e-13)
    h = gpu.Engine(m, gpu.EULER, order, **kw)
    h.project(gpu.IC_EULER_REGIONS, K.JUMP)
    with pytest.raises(gpu.SimulationError, match="^simulation: Euler cell average is not physical$"):
        h.advance(cfl, end)


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


def test_unserved_kinds_are_refused(gpu):
    """Serving takes the workload mix's two kinds (2D Euler, 1D advection);
    1D Euler and 2D advection cases are refused at submission, with a message,
    and run on their own engines."""
    from wasl_amd.serve import Server

    s = Server()
    lines = gpu.Mesh.from_input(MG.jittered_line(50))
    tris = gpu.Mesh.from_input(K.vortex_mesh(4))
    with pytest.raises(gpu.WaslHipError, match="serving takes"):
        s.submit(lines, gpu.EULER, 1, gpu.IC_EULER_REGIONS, K.SHOCK_TUBE, end_steps=5)
    with pytest.raises(gpu.WaslHipError, match="serving takes"):
        s.submit(tris, gpu.SCALAR_ADVECTION, 1, gpu.IC_SCALAR_WAVE, K.SCALAR_WAVE, end_steps=5)
    ok = s.submit(lines, gpu.SCALAR_ADVECTION, 1, gpu.IC_SCALAR_PULSE, (), end_steps=5, cfl=0.1)
    assert s.drain()[ok][2] == 5
    s.close()
