"""The five workload configurations at their stated sizes on the device.

  config 1  1D advection, P2, 200 jittered cells, CFL 0.1 to T = 1: the device
            run against the oracle's runner (1e-10) and the L2 error against
            the exact shifted solution (7-point Gauss, ssp_rk3_test.cpp:76-99),
            with the P2 convergence rate over 100-200-400 cells;
  config 2  2D Euler vortex, P2, 2x128x128 = 32768 triangles: three steps of
            per-step parity (1e-12) and a 20-step device loop against the
            oracle's runner (1e-10 at the end, history rows);
  config 3  limited radial rarefaction, P2, 2x256x256 = 131072 triangles,
            floors 1e-13: it loses a physical cell average in its 22nd step,
            in the reference too (NaN face fluxes next to the origin); the
            device stops with the reference's message, and before that the
            limiter and limited steps match the oracle on the full mesh
            (flags bit-exact, state within 1e-12);
  config 4  P3 vortex on 2x512x512 = 524288 triangles: 40 steps mixed against
            the FP64 engine (density L-inf <= 1e-5), finite states, mass
            conserved, and free-stream preservation of both stages.
Config 5 (the serving mix) is measured by bench.py; its parity is in
test_gpu_serve.py.
"""

import math
import time

import numpy as np
import pytest

import cases as K
from oracle import oracle as O
from wasl_amd import meshgen as MG

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)


# ------------------------------------------------------------- config 1 ----

def _pulse(x):
    return np.sin(2.0 * np.pi * x) + np.exp(-(((x - 0.5) / 0.05) ** 2))


def _line_l2_error(c, U, t, p):
    """L2 error of the AoS modal line state against u0(x - t) on periodic
    [0,1] with the 7-point Gauss rule, as ssp_rk3_test.cpp:76-99 computes it.
    The rule and the orthonormal Legendre basis sqrt((2j+1)/2) P_j come from
    numpy, independent of the oracle; only the element ends are the oracle's."""
    ex = c.topology()["element_x"]
    xq, wq = np.polynomial.legendre.leggauss(7)
    B = np.stack([math.sqrt((2 * j + 1) / 2.0) *
                  np.polynomial.legendre.Legendre.basis(j)(xq) for j in range(p + 1)], 1)
    U = np.asarray(U).reshape(c.E, p + 1)
    jac = (ex[:, 1] - ex[:, 0]) / 2.0
    x = (ex[:, 0] + ex[:, 1])[:, None] / 2.0 + jac[:, None] * xq[None, :]
    uh = U @ B.T
    diff = uh - _pulse(np.mod(x - t, 1.0))
    return math.sqrt(float(np.sum(wq[None, :] * jac[:, None] * diff * diff)))


def _config1(gpu, cells, p=2):
    line = MG.jittered_line(cells)
    o = O.line_scalar(line.node_tags, line.node_x, line.segments, p, 1.0)
    g = gpu.Engine(gpu.Mesh.from_input(line), gpu.SCALAR_ADVECTION, p, wavespeed=1.0)
    U0 = o.project(O.IC_SCALAR_PULSE)
    g.state = U0
    t, s, _ = g.advance(0.1, 1.0)
    return o, g, U0, t, s


def test_config1_advection_to_t1(gpu):
    o, g, U0, t, s = _config1(gpu, 200)
    Uo, so, to, _ = o.run(U0, 0.1, 1.0)
    assert s == so and t == to == 1.0
    err = K.rel_inf(g.state, Uo)
    e_dev = _line_l2_error(o, g.state, t, 2)
    e_ref = _line_l2_error(o, Uo, to, 2)
    e_0 = _line_l2_error(o, U0, 0.0, 2)
    _log(f"config 1: {s} steps to T = {t}, device vs oracle rel L-inf {err:.3e}; "
         f"L2 error vs exact: device {e_dev:.6e}, oracle {e_ref:.6e} (projection {e_0:.3e})")
    assert err <= K.RUN_TOL
    assert 0.0 < e_0 < e_ref and abs(e_dev - e_ref) <= 1e-8 * e_ref
    assert e_dev < 1e-3


def test_config1_convergence_rate(gpu):
    errs = []
    for cells in (100, 200, 400):
        o, g, _, t, _ = _config1(gpu, cells)
        errs.append(_line_l2_error(o, g.state, t, 2))
    rates = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    _log(f"config 1 P2 L2 errors {errs}, rates {rates}")
    assert min(rates) >= 2.5


# ------------------------------------------------------------- config 2 ----

@pytest.fixture(scope="module")
def config2(gpu):
    m = K.vortex_mesh(128)
    o = K.oracle_tri(m, 2)
    U0 = o.project(O.IC_VORTEX, K.VORTEX)
    return m, o, U0


def test_config2_steps(gpu, config2):
    m, o, U0 = config2
    g = K.engine_tri(gpu, m, 2)
    assert g.E == 32768
    U = U0
    g.state = U
    for k in range(3):
        dt = o.stable_dt(U, 0.15)
        assert abs(g.stable_time_step(0.15) - dt) <= 1e-15 * dt
        t0 = time.perf_counter()
        U = o.step(U, dt)
        g.step(dt)
        err = K.rel_inf_per_variable(g.state, U, 4)
        _log(f"config 2 step {k + 1}: rel L-inf {err:.3e} (oracle {time.perf_counter() - t0:.1f} s)")
        assert err <= K.STEP_TOL


def test_config2_device_loop(gpu, config2):
    m, o, U0 = config2
    end = 20.5 * o.stable_dt(U0, 0.15)
    Uo, so, to, ho = o.run(U0, 0.15, end, 10 ** 6, 5)
    for tile in ((1, 2), (1, 4)):
        g = K.engine_tri(gpu, m, 2, stage_kernel=tile[0], lanes=tile[1])
        g.state = U0
        tg, sg, hg = g.advance(0.15, end, history_interval=5, history_capacity=64)
        err = K.rel_inf_per_variable(g.state, Uo, 4)
        _log(f"config 2 tile {tile}: {sg} steps, rel L-inf {err:.3e} at the end")
        assert sg == so and abs(tg - to) <= 1e-13 * to
        assert err <= K.RUN_TOL
        np.testing.assert_array_equal(hg[:, 0], ho[:, 0])
        np.testing.assert_allclose(hg[:, 1:], ho[:, 1:], rtol=1e-12, atol=0)


# ------------------------------------------------------------- config 3 ----

def _flags_match(gflags, oflags, U, E, Np, hb, threshold):
    diff = np.nonzero(gflags != oflags)[0]
    if diff.size == 0:
        return 0
    d = np.asarray(U)[: Np * E].reshape(Np, E)
    ratio = np.sum(d[hb:] ** 2, axis=0) / np.sum(d * d, axis=0)
    for e in diff:
        _log(f"flag differs at element {e}: ratio {ratio[e]!r} threshold {threshold}")
        assert abs(ratio[e] - threshold) <= 1e-12 * threshold, "flags differ off the threshold"
    return diff.size


def _config3_run(gpu):
    """The device's run of config 3 step by step from the limited projection
    until a cell average stops being physical: (engine, states before each
    step, their dt, troubled count of each, failing step or None)."""
    F = 1e-13
    m = K.rarefaction_mesh(256)
    g = K.engine_tri(gpu, m, 2, boundary=gpu.OUTFLOW, limiting=True, floors=F)
    probe = K.engine_tri(gpu, m, 2, boundary=gpu.OUTFLOW, limiting=True, floors=F)
    assert g.E == 131072
    g.project(gpu.IC_RAREFACTION, K.RAREFACTION)
    g.limit()
    states, dts, troubled = [], [], []
    for k in range(1, 200):
        try:
            dt = g.stable_time_step(0.1)
        except gpu.SimulationError as exc:
            assert "not physical" in str(exc)
            return m, g, states, dts, troubled, k
        states.append(g.state)
        dts.append(dt)
        probe.state = states[-1]
        troubled.append(int(probe.limit().sum()))
        g.step(dt)
    return m, g, states, dts, troubled, None


def test_config3_limited_steps(gpu):
    """Config 3 as specified loses a physical cell average in its 22nd step,
    in the reference as in the engine: the projection of the velocity 2 r_hat,
    discontinuous at the origin, has negative pressures at face points next
    to it, HLLC's sound speed is NaN there, and the reference's std::min /
    std::max carry the NaN into the cell averages (test_config3_stops_like_
    the_reference). Before that, from the device's states where its limiter
    flags the most cells: the limiter and one limited step against the oracle
    on the full mesh, troubled flags bit-exact, state within 1e-12."""
    F = 1e-13
    m, g, states, dts, troubled, fail = _config3_run(gpu)
    _log(f"config 3: troubled cells per step {troubled}; not physical at step {fail}")
    o = K.oracle_tri(m, 2, boundary=O.OUTFLOW)
    # not the last state: its step produces the NaN cell averages
    order = sorted(range(len(states) - 1), key=lambda k: -troubled[k])[:2]
    for k in order:
        U = states[k]
        g.state = U
        gflags = g.limit()
        Ul, oflags = o.limit(U, 1e-3, F, F)
        near = _flags_match(gflags, oflags, U, g.E, 6, 3, 1e-3)
        err = K.rel_inf_per_variable(g.state, Ul, 4)
        g.state = U
        g.step(dts[k])
        err_s = K.rel_inf_per_variable(g.state, o.step(U, dts[k], True, 1e-3, F, F), 4)
        _log(f"config 3 before step {k + 1}: {int(oflags.sum())} troubled, {near} near-threshold, "
             f"limited state rel L-inf {err:.3e}, limited step rel L-inf {err_s:.3e}")
        assert err <= K.STEP_TOL and err_s <= K.STEP_TOL
    assert max(troubled) >= 1


def test_config3_stops_like_the_reference(gpu):
    """The device loop stops with the reference's message, and from the
    device's state before the failing step the oracle's own step yields the
    same unphysical cell average (the step before it matches within 1e-12)."""
    F = 1e-13
    m, g, states, dts, troubled, fail = _config3_run(gpu)
    assert fail is not None
    h = K.engine_tri(gpu, m, 2, boundary=gpu.OUTFLOW, limiting=True, floors=F)
    h.project(gpu.IC_RAREFACTION, K.RAREFACTION)
    h.limit()
    with pytest.raises(gpu.SimulationError, match="^simulation: Euler cell average is not physical$"):
        h.advance(0.1, 0.3)
    o = K.oracle_tri(m, 2, boundary=O.OUTFLOW)
    U = o.step(states[-2], dts[-2], True, 1e-3, F, F)
    err = K.rel_inf_per_variable(states[-1], U, 4)
    _log(f"config 3: device step {len(states) - 1} rel L-inf {err:.3e}; not physical after "
         f"step {len(states)} (t = {sum(dts):.6e})")
    assert err <= K.STEP_TOL
    U = o.step(states[-1], dts[-1], True, 1e-3, F, F)
    with pytest.raises(O.OracleError, match="Euler cell average is not physical"):
        o.stable_dt(U, 0.1)


# ------------------------------------------------------------- config 4 ----
#
# At 524288 triangles and P3 the oracle needs minutes per call (its periodic
# donor search and stencil setup alone take over a minute), so at this size
# the checks are size-independent properties and the mixed stage's distance
# from the FP64 engine, whose per-step parity with the oracle is covered on
# smaller meshes of the same kind (test_gpu_parity.py, P3, every tile).

def test_config4_mixed_against_fp64(gpu):
    m = K.vortex_mesh(512)
    mesh = gpu.Mesh.from_input(m)
    U0 = mesh.project(3, gpu.IC_VORTEX, K.VORTEX)
    runs = {}
    steps = 40
    for name, prec in (("fp64", gpu.FP64), ("mixed", gpu.MIXED)):
        g = gpu.Engine(mesh, gpu.EULER, 3, precision=prec)
        assert g.E == 524288
        g.state = U0
        if name == "fp64":
            dt = g.stable_time_step(0.15)
            m0 = g.mass()
        for _ in range(steps):
            g.step(dt)
        runs[name] = (g.state, g.mass())
        del g
    a, b = runs["mixed"][0].reshape(4, 10, -1), runs["fp64"][0].reshape(4, 10, -1)
    linf = float(np.max(np.abs(a[0] - b[0])))
    drift64 = abs(runs["fp64"][1] - m0) / m0
    drift32 = abs(runs["mixed"][1] - m0) / m0
    _log(f"config 4: {steps} steps of dt {dt:.3e}, density L-inf mixed vs FP64 {linf:.3e}, "
         f"mass drift FP64 {drift64:.3e}, mixed {drift32:.3e}")
    assert np.isfinite(a).all() and np.isfinite(b).all()
    assert linf <= 1e-5
    assert drift64 <= 1e-13
    assert drift32 <= 1e-9


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_config4_free_stream(gpu, precision):
    """A uniform state is a steady solution (euler_triangle_residual_test.cpp:
    185-212): its residual vanishes to round-off on the full mesh."""
    m = K.vortex_mesh(512)
    mesh = gpu.Mesh.from_input(m)
    prec = gpu.FP64 if precision == "fp64" else gpu.MIXED
    g = gpu.Engine(mesh, gpu.EULER, 3, precision=prec)
    U = np.zeros(g.state_size)
    rho, u, v, p = 1.0, 0.5, 0.25, 1.0
    cons = (rho, rho * u, rho * v, p / 0.4 + 0.5 * rho * (u * u + v * v))
    for k, c in enumerate(cons):
        U[k * 10 * g.E: k * 10 * g.E + g.E] = math.sqrt(2.0) * c
    R = np.abs(g.residual(U)).reshape(4, 10, g.E).max(axis=(0, 1))
    # the flux balance of an element, |F| perimeter / detJ, sets the scale a
    # rounding error of the residual is measured against
    prep = mesh.prepared()
    xy = prep["points"][:, :2][prep["element_nodes"]]               # [E][3][2]
    perimeter = sum(np.hypot(*(xy[:, (k + 1) % 3] - xy[:, k]).T) for k in range(3))
    flux = max(rho * u * u + p, rho * v * v + p, (cons[3] + p) * max(u, v))
    scale = flux * perimeter / prep["element_metrics"][:, 4]
    absolute = float(R.max())
    relative = float(np.max(R / scale))
    _log(f"config 4 free stream ({precision}): max |R| {absolute:.3e}, "
         f"max |R| / (|F| perimeter / detJ) {relative:.3e}")
    if precision == "fp64":
        assert absolute <= 3e-10      # the reference's free-stream bar
    else:
        assert relative <= 2e-5       # FP32 contractions: a few hundred FP32 ulps
