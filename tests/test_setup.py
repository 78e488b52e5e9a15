"""Host-side parts of the engine (no device): canonical topology, periodic
donors, limiter stencil and projection must equal the oracle bit for bit;
the C ABI exports every entry point include/wasl_hip.h declares."""

import ctypes
import os
import re

import numpy as np
import pytest

import cases as K
import refmeshes as RM
from conftest import ROOT
from oracle import oracle as O
from wasl_amd import meshgen as MG


@pytest.mark.parametrize("order", [0, 1, 2, 3])
@pytest.mark.parametrize("boundary", [O.PERIODIC, O.OUTFLOW])
def test_setup_bitwise_equals_oracle(wasl, order, boundary):
    m = K.vortex_mesh(11) if boundary == O.PERIODIC else K.rarefaction_mesh(11)
    mesh = wasl.Mesh.from_input(m)
    o = K.oracle_tri(m, order, boundary)
    t = o.topology()
    pr = mesh.prepared()
    assert np.array_equal(pr["face_elements"], t["face_elements"])
    assert np.array_equal(pr["face_local_edges"], t["face_edges"])
    assert np.array_equal(pr["face_normals"], t["face_normals"])
    assert np.array_equal(mesh.face_vertices(), t["face_vertices"])
    tags = pr["node_tags"][pr["element_nodes"]]
    assert np.array_equal(tags, t["element_nodes"])
    assert np.array_equal(pr["points"][pr["element_nodes"]][:, :, :2].reshape(-1, 6),
                          t["element_xy"])
    if boundary == O.PERIODIC:
        d1, b1 = mesh.periodic_trace(order)
        d2, b2 = o.trace()
        assert np.array_equal(d1, d2) and np.array_equal(b1, b2)
    for a, b in zip(mesh.limiter_stencil(order, boundary == O.PERIODIC), o.stencil()):
        assert np.array_equal(a, b)
    ic, prm = (O.IC_VORTEX, K.VORTEX) if boundary == O.PERIODIC else (O.IC_EULER_REGIONS,
                                                                       K.JUMP)
    assert np.array_equal(mesh.project(order, ic, prm), o.project(ic, prm))


@pytest.mark.parametrize("order", [0, 1, 2, 3])
def test_scalar_2d_projection_bitwise_equals_oracle(wasl, order):
    """2D scalar initial states (circles, plane wave) in U[element][mode]."""
    m = K.vortex_mesh(7)
    mesh = wasl.Mesh.from_input(m)
    o = K.oracle_tri_scalar(m, order)
    for ic, prm in ((O.IC_SCALAR_CIRCLES, K.SCALAR_DISC), (O.IC_SCALAR_WAVE, K.SCALAR_WAVE)):
        assert np.array_equal(mesh.project(order, ic, prm), o.project(ic, prm))
    bt, bxy, btri = RM.irregular_box(5)
    box = wasl.Mesh.triangles(bt, bxy, btri)
    ob = O.tri_scalar(bt, bxy, btri, order, 0.7, -0.4)
    prm = [0.25, 0.5, 0.4, 0.25, 2.0]
    assert np.array_equal(box.project(order, O.IC_SCALAR_CIRCLES, prm),
                          ob.project(O.IC_SCALAR_CIRCLES, prm))


def test_regression_meshes_match(wasl):
    """The reference test meshes with shuffled records and orientations."""
    bt, bxy, btri = RM.irregular_box(5)
    mesh = wasl.Mesh.triangles(bt, bxy, btri)
    o = O.tri_euler(bt, bxy, btri, 2)
    t = o.topology()
    pr = mesh.prepared()
    assert np.array_equal(pr["face_elements"], t["face_elements"])
    assert np.array_equal(pr["face_normals"], t["face_normals"])
    lt, lx, ls = RM.irregular_line(12)
    line = wasl.Mesh.lines(lt, lx, ls)
    ol = O.line_scalar(lt, lx, ls, 2, 0.7)
    lp = line.prepared()
    ot = ol.topology()
    assert np.array_equal(lp["face_elements"], ot["face_elements"])
    assert np.array_equal(lp["points"][lp["element_nodes"]][:, :, 0], ot["element_x"])
    assert np.array_equal(line.project(2, O.IC_SCALAR_REGIONS, [0.25, 1, 1.5, 0.2, 0.55, 1.0]),
                          ol.project(O.IC_SCALAR_REGIONS, [0.25, 1, 1.5, 0.2, 0.55, 1.0]))


def test_prepared_round_trip(wasl):
    m = K.vortex_mesh(7)
    mesh = wasl.Mesh.from_input(m)
    pr = mesh.prepared()
    again = wasl.Mesh.from_prepared(pr).prepared()
    for k in ("node_tags", "points", "element_nodes", "face_elements", "face_local_edges",
              "face_normals", "face_measures", "element_metrics"):
        assert np.array_equal(pr[k], again[k]), k
    assert pr["face_elements"].shape[0] == mesh.faces
    # element metrics are the prepared mesh's J entries and det J
    xy = pr["points"][pr["element_nodes"]]
    ax = xy[:, 1, 0] - xy[:, 0, 0]
    by = xy[:, 2, 1] - xy[:, 0, 1]
    assert np.array_equal(pr["element_metrics"][:, 0], ax / 2.0)
    assert np.array_equal(pr["element_metrics"][:, 3], by / 2.0)


def test_mesh_errors(wasl):
    with pytest.raises(wasl.WaslHipError, match="zero-area element"):
        wasl.Mesh.triangles([1, 2, 3], [(0, 0), (1, 0), (2, 0)], [[1, 2, 3]])
    with pytest.raises(wasl.WaslHipError, match="missing node 9"):
        wasl.Mesh.triangles([1, 2, 3], [(0, 0), (1, 0), (0, 1)], [[1, 2, 9]])
    with pytest.raises(wasl.WaslHipError, match="more than two adjacent elements"):
        wasl.Mesh.triangles([1, 2, 3, 4, 5], [(0, 0), (1, 0), (0, 1), (0, -1), (1, 1)],
                            [[1, 2, 3], [1, 4, 2], [1, 2, 5]])
    with pytest.raises(wasl.WaslHipError, match="gap or discontinuity"):
        wasl.Mesh.lines([1, 2, 3, 4], [0.0, 0.5, 0.6, 1.0], [[1, 2], [3, 4]])


def test_periodic_trace_requires_box(wasl):
    # a diamond: every boundary edge is diagonal
    mesh = wasl.Mesh.triangles([1, 2, 3, 4], [(0, -1), (1, 0), (0, 1), (-1, 0)],
                               [[1, 2, 3], [1, 3, 4]])
    with pytest.raises(wasl.WaslHipError, match="axis-aligned box"):
        mesh.periodic_trace(1)


def test_meshgen_deterministic_and_bounded():
    a = MG.jittered_triangles(16, (0, 0), (10, 10), 0.15, seed=7)
    b = MG.jittered_triangles(16, (0, 0), (10, 10), 0.15, seed=7)
    c = MG.jittered_triangles(16, (0, 0), (10, 10), 0.15, seed=8)
    assert np.array_equal(a.node_xy, b.node_xy) and np.array_equal(a.triangles, b.triangles)
    assert not np.array_equal(a.node_xy, c.node_xy)
    grid = np.stack(np.meshgrid(np.arange(17), np.arange(17), indexing="xy"), -1).reshape(-1, 2)
    ref = grid * 10.0 / 16
    d = np.abs(a.node_xy - ref)
    assert d.max() <= 0.15 * 10.0 / 16 + 1e-12
    onb = (grid == 0) | (grid == 16)
    assert np.all(d[onb] == 0.0)
    assert a.triangles.shape == (2 * 16 * 16, 3)
    # splitmix64 reference values for seed 0: the first outputs of the stream
    assert MG.splitmix64(0, 2).tolist() == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4]
    l1 = MG.jittered_line(200)
    h = 1.0 / 200
    assert np.abs(l1.node_x - np.arange(201) * h).max() <= 0.2 * h + 1e-15
    assert l1.node_x[0] == 0.0 and l1.node_x[-1] == 1.0


def _declared_symbols():
    """Entry points of the headers libwasl_hip.so implements (engine, serving)."""
    names = set()
    for header in ("wasl_hip.h", "wasl_hip_serve.h"):
        text = open(os.path.join(ROOT, "include", header)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(wasl_hip_\w+)\s*\(", text))
    return sorted(names)


def test_cabi_exports_every_declared_symbol(wasl):
    lib = ctypes.CDLL(os.path.join(ROOT, "wasl_amd", "libwasl_hip.so"))
    names = _declared_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_numerics_defaults(wasl):
    n = wasl.engine.Numerics()
    wasl.library().wasl_hip_numerics_default(ctypes.byref(n))
    assert n.sensor_threshold == 1e-3 and n.density_floor == 1e-10
    assert n.pressure_floor == 1e-10 and n.heat_capacity_ratio == 1.4
    assert n.velocity_y == 0.0
    assert ctypes.sizeof(n) == 104   # sizeof(wasl_hip_numerics) on x86-64


def test_server_config_defaults(wasl):
    """wasl_hip_server_config_default: the served FP64 tiles by order (P0 left
    to the autotuner; P1 point split 1 lane, P2 point split 2 lanes, P3
    variable split 4 lanes) and the struct layout the Python mirror assumes."""
    from wasl_amd.serve import ServerConfig, default_fp64_tiles

    assert default_fp64_tiles() == {0: (0, 0), 1: (1, 2), 2: (1, 2), 3: (2, 4)}
    c = ServerConfig()
    wasl.library().wasl_hip_server_config_default(ctypes.byref(c))
    assert (c.max_resident, c.steps_per_batch, c.setup_threads) == (256, 16, 4)
    assert (c.fp64_max_order, c.fp64_limited, c.use_graphs) == (1, 1, 1)
    assert ctypes.sizeof(c) == 72   # sizeof(wasl_hip_server_config) on x86-64


@pytest.mark.parametrize("order", [0, 1, 2, 3])
def test_density_lattice_basis_equals_oracle(order):
    """The serving bench samples density fields through its own Dubiner
    evaluation; it agrees with the oracle's basis at every lattice point,
    vertices (the collapsed top vertex included) and edges too."""
    from wasl_amd import workload as WL

    R = O.Routines()
    n = 2 * order + 2
    pts = [(-1.0 + 2.0 * a / n, -1.0 + 2.0 * b / n) for b in range(n + 1) for a in range(n + 1 - b)]
    ref = np.array([R.dubiner(x, y, order) for x, y in pts])
    B = WL.dubiner_lattice(order)
    assert B.shape == ref.shape == (len(pts), (order + 1) * (order + 2) // 2)
    np.testing.assert_allclose(B, ref, rtol=1e-13, atol=1e-13)
