"""The on-disk formats against the reference's own readers and writers.

Prepared meshes (src/io/prepared_mesh_file.cpp:32-97) and checkpoints
(src/io/checkpoint_file.cpp:17-71), HDF5 version 2:
  - files the reference wrote (tests/golden/ref_prepared_box.h5,
    ref_checkpoint_p1.h5, from tests/golden/make_fixtures.py) are read by the
    engine's case layer: the mesh equals the engine's own canonical mesh of
    the same input, the checkpoint's fields and coefficients are the ones the
    reference stored, bit for bit;
  - files the engine's case layer writes are read by the reference
    (oracle/_ref/libwaslref_io.so, built from the reference's IO sources),
    and what the reference reads and re-emits reads back identically.
No GPU is needed: the case layer's file calls are host code.
"""

import os

import numpy as np
import pytest

import cases as K
from conftest import GOLDEN
from oracle import oracle as O

KEYS = ("node_tags", "points", "element_nodes", "element_metrics", "face_nodes",
        "face_elements", "face_local_edges", "face_normals", "face_measures")


@pytest.fixture(scope="module")
def case(wasl):
    from wasl_amd import case as C

    C.case_library()
    return C


@pytest.fixture(scope="module")
def reference_io():
    if not O.ReferenceIO.available():
        pytest.skip("oracle/_ref/libwaslref_io.so is not built here (no /root/reference)")
    return O.ReferenceIO()


def _same_mesh(a, b):
    pa, pb = a.prepared(), b.prepared()
    for k in KEYS:
        assert np.array_equal(pa[k], pb[k]), k


def test_engine_reads_reference_prepared_mesh(wasl, case):
    from golden.make_fixtures import FORMAT_MESH

    got = case.read_prepared_mesh(os.path.join(GOLDEN, FORMAT_MESH))
    own = wasl.Mesh.from_input(K.vortex_mesh(4))
    assert got.elements == own.elements == 32 and got.faces == own.faces
    _same_mesh(got, own)
    summary = case.inspect_prepared_mesh(os.path.join(GOLDEN, FORMAT_MESH))
    assert summary == "2D mesh with 32 elements"


def test_engine_reads_reference_checkpoint(case):
    from golden.make_fixtures import FORMAT_CHECKPOINT, format_checkpoint_fields

    want = format_checkpoint_fields()
    got = case.read_checkpoint(os.path.join(GOLDEN, FORMAT_CHECKPOINT))
    for k in ("time", "step", "variable_count", "mode_count", "element_count"):
        assert got[k] == want[k], k
    assert np.array_equal(got["coefficients"], want["coefficients"])


@pytest.mark.parametrize("build", ["triangles", "lines"])
def test_reference_reads_engine_prepared_mesh(wasl, case, reference_io, tmp_path, build):
    from wasl_amd import meshgen as MG

    if build == "triangles":
        m = wasl.Mesh.from_input(K.rarefaction_mesh(6))
    else:
        m = wasl.Mesh.from_input(MG.jittered_line(40))
    ours = tmp_path / "engine.h5"
    case.prepare_mesh(m, ours)
    back = tmp_path / "reference.h5"
    dim, nodes, elements, faces = reference_io.copy_prepared(ours, back)
    assert (dim, elements, faces) == (m.dimension, m.elements, m.faces) and nodes > 0
    _same_mesh(case.read_prepared_mesh(back), m)


def test_reference_prepared_mesh_matches_engine_prepare(wasl, case, reference_io, tmp_path):
    """The reference's prepare_mesh + writer and the engine's, on one input,
    give files whose meshes are identical."""
    m = K.vortex_mesh(5)
    ref = tmp_path / "reference.h5"
    reference_io.write_prepared(m.node_tags, m.node_xy, m.triangles, ref)
    ours = tmp_path / "engine.h5"
    case.prepare_mesh(wasl.Mesh.from_input(m), ours)
    _same_mesh(case.read_prepared_mesh(ref), case.read_prepared_mesh(ours))


def test_reference_reads_engine_checkpoint(case, reference_io, tmp_path):
    rng = np.random.default_rng(3)
    U = rng.standard_normal(4 * 6 * 50)
    path = tmp_path / "ck.h5"
    case.write_checkpoint(path, 0.75, 123, 4, 6, 50, "some/prepared.h5", U)
    got = reference_io.read_checkpoint(path)
    assert (got["time"], got["step"], got["variable_count"], got["mode_count"],
            got["element_count"], got["partition_rank_count"]) == (0.75, 123, 4, 6, 50, 1)
    assert got["prepared_mesh"] == "some/prepared.h5"
    assert np.array_equal(got["coefficients"], U)
    back = tmp_path / "ref.h5"
    reference_io.write_checkpoint(back, got["time"], got["step"], 4, 6, 50, got["prepared_mesh"],
                                  got["coefficients"])
    again = case.read_checkpoint(back)
    assert again["step"] == 123 and np.array_equal(again["coefficients"], U)


def test_exported_vtu_is_the_reference_bytes(case, reference_io, tmp_path):
    """export_checkpoint of the reference's checkpoint on the reference's
    mesh: the engine's VTU file and the reference's (bridge.cpp:107-126 with
    its write_vtu) are byte for byte the same."""
    from golden.make_fixtures import FORMAT_CHECKPOINT, FORMAT_MESH

    ck, mesh = os.path.join(GOLDEN, FORMAT_CHECKPOINT), os.path.join(GOLDEN, FORMAT_MESH)
    ours, theirs = tmp_path / "engine.vtu", tmp_path / "reference.vtu"
    assert case.export_checkpoint(ck, mesh, ours) == "exported 32 cells"
    reference_io.export(ck, mesh, theirs)
    assert ours.read_bytes() == theirs.read_bytes()
