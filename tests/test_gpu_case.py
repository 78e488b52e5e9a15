"""The bridge's coarse calls on the device, checked the way the reference's
own case-regression test checks its runs (cpp/tests/runtime/
case_regression_test.cpp:131-202): prepared meshes written to HDF5, each case
run through run_scalar_advection_case / run_euler_case, then step, time, the
sum and L2 norm of the final checkpoint's coefficients and the last history
mass against the stored reference rows at relative 1e-10."""

import math
import os

import numpy as np
import pytest

import refmeshes as RM
from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _measure(case, summary):
    m = dict(kv.split("=", 1) for kv in summary.split(" "))
    ck = case.read_checkpoint(m["checkpoint"])
    c = ck["coefficients"]
    with open(os.path.join(os.path.dirname(m["checkpoint"]), "history.csv")) as fh:
        lines = [ln for ln in fh.read().splitlines() if ln]
    assert lines[0] == "step,time,dt,mass"
    mass = float(lines[-1].rsplit(",", 1)[1])
    return (int(m["step"]), float(m["time"]), float(c.sum()), math.sqrt(float((c * c).sum())),
            mass), m["checkpoint"]


def test_coarse_calls_reproduce_case_regression_rows(gpu, tmp_path):
    from wasl_amd import case

    line_path, box_path = tmp_path / "line.h5", tmp_path / "box.h5"
    case.prepare_mesh(gpu.Mesh.lines(*RM.irregular_line(12)), line_path)
    case.prepare_mesh(gpu.Mesh.triangles(*RM.irregular_box(5)), box_path)
    got = {}
    for p in range(4):
        name = f"scalar-d1-p{p}"
        args = dict(mpi=False, dimensions=1, polynomial_order=p, velocity_x=0.7, velocity_y=0.0,
                    cfl=0.3, start_time=0.0, max_steps=1000, history_interval=3,
                    visualization_interval=4, checkpoint_interval=5, keep_checkpoints=2,
                    initial_default=0.25, initial_sine=(p == 2), initial_sine_amplitude=1.5,
                    intervals=[0.2, 0.55, 1.0], circles=[0.5, 0.4, 0.25, 2.0])
        s = case.run_scalar_advection_case(line_path, tmp_path / name, "", end_time=0.2, **args)
        got[name], ck = _measure(case, s)
        s = case.run_scalar_advection_case(line_path, tmp_path / (name + "-restart"), ck,
                                           end_time=0.3, **args)
        got[name + "-restart"], _ = _measure(case, s)
    for p in range(4):
        name = f"scalar-d2-p{p}"
        args = dict(mpi=False, dimensions=2, polynomial_order=p, velocity_x=0.7, velocity_y=-0.4,
                    cfl=0.3, start_time=0.0, max_steps=1000, history_interval=3,
                    visualization_interval=4, checkpoint_interval=5, keep_checkpoints=2,
                    initial_default=0.25, initial_sine=False, initial_sine_amplitude=1.5,
                    intervals=[0.2, 0.55, 1.0], circles=[0.5, 0.4, 0.25, 2.0])
        s = case.run_scalar_advection_case(box_path, tmp_path / name, "", end_time=0.2, **args)
        got[name], ck = _measure(case, s)
        s = case.run_scalar_advection_case(box_path, tmp_path / (name + "-restart"), ck,
                                           end_time=0.3, **args)
        got[name + "-restart"], _ = _measure(case, s)
    for p in range(3):
        for lim in (False, True):
            name = f"euler-d1-p{p}" + ("-limited" if lim else "")
            inside = (0.2, 0.0, 0.0, 0.15) if lim else (0.9, 0.25, 0.0, 0.95)
            s = case.run_euler_case(
                line_path, tmp_path / name, "", lim, 1e-4, 1e-10, 1e-10, 1, p, 1.4, 0.2, 0.0,
                0.03, 1000, 2, 3, 4, 1,
                regions=[0, 0, 0, 0, 0, 0, 1.0, 0.3, 0.0, 1.0, 1, 0.3, 0.6, 0, 0, 0, *inside])
            got[name], _ = _measure(case, s)
    for p in range(3):
        for lim in (False, True):
            name = f"euler-d2-p{p}" + ("-limited" if lim else "")
            inside = (0.2, 0.0, 0.0, 0.15) if lim else (0.9, 0.25, 0.1, 0.95)
            s = case.run_euler_case(
                box_path, tmp_path / name, "", lim, 1e-4, 1e-10, 1e-10, 2, p, 1.4, 0.2, 0.0,
                0.03, 1000, 2, 3, 4, 1,
                regions=[0, 0, 0, 0, 0, 0, 1.0, 0.3, 0.2, 1.0, 2, 0, 0, 0.5, 0.5, 0.25, *inside])
            got[name], _ = _measure(case, s)
    ref = RM.read_case_regression(os.path.join(GOLDEN, "case_regression_rows.txt"))
    assert set(got) == set(ref)
    for name, exp in sorted(ref.items()):
        act = got[name]
        rel = [abs(a - e) / max(1.0, abs(e)) for a, e in zip(act[1:], exp[1:])]
        print(f"{name}: steps {act[0]}/{exp[0]} max rel {max(rel):.2e}")
        assert act[0] == exp[0], name
        assert max(rel) <= 1e-10, (name, act, exp)


def test_runner_outputs_follow_reference_cadence(gpu, tmp_path):
    """history rows at every history_interval-th step and the last; snapshots
    and rotated checkpoints named prefix-NNNNNN (serial_case_runner.cpp:123-148)."""
    from wasl_amd import case

    box_path = tmp_path / "box.h5"
    case.prepare_mesh(gpu.Mesh.triangles(*RM.irregular_box(5)), box_path)
    out = tmp_path / "run"
    s = case.run_euler_case(box_path, out, "", False, 1e-3, 1e-10, 1e-10, 2, 1, 1.4, 0.2, 0.0,
                            0.03, 1000, 2, 3, 4, 2,
                            regions=[0, 0, 0, 0, 0, 0, 1.0, 0.3, 0.2, 1.0])
    step = int(s.split()[0].split("=")[1])
    rows = np.loadtxt(out / "history.csv", delimiter=",", skiprows=1)
    expect = [0] + [k for k in range(1, step + 1) if k % 2 == 0 or k == step]
    assert rows[:, 0].astype(int).tolist() == expect
    assert rows[0, 2] == 0.0 and np.all(rows[1:, 2] > 0)
    assert abs(rows[:, 1][-1] - 0.03) <= 1e-15
    vtus = sorted(p.name for p in out.glob("solution-*.vtu"))
    assert vtus == sorted({f"solution-{k:06d}.vtu" for k in [0] + [k for k in range(1, step + 1)
                                                                   if k % 3 == 0 or k == step]})
    cks = sorted(p.name for p in out.glob("checkpoint-*.h5"))
    written = [k for k in range(1, step + 1) if k % 4 == 0 or k == step]
    assert cks == [f"checkpoint-{k:06d}.h5" for k in written[-2:]]
    assert s.endswith(f"checkpoint={out / f'checkpoint-{step:06d}.h5'}")
    assert (out / "solution.pvd").read_text().count("<DataSet") == len(vtus)


def test_runs_on_reference_files(gpu, tmp_path):
    """A case run entirely on files of the reference's own formats: the mesh
    is written by the reference's prepare_mesh and writer, the run's
    checkpoint is read by the reference's reader, and a restart from a
    checkpoint the reference wrote (same fields, its own writer) ends where
    the uninterrupted run ends."""
    from oracle import oracle as O
    from wasl_amd import case

    if not O.ReferenceIO.available():
        pytest.skip("oracle/_ref/libwaslref_io.so is not built here (no /root/reference)")
    import cases as K

    io = O.ReferenceIO()
    m = K.vortex_mesh(6)
    mesh_path = tmp_path / "box.h5"
    io.write_prepared(m.node_tags, m.node_xy, m.triangles, mesh_path)
    regions = [0, 0, 0, 0, 0, 0, 1.0, 0.3, 0.2, 1.0,
               2, 0, 0, 5.0, 5.0, 2.0, 0.5, 0.0, 0.0, 0.7]   # a limited disc
    args = (True, 1e-3, 1e-10, 1e-10, 2, 2, 1.4, 0.2)

    def run(out, restart):
        return dict(kv.split("=", 1) for kv in case.run_euler_case(
            mesh_path, out, restart, *args, 0.0, 0.2, 10 ** 6, 1, 5, 5, 100,
            regions=regions).split(" "))

    full = run(tmp_path / "full", "")
    mid = tmp_path / "full" / "checkpoint-000005.h5"
    ck = io.read_checkpoint(mid)
    assert ck["step"] == 5 and ck["element_count"] == 72 and ck["partition_rank_count"] == 1
    assert os.path.samefile(ck["prepared_mesh"], mesh_path)
    final = io.read_checkpoint(full["checkpoint"])
    assert final["step"] == int(full["step"]) > 5 and final["time"] == float(full["time"]) == 0.2
    theirs = tmp_path / "reference_checkpoint.h5"
    io.write_checkpoint(theirs, ck["time"], ck["step"], ck["variable_count"], ck["mode_count"],
                        ck["element_count"], ck["prepared_mesh"], ck["coefficients"])
    resumed = run(tmp_path / "resumed", str(theirs))
    assert resumed["step"] == full["step"] and resumed["time"] == full["time"]
    b = io.read_checkpoint(resumed["checkpoint"])["coefficients"]
    assert np.array_equal(final["coefficients"], b)

    # the run's text outputs are the bytes the reference's writers emit for
    # the same rows, snapshots and cell averages
    import re

    out = tmp_path / "full"
    rows = np.loadtxt(out / "history.csv", delimiter=",", skiprows=1, ndmin=2)
    io.write_history(tmp_path / "history.csv", rows)
    assert (out / "history.csv").read_bytes() == (tmp_path / "history.csv").read_bytes()
    pvd = (out / "solution.pvd").read_text()
    snaps = [(float(t), f) for t, f in re.findall(r'timestep="([^"]+)"[^>]*file="([^"]+)"', pvd)]
    assert [f for _, f in snaps[:2]] == ["solution-000000.vtu", "solution-000005.vtu"]
    io.write_pvd(tmp_path / "solution.pvd", snaps)
    assert (out / "solution.pvd").read_bytes() == (tmp_path / "solution.pvd").read_bytes()
    io.export(mid, mesh_path, tmp_path / "step5.vtu")
    assert (out / "solution-000005.vtu").read_bytes() == (tmp_path / "step5.vtu").read_bytes()
