"""Writes the golden fixtures under tests/golden/ (run in a container where
the reference is present: `python tests/golden/make_fixtures.py`).

  case_regression_rows.txt  the 28 on-path rows (euler-d1-*, euler-d2-*,
                            scalar-d1-*, scalar-d2-*), i.e. every row of
                            the reference's stored case-regression metrics
                            (cpp/tests/runtime/case_regression_reference.txt),
                            copied verbatim as data;
  ref_prepared_box.h5,      a prepared mesh and a checkpoint written by the
  ref_checkpoint_p1.h5      reference's own HDF5 writers (format checks);
  reference_vectors.npz     inputs and outputs of the reference's own code
                            (oracle/_ref, compiled from the reference sources)
                            on small seeded meshes: residuals P0-P3, limiter
                            flags and states, stepping trajectories, 1D
                            advection, 2D scalar advection residuals and
                            steps, 1D Euler residuals, limiter and limited
                            steps.

The reference cannot travel to the GPU box; these vectors do.
"""

from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

REFERENCE_TXT = "/root/reference/cpp/tests/runtime/case_regression_reference.txt"


def regression_rows() -> None:
    keep = []
    with open(REFERENCE_TXT) as fh:
        for line in fh:
            if line.startswith(("#", "euler-d1-", "euler-d2-", "scalar-d1-", "scalar-d2-")):
                keep.append(line)
    with open(os.path.join(HERE, "case_regression_rows.txt"), "w") as fh:
        fh.writelines(keep)


def vectors() -> None:
    import cases as K
    from oracle import oracle as O
    from wasl_amd import meshgen as MG

    out = {}
    m = K.vortex_mesh(8)
    for p in range(4):
        c = K.oracle_tri(m, p, kind="reference")
        U = c.project(O.IC_VORTEX, K.VORTEX)
        rng = np.random.default_rng(100 + p)
        U = U * (1.0 + 1e-3 * rng.standard_normal(U.size))
        out[f"vortex_p{p}_state"] = U
        out[f"vortex_p{p}_residual"] = c.residual(U)
        out[f"vortex_p{p}_dt"] = np.array([c.stable_dt(U, 0.15)])
    c = K.oracle_tri(m, 2, kind="reference")
    U = c.project(O.IC_VORTEX, K.VORTEX)
    traj = [U]
    for _ in range(3):
        U = c.step(U, c.stable_dt(U, 0.15))
        traj.append(U)
    out["vortex_p2_steps"] = np.stack(traj)

    r = K.rarefaction_mesh(8)
    for p in (1, 2, 3):
        c = K.oracle_tri(r, p, boundary=O.OUTFLOW, kind="reference")
        U = c.project(O.IC_EULER_REGIONS, K.JUMP)
        Ul, flags = c.limit(U, 1e-3, 1e-13, 1e-13)
        out[f"jump_p{p}_state"] = U
        out[f"jump_p{p}_limited"] = Ul
        out[f"jump_p{p}_flags"] = flags
        out[f"jump_p{p}_smooth_residual"] = c.residual(c.project(O.IC_SMOOTH, K.SMOOTH))
    c = K.oracle_tri(r, 2, boundary=O.OUTFLOW, kind="reference")
    U = c.project(O.IC_RAREFACTION, K.RAREFACTION)
    U, _ = c.limit(U, 1e-3, 1e-13, 1e-13)
    traj = [U]
    for _ in range(3):
        U = c.step(U, c.stable_dt(U, 0.1), True, 1e-3, 1e-13, 1e-13)
        traj.append(U)
    out["rarefaction_p2_limited_steps"] = np.stack(traj)

    line = MG.jittered_line(200)
    c = O.line_scalar(line.node_tags, line.node_x, line.segments, 2, 1.0, kind="reference")
    U = c.project(O.IC_SCALAR_PULSE)
    dt = c.stable_dt(U, 0.1)
    traj = [U]
    for _ in range(10):
        U = c.step(U, dt)
        traj.append(U)
    out["pulse_p2_steps"] = np.stack(traj)
    out["pulse_p2_dt"] = np.array([dt])

    for p in range(4):
        c = O.tri_scalar(m.node_tags, m.node_xy, m.triangles, p, 0.7, -0.4, kind="reference")
        U = c.project(O.IC_SCALAR_WAVE, K.SCALAR_WAVE)
        rng = np.random.default_rng(200 + p)
        U = U + 1e-3 * rng.standard_normal(U.size)
        out[f"advect2d_p{p}_state"] = U
        out[f"advect2d_p{p}_residual"] = c.residual(U)
        out[f"advect2d_p{p}_dt"] = np.array([c.stable_dt(U, 0.3)])
    c = O.tri_scalar(m.node_tags, m.node_xy, m.triangles, 2, 0.7, -0.4, kind="reference")
    U = c.project(O.IC_SCALAR_CIRCLES, K.SCALAR_DISC)
    dt = c.stable_dt(U, 0.3)
    traj = [U]
    for _ in range(5):
        U = c.step(U, dt)
        traj.append(U)
    out["advect2d_p2_steps"] = np.stack(traj)

    import refmeshes as RM

    lt, lx, ls = RM.irregular_line(12)
    for p in range(4):
        for bc, tag in ((O.PERIODIC, "periodic"), (O.OUTFLOW, "outflow")):
            c = O.line_euler(lt, lx, ls, p, left=bc, right=bc, kind="reference")
            U = c.project(O.IC_EULER_REGIONS, K.SHOCK_TUBE)
            Ul, flags = c.limit(U, 1e-3)
            out[f"euler1d_p{p}_{tag}_state"] = U
            out[f"euler1d_p{p}_{tag}_residual"] = c.residual(U)
            out[f"euler1d_p{p}_{tag}_limited"] = Ul
            out[f"euler1d_p{p}_{tag}_flags"] = flags
            out[f"euler1d_p{p}_{tag}_dt"] = np.array([c.stable_dt(Ul, 0.2)])
            traj = [Ul]
            for _ in range(3):
                traj.append(c.step(traj[-1], c.stable_dt(traj[-1], 0.2), True, 1e-3))
            out[f"euler1d_p{p}_{tag}_steps"] = np.stack(traj)
    positivity_vectors(out)
    np.savez_compressed(os.path.join(HERE, "reference_vectors.npz"), **out)


def positivity_meshes():
    import cases as K
    from oracle import oracle as O

    return (("transmissive", O.OUTFLOW, K.rarefaction_mesh(8)),
            ("periodic", O.PERIODIC, K.vortex_mesh(8)))


# The floor-error input: positivity_state at P2 on the transmissive mesh with
# element FLOOR_ERROR_ELEMENTS[0]'s density mean and element [1]'s pressure
# mean pushed below the floors. The reference throws at the first of them.
FLOOR_ERROR_ELEMENTS = (37, 90)


def floor_error_state(c):
    import cases as K

    U = K.positivity_state(c, 2, 42).reshape(4, c.Np, c.E)
    s2 = np.sqrt(2.0)
    a, b = FLOOR_ERROR_ELEMENTS
    U[0, 0, a] = 0.5 * K.POSITIVITY_FLOORS * s2
    rho, mx, my = U[0, 0, b] / s2, U[1, 0, b] / s2, U[2, 0, b] / s2
    U[3, 0, b] = (0.5 * (mx * mx + my * my) / rho - 1e-3) * s2   # pressure -4e-4
    return U.ravel()


def positivity_vectors(out) -> None:
    """Limiter inputs on which the Zhang-Shu positivity limiter acts
    (tests/cases.py positivity_state): per order and boundary the input, the
    reference's limited state and flags at threshold 1e-3, and two limited
    SSP-RK3 steps at threshold 1e-10 (every cell troubled) from the limited
    state; and the floor-error input with the reference's partial state."""
    import cases as K

    F = K.POSITIVITY_FLOORS
    for tag, bnd, mesh in positivity_meshes():
        for p in (1, 2, 3):
            c = K.oracle_tri(mesh, p, boundary=bnd, kind="reference")
            U = K.positivity_state(c, p, 40 + p)
            Ul, flags = c.limit(U, 1e-3, F, F)
            key = f"positivity_{tag}_p{p}"
            out[key + "_state"] = U
            out[key + "_limited"] = Ul
            out[key + "_flags"] = flags
            U0, _ = c.limit(U, 1e-10, F, F)
            traj = [U0]
            for _ in range(2):
                traj.append(c.step(traj[-1], c.stable_dt(traj[-1], 0.1), True, 1e-10, F, F))
            out[key + "_steps"] = np.stack(traj)
    _, bnd, mesh = positivity_meshes()[0]
    c = K.oracle_tri(mesh, 2, boundary=bnd, kind="reference")
    U = floor_error_state(c)
    part, flags, msg = c.limit(U, 1e-3, F, F, partial=True)
    assert msg and "below a floor" in msg, msg
    out["floor_error_p2_state"] = U
    out["floor_error_p2_partial"] = part


# On-disk formats written by the reference's own writers (oracle/_ref's
# libwaslref_io.so): a prepared mesh of a 2x4x4 jittered periodic box and a
# checkpoint of the projected P1 vortex on it (the mesh path stored as the
# file name alone).
FORMAT_MESH = "ref_prepared_box.h5"
FORMAT_CHECKPOINT = "ref_checkpoint_p1.h5"


def format_checkpoint_fields():
    import cases as K
    from oracle import oracle as O

    c = K.oracle_tri(K.vortex_mesh(4), 1)
    return dict(time=0.125, step=42, variable_count=4, mode_count=3, element_count=c.E,
                prepared_mesh=FORMAT_MESH, coefficients=c.project(O.IC_VORTEX, K.VORTEX))


def formats() -> None:
    import cases as K
    from oracle import oracle as O

    io = O.ReferenceIO()
    m = K.vortex_mesh(4)
    io.write_prepared(m.node_tags, m.node_xy, m.triangles, os.path.join(HERE, FORMAT_MESH))
    f = format_checkpoint_fields()
    io.write_checkpoint(os.path.join(HERE, FORMAT_CHECKPOINT), f["time"], f["step"],
                        f["variable_count"], f["mode_count"], f["element_count"],
                        f["prepared_mesh"], f["coefficients"])


if __name__ == "__main__":
    regression_rows()
    vectors()
    formats()
    print("fixtures written to", HERE)
