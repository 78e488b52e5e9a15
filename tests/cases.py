"""Matched cases for parity tests: the same raw mesh handed to the oracle and
to the HIP engine, plus tolerance helpers.

Tolerances (the parity contract of the plain FP64 path):
  per residual / per step   relative L-inf <= 1e-12 (max-norm scaled)
  end of run                relative L-inf <= 1e-10
  connectivity, donors, stencil, troubled flags: bit-exact
"""

from __future__ import annotations

import numpy as np

from oracle import oracle as O
from wasl_amd import meshgen as MG

GAMMA = 1.4
STEP_TOL = 1e-12
RUN_TOL = 1e-10

VORTEX = (5.0, 5.0, 5.0)
RAREFACTION = (1.0, 0.4, 2.0)
# rho = 1 + 0.1 sum_k a_k sin(2 pi (m_k x + n_k y) / L + phase_k), p = 1, u = v = 0.5
SMOOTH = (2.0, 0.5, 1, 2, 0.3, -0.7, 2, 1, 1.1, 0.4, 3, 1, 2.0, -0.2, 1, 3, 0.5)
# default region plus a low-density, low-pressure disc: troubled cells on its rim
JUMP = (0, 0, 0, 0, 0, 0, 1.0, 0.3, 0.2, 1.0,
        2, 0, 0, 0.1, -0.2, 0.45, 0.2, 0.0, 0.0, 0.15)


# 1D Euler regions (width 10): a default state and an interval [0.3, 0.6] of
# low density and pressure, the case-regression test's limited shock tube
SHOCK_TUBE = (0, 0, 0, 0, 0, 0, 1.0, 0.3, 0.0, 1.0,
              1, 0.3, 0.6, 0, 0, 0, 0.2, 0.0, 0.0, 0.15)
# Sod (cases/tutorials/euler/sod-shock-tube/initial/state.toml)
SOD = (0, 0, 0, 0, 0, 0, 1.0, 0.0, 0.0, 1.0,
       1, 0.5, 1.0, 0, 0, 0, 0.125, 0.0, 0.0, 0.1)

# 2D scalar advection: plane wave sin(2 pi (kx x + ky y) / L) on [0,10]^2 and
# a disc of value 2 on a 0.25 background (circles: default, cx, cy, r, value)
SCALAR_WAVE = (10.0, 1.0, 2.0)
SCALAR_DISC = (0.25, 5.0, 4.0, 2.5, 2.0)
ADVECTION = (0.7, -0.4)


def rel_inf(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale)


def rel_inf_per_variable(a, b, V: int) -> float:
    """Max over variables of the max-norm-scaled difference (U[variable] blocks)."""
    a = np.asarray(a).reshape(V, -1)
    b = np.asarray(b).reshape(V, -1)
    return max(rel_inf(a[v], b[v]) for v in range(V))


def vortex_mesh(n=12, seed=MG.DEFAULT_SEED):
    return MG.jittered_triangles(n, (0.0, 0.0), (10.0, 10.0), 0.15, seed)


def nonconforming_vortex_mesh(n=10, seed=MG.DEFAULT_SEED):
    """vortex_mesh(n) with the nodes of the x = 10 boundary (corners kept)
    moved along it by a quarter of the spacing. The periodic images of the
    x = 0 boundary's Gauss points then fall between the x = 10 boundary's
    Gauss points (and the other way round), so those 2 n faces keep the
    per-point donor lookup, while the n conforming y = 0 / y = 10 pairs are
    recorded as interior faces."""
    m = vortex_mesh(n, seed)
    xy = m.node_xy.copy()
    on = (xy[:, 0] == 10.0) & (xy[:, 1] > 0.0) & (xy[:, 1] < 10.0)
    assert on.sum() == n - 1
    xy[on, 1] += 0.25 * 10.0 / n
    return MG.TriangleMeshInput(m.node_tags, xy, m.triangles, m.box)


def rarefaction_mesh(n=10, seed=MG.DEFAULT_SEED):
    return MG.jittered_triangles(n, (-1.0, -1.0), (1.0, 1.0), 0.15, seed)


def odd_mesh(n=9, seed=MG.DEFAULT_SEED):
    """rarefaction_mesh(n) with one boundary triangle split at the midpoint of
    its edge on y = -1: 2 n^2 + 1 triangles, an odd count (transmissive
    boundaries only)."""
    m = rarefaction_mesh(n, seed)
    tags, xy, tris = m.node_tags, m.node_xy, m.triangles.copy()
    pos = {int(t): k for k, t in enumerate(tags)}
    t = tris[0]
    a, b = int(t[0]), int(t[1])                  # quad (0, 0): edge ll-lr lies on y = -1
    assert xy[pos[a], 1] == -1.0 and xy[pos[b], 1] == -1.0
    mid_tag = int(tags.max()) + 1
    mid = 0.5 * (xy[pos[a]] + xy[pos[b]])
    tris = np.vstack([tris, [[mid_tag, b, int(t[2])]]])
    tris[0] = [a, mid_tag, int(t[2])]
    return MG.TriangleMeshInput(np.append(tags, mid_tag), np.vstack([xy, mid]), tris, m.box)


def oracle_tri(m, order, boundary=O.PERIODIC, kind="port"):
    return O.tri_euler(m.node_tags, m.node_xy, m.triangles, order, boundary=boundary, kind=kind)


def oracle_tri_scalar(m, order, kind="port"):
    return O.tri_scalar(m.node_tags, m.node_xy, m.triangles, order, *ADVECTION, kind=kind)


def engine_tri(wasl, m, order, boundary=0, limiting=False, threshold=1e-3, floors=1e-10,
               lanes=0, precision=0, stage_kernel=0):
    mesh = wasl.Mesh.from_input(m)
    return wasl.Engine(mesh, wasl.EULER, order, boundary=(boundary, boundary), limiting=limiting,
                       sensor_threshold=threshold, density_floor=floors, pressure_floor=floors,
                       lanes=lanes, precision=precision, stage_kernel=stage_kernel)


# Positivity-engaged limiter inputs. A low-density, low-pressure background
# (rho 0.05, u 1, v 0.5, p 0.01) with seeded random high modes of amplitude
# POSITIVITY_AMPLITUDE[p] x |mean| on every variable: the characteristic WENO
# blends oscillatory candidates, so most troubled cells reach the Zhang-Shu
# limiter with density or pressure below the floor at some check points
# (euler_positivity.hpp:63-102), some with density only scaled, some with the
# pressure bisection as well, some untouched.
POSITIVITY_BACKGROUND = (0, 0, 0, 0, 0, 0, 0.05, 1.0, 0.5, 0.01)
POSITIVITY_AMPLITUDE = {1: 0.3, 2: 0.15, 3: 0.08}
POSITIVITY_FLOORS = 1e-13


def positivity_state(c, order, seed):
    U = c.project(O.IC_EULER_REGIONS, POSITIVITY_BACKGROUND).reshape(4, c.Np, c.E)
    rng = np.random.default_rng(seed)
    U[:, 1:, :] = (POSITIVITY_AMPLITUDE[order] * np.abs(U[:, :1, :]) *
                   rng.standard_normal((4, c.Np - 1, c.E)))
    return U.ravel()


def positivity_scales(c, U, threshold, density_floor, pressure_floor, routines):
    """Per element, for the troubled cells of one limiter call: whether the
    characteristic-WENO output had a check-point density below the floor
    (the density pass scales), and the total high-mode scale the positivity
    routine applied (1 where it left the cell alone). The call is recomposed
    from the reference's single routines -- candidates T_k U_nb with the
    target's own mean (euler_modal_limiter.cpp:43-66, summed in the same
    order), euler_characteristic_weno, limit_euler_positivity -- and every
    recomposed element must equal the case-level limiter's output bit for
    bit. Returns (flags, density_engaged[E], scale[E], limited state)."""
    nb, tr, nrm, cb = c.stencil()
    E, Np = c.E, c.Np
    Ul, flags = c.limit(U, threshold, density_floor, pressure_floor)
    Uv = np.asarray(U, np.float64).reshape(4, Np, E)
    Lv = Ul.reshape(4, Np, E)
    dens = np.zeros(E, bool)
    scale = np.ones(E)
    for e in np.nonzero(flags)[0]:
        t = Uv[:, :, e].copy()
        cands = []
        for k in range(nb.shape[1]):
            n = int(nb[e, k])
            if n < 0:
                continue
            cv = np.zeros((4, Np))
            for v in range(4):
                for tm in range(Np):
                    x = 0.0
                    for sm in range(Np):
                        x += tr[e, k, tm, sm] * Uv[v, sm, n]
                    cv[v, tm] = x
                cv[v, 0] = t[v, 0]
            cands.append(cv)
        avg = t[:, 0] / np.sqrt(2.0)
        lim = routines.characteristic_weno(GAMMA, avg, nrm[e, 0], nrm[e, 1], t,
                                           np.array(cands) if cands else np.zeros(1),
                                           len(cands), Np).reshape(4, Np)
        dens[e] = bool(np.any(cb @ lim[0] < density_floor))
        out, s = routines.positivity(GAMMA, lim, Np, cb, density_floor, pressure_floor)
        scale[e] = s
        assert np.array_equal(out.reshape(4, Np), Lv[:, :, e]), f"element {e}"
    return flags, dens, scale, Ul
