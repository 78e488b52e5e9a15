"""ctypes wrapper over the oracle C API (oracle_api.h).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the checker, never as the thing
measured on the GPU. Two backends implement the same API:

  "port"       oracle/liboracle.so, the project's C++ restatement;
  "reference"  oracle/_ref/libwaslref.so, the reference's own sources
               compiled by oracle/Makefile (present only where it was built);
  "reference_fast"  the same sources built for speed, only for bench.py's
               timed CPU baseline.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PATHS = {
    "port": os.path.join(HERE, "liboracle.so"),
    "reference": os.path.join(HERE, "_ref", "libwaslref.so"),
    # the same reference sources built -O3 -march=x86-64-v3 (AVX2 and FMA,
    # contraction on): the timed CPU baseline, never a parity checker
    "reference_fast": os.path.join(HERE, "_ref", "libwaslref_fast.so"),
}

PERIODIC, INFLOW, OUTFLOW = 0, 1, 2
IC_EULER_REGIONS, IC_VORTEX, IC_RAREFACTION, IC_SMOOTH = 0, 1, 2, 3
IC_SCALAR_PULSE, IC_SCALAR_REGIONS, IC_SCALAR_CIRCLES, IC_SCALAR_WAVE = 10, 11, 12, 13

_libs: dict[str, C.CDLL] = {}

_P = C.c_void_p
_i64 = C.c_int64
_d = C.c_double
_pd = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_pi64 = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_pi32 = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_pu8 = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_pu64 = C.POINTER(C.c_uint64)


def available(kind: str) -> bool:
    return os.path.exists(PATHS[kind])


def load(kind: str = "port") -> C.CDLL:
    if kind in _libs:
        return _libs[kind]
    path = PATHS[kind]
    if not os.path.exists(path):
        raise FileNotFoundError(f"oracle library {path} is not built (make -C oracle)")
    lib = C.CDLL(path)
    lib.orc_error.restype = C.c_char_p
    lib.orc_tri_euler_create.restype = _P
    lib.orc_tri_euler_create.argtypes = [_i64, _pi64, _pd, _i64, _pi64, C.c_int, _d, C.c_int]
    lib.orc_tri_scalar_create.restype = _P
    lib.orc_tri_scalar_create.argtypes = [_i64, _pi64, _pd, _i64, _pi64, C.c_int, _d, _d]
    lib.orc_line_euler_create.restype = _P
    lib.orc_line_euler_create.argtypes = [_i64, _pi64, _pd, _i64, _pi64, C.c_int, _d, C.c_int,
                                          C.c_int]
    lib.orc_line_scalar_create.restype = _P
    lib.orc_line_scalar_create.argtypes = [_i64, _pi64, _pd, _i64, _pi64, C.c_int, _d, C.c_int,
                                           _d, C.c_int, _d]
    lib.orc_destroy.argtypes = [_P]
    lib.orc_size.restype = _i64
    lib.orc_size.argtypes = [_P, C.c_int]
    lib.orc_tri_topology.argtypes = [_P, _pi64, _pd, _pi32, _pi32, _pi64, _pd]
    lib.orc_line_topology.argtypes = [_P, _pd, _pi32, _pd]
    lib.orc_tri_trace.argtypes = [_P, _pi32, _pd]
    lib.orc_stencil_check_count.restype = _i64
    lib.orc_stencil_check_count.argtypes = [_P]
    lib.orc_stencil.argtypes = [_P, _pi32, _pd, _pd, _pd]
    lib.orc_residual.argtypes = [_P, _pd, _pd]
    lib.orc_limit.argtypes = [_P, _d, _d, _d, _pd, _pu8]
    lib.orc_stable_dt.argtypes = [_P, _pd, _d, C.POINTER(_d)]
    lib.orc_step.argtypes = [_P, _d, C.c_int, _d, _d, _d, _pd]
    lib.orc_mass.argtypes = [_P, _pd, C.POINTER(_d)]
    lib.orc_project.argtypes = [_P, C.c_int, _pd, _i64, _pd]
    lib.orc_run.argtypes = [_P, _d, C.c_int, _d, _d, _d, _d, C.c_uint64, _d, C.c_uint64,
                            C.c_uint64, _pd, _pu64, C.POINTER(_d), _pd, _i64, C.POINTER(_i64)]
    lib.orc_basis.argtypes = [C.c_int, C.c_int, _d, _d, C.c_int, _pd, _pd, _pd]
    lib.orc_rule.argtypes = [C.c_int, C.c_int, C.c_int, _pd, _pd, _pd]
    lib.orc_euler_flux.argtypes = [_d, _pd, _pd, _pd]
    lib.orc_hllc.argtypes = [_d, _pd, _pd, _d, _d, _pd]
    lib.orc_conserved.argtypes = [_d, _pd, _pd]
    lib.orc_characteristic.argtypes = [_d, _pd, _pd, _d, _d, _pd, _pd]
    lib.orc_simple_weno.argtypes = [_pd, _pd, C.c_int, C.c_int, _pd]
    lib.orc_characteristic_weno.argtypes = [_d, _pd, _d, _d, _pd, _pd, C.c_int, C.c_int, _pd]
    lib.orc_positivity.argtypes = [_d, _pd, C.c_int, _pd, C.c_int, _d, _d, C.POINTER(_d)]
    lib.orc_modal_decay_troubled.argtypes = [_pd, C.c_int, _d, C.POINTER(C.c_int)]
    for name in ("orc_tri_topology", "orc_line_topology", "orc_tri_trace", "orc_stencil",
                 "orc_residual", "orc_limit", "orc_stable_dt", "orc_step", "orc_mass",
                 "orc_project", "orc_run", "orc_basis", "orc_rule", "orc_euler_flux",
                 "orc_hllc", "orc_conserved", "orc_characteristic", "orc_simple_weno",
                 "orc_characteristic_weno", "orc_positivity", "orc_modal_decay_troubled"):
        getattr(lib, name).restype = C.c_int
    _libs[kind] = lib
    return lib


class OracleError(RuntimeError):
    pass


class Routines:
    """Single reference routines of one backend (known-answer tests)."""

    def __init__(self, kind: str = "port"):
        self.lib = load(kind)

    def _check(self, rc):
        if rc != 0:
            raise OracleError(self.lib.orc_error().decode())

    @staticmethod
    def _a(x):
        return np.ascontiguousarray(x, np.float64).ravel()

    def legendre(self, xi, p, gradient=False):
        v, g, d = np.zeros(p + 1), np.zeros(p + 1), np.zeros(1)
        self._check(self.lib.orc_basis(0, p, xi, 0.0, int(gradient), v, g, d))
        return (v, g) if gradient else v

    def dubiner(self, xi, eta, p, gradient=False):
        n = (p + 1) * (p + 2) // 2
        v, gx, ge = np.zeros(n), np.zeros(n), np.zeros(n)
        self._check(self.lib.orc_basis(1, p, xi, eta, int(gradient), v, gx, ge))
        return (v, gx, ge) if gradient else v

    def rule(self, kind, n, edge=0):
        a, b, w = np.zeros(64), np.zeros(64), np.zeros(64)
        count = self.lib.orc_rule(kind, n, edge, a, b, w)
        if count < 0:
            raise OracleError(self.lib.orc_error().decode())
        return a[:count], b[:count], w[:count]

    def euler_flux(self, gamma, state):
        fx, fy = np.zeros(4), np.zeros(4)
        self._check(self.lib.orc_euler_flux(gamma, self._a(state), fx, fy))
        return fx, fy

    def hllc(self, gamma, left, right, nx, ny):
        f = np.zeros(4)
        self._check(self.lib.orc_hllc(gamma, self._a(left), self._a(right), nx, ny, f))
        return f

    def conserved(self, gamma, prim):
        s = np.zeros(4)
        self._check(self.lib.orc_conserved(gamma, self._a(prim), s))
        return s

    def characteristic(self, gamma, average, change, nx, ny):
        w, b = np.zeros(4), np.zeros(4)
        self._check(self.lib.orc_characteristic(gamma, self._a(average), self._a(change), nx,
                                                ny, w, b))
        return w, b

    def simple_weno(self, target, candidates, K, n):
        out = np.zeros(n)
        cand = self._a(candidates) if K else np.zeros(1)
        self._check(self.lib.orc_simple_weno(self._a(target), cand, K, n, out))
        return out

    def characteristic_weno(self, gamma, average, nx, ny, target, candidates, K, n):
        out = np.zeros(4 * n)
        cand = self._a(candidates) if K else np.zeros(1)
        self._check(self.lib.orc_characteristic_weno(gamma, self._a(average), nx, ny,
                                                     self._a(target), cand, K, n, out))
        return out

    def positivity(self, gamma, coefficients, n, basis, density_floor, pressure_floor):
        c = np.array(coefficients, np.float64).ravel()
        b = self._a(basis)
        s = _d()
        self._check(self.lib.orc_positivity(gamma, c, n, b, b.size // n, density_floor,
                                            pressure_floor, C.byref(s)))
        return c, s.value

    def modal_decay_troubled(self, coefficients, p, threshold):
        t = C.c_int()
        self._check(self.lib.orc_modal_decay_troubled(self._a(coefficients), p, threshold,
                                                      C.byref(t)))
        return bool(t.value)


class Case:
    """One oracle case (mesh + numerics), mirroring the C API."""

    def __init__(self, lib: C.CDLL, handle: int, dim: int):
        self.lib = lib
        self.h = handle
        self.dim = dim
        self.E = lib.orc_size(handle, 0)
        self.F = lib.orc_size(handle, 1)
        self.Np = lib.orc_size(handle, 2)
        self.V = lib.orc_size(handle, 3)
        self.P = lib.orc_size(handle, 4)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    def _check(self, rc: int):
        if rc != 0:
            raise OracleError(self.lib.orc_error().decode())

    @property
    def state_size(self) -> int:
        return int(self.V * self.Np * self.E)

    def topology(self):
        if self.dim == 2:
            en = np.zeros((self.E, 3), np.int64)
            exy = np.zeros((self.E, 6))
            fe = np.zeros((self.F, 2), np.int32)
            fd = np.zeros((self.F, 2), np.int32)
            fv = np.zeros((self.F, 4), np.int64)
            fn = np.zeros((self.F, 2))
            self._check(self.lib.orc_tri_topology(self.h, en, exy, fe, fd, fv, fn))
            return dict(element_nodes=en, element_xy=exy, face_elements=fe, face_edges=fd,
                        face_vertices=fv, face_normals=fn)
        ex = np.zeros((self.E, 2))
        fe = np.zeros((self.F, 2), np.int32)
        fn = np.zeros(self.F)
        self._check(self.lib.orc_line_topology(self.h, ex, fe, fn))
        return dict(element_x=ex, face_elements=fe, face_normals=fn)

    def trace(self):
        d = np.zeros(self.F * self.P, np.int32)
        b = np.zeros(self.F * self.P * self.Np)
        self._check(self.lib.orc_tri_trace(self.h, d, b))
        return d.reshape(self.F, self.P), b.reshape(self.F, self.P, self.Np)

    def stencil(self):
        K = 3 if self.dim == 2 else 2
        C_ = self.lib.orc_stencil_check_count(self.h)
        nb = np.zeros(self.E * K, np.int32)
        tr = np.zeros(self.E * K * self.Np * self.Np)
        nrm = np.zeros(self.E * 2)
        cb = np.zeros(max(C_, 1) * self.Np)
        self._check(self.lib.orc_stencil(self.h, nb, tr, nrm, cb))
        return (nb.reshape(self.E, K), tr.reshape(self.E, K, self.Np, self.Np),
                nrm.reshape(self.E, 2), cb[: C_ * self.Np].reshape(C_, self.Np))

    def residual(self, U):
        U = np.ascontiguousarray(U, np.float64).ravel()
        out = np.zeros_like(U)
        self._check(self.lib.orc_residual(self.h, U, out))
        return out

    def limit(self, U, threshold, density_floor=1e-10, pressure_floor=1e-10, partial=False):
        """Limited state and troubled flags. With partial=True a failure
        (a cell average below a floor) returns (state, flags, message): the
        state holds the elements limited before the reference threw, in
        ascending element order (euler_modal_limiter.cpp:81)."""
        U = np.array(U, np.float64).ravel()
        flags = np.zeros(self.E, np.uint8)
        rc = self.lib.orc_limit(self.h, threshold, density_floor, pressure_floor, U, flags)
        if partial:
            return U, flags, (self.lib.orc_error().decode() if rc != 0 else None)
        self._check(rc)
        return U, flags

    def stable_dt(self, U, cfl):
        U = np.ascontiguousarray(U, np.float64).ravel()
        dt = _d()
        self._check(self.lib.orc_stable_dt(self.h, U, cfl, C.byref(dt)))
        return dt.value

    def step(self, U, dt, limiting=False, threshold=1e-3, density_floor=1e-10,
             pressure_floor=1e-10):
        U = np.array(U, np.float64).ravel()
        self._check(self.lib.orc_step(self.h, dt, int(limiting), threshold, density_floor,
                                      pressure_floor, U))
        return U

    def mass(self, U):
        U = np.ascontiguousarray(U, np.float64).ravel()
        m = _d()
        self._check(self.lib.orc_mass(self.h, U, C.byref(m)))
        return m.value

    def project(self, ic, params=()):
        prm = np.ascontiguousarray(np.asarray(params, np.float64).ravel())
        if prm.size == 0:
            prm = np.zeros(1)
            n = 0
        else:
            n = prm.size
        U = np.zeros(self.state_size)
        self._check(self.lib.orc_project(self.h, ic, prm, n, U))
        return U

    def run(self, U, cfl, end_time, max_steps=10 ** 9, history_interval=10 ** 9, limiting=False,
            threshold=1e-3, density_floor=1e-10, pressure_floor=1e-10, start_time=0.0,
            start_step=0, history_capacity=4096):
        U = np.array(U, np.float64).ravel()
        step = C.c_uint64()
        time = _d()
        hist = np.zeros(history_capacity * 4)
        rows = _i64()
        self._check(self.lib.orc_run(self.h, cfl, int(limiting), threshold, density_floor,
                                     pressure_floor, start_time, start_step, end_time, max_steps,
                                     history_interval, U, C.byref(step), C.byref(time), hist,
                                     history_capacity, C.byref(rows)))
        n = min(rows.value, history_capacity)
        return U, step.value, time.value, hist[: 4 * n].reshape(n, 4)


def tri_euler(node_tags, node_xy, triangles, order, gamma=1.4, boundary=PERIODIC,
              kind="port") -> Case:
    lib = load(kind)
    tags = np.ascontiguousarray(node_tags, np.int64)
    xy = np.ascontiguousarray(node_xy, np.float64).reshape(-1)
    tri = np.ascontiguousarray(triangles, np.int64).reshape(-1)
    h = lib.orc_tri_euler_create(tags.size, tags, xy, tri.size // 3, tri, order, gamma, boundary)
    if not h:
        raise OracleError(lib.orc_error().decode())
    return Case(lib, h, 2)


def tri_scalar(node_tags, node_xy, triangles, order, velocity_x, velocity_y,
               kind="port") -> Case:
    """2D scalar advection on periodic triangles, state U[element][mode]."""
    lib = load(kind)
    tags = np.ascontiguousarray(node_tags, np.int64)
    xy = np.ascontiguousarray(node_xy, np.float64).reshape(-1)
    tri = np.ascontiguousarray(triangles, np.int64).reshape(-1)
    h = lib.orc_tri_scalar_create(tags.size, tags, xy, tri.size // 3, tri, order, velocity_x,
                                  velocity_y)
    if not h:
        raise OracleError(lib.orc_error().decode())
    return Case(lib, h, 2)


def line_euler(node_tags, node_x, segments, order, gamma=1.4, left=PERIODIC, right=PERIODIC,
               kind="port") -> Case:
    """1D Euler on a line mesh, state U[variable][mode][element]."""
    lib = load(kind)
    tags = np.ascontiguousarray(node_tags, np.int64)
    x = np.ascontiguousarray(node_x, np.float64)
    seg = np.ascontiguousarray(segments, np.int64).reshape(-1)
    h = lib.orc_line_euler_create(tags.size, tags, x, seg.size // 2, seg, order, gamma, left,
                                  right)
    if not h:
        raise OracleError(lib.orc_error().decode())
    return Case(lib, h, 1)


def line_scalar(node_tags, node_x, segments, order, wavespeed, left=(PERIODIC, 0.0),
                right=(PERIODIC, 0.0), kind="port") -> Case:
    lib = load(kind)
    tags = np.ascontiguousarray(node_tags, np.int64)
    x = np.ascontiguousarray(node_x, np.float64)
    seg = np.ascontiguousarray(segments, np.int64).reshape(-1)
    h = lib.orc_line_scalar_create(tags.size, tags, x, seg.size // 2, seg, order, wavespeed,
                                   left[0], left[1], right[0], right[1])
    if not h:
        raise OracleError(lib.orc_error().decode())
    return Case(lib, h, 1)


class ReferenceIO:
    """The reference's prepared-mesh and checkpoint files through
    oracle/_ref/libwaslref_io.so (oracle/ref_io.cpp); test infrastructure."""

    PATH = os.path.join(HERE, "_ref", "libwaslref_io.so")

    @classmethod
    def available(cls) -> bool:
        return os.path.exists(cls.PATH)

    def __init__(self):
        lib = C.CDLL(self.PATH)
        lib.orc_io_error.restype = C.c_char_p
        lib.orc_io_prepared_write.argtypes = [_i64, _pi64, _pd, _i64, _pi64, C.c_char_p]
        lib.orc_io_prepared_copy.argtypes = [C.c_char_p, C.c_char_p, _pi64]
        lib.orc_io_checkpoint_write.argtypes = [C.c_char_p, _d, C.c_uint64, C.c_uint64,
                                                C.c_uint64, C.c_uint64, C.c_char_p, _pd]
        lib.orc_io_checkpoint_read.argtypes = [C.c_char_p, C.POINTER(_d), _pu64,
                                               np.ctypeslib.ndpointer(np.uint64), _P, _i64,
                                               C.c_char_p, _i64]
        lib.orc_io_export.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p]
        lib.orc_io_write_history.argtypes = [C.c_char_p, _i64, _pd]
        lib.orc_io_write_pvd.argtypes = [C.c_char_p, _i64, _pd, C.c_char_p]
        for name in ("orc_io_prepared_write", "orc_io_prepared_copy", "orc_io_checkpoint_write",
                     "orc_io_checkpoint_read", "orc_io_export", "orc_io_write_history",
                     "orc_io_write_pvd"):
            getattr(lib, name).restype = C.c_int
        self.lib = lib

    def _check(self, rc):
        if rc != 0:
            raise OracleError(self.lib.orc_io_error().decode())

    def write_prepared(self, node_tags, node_xy, triangles, path):
        tags = np.ascontiguousarray(node_tags, np.int64)
        xy = np.ascontiguousarray(node_xy, np.float64).ravel()
        tri = np.ascontiguousarray(triangles, np.int64).ravel()
        self._check(self.lib.orc_io_prepared_write(tags.size, tags, xy, tri.size // 3, tri,
                                                   str(path).encode()))

    def copy_prepared(self, src, dst=""):
        """Reads src with the reference (and writes it to dst when given);
        returns (dimension, nodes, elements, faces)."""
        counts = np.zeros(4, np.int64)
        self._check(self.lib.orc_io_prepared_copy(str(src).encode(), str(dst).encode(), counts))
        return tuple(int(x) for x in counts)

    def write_checkpoint(self, path, time, step, V, Np, E, prepared_mesh, coefficients):
        c = np.ascontiguousarray(coefficients, np.float64).ravel()
        self._check(self.lib.orc_io_checkpoint_write(str(path).encode(), time, step, V, Np, E,
                                                     str(prepared_mesh).encode(), c))

    def read_checkpoint(self, path):
        t = _d()
        s = C.c_uint64()
        layout = np.zeros(4, np.uint64)
        name = C.create_string_buffer(4096)
        self._check(self.lib.orc_io_checkpoint_read(str(path).encode(), C.byref(t), C.byref(s),
                                                    layout, None, 0, name, 4096))
        n = int(layout[0] * layout[1] * layout[2])
        coeff = np.zeros(n)
        self._check(self.lib.orc_io_checkpoint_read(str(path).encode(), None, None, layout,
                                                    coeff.ctypes.data, n, None, 0))
        return {"time": t.value, "step": s.value, "variable_count": int(layout[0]),
                "mode_count": int(layout[1]), "element_count": int(layout[2]),
                "partition_rank_count": int(layout[3]), "prepared_mesh": name.value.decode(),
                "coefficients": coeff}

    def export(self, checkpoint_path, mesh_path, out):
        """The bridge's export_checkpoint with the reference's VTU writer."""
        self._check(self.lib.orc_io_export(str(checkpoint_path).encode(),
                                           str(mesh_path).encode(), str(out).encode()))

    def write_history(self, path, rows):
        r = np.ascontiguousarray(rows, np.float64).reshape(-1, 4)
        self._check(self.lib.orc_io_write_history(str(path).encode(), r.shape[0], r.ravel()))

    def write_pvd(self, path, snapshots):
        times = np.ascontiguousarray([t for t, _ in snapshots], np.float64)
        names = "\n".join(n for _, n in snapshots).encode()
        self._check(self.lib.orc_io_write_pvd(str(path).encode(), len(snapshots),
                                              times if times.size else np.zeros(1), names))
