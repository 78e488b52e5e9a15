"""Python mirror of the engine's C ABI (include/wasl_hip.h) over ctypes.

Names follow the reference's operator interface: residual()
(euler_triangle_residual / scalar_advection_residual), limit()
(limit_euler_modal), stable_time_step() (euler_cfl_timestep /
advection_cfl_timestep), step() (ssp_rk3_step), advance() (the stepping loop
of run_serial_case), mass(), cell_averages().

The shared library is required: every entry point raises when
wasl_amd/libwasl_hip.so is missing; there is no CPU fallback.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwasl_hip.so")

PERIODIC, INFLOW, OUTFLOW = 0, 1, 2
SCALAR_ADVECTION, EULER = 0, 1
FP64, MIXED, BF16 = 0, 1, 2
IC_EULER_REGIONS, IC_VORTEX, IC_RAREFACTION, IC_SMOOTH = 0, 1, 2, 3
IC_SCALAR_PULSE, IC_SCALAR_REGIONS, IC_SCALAR_CIRCLES, IC_SCALAR_WAVE = 10, 11, 12, 13

OK, ERROR, SIMULATION, DEVICE = 0, -1, -2, -3


class WaslHipError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


class SimulationError(WaslHipError):
    """A numerical failure the reference throws for (same message text)."""


class Numerics(C.Structure):
    _fields_ = [
        ("equation", C.c_int32),
        ("polynomial_order", C.c_int32),
        ("heat_capacity_ratio", C.c_double),
        ("wavespeed", C.c_double),
        ("boundary_left", C.c_int32),
        ("boundary_right", C.c_int32),
        ("boundary_value_left", C.c_double),
        ("boundary_value_right", C.c_double),
        ("limiting", C.c_int32),
        ("sensor_threshold", C.c_double),
        ("density_floor", C.c_double),
        ("pressure_floor", C.c_double),
        ("precision", C.c_int32),
        ("lanes_per_element", C.c_int32),
        ("stage_kernel", C.c_int32),
        ("velocity_y", C.c_double),
    ]


_lib = None
_P = C.c_void_p
_i64 = C.c_int64
_d = C.c_double
_pd = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_pi64 = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_pi32 = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_pu8 = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


def library() -> C.CDLL:
    """Load libwasl_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with "
                          "`python -m wasl_amd.build` (hipcc, gfx950)")
    lib = C.CDLL(LIB_PATH)
    lib.wasl_hip_last_error.restype = C.c_char_p
    sig = {
        "wasl_hip_mesh_build": [C.c_int, _i64, _pi64, _pd, _i64, _pi64, C.POINTER(_P)],
        "wasl_hip_mesh_from_prepared": [C.c_int, _i64, _pi64, _pd, _i64, _pi64, _i64, _pi64,
                                        _pi32, _pi32, _pd, C.POINTER(_P)],
        "wasl_hip_mesh_export": [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P],
        "wasl_hip_mesh_face_vertices": [_P, _pi64],
        "wasl_hip_mesh_periodic_trace": [_P, C.c_int, _pi32, _pd],
        "wasl_hip_mesh_limiter_stencil": [_P, C.c_int, C.c_int, _pi32, _pd, _pd, _pd],
        "wasl_hip_mesh_project": [_P, C.c_int, _d, C.c_int, _pd, _i64, _pd],
        "wasl_hip_engine_create": [_P, C.POINTER(Numerics), C.POINTER(_P)],
        "wasl_hip_engine_trace": [_P, _pi32, _pd],
        "wasl_hip_engine_stencil": [_P, _pi32, _pd, _pd, _pd],
        "wasl_hip_project": [_P, C.c_int, _pd, _i64, _pd],
        "wasl_hip_upload_state": [_P, _pd],
        "wasl_hip_download_state": [_P, _pd],
        "wasl_hip_residual": [_P, _pd, _pd],
        "wasl_hip_limit": [_P, _P],
        "wasl_hip_stable_dt": [_P, _d, C.POINTER(_d)],
        "wasl_hip_step": [_P, _d],
        "wasl_hip_advance": [_P, _d, _d, C.c_uint64, C.c_uint64, C.POINTER(_d),
                             C.POINTER(C.c_uint64), _P, _i64, C.POINTER(_i64)],
        "wasl_hip_mass": [_P, C.POINTER(_d)],
        "wasl_hip_cell_averages": [_P, _pd],
        "wasl_hip_last_timing": [_P, C.POINTER(_d), C.POINTER(_d), C.POINTER(_i64)],
        "wasl_hip_set_timing": [_P, C.c_int],
        "wasl_hip_capture_graphs": [_P],
        "wasl_hip_time_stages": [_P, C.c_int, _d, _pd],
        "wasl_hip_fp64_ceilings": [C.POINTER(_d), C.POINTER(_d)],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.wasl_hip_mesh_count.argtypes = [_P, C.c_int]
    lib.wasl_hip_mesh_count.restype = _i64
    lib.wasl_hip_engine_size.argtypes = [_P, C.c_int]
    lib.wasl_hip_engine_size.restype = _i64
    lib.wasl_hip_mesh_destroy.argtypes = [_P]
    lib.wasl_hip_mesh_destroy.restype = None
    lib.wasl_hip_engine_destroy.argtypes = [_P]
    lib.wasl_hip_engine_destroy.restype = None
    lib.wasl_hip_numerics_default.argtypes = [C.POINTER(Numerics)]
    lib.wasl_hip_numerics_default.restype = None
    _lib = lib
    return lib


def _check(rc: int) -> None:
    if rc != OK:
        msg = library().wasl_hip_last_error().decode()
        raise (SimulationError if rc == SIMULATION else WaslHipError)(rc, msg)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class Mesh:
    """Canonical mesh (build_triangle_mesh / build_line_mesh)."""

    def __init__(self, handle: int, dimension: int):
        self.h = handle
        self.dimension = dimension
        lib = library()
        self.elements = lib.wasl_hip_mesh_count(handle, 0)
        self.faces = lib.wasl_hip_mesh_count(handle, 1)
        self.boundary_faces = lib.wasl_hip_mesh_count(handle, 2)
        self.nodes = lib.wasl_hip_mesh_count(handle, 3)

    def __del__(self):
        if getattr(self, "h", None):
            library().wasl_hip_mesh_destroy(self.h)
            self.h = None

    @classmethod
    def triangles(cls, node_tags, node_xy, triangles) -> "Mesh":
        tags = np.ascontiguousarray(node_tags, np.int64)
        xy = np.asarray(node_xy, np.float64).reshape(-1, 2)
        xyz = np.ascontiguousarray(np.concatenate([xy, np.zeros((xy.shape[0], 1))], axis=1))
        tri = np.ascontiguousarray(triangles, np.int64).reshape(-1)
        h = _P()
        _check(library().wasl_hip_mesh_build(2, tags.size, tags, xyz.reshape(-1), tri.size // 3,
                                             tri, C.byref(h)))
        return cls(h.value, 2)

    @classmethod
    def lines(cls, node_tags, node_x, segments) -> "Mesh":
        tags = np.ascontiguousarray(node_tags, np.int64)
        x = np.asarray(node_x, np.float64)
        xyz = np.ascontiguousarray(np.stack([x, np.zeros_like(x), np.zeros_like(x)], axis=1))
        seg = np.ascontiguousarray(segments, np.int64).reshape(-1)
        h = _P()
        _check(library().wasl_hip_mesh_build(1, tags.size, tags, xyz.reshape(-1), seg.size // 2,
                                             seg, C.byref(h)))
        return cls(h.value, 1)

    @classmethod
    def from_input(cls, mesh_input) -> "Mesh":
        """From a wasl_amd.meshgen input (triangles or lines)."""
        if hasattr(mesh_input, "triangles"):
            return cls.triangles(mesh_input.node_tags, mesh_input.node_xy, mesh_input.triangles)
        return cls.lines(mesh_input.node_tags, mesh_input.node_x, mesh_input.segments)

    @classmethod
    def from_prepared(cls, prepared: dict) -> "Mesh":
        d = prepared["dimension"]
        tags = np.ascontiguousarray(prepared["node_tags"], np.int64)
        pts = np.ascontiguousarray(prepared["points"], np.float64).reshape(-1)
        en = np.ascontiguousarray(prepared["element_nodes"], np.int64).reshape(-1)
        fnodes = np.ascontiguousarray(prepared["face_nodes"], np.int64).reshape(-1)
        fe = np.ascontiguousarray(prepared["face_elements"], np.int32).reshape(-1)
        fl = np.ascontiguousarray(prepared["face_local_edges"], np.int32).reshape(-1)
        fn = np.ascontiguousarray(prepared["face_normals"], np.float64).reshape(-1)
        h = _P()
        _check(library().wasl_hip_mesh_from_prepared(d, tags.size, tags, pts, en.size // (d + 1),
                                                     en, fe.size // 2, fnodes, fe, fl, fn,
                                                     C.byref(h)))
        return cls(h.value, d)

    def prepared(self) -> dict:
        """PreparedMesh arrays of the canonical mesh (prepare_mesh)."""
        d, E, F, N = self.dimension, self.elements, self.faces, self.nodes
        npe = d + 1
        out = dict(
            dimension=d,
            node_tags=np.zeros(N, np.int64), points=np.zeros((N, 3)),
            element_nodes=np.zeros((E, npe), np.int64),
            element_metrics=np.zeros((E, 5 if d == 2 else 1)),
            face_nodes=np.zeros((F, 2), np.int64), face_elements=np.zeros((F, 2), np.int32),
            face_local_edges=np.zeros((F, 2), np.int32),
            face_normals=np.zeros((F, d) if d == 2 else F),
            face_measures=np.zeros(F))
        keys = ("node_tags", "points", "element_nodes", "element_metrics", "face_nodes",
                "face_elements", "face_local_edges", "face_normals", "face_measures")
        _check(library().wasl_hip_mesh_export(self.h, *[_ptr(out[k]) for k in keys]))
        return out

    def face_vertices(self) -> np.ndarray:
        v = np.zeros((self.faces, 4), np.int64)
        _check(library().wasl_hip_mesh_face_vertices(self.h, v.reshape(-1)))
        return v

    # Host-side setup (no device): the reference's periodic_triangle_trace,
    # make_euler_limiter_stencil and initial projection.
    def periodic_trace(self, order: int):
        P, Np = order + 1, (order + 1) * (order + 2) // 2
        d = np.zeros(self.faces * P, np.int32)
        b = np.zeros(self.faces * P * Np)
        _check(library().wasl_hip_mesh_periodic_trace(self.h, order, d, b))
        return d.reshape(self.faces, P), b.reshape(self.faces, P, Np)

    def limiter_stencil(self, order: int, periodic: bool = True):
        Np = (order + 1) * (order + 2) // 2
        C_ = (order + 1) ** 2 + 3 * (order + 1)
        E = self.elements
        nb = np.zeros(E * 3, np.int32)
        tr = np.zeros(E * 3 * Np * Np)
        nrm = np.zeros(E * 2)
        cb = np.zeros(C_ * Np)
        _check(library().wasl_hip_mesh_limiter_stencil(self.h, order, int(periodic), nb, tr, nrm,
                                                       cb))
        return nb.reshape(E, 3), tr.reshape(E, 3, Np, Np), nrm.reshape(E, 2), cb.reshape(C_, Np)

    def project(self, order: int, ic: int, params=(), gamma: float = 1.4) -> np.ndarray:
        prm = np.ascontiguousarray(np.asarray(params, np.float64).ravel())
        count = prm.size
        if count == 0:
            prm = np.zeros(1)
        V = 1 if ic >= IC_SCALAR_PULSE else 4
        if self.dimension == 2:
            n = V * (order + 1) * (order + 2) // 2 * self.elements
        else:
            n = V * (order + 1) * self.elements
        out = np.zeros(n)
        _check(library().wasl_hip_mesh_project(self.h, order, gamma, ic, prm, count, out))
        return out


def fp64_ceilings() -> tuple[float, float]:
    """(vector ALU, matrix core) FP64 TFLOP/s that the current device sustains
    (wasl_hip_fp64_ceilings)."""
    valu, mfma = C.c_double(0.0), C.c_double(0.0)
    _check(library().wasl_hip_fp64_ceilings(C.byref(valu), C.byref(mfma)))
    return valu.value, mfma.value


class Engine:
    """One resident case on the GPU."""

    def __init__(self, mesh: Mesh, equation: int = EULER, order: int = 2, gamma: float = 1.4,
                 wavespeed: float = 0.0, boundary=(PERIODIC, PERIODIC),
                 boundary_values=(0.0, 0.0), limiting: bool = False,
                 sensor_threshold: float = 1e-3, density_floor: float = 1e-10,
                 pressure_floor: float = 1e-10, precision: int = FP64, lanes: int = 0,
                 stage_kernel: int = 0, velocity_y: float = 0.0):
        lib = library()
        n = Numerics()
        lib.wasl_hip_numerics_default(C.byref(n))
        n.equation = equation
        n.polynomial_order = order
        n.heat_capacity_ratio = gamma
        n.wavespeed = wavespeed
        n.boundary_left, n.boundary_right = boundary
        n.boundary_value_left, n.boundary_value_right = boundary_values
        n.limiting = int(limiting)
        n.sensor_threshold = sensor_threshold
        n.density_floor = density_floor
        n.pressure_floor = pressure_floor
        n.precision = precision
        n.lanes_per_element = lanes
        n.stage_kernel = stage_kernel
        n.velocity_y = velocity_y
        self.numerics = n
        h = _P()
        _check(lib.wasl_hip_engine_create(mesh.h, C.byref(n), C.byref(h)))
        self.h = h.value
        self.E = lib.wasl_hip_engine_size(self.h, 0)
        self.Np = lib.wasl_hip_engine_size(self.h, 1)
        self.V = lib.wasl_hip_engine_size(self.h, 2)
        self.F = lib.wasl_hip_engine_size(self.h, 4)
        self.lanes = lib.wasl_hip_engine_size(self.h, 6)
        self.stage_kernel = lib.wasl_hip_engine_size(self.h, 8)
        self.order = order
        self.dimension = mesh.dimension

    def __del__(self):
        if getattr(self, "h", None):
            library().wasl_hip_engine_destroy(self.h)
            self.h = None

    @property
    def state_size(self) -> int:
        return int(self.V * self.Np * self.E)

    @property
    def device_bytes(self) -> int:
        return int(library().wasl_hip_engine_size(self.h, 5))

    @property
    def boundary_faces(self) -> int:
        """Faces kept as boundary faces: transmissive, or periodic with a
        donor looked up per Gauss point."""
        return int(library().wasl_hip_engine_size(self.h, 9))

    @property
    def periodic_pairs(self) -> int:
        """Conforming periodic face pairs recorded as interior faces."""
        return int(library().wasl_hip_engine_size(self.h, 10))

    def trace(self):
        P = self.order + 1
        d = np.zeros(self.F * P, np.int32)
        b = np.zeros(self.F * P * self.Np)
        _check(library().wasl_hip_engine_trace(self.h, d, b))
        return d.reshape(self.F, P), b.reshape(self.F, P, self.Np)

    def stencil(self):
        C_ = library().wasl_hip_engine_size(self.h, 7)
        nb = np.zeros(self.E * 3, np.int32)
        tr = np.zeros(self.E * 3 * self.Np * self.Np)
        nrm = np.zeros(self.E * 2)
        cb = np.zeros(max(C_, 1) * self.Np)
        _check(library().wasl_hip_engine_stencil(self.h, nb, tr, nrm, cb))
        return (nb.reshape(self.E, 3), tr.reshape(self.E, 3, self.Np, self.Np),
                nrm.reshape(self.E, 2), cb[: C_ * self.Np].reshape(C_, self.Np))

    def project(self, ic: int, params=()) -> np.ndarray:
        prm = np.ascontiguousarray(np.asarray(params, np.float64).ravel())
        count = prm.size
        if count == 0:
            prm = np.zeros(1)
        out = np.zeros(self.state_size)
        _check(library().wasl_hip_project(self.h, ic, prm, count, out))
        return out

    @property
    def state(self) -> np.ndarray:
        out = np.zeros(self.state_size)
        _check(library().wasl_hip_download_state(self.h, out))
        return out

    @state.setter
    def state(self, U) -> None:
        U = np.ascontiguousarray(U, np.float64).ravel()
        if U.size != self.state_size:
            raise ValueError(f"state has {U.size} values, expected {self.state_size}")
        _check(library().wasl_hip_upload_state(self.h, U))

    def residual(self, U) -> np.ndarray:
        U = np.ascontiguousarray(U, np.float64).ravel()
        if U.size != self.state_size:
            raise ValueError(f"state has {U.size} values, expected {self.state_size}")
        out = np.zeros_like(U)
        _check(library().wasl_hip_residual(self.h, U, out))
        return out

    def limit(self) -> np.ndarray:
        flags = np.zeros(self.E, np.uint8)
        _check(library().wasl_hip_limit(self.h, _ptr(flags)))
        return flags

    def stable_time_step(self, cfl: float) -> float:
        dt = _d()
        _check(library().wasl_hip_stable_dt(self.h, cfl, C.byref(dt)))
        return dt.value

    def step(self, dt: float) -> None:
        _check(library().wasl_hip_step(self.h, dt))

    def advance(self, cfl: float, end_time: float, max_steps: int = 2 ** 62,
                history_interval: int = 0, time: float = 0.0, step: int = 0,
                history_capacity: int = 0, stop_at_max_steps: bool = False):
        """Device-driven stepping loop; returns (time, step, history rows).

        Like run_serial_case, reaching max_steps before end_time raises
        SimulationError unless stop_at_max_steps is set (fixed step counts)."""
        t = _d(time)
        s = C.c_uint64(step)
        rows = _i64(0)
        hist = np.zeros(max(history_capacity, 1) * 4) if history_capacity else None
        rc = library().wasl_hip_advance(self.h, cfl, end_time, max_steps, history_interval,
                                        C.byref(t), C.byref(s), _ptr(hist), history_capacity,
                                        C.byref(rows))
        if not (stop_at_max_steps and rc == SIMULATION and s.value == max_steps
                and b"max_steps" in library().wasl_hip_last_error()):
            _check(rc)
        h = hist[: 4 * min(rows.value, history_capacity)].reshape(-1, 4) if hist is not None \
            else None
        return t.value, s.value, h

    def mass(self) -> float:
        m = _d()
        _check(library().wasl_hip_mass(self.h, C.byref(m)))
        return m.value

    def cell_averages(self) -> np.ndarray:
        out = np.zeros(self.E * self.V)
        _check(library().wasl_hip_cell_averages(self.h, out))
        return out.reshape(self.E, -1)

    def capture_graphs(self) -> None:
        """Capture every step graph of advance() now (setup)."""
        _check(library().wasl_hip_capture_graphs(self.h))

    def time_stages(self, reps: int, dt: float) -> np.ndarray:
        """Average duration (us) of each stage kernel launched reps times back
        to back; leaves the resident state not meaningful."""
        out = np.zeros(3)
        _check(library().wasl_hip_time_stages(self.h, reps, dt, out))
        return out

    def set_timing(self, on: bool) -> None:
        _check(library().wasl_hip_set_timing(self.h, int(on)))

    def last_timing(self):
        w, s, n = _d(), _d(), _i64()
        _check(library().wasl_hip_last_timing(self.h, C.byref(w), C.byref(s), C.byref(n)))
        return w.value, s.value, n.value
