"""Python mirror of the serving C ABI (include/wasl_hip_serve.h): many
resident cases advanced together by batched stage kernels, with continuous
admission as device slots and memory free up."""

from __future__ import annotations

import ctypes as C

import numpy as np

from .engine import (MIXED, EULER, Mesh, Numerics, SimulationError, WaslHipError, library,
                     _check)

_P = C.c_void_p
_d = C.c_double
_u64 = C.c_uint64
_i64 = C.c_int64


class ServerConfig(C.Structure):
    _fields_ = [
        ("precision", C.c_int32),
        ("max_resident", C.c_int32),
        ("memory_budget", C.c_uint64),
        ("steps_per_batch", C.c_int32),
        ("use_graphs", C.c_int32),
        ("setup_threads", C.c_int32),
        ("fp64_max_order", C.c_int32),
        ("fp64_limited", C.c_int32),
        ("fp64_tile", C.c_int32 * 8),
    ]


_bound = False


def _lib() -> C.CDLL:
    global _bound
    lib = library()
    if not _bound:
        lib.wasl_hip_server_config_default.argtypes = [C.POINTER(ServerConfig)]
        lib.wasl_hip_server_config_default.restype = None
        lib.wasl_hip_server_create.argtypes = [C.POINTER(ServerConfig), C.POINTER(_P)]
        lib.wasl_hip_server_create.restype = C.c_int
        lib.wasl_hip_server_destroy.argtypes = [_P]
        lib.wasl_hip_server_destroy.restype = None
        lib.wasl_hip_server_submit.argtypes = [_P, _P, C.POINTER(Numerics), C.c_int, _P, _i64, _d,
                                               _d, _u64, _u64, C.POINTER(_u64)]
        lib.wasl_hip_server_submit.restype = C.c_int
        lib.wasl_hip_server_pump.argtypes = [_P, _P, _i64, C.POINTER(_i64)]
        lib.wasl_hip_server_pump.restype = C.c_int
        lib.wasl_hip_server_result.argtypes = [_P, _u64, _P, _i64, C.POINTER(_d), C.POINTER(_u64),
                                               C.POINTER(C.c_int)]
        lib.wasl_hip_server_result.restype = C.c_int
        lib.wasl_hip_server_state_size.argtypes = [_P, _u64]
        lib.wasl_hip_server_state_size.restype = _i64
        lib.wasl_hip_server_count.argtypes = [_P, C.c_int]
        lib.wasl_hip_server_count.restype = _i64
        _bound = True
    return lib


def default_fp64_tiles() -> dict:
    """{order: (stage_kernel, lanes_per_element)} of the served FP64 2D Euler
    classes in the library's default configuration; (0, 0) is the engine's
    autotuned tile."""
    cfg = ServerConfig()
    _lib().wasl_hip_server_config_default(C.byref(cfg))
    return {p: (cfg.fp64_tile[2 * p], cfg.fp64_tile[2 * p + 1]) for p in range(4)}


class Server:
    """A serving engine. submit() queues a case; pump() runs one scheduling
    round and returns the ids finished in it; result() hands back a finished
    case (final state, time, step) and raises its error if it failed."""

    def __init__(self, precision: int = MIXED, max_resident: int = 256, memory_budget: int = 0,
                 steps_per_batch: int = 16, use_graphs: bool = True, setup_threads: int = 4,
                 fp64_max_order: int = 1, fp64_limited: bool = True, fp64_tiles=None):
        lib = _lib()
        cfg = ServerConfig()
        lib.wasl_hip_server_config_default(C.byref(cfg))
        cfg.precision = precision
        cfg.max_resident = max_resident
        cfg.memory_budget = memory_budget
        cfg.steps_per_batch = steps_per_batch
        cfg.use_graphs = int(bool(use_graphs))
        cfg.setup_threads = setup_threads
        cfg.fp64_max_order = fp64_max_order
        cfg.fp64_limited = int(bool(fp64_limited))
        # {order: (stage_kernel, lanes_per_element)} for FP64 2D Euler classes
        for order, (kernel, lanes) in (fp64_tiles or {}).items():
            cfg.fp64_tile[2 * order] = kernel
            cfg.fp64_tile[2 * order + 1] = lanes
        h = _P()
        _check(lib.wasl_hip_server_create(C.byref(cfg), C.byref(h)))
        self.h = h.value

    def close(self) -> None:
        if getattr(self, "h", None):
            _lib().wasl_hip_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def submit(self, mesh: Mesh, equation: int, order: int, ic: int, params=(), cfl: float = 0.15,
               end_time: float = 0.0, end_steps: int = 0, max_steps: int = 10 ** 9,
               gamma: float = 1.4, wavespeed: float = 1.0, boundary=(0, 0),
               boundary_values=(0.0, 0.0), limiting: bool = False,
               sensor_threshold: float = 1e-3, density_floor: float = 1e-10,
               pressure_floor: float = 1e-10) -> int:
        lib = _lib()
        n = Numerics()
        library().wasl_hip_numerics_default(C.byref(n))
        n.equation = equation
        n.polynomial_order = order
        n.heat_capacity_ratio = gamma
        n.wavespeed = wavespeed
        n.boundary_left, n.boundary_right = boundary
        n.boundary_value_left, n.boundary_value_right = boundary_values
        n.limiting = int(bool(limiting))
        n.sensor_threshold = sensor_threshold
        n.density_floor = density_floor
        n.pressure_floor = pressure_floor
        prm = np.ascontiguousarray(params, np.float64)
        rid = _u64()
        _check(lib.wasl_hip_server_submit(self.h, mesh.h, C.byref(n), ic,
                                          prm.ctypes.data if prm.size else None, prm.size, cfl,
                                          end_time, end_steps, max_steps, C.byref(rid)))
        return rid.value

    def pump(self, capacity: int = 4096) -> list[int]:
        out = np.zeros(capacity, np.uint64)
        cnt = _i64()
        _check(_lib().wasl_hip_server_pump(self.h, out.ctypes.data, capacity, C.byref(cnt)))
        return [int(x) for x in out[: cnt.value]]

    def result(self, rid: int):
        """(state, time, step) of a finished request; raises on its failure."""
        lib = _lib()
        n = lib.wasl_hip_server_state_size(self.h, rid)
        state = np.empty(max(n, 0), np.float64)
        t, s, st = _d(), _u64(), C.c_int()
        _check(lib.wasl_hip_server_result(self.h, rid, state.ctypes.data if n > 0 else None,
                                          max(n, 0), C.byref(t), C.byref(s), C.byref(st)))
        if st.value != 0:
            msg = lib.wasl_hip_last_error().decode()
            raise (SimulationError if st.value == -2 else WaslHipError)(st.value, msg)
        return state, t.value, s.value

    def count(self, what: int) -> int:
        return int(_lib().wasl_hip_server_count(self.h, what))

    @property
    def queued(self) -> int:
        return self.count(0)

    @property
    def preparing(self) -> int:
        return self.count(1)

    @property
    def resident(self) -> int:
        return self.count(2)

    @property
    def idle(self) -> bool:
        return self.count(0) == 0 and self.count(1) == 0 and self.count(2) == 0

    def drain(self, ids=None) -> dict:
        """Pump until every submitted request has finished; returns
        {id: (state, time, step) or the exception}."""
        out = {}
        while not self.idle or self.count(3):
            for rid in self.pump():
                try:
                    out[rid] = self.result(rid)
                except WaslHipError as exc:
                    out[rid] = exc
        return out


__all__ = ["Server", "ServerConfig", "EULER"]
