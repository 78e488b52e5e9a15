"""Python mirror of the case layer's C ABI (include/wasl_hip_case.h).

The functions are the bridge's coarse calls (crates/wasl-bridge/include/
bridge.hpp:7-32) with the same argument order and packing: run_euler_case,
run_scalar_advection_case, prepare_mesh, inspect_prepared_mesh,
inspect_checkpoint, export_checkpoint. Each returns the bridge's summary
string or raises CaseError with the reference's message.
"""

from __future__ import annotations

import ctypes as C
import os
from typing import Sequence

import numpy as np

from .engine import Mesh, library

HERE = os.path.dirname(os.path.abspath(__file__))
CASE_LIB_PATH = os.path.join(HERE, "libwasl_hip_case.so")

_lib = None
_P = C.c_void_p
_d = C.c_double
_u64 = C.c_uint64
_sz = C.c_size_t
_str = C.c_char_p
_pstr = C.POINTER(C.c_char_p)


class CaseError(RuntimeError):
    """A failed coarse call; the message is the reference's text."""


def case_library() -> C.CDLL:
    """Load libwasl_hip_case.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    library()  # the engine library it links against
    if not os.path.exists(CASE_LIB_PATH):
        raise ImportError(f"{CASE_LIB_PATH} is missing: build it with "
                          "`python -m wasl_amd.build`")
    lib = C.CDLL(CASE_LIB_PATH)
    lib.wasl_hip_case_last_error.restype = C.c_char_p
    lib.wasl_hip_string_free.argtypes = [_P]
    lib.wasl_hip_string_free.restype = None
    sig = {
        "wasl_hip_run_euler_case": [
            _str, _str, _str, C.c_int, _d, _d, _d, C.c_int, C.c_int, _d, _d, _d, _d, _u64, _u64,
            _u64, _u64, _u64, _P, _sz, _P, _P, _P, _sz, C.POINTER(_P)],
        "wasl_hip_run_scalar_advection_case": [
            _str, _str, _str, C.c_int, C.c_int, C.c_int, _d, _d, _d, _d, _d, _u64, _u64, _u64,
            _u64, _u64, _d, C.c_int, _d, _P, _sz, _P, _sz, _P, _P, _P, _sz, C.POINTER(_P)],
        "wasl_hip_prepare_mesh": [_P, _str, C.POINTER(_P)],
        "wasl_hip_prepare_mesh_tagged": [_P, _P, _P, _sz, _P, _P, _P, _str, C.POINTER(_P)],
        "wasl_hip_inspect_prepared_mesh": [_str, C.POINTER(_P)],
        "wasl_hip_inspect_checkpoint": [_str, C.POINTER(_P)],
        "wasl_hip_export_checkpoint": [_str, _str, _str, C.POINTER(_P)],
        "wasl_hip_mesh_read_prepared": [_str, C.POINTER(_P)],
        "wasl_hip_checkpoint_read": [_str, C.POINTER(_d), C.POINTER(_u64), C.POINTER(_u64),
                                     C.POINTER(_u64), C.POINTER(_u64), _P],
        "wasl_hip_checkpoint_write": [_str, _d, _u64, _u64, _u64, _u64, _str, _P],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    _lib = lib
    return lib


def _check(status: int) -> None:
    if status != 0:
        raise CaseError(case_library().wasl_hip_case_last_error().decode())


def _text(fn, *args) -> str:
    out = _P()
    _check(fn(*args, C.byref(out)))
    try:
        return C.string_at(out.value).decode()
    finally:
        case_library().wasl_hip_string_free(out)


def _b(s) -> bytes:
    return os.fsencode(s) if s is not None else b""


def _doubles(values) -> tuple[np.ndarray, int]:
    a = np.ascontiguousarray(values if values is not None else [], np.float64).reshape(-1)
    return a, a.size


def _boundaries(names, kinds, values):
    names = list(names or [])
    kinds = np.ascontiguousarray(kinds if kinds is not None else [], np.uint8)
    values = np.ascontiguousarray(values if values is not None else [], np.float64)
    arr = (C.c_char_p * max(len(names), 1))(*[_b(n) for n in names])
    if kinds.size != len(names) or values.size != len(names):
        raise CaseError("boundaries: invalid boundary data")
    return arr, kinds, values, len(names)


def run_euler_case(mesh_path, output_path, restart_path, limiting: bool,
                   sensor_threshold: float, density_floor: float, pressure_floor: float,
                   dimensions: int, polynomial_order: int, heat_capacity_ratio: float, cfl: float,
                   start_time: float, end_time: float, max_steps: int, history_interval: int,
                   visualization_interval: int, checkpoint_interval: int, keep_checkpoints: int,
                   regions: Sequence[float], boundary_names=(), boundary_kinds=(),
                   boundary_values=()) -> str:
    """run_euler_case (crates/wasl-bridge/src/bridge.cpp:173-216): regions
    packed 10 per region (shape, lower, upper, cx, cy, r, rho, u, v, p)."""
    lib = case_library()
    reg, nreg = _doubles(regions)
    names, kinds, values, nb = _boundaries(boundary_names, boundary_kinds, boundary_values)
    return _text(lib.wasl_hip_run_euler_case, _b(mesh_path), _b(output_path), _b(restart_path),
                 int(bool(limiting)), sensor_threshold, density_floor, pressure_floor,
                 dimensions, polynomial_order, heat_capacity_ratio, cfl, start_time, end_time,
                 max_steps, history_interval, visualization_interval, checkpoint_interval,
                 keep_checkpoints, reg.ctypes.data, nreg, C.cast(names, _P),
                 kinds.ctypes.data, values.ctypes.data, nb)


def run_scalar_advection_case(mesh_path, output_path, restart_path, mpi: bool, dimensions: int,
                              polynomial_order: int, velocity_x: float, velocity_y: float,
                              cfl: float, start_time: float, end_time: float, max_steps: int,
                              history_interval: int, visualization_interval: int,
                              checkpoint_interval: int, keep_checkpoints: int,
                              initial_default: float, initial_sine: bool,
                              initial_sine_amplitude: float, intervals: Sequence[float] = (),
                              circles: Sequence[float] = (), boundary_names=(),
                              boundary_kinds=(), boundary_values=()) -> str:
    """run_scalar_advection_case (crates/wasl-bridge/src/bridge.cpp:132-171)."""
    lib = case_library()
    iv, niv = _doubles(intervals)
    ci, nci = _doubles(circles)
    names, kinds, values, nb = _boundaries(boundary_names, boundary_kinds, boundary_values)
    return _text(lib.wasl_hip_run_scalar_advection_case, _b(mesh_path), _b(output_path),
                 _b(restart_path), int(bool(mpi)), dimensions, polynomial_order, velocity_x,
                 velocity_y, cfl, start_time, end_time, max_steps, history_interval,
                 visualization_interval, checkpoint_interval, keep_checkpoints, initial_default,
                 int(bool(initial_sine)), initial_sine_amplitude, iv.ctypes.data, niv,
                 ci.ctypes.data, nci, C.cast(names, _P), kinds.ctypes.data, values.ctypes.data,
                 nb)


def prepare_mesh(mesh: Mesh, output_path, element_physical_tags=None, face_physical_tags=None,
                 physical_names=None) -> str:
    """prepare_mesh + write_prepared_mesh for a built mesh; physical_names is
    a list of (dimension, tag, name)."""
    lib = case_library()
    if element_physical_tags is None and face_physical_tags is None and not physical_names:
        return _text(lib.wasl_hip_prepare_mesh, mesh.h, _b(output_path))
    et = None if element_physical_tags is None else np.ascontiguousarray(
        element_physical_tags, np.int32)
    ft = None if face_physical_tags is None else np.ascontiguousarray(face_physical_tags, np.int32)
    if et is not None and et.size != mesh.elements:
        raise CaseError("prepare mesh: physical-tag count mismatch")
    if ft is not None and ft.size != mesh.faces:
        raise CaseError("prepare mesh: physical-tag count mismatch")
    physical_names = list(physical_names or [])
    dims = np.ascontiguousarray([d for d, _, _ in physical_names] or [0], np.int32)
    tags = np.ascontiguousarray([t for _, t, _ in physical_names] or [0], np.int32)
    names = (C.c_char_p * max(len(physical_names), 1))(*[_b(n) for _, _, n in physical_names])
    return _text(lib.wasl_hip_prepare_mesh_tagged, mesh.h,
                 None if et is None else et.ctypes.data, None if ft is None else ft.ctypes.data,
                 len(physical_names), dims.ctypes.data, tags.ctypes.data, C.cast(names, _P),
                 _b(output_path))


def inspect_prepared_mesh(mesh_path) -> str:
    return _text(case_library().wasl_hip_inspect_prepared_mesh, _b(mesh_path))


def inspect_checkpoint(checkpoint_path) -> str:
    return _text(case_library().wasl_hip_inspect_checkpoint, _b(checkpoint_path))


def export_checkpoint(checkpoint_path, mesh_path, output_path) -> str:
    return _text(case_library().wasl_hip_export_checkpoint, _b(checkpoint_path), _b(mesh_path),
                 _b(output_path))


def read_prepared_mesh(mesh_path) -> Mesh:
    """The canonical mesh stored in a prepared-mesh file."""
    h = _P()
    _check(case_library().wasl_hip_mesh_read_prepared(_b(mesh_path), C.byref(h)))
    return Mesh(h.value, int(library().wasl_hip_mesh_count(h.value, 4)))


def read_checkpoint(checkpoint_path) -> dict:
    """time, step, variable_count, mode_count, element_count and the
    coefficients [variable][mode][element] of a checkpoint file."""
    lib = case_library()
    t, s, v, m, e = _d(), _u64(), _u64(), _u64(), _u64()
    _check(lib.wasl_hip_checkpoint_read(_b(checkpoint_path), C.byref(t), C.byref(s), C.byref(v),
                                        C.byref(m), C.byref(e), None))
    coeff = np.empty(v.value * m.value * e.value, np.float64)
    _check(lib.wasl_hip_checkpoint_read(_b(checkpoint_path), None, None, None, None, None,
                                        coeff.ctypes.data))
    return {"time": t.value, "step": s.value, "variable_count": v.value, "mode_count": m.value,
            "element_count": e.value, "coefficients": coeff}


def write_checkpoint(checkpoint_path, time: float, step: int, variable_count: int,
                     mode_count: int, element_count: int, prepared_mesh,
                     coefficients) -> None:
    c = np.ascontiguousarray(coefficients, np.float64).reshape(-1)
    if c.size != variable_count * mode_count * element_count:
        raise CaseError("checkpoint: coefficient size does not match declared layout")
    _check(case_library().wasl_hip_checkpoint_write(
        _b(checkpoint_path), time, step, variable_count, mode_count, element_count,
        _b(prepared_mesh), c.ctypes.data))
