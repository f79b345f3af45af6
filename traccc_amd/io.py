"""Measurement ingest in traccc's CSV format, over the C ABI (include/traccc_amd/io.h).

    meas, new_index = read_measurements_csv(path, do_sort=True)
    write_measurements_csv(path, event)

mirrors traccc::io::csv::read_measurements (io/src/csv/read_measurements.cpp:20-99):
columns by header name, local_key as one raw byte selecting loc0 / loc1, and
with do_sort the traccc measurement ordering (surface, loc0, loc1, var0,
var1) plus the old-row -> sorted-row map it returns.  Host only.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _ffi as F
from .ckf import _check, _lib as _ckf_lib


def _lib() -> C.CDLL:
    lib = _ckf_lib()
    if getattr(lib, "_io_typed", False):
        return lib
    vp = C.c_void_p
    lib.ckf_read_measurements_csv.argtypes = [C.c_char_p, C.c_int, C.POINTER(vp)]
    lib.ckf_measurement_buffer_view.argtypes = [vp, C.POINTER(F.Measurements),
                                                C.POINTER(F.c_u32_p)]
    lib.ckf_measurement_buffer_destroy.argtypes = [vp]
    lib.ckf_write_measurements_csv.argtypes = [C.c_char_p, C.POINTER(F.Measurements)]
    for name in ("ckf_read_measurements_csv", "ckf_measurement_buffer_view",
                 "ckf_measurement_buffer_destroy", "ckf_write_measurements_csv"):
        getattr(lib, name).restype = C.c_int
    lib._io_typed = True
    return lib


@dataclass
class Measurements:
    """traccc measurement_collection columns (host copy)."""
    local_position: np.ndarray   # [n][2] float32
    local_variance: np.ndarray   # [n][2] float32
    dimensions: np.ndarray       # [n] uint32
    time: np.ndarray             # [n] float32
    identifier: np.ndarray       # [n] uint32 (the file's measurement_id)
    surface_link: np.ndarray     # [n] uint64
    subspace: np.ndarray         # [n][2] uint8

    @property
    def n_meas(self) -> int:
        return len(self.dimensions)

    def measurements_struct(self) -> F.Measurements:
        m = F.Measurements()
        m.size = self.n_meas
        m.local_position = F.ptr(self.local_position, C.c_float)
        m.local_variance = F.ptr(self.local_variance, C.c_float)
        m.dimensions = F.ptr(self.dimensions, C.c_uint32)
        m.time = F.ptr(self.time, C.c_float)
        m.identifier = F.ptr(self.identifier, C.c_uint32)
        m.surface_link = F.ptr(self.surface_link, C.c_uint64)
        m.subspace = F.ptr(self.subspace, C.c_uint8)
        return m


def read_measurements_csv(path: str, do_sort: bool = True) -> tuple[Measurements, np.ndarray]:
    """(measurements, new_index) where new_index[old row] = row after sorting."""
    lib = _lib()
    h = C.c_void_p()
    _check(lib.ckf_read_measurements_csv(str(path).encode(), int(bool(do_sort)), C.byref(h)))
    try:
        v = F.Measurements()
        idx = F.c_u32_p()
        _check(lib.ckf_measurement_buffer_view(h, C.byref(v), C.byref(idx)))
        n = v.size
        m = Measurements(
            local_position=F.np_from(v.local_position, 2 * n, np.float32, (n, 2)),
            local_variance=F.np_from(v.local_variance, 2 * n, np.float32, (n, 2)),
            dimensions=F.np_from(v.dimensions, n, np.uint32),
            time=F.np_from(v.time, n, np.float32),
            identifier=F.np_from(v.identifier, n, np.uint32),
            surface_link=F.np_from(v.surface_link, n, np.uint64),
            subspace=F.np_from(v.subspace, 2 * n, np.uint8, (n, 2)))
        new_index = F.np_from(idx, n, np.uint32)
    finally:
        lib.ckf_measurement_buffer_destroy(h)
    return m, new_index


def write_measurements_csv(path: str, measurements) -> None:
    """Write any object with measurements_struct() (an Event, Measurements)."""
    m = measurements.measurements_struct()
    _check(_lib().ckf_write_measurements_csv(str(path).encode(), C.byref(m)))
