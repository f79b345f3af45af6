"""Python host mirror of traccc's greedy ambiguity resolution, over the C ABI.

    resolver = greedy_ambiguity_resolution_algorithm(ambiguity_resolution_config())
    keep = resolver(offsets, measurement_ids, pvals)      # host arrays
    keep = resolver.run_device(d_offsets, d_ids, d_pvals, id_bound)  # torch tensors in HBM
    keep = resolver.run_fitted(devents, fitted)           # ckf_fit_tracks output

mirrors traccc::host::greedy_ambiguity_resolution_algorithm
(core/include/traccc/ambiguity_resolution/greedy_ambiguity_resolution_algorithm.hpp:26-60,
core/src/ambiguity_resolution/greedy_ambiguity_resolution_algorithm.cpp:19-254)
and its CUDA port (device/cuda/src/ambiguity_resolution/greedy_ambiguity_resolution_algorithm.cu):
an ambiguity_resolution_config at construction; the operator takes tracks
(their p-values and measurement identifiers) and keeps a subset.  Here the
operator returns the indices of the kept tracks, in input order (the host
reference's order).  max_shared_meas = 0 raises ValueError (the reference
asserts it is positive).  The HIP library is required: there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _ffi as F
from .ckf import DeviceEvents, _check, _lib as _ckf_lib


class AmbiguityConfig(C.Structure):
    """include/traccc_amd/ambiguity.h ckf_ambiguity_config."""
    _fields_ = [
        ("min_meas_per_track", C.c_uint32),
        ("max_iterations", C.c_uint32),
        ("max_shared_meas", C.c_uint32),
    ]


class ResolverTracks(C.Structure):
    _fields_ = [
        ("n_tracks", C.c_uint32),
        ("constituent_offsets", C.c_void_p),
        ("measurement_id", C.c_void_p),
        ("pval", C.c_void_p),
        ("id_bound", C.c_uint32),
    ]


class ResolverStats(C.Structure):
    _fields_ = [
        ("n_tracks", C.c_uint32),
        ("n_candidates", C.c_uint32),
        ("n_removed", C.c_uint32),
        ("n_accepted", C.c_uint32),
        ("n_rounds", C.c_uint32),
        ("sequential", C.c_uint32),
    ]


def _lib() -> C.CDLL:
    lib = _ckf_lib()
    if getattr(lib, "_amb_typed", False):
        return lib
    vp = C.c_void_p
    pvp = C.POINTER(C.c_void_p)
    u8p = C.POINTER(C.c_uint8)
    lib.ckf_ambiguity_config_default.argtypes = [C.POINTER(AmbiguityConfig)]
    lib.ckf_ambiguity_config_default.restype = None
    lib.ckf_resolver_create.argtypes = [C.POINTER(AmbiguityConfig), pvp]
    lib.ckf_resolver_destroy.argtypes = [vp]
    lib.ckf_resolver_set_stream.argtypes = [vp, vp]
    lib.ckf_resolver_last_stats.argtypes = [vp, C.POINTER(ResolverStats)]
    lib.ckf_resolve_ambiguities.argtypes = [vp, C.POINTER(ResolverTracks), u8p, F.c_u32_p]
    lib.ckf_resolve_ambiguities_device.argtypes = [vp, C.POINTER(ResolverTracks), vp, F.c_u32_p]
    lib.ckf_resolve_fitted.argtypes = [vp, vp, vp, u8p, F.c_u32_p]
    for name in ("ckf_resolver_create", "ckf_resolver_destroy", "ckf_resolver_set_stream",
                 "ckf_resolver_last_stats", "ckf_resolve_ambiguities",
                 "ckf_resolve_ambiguities_device", "ckf_resolve_fitted"):
        getattr(lib, name).restype = C.c_int
    lib._amb_typed = True
    return lib


def ambiguity_resolution_config(**kw) -> AmbiguityConfig:
    """traccc::ambiguity_resolution_config with its defaults
    (min_meas_per_track 3, max_iterations UINT32_MAX, max_shared_meas 1),
    overridden by keyword."""
    cfg = AmbiguityConfig()
    _lib().ckf_ambiguity_config_default(C.byref(cfg))
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


@dataclass
class ResolveStats:
    n_tracks: int
    n_candidates: int
    n_removed: int
    n_accepted: int
    n_rounds: int
    sequential: bool


class greedy_ambiguity_resolution_algorithm:  # noqa: N801  (traccc's class name)
    """traccc's greedy ambiguity resolution on the MI355X."""

    def __init__(self, config: AmbiguityConfig | None = None, stream: int | None = None):
        self.config = config if config is not None else ambiguity_resolution_config()
        h = C.c_void_p()
        _check(_lib().ckf_resolver_create(C.byref(self.config), C.byref(h)))
        self._h = h
        self._L = _lib()
        if stream is not None:
            _check(_lib().ckf_resolver_set_stream(self._h, C.c_void_p(stream)))

    def __del__(self):
        lib = getattr(self, "_L", None)
        if getattr(self, "_h", None) and lib is not None:
            lib.ckf_resolver_destroy(self._h)
            self._h = None

    def __call__(self, offsets, measurement_ids, pvals, id_bound: int | None = None) -> np.ndarray:
        """Host arrays: track t owns measurement_ids[offsets[t]:offsets[t+1]].
        Returns the kept track indices, ascending."""
        off = np.ascontiguousarray(offsets, np.uint32)
        ids = np.ascontiguousarray(measurement_ids, np.uint32)
        pv = np.ascontiguousarray(pvals, np.float32)
        n = len(off) - 1
        if n < 0 or len(pv) != n:
            raise ValueError("offsets must have one entry more than pvals")
        if n > 0 and len(ids) < int(off[-1]):
            raise ValueError("fewer measurement identifiers than offsets[-1]")
        if id_bound is None:
            id_bound = int(ids.max()) + 1 if len(ids) else 0
        acc = np.zeros(max(n, 1), np.uint8)
        t = ResolverTracks(n, off.ctypes.data, ids.ctypes.data if len(ids) else None,
                           pv.ctypes.data, id_bound)
        na = C.c_uint32()
        _check(_lib().ckf_resolve_ambiguities(self._h, C.byref(t),
                                              acc.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(na)))
        return np.nonzero(acc[:n])[0].astype(np.int64)

    def run_device(self, d_offsets, d_ids, d_pvals, id_bound: int, d_accepted=None):
        """Torch tensors already in HBM (uint32 / int32 offsets and
        identifiers, float32 p-values).  Writes the uint8 accepted flags to
        d_accepted (allocated when None) and returns it with the count."""
        import torch
        n = d_pvals.numel()
        if d_offsets.numel() != n + 1:
            raise ValueError("offsets must have one entry more than pvals")
        for x in (d_offsets, d_ids, d_pvals):
            if not x.is_cuda or not x.is_contiguous():
                raise ValueError("inputs must be contiguous device tensors")
        if d_accepted is None:
            d_accepted = torch.empty(max(n, 1), dtype=torch.uint8, device=d_pvals.device)
        t = ResolverTracks(n, d_offsets.data_ptr(), d_ids.data_ptr() if d_ids.numel() else None,
                           d_pvals.data_ptr(), id_bound)
        na = C.c_uint32()
        _check(_lib().ckf_resolve_ambiguities_device(self._h, C.byref(t), d_accepted.data_ptr(),
                                                     C.byref(na)))
        return d_accepted, na.value

    def run_fitted(self, devents: DeviceEvents, fitted) -> np.ndarray:
        """Fitted tracks (kalman_fitting_algorithm.run_device) of the events:
        their p-values and their candidates' measurements.  Returns the kept
        track indices, ascending."""
        nt, _ = fitted.sizes()
        acc = np.zeros(max(nt, 1), np.uint8)
        na = C.c_uint32()
        _check(_lib().ckf_resolve_fitted(self._h, devents._h, fitted._h,
                                         acc.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(na)))
        return np.nonzero(acc[:nt])[0].astype(np.int64)

    def last_stats(self) -> ResolveStats:
        s = ResolverStats()
        _check(_lib().ckf_resolver_last_stats(self._h, C.byref(s)))
        return ResolveStats(s.n_tracks, s.n_candidates, s.n_removed, s.n_accepted, s.n_rounds,
                            bool(s.sequential))
