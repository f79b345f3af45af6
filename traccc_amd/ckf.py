"""Python host mirror of traccc's track-finding operator, over the C ABI.

    alg = combinatorial_kalman_filter_algorithm(config)
    tracks = alg(detector, field, measurements, seeds)

mirrors traccc::host/device::combinatorial_kalman_filter_algorithm
(core/include/traccc/finding/combinatorial_kalman_filter_algorithm.hpp:34-72):
a finding_config at construction, the operator takes the detector, the
magnetic field, the measurement collection and the seed collection, and
returns a track container (track fields + constituent measurement links).
Invalid configurations raise ValueError where traccc throws
std::invalid_argument.  The HIP library is required: there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _ffi as F

LIB = "libtraccc_amd.so"

CKF_OK = 0
_STATUS = {1: "invalid argument", 2: "unsupported", 3: "device error", 4: "out of memory",
           5: "precondition violated", 6: "internal error"}


class CKFError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{_STATUS.get(code, code)}: {msg}")
        self.code = code


def _lib() -> C.CDLL:
    lib = F.load_library(LIB)
    if getattr(lib, "_ckf_typed", False):
        return lib
    vp = C.c_void_p
    pvp = C.POINTER(C.c_void_p)
    lib.ckf_abi_version.restype = C.c_uint32
    lib.ckf_last_error.restype = C.c_char_p
    lib.ckf_finding_config_default.argtypes = [C.POINTER(F.FindingConfig)]
    lib.ckf_algorithm_create.argtypes = [C.POINTER(F.FindingConfig), pvp]
    lib.ckf_algorithm_destroy.argtypes = [vp]
    lib.ckf_algorithm_set_stream.argtypes = [vp, vp]
    lib.ckf_algorithm_last_stats.argtypes = [vp, C.POINTER(F.RunStats)]
    lib.ckf_device_detector_create.argtypes = [C.POINTER(F.Detector), C.POINTER(F.BField), pvp]
    lib.ckf_device_detector_destroy.argtypes = [vp]
    lib.ckf_device_event_create.argtypes = [vp, C.POINTER(F.Measurements),
                                            C.POINTER(F.BoundParams), C.c_uint32, pvp]
    lib.ckf_device_event_destroy.argtypes = [vp]
    lib.ckf_find_tracks.argtypes = [vp, vp, vp, pvp]
    lib.ckf_algorithm_set_scorer.argtypes = [vp, vp, C.c_float]
    lib.ckf_find_tracks_logged.argtypes = [vp, vp, vp, vp, vp, C.c_uint32, F.c_u32_p, pvp]
    lib.ckf_find_tracks_host.argtypes = [vp, C.POINTER(F.Detector), C.POINTER(F.BField),
                                         C.POINTER(F.Measurements), C.POINTER(F.BoundParams), pvp]
    lib.ckf_tracks_size.argtypes = [vp, F.c_u32_p, F.c_u32_p]
    lib.ckf_tracks_event_offsets.argtypes = [vp, F.c_u32_p, F.c_u32_p, F.c_u32_p]
    lib.ckf_tracks_download.argtypes = [vp, C.POINTER(F.TrackCandidates)]
    lib.ckf_tracks_destroy.argtypes = [vp]
    for name in ("ckf_algorithm_create", "ckf_algorithm_destroy", "ckf_algorithm_set_stream",
                 "ckf_algorithm_last_stats", "ckf_device_detector_create",
                 "ckf_device_detector_destroy", "ckf_device_event_create",
                 "ckf_device_event_destroy", "ckf_find_tracks", "ckf_find_tracks_host",
                 "ckf_tracks_size", "ckf_tracks_event_offsets", "ckf_tracks_download",
                 "ckf_tracks_destroy"):
        getattr(lib, name).restype = C.c_int
    lib._ckf_typed = True
    return lib


def _check(rc: int) -> None:
    if rc != CKF_OK:
        msg = _lib().ckf_last_error().decode()
        if rc == 1:
            raise ValueError(msg)
        raise CKFError(rc, msg)


def finding_config() -> F.FindingConfig:
    """traccc::finding_config with its defaults."""
    cfg = F.FindingConfig()
    _lib().ckf_finding_config_default(C.byref(cfg))
    return cfg


@dataclass
class TrackContainer:
    """traccc track_container fields (host copy).  For a batch, event e's
    tracks are [event_track_offsets[e], event_track_offsets[e+1])."""
    fit_outcome: np.ndarray
    params_surface_link: np.ndarray
    params_vector: np.ndarray
    params_covariance: np.ndarray
    ndf: np.ndarray
    chi2: np.ndarray
    pval: np.ndarray
    nholes: np.ndarray
    offsets: np.ndarray        # into `constituents`
    constituents: np.ndarray   # measurement indices (event-relative)
    event_track_offsets: np.ndarray
    event_constituent_offsets: np.ndarray

    def __len__(self) -> int:
        return len(self.ndf)

    def track(self, i: int) -> np.ndarray:
        return self.constituents[self.offsets[i]:self.offsets[i + 1]]

    def event(self, e: int) -> "TrackContainer":
        a, b = int(self.event_track_offsets[e]), int(self.event_track_offsets[e + 1])
        ca = int(self.event_constituent_offsets[e])
        cb = int(self.event_constituent_offsets[e + 1])
        return TrackContainer(
            self.fit_outcome[a:b], self.params_surface_link[a:b], self.params_vector[a:b],
            self.params_covariance[a:b], self.ndf[a:b], self.chi2[a:b], self.pval[a:b],
            self.nholes[a:b], self.offsets[a:b + 1] - ca, self.constituents[ca:cb],
            np.array([0, b - a], dtype=np.uint32), np.array([0, cb - ca], dtype=np.uint32))


class DeviceTracks:
    """Track candidates resident in HBM (ckf_device_tracks)."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle
        self._L = _lib()

    def __del__(self):
        # the library handle is held by the object: module globals may
        # already be gone when this runs at interpreter shutdown
        lib = getattr(self, "_L", None)
        if getattr(self, "_h", None) and lib is not None:
            lib.ckf_tracks_destroy(self._h)
            self._h = None

    def sizes(self) -> tuple[int, int]:
        nt, nc = C.c_uint32(), C.c_uint32()
        _check(_lib().ckf_tracks_size(self._h, C.byref(nt), C.byref(nc)))
        return nt.value, nc.value

    def event_offsets(self) -> np.ndarray:
        """[n_events + 1] first track of each event in the batch."""
        lib = _lib()
        ne = C.c_uint32()
        _check(lib.ckf_tracks_event_offsets(self._h, C.byref(ne), None, None))
        eto = np.zeros(ne.value + 1, dtype=np.uint32)
        eco = np.zeros(ne.value + 1, dtype=np.uint32)
        _check(lib.ckf_tracks_event_offsets(self._h, C.byref(ne), F.ptr(eto, C.c_uint32),
                                            F.ptr(eco, C.c_uint32)))
        return eto

    def download(self) -> TrackContainer:
        lib = _lib()
        nt, nc = self.sizes()
        ne = C.c_uint32()
        _check(lib.ckf_tracks_event_offsets(self._h, C.byref(ne), None, None))
        eto = np.zeros(ne.value + 1, dtype=np.uint32)
        eco = np.zeros(ne.value + 1, dtype=np.uint32)
        _check(lib.ckf_tracks_event_offsets(self._h, C.byref(ne), F.ptr(eto, C.c_uint32),
                                            F.ptr(eco, C.c_uint32)))
        out = TrackContainer(
            fit_outcome=np.zeros(nt, np.uint16), params_surface_link=np.zeros(nt, np.uint64),
            params_vector=np.zeros((nt, 6), np.float32),
            params_covariance=np.zeros((nt, 6, 6), np.float32), ndf=np.zeros(nt, np.float32),
            chi2=np.zeros(nt, np.float32), pval=np.zeros(nt, np.float32),
            nholes=np.zeros(nt, np.uint32), offsets=np.zeros(nt + 1, np.uint32),
            constituents=np.zeros(nc, np.uint32), event_track_offsets=eto,
            event_constituent_offsets=eco)
        ctype = np.zeros(max(nc, 1), np.uint16)
        tc = F.TrackCandidates(
            nt, nc, F.ptr(out.fit_outcome, C.c_uint16), F.ptr(out.params_surface_link, C.c_uint64),
            F.ptr(out.params_vector, C.c_float), F.ptr(out.params_covariance, C.c_float),
            F.ptr(out.ndf, C.c_float), F.ptr(out.chi2, C.c_float), F.ptr(out.pval, C.c_float),
            F.ptr(out.nholes, C.c_uint32), F.ptr(out.offsets, C.c_uint32),
            F.ptr(ctype, C.c_uint16), F.ptr(out.constituents, C.c_uint32))
        _check(lib.ckf_tracks_download(self._h, C.byref(tc)))
        return out


class DeviceDetector:
    """Detector geometry + magnetic field in HBM (traccc detector_buffer)."""

    def __init__(self, detector, field):
        h = C.c_void_p()
        self._keep = (detector, field)
        _check(_lib().ckf_device_detector_create(C.byref(detector.struct), C.byref(field.struct),
                                                 C.byref(h)))
        self._h = h
        self._L = _lib()

    def __del__(self):
        # the library handle is held by the object: module globals may
        # already be gone when this runs at interpreter shutdown
        lib = getattr(self, "_L", None)
        if getattr(self, "_h", None) and lib is not None:
            lib.ckf_device_detector_destroy(self._h)
            self._h = None


class DeviceEvents:
    """One or more events' measurements + seeds in HBM."""

    def __init__(self, ddet: DeviceDetector, events):
        if not isinstance(events, (list, tuple)):
            events = [events]
        n = len(events)
        ms = (F.Measurements * n)(*[e.measurements_struct() for e in events])
        ss = (F.BoundParams * n)(*[e.seeds_struct() for e in events])
        h = C.c_void_p()
        _check(_lib().ckf_device_event_create(ddet._h, ms, ss, n, C.byref(h)))
        self._h = h
        self._L = _lib()
        self.n_events = n
        self.n_seeds = sum(e.n_seeds for e in events)
        self.event_n_meas = [e.n_meas for e in events]

    def __del__(self):
        # the library handle is held by the object: module globals may
        # already be gone when this runs at interpreter shutdown
        lib = getattr(self, "_L", None)
        if getattr(self, "_h", None) and lib is not None:
            lib.ckf_device_event_destroy(self._h)
            self._h = None


class combinatorial_kalman_filter_algorithm:  # noqa: N801  (traccc's class name)
    """traccc's CKF operator on the MI355X."""

    def __init__(self, config: F.FindingConfig | None = None, stream: int | None = None):
        self.config = config if config is not None else finding_config()
        h = C.c_void_p()
        _check(_lib().ckf_algorithm_create(C.byref(self.config), C.byref(h)))
        self._h = h
        self._L = _lib()
        if stream is not None:
            _check(_lib().ckf_algorithm_set_stream(self._h, C.c_void_p(stream)))

    def __del__(self):
        # the library handle is held by the object: module globals may
        # already be gone when this runs at interpreter shutdown
        lib = getattr(self, "_L", None)
        if getattr(self, "_h", None) and lib is not None:
            lib.ckf_algorithm_destroy(self._h)
            self._h = None

    def run_device(self, ddet: DeviceDetector, devents: DeviceEvents) -> DeviceTracks:
        """Operator on HBM-resident inputs; output stays in HBM."""
        out = C.c_void_p()
        _check(_lib().ckf_find_tracks(self._h, ddet._h, devents._h, C.byref(out)))
        return DeviceTracks(out)

    def set_scorer(self, scorer, threshold: float = 0.0) -> None:
        """Prune candidates whose scorer logit is below `threshold` before the
        Kalman update (scorer None: the plain CKF).  The scorer object must
        outlive its use."""
        self._scorer = scorer  # keeps it alive while the algorithm uses it
        _check(_lib().ckf_algorithm_set_scorer(self._h, scorer.handle if scorer else None,
                                               float(threshold)))

    def run_device_logged(self, ddet: DeviceDetector, devents: DeviceEvents, capacity: int):
        """Plain CKF with its branches logged: (tracks, features [n][12],
        labels [n], n) with features / labels as torch tensors on the GPU."""
        import torch
        feats = torch.empty((max(capacity, 1), 12), dtype=torch.float32, device="cuda")
        labels = torch.empty(max(capacity, 1), dtype=torch.uint8, device="cuda")
        n = C.c_uint32()
        out = C.c_void_p()
        _check(_lib().ckf_find_tracks_logged(self._h, ddet._h, devents._h, feats.data_ptr(),
                                             labels.data_ptr(), capacity, C.byref(n),
                                             C.byref(out)))
        k = min(n.value, capacity)
        return DeviceTracks(out), feats[:k], labels[:k], n.value

    def __call__(self, detector, field, measurements, seeds=None) -> TrackContainer:
        """traccc-shaped operator(det, field, measurements, seeds).  With
        `seeds` None, `measurements` is an Event carrying both."""
        ev = measurements
        m = ev.measurements_struct()
        s = (seeds if seeds is not None else ev).seeds_struct()
        out = C.c_void_p()
        _check(_lib().ckf_find_tracks_host(self._h, C.byref(detector.struct),
                                           C.byref(field.struct), C.byref(m), C.byref(s),
                                           C.byref(out)))
        return DeviceTracks(out).download()

    def last_stats(self) -> F.RunStats:
        st = F.RunStats()
        _check(_lib().ckf_algorithm_last_stats(self._h, C.byref(st)))
        return st
