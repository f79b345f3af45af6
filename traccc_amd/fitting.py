"""Python host mirror of traccc's track-fitting operator, over the C ABI.

    fitter = kalman_fitting_algorithm(fitting_config())
    fitted = fitter.run_device(ddet, devents, candidates)   # all in HBM
    tracks = fitted.download()

mirrors traccc::device::kalman_fitting_algorithm
(device/common/include/traccc/fitting/device/kalman_fitting_algorithm.hpp:91-101):
a fitting_config at construction; the operator takes the detector (with its
field), and the track candidates the CKF found, and returns fitted tracks
with their track states (filtered and smoothed parameters per measurement).
Invalid configurations raise ValueError where traccc throws
std::invalid_argument.  The HIP library is required: there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _ffi as F
from .ckf import DeviceDetector, DeviceEvents, DeviceTracks, _check, _lib as _ckf_lib

# traccc track_fit_outcome (core/include/traccc/edm/track_fit_outcome.hpp:16-27)
OUTCOME_UNKNOWN = 0
OUTCOME_SUCCESS = 1
OUTCOME_NAMES = ["UNKNOWN", "SUCCESS", "FAILURE_NON_POSITIVE_NDF", "FAILURE_NOT_ALL_FITTED",
                 "FAILURE_NOT_ALL_SMOOTHED", "FAILURE_FITTER", "FAILURE_SMOOTHER",
                 "FAILURE_FORWARD_PROPAGATION", "FAILURE_BACKWARD_PROPAGATION"]
STATE_HOLE = 0x01
STATE_SMOOTHED = 0x02


def _lib() -> C.CDLL:
    lib = _ckf_lib()
    if getattr(lib, "_fit_typed", False):
        return lib
    vp = C.c_void_p
    pvp = C.POINTER(C.c_void_p)
    lib.ckf_fitting_config_default.argtypes = [C.POINTER(F.FittingConfig)]
    lib.ckf_fitter_create.argtypes = [C.POINTER(F.FittingConfig), pvp]
    lib.ckf_fitter_destroy.argtypes = [vp]
    lib.ckf_fitter_set_stream.argtypes = [vp, vp]
    lib.ckf_fit_tracks.argtypes = [vp, vp, vp, vp, pvp]
    lib.ckf_fitted_size.argtypes = [vp, F.c_u32_p, F.c_u32_p]
    lib.ckf_fitted_download.argtypes = [vp, C.POINTER(F.FittedTracks)]
    lib.ckf_fitted_destroy.argtypes = [vp]
    for name in ("ckf_fitter_create", "ckf_fitter_destroy", "ckf_fitter_set_stream",
                 "ckf_fit_tracks", "ckf_fitted_size", "ckf_fitted_download", "ckf_fitted_destroy"):
        getattr(lib, name).restype = C.c_int
    lib._fit_typed = True
    return lib


def fitting_config(finding: F.FindingConfig | None = None) -> F.FittingConfig:
    """traccc::fitting_config with its defaults; with a finding_config, its
    propagation settings, momentum cuts and particle hypothesis."""
    from . import simulation  # the host library: no HIP device needed for defaults
    cfg = F.FittingConfig()
    lib = simulation._lib()
    lib.ckf_fitting_config_default.argtypes = [C.POINTER(F.FittingConfig)]
    lib.ckf_fitting_config_default(C.byref(cfg))
    if finding is not None:
        cfg.propagation = finding.propagation
        cfg.min_p, cfg.min_pT = finding.min_p, finding.min_pT
        cfg.ptc_pdg, cfg.ptc_mass, cfg.ptc_charge = (finding.ptc_pdg, finding.ptc_mass,
                                                     finding.ptc_charge)
    return cfg


@dataclass
class FittedTracks:
    """traccc track_collection (fitted) + track_state_collection, host copy.
    Track t owns states [state_offsets[t], state_offsets[t+1])."""
    fit_outcome: np.ndarray
    params_surface_link: np.ndarray
    params_vector: np.ndarray
    params_covariance: np.ndarray
    ndf: np.ndarray
    chi2: np.ndarray
    pval: np.ndarray
    nholes: np.ndarray
    state_offsets: np.ndarray
    state_flags: np.ndarray
    filtered_chi2: np.ndarray
    smoothed_chi2: np.ndarray
    backward_chi2: np.ndarray
    state_surface_link: np.ndarray
    filtered_vector: np.ndarray
    filtered_covariance: np.ndarray
    smoothed_vector: np.ndarray
    smoothed_covariance: np.ndarray
    measurement_index: np.ndarray

    def __len__(self) -> int:
        return len(self.fit_outcome)


class DeviceFitted:
    """Fitted tracks and states resident in HBM (ckf_device_fitted)."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle
        self._L = _lib()

    def __del__(self):
        lib = getattr(self, "_L", None)
        if getattr(self, "_h", None) and lib is not None:
            lib.ckf_fitted_destroy(self._h)
            self._h = None

    def sizes(self) -> tuple[int, int]:
        nt, ns = C.c_uint32(), C.c_uint32()
        _check(_lib().ckf_fitted_size(self._h, C.byref(nt), C.byref(ns)))
        return nt.value, ns.value

    def download(self) -> FittedTracks:
        nt, ns = self.sizes()
        f32 = np.float32
        o = FittedTracks(
            fit_outcome=np.zeros(nt, np.uint16), params_surface_link=np.zeros(nt, np.uint64),
            params_vector=np.zeros((nt, 6), f32), params_covariance=np.zeros((nt, 6, 6), f32),
            ndf=np.zeros(nt, f32), chi2=np.zeros(nt, f32), pval=np.zeros(nt, f32),
            nholes=np.zeros(nt, np.uint32), state_offsets=np.zeros(nt + 1, np.uint32),
            state_flags=np.zeros(ns, np.uint8), filtered_chi2=np.zeros(ns, f32),
            smoothed_chi2=np.zeros(ns, f32), backward_chi2=np.zeros(ns, f32),
            state_surface_link=np.zeros(ns, np.uint64), filtered_vector=np.zeros((ns, 6), f32),
            filtered_covariance=np.zeros((ns, 6, 6), f32), smoothed_vector=np.zeros((ns, 6), f32),
            smoothed_covariance=np.zeros((ns, 6, 6), f32),
            measurement_index=np.zeros(ns, np.uint32))
        s = F.FittedTracks(
            nt, ns, F.ptr(o.fit_outcome, C.c_uint16), F.ptr(o.params_surface_link, C.c_uint64),
            F.ptr(o.params_vector, C.c_float), F.ptr(o.params_covariance, C.c_float),
            F.ptr(o.ndf, C.c_float), F.ptr(o.chi2, C.c_float), F.ptr(o.pval, C.c_float),
            F.ptr(o.nholes, C.c_uint32), F.ptr(o.state_offsets, C.c_uint32),
            F.ptr(o.state_flags, C.c_uint8), F.ptr(o.filtered_chi2, C.c_float),
            F.ptr(o.smoothed_chi2, C.c_float), F.ptr(o.backward_chi2, C.c_float),
            F.ptr(o.state_surface_link, C.c_uint64), F.ptr(o.filtered_vector, C.c_float),
            F.ptr(o.filtered_covariance, C.c_float), F.ptr(o.smoothed_vector, C.c_float),
            F.ptr(o.smoothed_covariance, C.c_float), F.ptr(o.measurement_index, C.c_uint32))
        _check(_lib().ckf_fitted_download(self._h, C.byref(s)))
        return o


class kalman_fitting_algorithm:  # noqa: N801  (traccc's class name)
    """traccc's Kalman fitting operator (two-filter smoother) on the MI355X."""

    def __init__(self, config: F.FittingConfig | None = None, stream: int | None = None):
        self.config = config if config is not None else fitting_config()
        h = C.c_void_p()
        _check(_lib().ckf_fitter_create(C.byref(self.config), C.byref(h)))
        self._h = h
        self._L = _lib()
        if stream is not None:
            _check(_lib().ckf_fitter_set_stream(self._h, C.c_void_p(stream)))

    def __del__(self):
        lib = getattr(self, "_L", None)
        if getattr(self, "_h", None) and lib is not None:
            lib.ckf_fitter_destroy(self._h)
            self._h = None

    def run_device(self, ddet: DeviceDetector, devents: DeviceEvents,
                   candidates: DeviceTracks) -> DeviceFitted:
        """Operator on HBM-resident candidates; the output stays in HBM."""
        out = C.c_void_p()
        _check(_lib().ckf_fit_tracks(self._h, ddet._h, devents._h, candidates._h, C.byref(out)))
        return DeviceFitted(out)
