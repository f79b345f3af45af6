"""CPU oracle for the CKF path -- TEST INFRASTRUCTURE ONLY.

Loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, as the checker / the timed CPU baseline; the product never imports it.

liboracle_ckf.so is this project's restatement of traccc's host CKF
(core/include/traccc/finding/details/combinatorial_kalman_filter.hpp:59-718),
templated on the scalar: FP64 = oracle, FP32 = timed CPU baseline.

Pinning status of each function (what fixes it independently of the
restatement itself) is listed in tests/test_oracle_pins.py.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle_ckf.so")

from traccc_amd import _ffi as F  # noqa: E402  (struct layouts only)


class OracleTracks(C.Structure):
    _fields_ = [
        ("n_tracks", C.c_uint32),
        ("n_constituents", C.c_uint32),
        ("params_surface_link", F.c_u64_p),
        ("params_vector", F.c_float_p),
        ("params_covariance", F.c_float_p),
        ("ndf", C.POINTER(C.c_double)),
        ("chi2", C.POINTER(C.c_double)),
        ("pval", C.POINTER(C.c_double)),
        ("nholes", F.c_u32_p),
        ("constituent_offsets", F.c_u32_p),
        ("constituent_index", F.c_u32_p),
        ("n_steps", C.c_uint32),
        ("n_links", C.c_uint64),
        ("n_propagations", C.c_uint64),
        ("n_tips", C.c_uint64),
    ]


class OracleSeeds(C.Structure):
    _fields_ = [("n_seeds", C.c_uint32), ("n_phi_bins", C.c_uint32), ("n_z_bins", C.c_uint32),
                ("bottom", F.c_u32_p), ("middle", F.c_u32_p), ("top", F.c_u32_p),
                ("weight", F.c_float_p)]


class OracleFitted(C.Structure):
    _d = C.POINTER(C.c_double)
    _fields_ = [
        ("n_tracks", C.c_uint32),
        ("n_states", C.c_uint32),
        ("outcome", F.c_u16_p),
        ("params_surface", F.c_u32_p),
        ("params_vector", _d),
        ("params_covariance", _d),
        ("ndf", _d),
        ("chi2", _d),
        ("pval", _d),
        ("nholes", F.c_u32_p),
        ("state_offsets", F.c_u32_p),
        ("state_flags", F.c_u8_p),
        ("filtered_chi2", _d),
        ("smoothed_chi2", _d),
        ("backward_chi2", _d),
        ("filtered_vector", _d),
        ("filtered_covariance", _d),
        ("smoothed_vector", _d),
        ("smoothed_covariance", _d),
    ]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} not built (make -C oracle)")
        L = C.CDLL(LIB)
        dptr = C.POINTER(C.c_double)
        L.oracle_ckf_run.argtypes = [C.POINTER(F.FindingConfig), C.POINTER(F.Detector),
                                     C.POINTER(F.BField), C.POINTER(F.Measurements),
                                     C.POINTER(F.BoundParams), C.c_int, C.POINTER(OracleTracks)]
        L.oracle_ckf_run.restype = C.c_int
        L.oracle_tracks_free.argtypes = [C.POINTER(OracleTracks)]
        L.oracle_ckf_run_masked.argtypes = [C.POINTER(F.FindingConfig), C.POINTER(F.Detector),
                                            C.POINTER(F.BField), C.POINTER(F.Measurements),
                                            C.POINTER(F.BoundParams), C.c_int, C.c_uint32,
                                            F.c_u64_p, C.c_uint64, C.POINTER(OracleTracks)]
        L.oracle_ckf_run_masked.restype = C.c_int
        L.oracle_ckf_run_many.argtypes = [C.POINTER(F.FindingConfig), C.POINTER(F.Detector),
                                          C.POINTER(F.BField), C.POINTER(F.Measurements),
                                          C.POINTER(F.BoundParams), C.c_uint32, C.c_int,
                                          C.c_uint32, C.POINTER(C.c_uint64), dptr]
        L.oracle_ckf_run_many.restype = C.c_int
        L.oracle_prob.argtypes = [C.c_double, C.c_double]
        L.oracle_prob.restype = C.c_double
        L.oracle_prob_f.argtypes = [C.c_float, C.c_float]
        L.oracle_prob_f.restype = C.c_float
        L.oracle_kf_update.argtypes = [dptr, dptr, F.c_float_p, F.c_float_p, C.c_uint32,
                                       F.c_u8_p, dptr, dptr, dptr]
        L.oracle_kf_update.restype = C.c_int
        L.oracle_propagate.argtypes = [C.POINTER(F.FindingConfig), C.POINTER(F.Detector),
                                       C.POINTER(F.BField), C.c_uint32, dptr, dptr, C.c_int,
                                       F.c_u32_p, dptr, dptr, dptr, F.c_u32_p]
        L.oracle_propagate.restype = C.c_int
        L.oracle_fit_run.argtypes = [C.POINTER(F.FittingConfig), C.POINTER(F.Detector),
                                     C.POINTER(F.BField), C.POINTER(F.Measurements), C.c_uint32,
                                     F.c_u64_p, F.c_float_p, F.c_float_p, F.c_u32_p, F.c_u32_p,
                                     C.c_int, C.c_uint32, C.POINTER(OracleFitted), dptr]
        L.oracle_fit_run.restype = C.c_int
        L.oracle_fitted_free.argtypes = [C.POINTER(OracleFitted)]
        L.oracle_propagate_direct.argtypes = [C.POINTER(F.FindingConfig), C.POINTER(F.Detector),
                                              C.POINTER(F.BField), C.c_uint32, C.c_uint32, dptr,
                                              dptr, dptr, dptr]
        L.oracle_propagate_direct.restype = C.c_int
        L.oracle_rk_step.argtypes = [C.POINTER(F.BField), C.c_double, C.c_double, C.c_double,
                                     dptr, dptr, dptr]
        L.oracle_greedy_resolve.argtypes = [C.c_uint32, F.c_u32_p, F.c_u32_p, F.c_float_p,
                                            C.c_uint32, C.c_uint32, C.c_uint32,
                                            C.POINTER(C.c_uint8), C.POINTER(C.c_uint32)]
        L.oracle_greedy_resolve.restype = C.c_int
        from traccc_amd import seeding as SD  # struct layouts only
        L.oracle_seedfinder_setup.argtypes = [C.POINTER(SD.SeedfinderConfig)]
        L.oracle_seedfinder_setup.restype = None
        L.oracle_seeding.argtypes = [C.POINTER(SD.SeedfinderConfig),
                                     C.POINTER(SD.SpacepointGridConfig),
                                     C.POINTER(SD.SeedfilterConfig), C.c_uint32, F.c_float_p,
                                     F.c_float_p, F.c_float_p, C.POINTER(OracleSeeds)]
        L.oracle_seeding.restype = C.c_int
        L.oracle_seeds_free.argtypes = [C.POINTER(OracleSeeds)]
        L.oracle_track_params.argtypes = [C.POINTER(SD.TrackParamsEstimationConfig), dptr, dptr,
                                          dptr, dptr, dptr, dptr, dptr]
        L.oracle_track_params.restype = C.c_int
        _lib = L
    return _lib


@dataclass
class Tracks:
    """Track candidates (traccc track_container fields) from the oracle."""
    params_surface_link: np.ndarray
    params_vector: np.ndarray
    ndf: np.ndarray
    chi2: np.ndarray
    pval: np.ndarray
    nholes: np.ndarray
    offsets: np.ndarray
    constituents: np.ndarray
    n_steps: int
    n_links: int
    n_propagations: int
    n_tips: int

    def __len__(self):
        return len(self.ndf)

    def track(self, i: int) -> np.ndarray:
        return self.constituents[self.offsets[i]:self.offsets[i + 1]]


def _tracks(out) -> Tracks:
    L = lib()
    try:
        n, nc = out.n_tracks, out.n_constituents
        t = Tracks(
            params_surface_link=F.np_from(out.params_surface_link, n, np.uint64),
            params_vector=F.np_from(out.params_vector, 6 * n, np.float32, (n, 6)),
            ndf=F.np_from(out.ndf, n, np.float64),
            chi2=F.np_from(out.chi2, n, np.float64),
            pval=F.np_from(out.pval, n, np.float64),
            nholes=F.np_from(out.nholes, n, np.uint32),
            offsets=F.np_from(out.constituent_offsets, n + 1, np.uint32),
            constituents=F.np_from(out.constituent_index, nc, np.uint32),
            n_steps=out.n_steps, n_links=out.n_links, n_propagations=out.n_propagations,
            n_tips=out.n_tips)
    finally:
        L.oracle_tracks_free(C.byref(out))
    return t


def run(cfg, det, fld, event, precision: int = 64) -> Tracks:
    L = lib()
    out = OracleTracks()
    m = event.measurements_struct()
    s = event.seeds_struct()
    rc = L.oracle_ckf_run(C.byref(cfg), C.byref(det.struct), C.byref(fld.struct), C.byref(m),
                          C.byref(s), precision, C.byref(out))
    if rc != 0:
        raise ValueError(f"oracle_ckf_run failed ({rc})")
    return _tracks(out)


def run_masked(cfg, det, fld, event, list_len: int, kept_keys, precision: int = 64) -> Tracks:
    """The CKF pruned by given candidate verdicts: per input the list_len best
    candidates by (chi2, index) passing the cut, of which only those whose key
    is in kept_keys are updated (in order, until max_num_branches_per_surface
    succeed).  A key (low 63 bits) is splitmix64 chained over the seed, the
    step, the branch's earlier measurements (latest first, holes 0xffffffff)
    and the candidate's measurement, as the device's verdict log writes it.
    kept_keys None keeps every listed candidate."""
    L = lib()
    out = OracleTracks()
    m = event.measurements_struct()
    s = event.seeds_struct()
    if kept_keys is None:   # keep every listed candidate
        kp, nk = None, 0xffffffffffffffff
    else:
        keys = np.ascontiguousarray(np.unique(np.asarray(kept_keys, np.uint64)))
        kp, nk = F.ptr(keys, C.c_uint64), len(keys)
    rc = L.oracle_ckf_run_masked(C.byref(cfg), C.byref(det.struct), C.byref(fld.struct),
                                 C.byref(m), C.byref(s), precision, list_len, kp, nk,
                                 C.byref(out))
    if rc != 0:
        raise ValueError(f"oracle_ckf_run_masked failed ({rc})")
    return _tracks(out)


def run_many(cfg, det, fld, events, precision: int = 32, threads: int = 1):
    """Time the CPU CKF over events on `threads` host threads.
    Returns (seconds, n_tracks_total)."""
    L = lib()
    n = len(events)
    ms = (F.Measurements * n)(*[e.measurements_struct() for e in events])
    ss = (F.BoundParams * n)(*[e.seeds_struct() for e in events])
    tot = C.c_uint64()
    sec = C.c_double()
    rc = L.oracle_ckf_run_many(C.byref(cfg), C.byref(det.struct), C.byref(fld.struct), ms, ss,
                               n, precision, threads, C.byref(tot), C.byref(sec))
    if rc != 0:
        raise ValueError("oracle_ckf_run_many failed")
    return sec.value, tot.value


def prob(chi2: float, ndf: float) -> float:
    return lib().oracle_prob(chi2, ndf)


def prob_f(chi2: float, ndf: float) -> float:
    return lib().oracle_prob_f(chi2, ndf)


def kf_update(vec, cov, meas, var, dims=2, subspace=(0, 1)):
    """Returns (status, chi2, filtered_vec, filtered_cov) in FP64."""
    L = lib()
    v = np.ascontiguousarray(vec, dtype=np.float64)
    c = np.ascontiguousarray(cov, dtype=np.float64).reshape(36)
    m = np.ascontiguousarray(meas, dtype=np.float32)
    va = np.ascontiguousarray(var, dtype=np.float32)
    su = np.ascontiguousarray(subspace, dtype=np.uint8)
    chi2 = C.c_double()
    ov = np.zeros(6)
    oc = np.zeros(36)
    dp = C.POINTER(C.c_double)
    st = L.oracle_kf_update(v.ctypes.data_as(dp), c.ctypes.data_as(dp), F.ptr(m, C.c_float),
                            F.ptr(va, C.c_float), dims, F.ptr(su, C.c_uint8), C.byref(chi2),
                            ov.ctypes.data_as(dp), oc.ctypes.data_as(dp))
    return st, chi2.value, ov, oc.reshape(6, 6)


def propagate(cfg, det, fld, start: int, vec, cov, apply_start_material=False):
    """Returns (status, end_surface, vec, cov, path, n_steps) in FP64."""
    L = lib()
    dp = C.POINTER(C.c_double)
    v = np.ascontiguousarray(vec, dtype=np.float64)
    c = np.ascontiguousarray(cov, dtype=np.float64).reshape(36)
    es = C.c_uint32()
    ov = np.zeros(6)
    oc = np.zeros(36)
    path = C.c_double()
    ns = C.c_uint32()
    st = L.oracle_propagate(C.byref(cfg), C.byref(det.struct), C.byref(fld.struct), start,
                            v.ctypes.data_as(dp), c.ctypes.data_as(dp), int(apply_start_material),
                            C.byref(es), ov.ctypes.data_as(dp), oc.ctypes.data_as(dp),
                            C.byref(path), C.byref(ns))
    return st, es.value, ov, oc.reshape(6, 6), path.value, ns.value


def propagate_direct(cfg, det, fld, start: int, target: int, vec, cov):
    """Returns (status, vec, cov) in FP64 on surface `target`."""
    L = lib()
    dp = C.POINTER(C.c_double)
    v = np.ascontiguousarray(vec, dtype=np.float64)
    c = np.ascontiguousarray(cov, dtype=np.float64).reshape(36)
    ov = np.zeros(6)
    oc = np.zeros(36)
    st = L.oracle_propagate_direct(C.byref(cfg), C.byref(det.struct), C.byref(fld.struct), start,
                                   target, v.ctypes.data_as(dp), c.ctypes.data_as(dp),
                                   ov.ctypes.data_as(dp), oc.ctypes.data_as(dp))
    return st, ov, oc.reshape(6, 6)


def rk_step(fld, mass, abs_q, h, free_in):
    L = lib()
    dp = C.POINTER(C.c_double)
    fi = np.ascontiguousarray(free_in, dtype=np.float64)
    fo = np.zeros(8)
    D = np.zeros(64)
    L.oracle_rk_step(C.byref(fld.struct), mass, abs_q, h, fi.ctypes.data_as(dp),
                     fo.ctypes.data_as(dp), D.ctypes.data_as(dp))
    return fo, D.reshape(8, 8)


@dataclass
class Fitted:
    """Fitted tracks and their states (traccc track_collection +
    track_state_collection fields), FP64."""
    outcome: np.ndarray
    params_surface: np.ndarray
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
    filtered_vector: np.ndarray
    filtered_covariance: np.ndarray
    smoothed_vector: np.ndarray
    smoothed_covariance: np.ndarray
    seconds: float

    def __len__(self):
        return len(self.outcome)


def fit(fcfg, det, fld, event, seed_surface_link, seed_vector, seed_covariance, offsets,
        meas_index, precision: int = 64, threads: int = 1) -> Fitted:
    """Kalman fit + two-filter smoother of candidate tracks of one event."""
    L = lib()
    m = event.measurements_struct()
    n = len(offsets) - 1
    ssl = np.ascontiguousarray(seed_surface_link, np.uint64)
    sv = np.ascontiguousarray(seed_vector, np.float32).reshape(-1)
    sc = np.ascontiguousarray(seed_covariance, np.float32).reshape(-1)
    off = np.ascontiguousarray(offsets, np.uint32)
    idx = np.ascontiguousarray(meas_index, np.uint32)
    if len(idx) == 0:
        idx = np.zeros(1, np.uint32)
    out = OracleFitted()
    sec = C.c_double()
    rc = L.oracle_fit_run(C.byref(fcfg), C.byref(det.struct), C.byref(fld.struct), C.byref(m), n,
                          F.ptr(ssl, C.c_uint64), F.ptr(sv, C.c_float), F.ptr(sc, C.c_float),
                          F.ptr(off, C.c_uint32), F.ptr(idx, C.c_uint32), precision, threads,
                          C.byref(out), C.byref(sec))
    if rc != 0:
        raise ValueError(f"oracle_fit_run failed ({rc})")
    try:
        ns = out.n_states
        d = np.float64
        r = Fitted(
            outcome=F.np_from(out.outcome, n, np.uint16),
            params_surface=F.np_from(out.params_surface, n, np.uint32),
            params_vector=F.np_from(out.params_vector, 6 * n, d, (n, 6)),
            params_covariance=F.np_from(out.params_covariance, 36 * n, d, (n, 6, 6)),
            ndf=F.np_from(out.ndf, n, d), chi2=F.np_from(out.chi2, n, d),
            pval=F.np_from(out.pval, n, d), nholes=F.np_from(out.nholes, n, np.uint32),
            state_offsets=F.np_from(out.state_offsets, n + 1, np.uint32),
            state_flags=F.np_from(out.state_flags, ns, np.uint8),
            filtered_chi2=F.np_from(out.filtered_chi2, ns, d),
            smoothed_chi2=F.np_from(out.smoothed_chi2, ns, d),
            backward_chi2=F.np_from(out.backward_chi2, ns, d),
            filtered_vector=F.np_from(out.filtered_vector, 6 * ns, d, (ns, 6)),
            filtered_covariance=F.np_from(out.filtered_covariance, 36 * ns, d, (ns, 6, 6)),
            smoothed_vector=F.np_from(out.smoothed_vector, 6 * ns, d, (ns, 6)),
            smoothed_covariance=F.np_from(out.smoothed_covariance, 36 * ns, d, (ns, 6, 6)),
            seconds=sec.value)
    finally:
        L.oracle_fitted_free(C.byref(out))
    return r


def greedy_resolve(offsets, meas_ids, pval, min_meas_per_track: int = 3,
                   max_iterations: int = 0xffffffff, max_shared_meas: int = 1):
    """traccc's host greedy ambiguity resolution restated
    (core/src/ambiguity_resolution/greedy_ambiguity_resolution_algorithm.cpp:19-254).
    Track t owns the measurement identifiers meas_ids[offsets[t]:offsets[t+1]].
    Returns (accepted flags as a bool array, number of tracks removed)."""
    L = lib()
    off = np.ascontiguousarray(offsets, np.uint32)
    ids = np.ascontiguousarray(meas_ids, np.uint32)
    pv = np.ascontiguousarray(pval, np.float32)
    n = len(off) - 1
    if len(pv) != n or (n > 0 and len(ids) < off[-1]):
        raise ValueError("inconsistent track arrays")
    acc = np.zeros(max(n, 1), np.uint8)
    nr = C.c_uint32()
    rc = L.oracle_greedy_resolve(n, F.ptr(off, C.c_uint32), F.ptr(ids, C.c_uint32),
                                 F.ptr(pv, C.c_float), min_meas_per_track, max_iterations,
                                 max_shared_meas, acc.ctypes.data_as(C.POINTER(C.c_uint8)),
                                 C.byref(nr))
    if rc != 0:
        raise ValueError(f"oracle_greedy_resolve failed ({rc})")
    return acc[:n].astype(bool), nr.value


def seedfinder_setup(cfg):
    """traccc seedfinder_config::setup restated (the derived fields)."""
    lib().oracle_seedfinder_setup(C.byref(cfg))
    return cfg


def seeding(finder, grid, filt, global_xyz, var_r=None, var_z=None):
    """traccc's host seeding_algorithm restated in single precision
    (core/src/seeding/seed_finding.cpp:57-121 and its helpers).  Returns
    (bottom, middle, top, weight, (n_phi_bins, n_z_bins))."""
    L = lib()
    g = np.ascontiguousarray(global_xyz, np.float32).reshape(-1, 3)
    vr = None if var_r is None else np.ascontiguousarray(var_r, np.float32)
    vz = None if var_z is None else np.ascontiguousarray(var_z, np.float32)
    out = OracleSeeds()
    rc = L.oracle_seeding(C.byref(finder), C.byref(grid), C.byref(filt), len(g),
                          F.ptr(g, C.c_float), None if vr is None else F.ptr(vr, C.c_float),
                          None if vz is None else F.ptr(vz, C.c_float), C.byref(out))
    if rc == 2:
        raise ValueError("grid configuration rejected (the reference throws std::domain_error)")
    if rc != 0:
        raise ValueError(f"oracle_seeding failed ({rc})")
    try:
        n = out.n_seeds
        r = (F.np_from(out.bottom, n, np.uint32), F.np_from(out.middle, n, np.uint32),
             F.np_from(out.top, n, np.uint32), F.np_from(out.weight, n, np.float32),
             (out.n_phi_bins, out.n_z_bins))
    finally:
        L.oracle_seeds_free(C.byref(out))
    return r


def track_params(cfg, bfield, p0, p1, p2, bottom_local=(0.0, 0.0)):
    """traccc's track parameters of one seed restated in double precision
    (track_params_estimation_helper.hpp:47-125, track_params_estimation.cpp:60-87).
    Returns (vector [6], covariance diagonal [6])."""
    L = lib()
    d = C.c_double * 3
    vec = (C.c_double * 6)()
    cov = (C.c_double * 6)()
    rc = L.oracle_track_params(C.byref(cfg), d(*bfield), d(*p0), d(*p1), d(*p2),
                               (C.c_double * 2)(*bottom_local), vec, cov)
    if rc != 0:
        raise ValueError("oracle_track_params failed")
    return np.array(vec[:]), np.array(cov[:])
