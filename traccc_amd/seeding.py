"""Python host mirror of traccc's seeding chain, over the C ABI.

    finder = seedfinder_config(); grid = spacepoint_grid_config(finder)
    seeder = seeding_algorithm(finder, grid, seedfilter_config())
    seeds = seeder(global_xyz)                      # (bottom, middle, top, weight)
    sp, midx = spacepoint_formation(ddet, devents, event=0)
    params = track_params_estimation()(sp, midx, seeds, event, bfield)

mirrors traccc::host::seeding_algorithm
(core/include/traccc/seeding/seeding_algorithm.hpp:26-66),
silicon_pixel_spacepoint_formation_algorithm and track_params_estimation
(core/include/traccc/seeding/track_params_estimation.hpp:31-67).  Units are
traccc's: mm, GeV, and the field in GeV / (e mm) (TESLA below converts).
The HIP library is required for the algorithms: there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _ffi as F

TESLA = 0.000299792458  # GeV / (e mm)
INVALID_MEASUREMENT_INDEX = 0xFFFFFFFF


class SeedfinderConfig(C.Structure):
    """include/traccc_amd/seeding.h ckf_seedfinder_config."""
    _fields_ = [(n, C.c_float) for n in (
        "zMin", "zMax", "rMax", "rMin", "collisionRegionMin", "collisionRegionMax", "phiMin",
        "phiMax", "minPt", "cotThetaMax", "deltaRMin", "deltaRMax", "deltaZMax", "impactMax",
        "sigmaScattering", "maxPtScattering")] + [
        ("maxSeedsPerSpM", C.c_uint32), ("bFieldInZ", C.c_float), ("beamPos", C.c_float * 2),
        ("radLengthPerSeed", C.c_float), ("zAlign", C.c_float), ("rAlign", C.c_float),
        ("sigmaError", C.c_float), ("phiBinDeflectionCoverage", C.c_int32),
        ("neighbor_scope", C.c_uint32 * 2)] + [(n, C.c_float) for n in (
        "highland", "maxScatteringAngle2", "pTPerHelixRadius", "minHelixDiameter2",
        "minHelixRadius", "pT2perRadius")]


class SpacepointGridConfig(C.Structure):
    _fields_ = [(n, C.c_float) for n in (
        "bFieldInZ", "minPt", "rMax", "zMax", "zMin", "deltaRMax", "cotThetaMax", "impactMax",
        "phiMin", "phiMax")] + [("phiBinDeflectionCoverage", C.c_int32)]


class SeedfilterConfig(C.Structure):
    _fields_ = [("deltaInvHelixDiameter", C.c_float), ("impactWeightFactor", C.c_float),
                ("compatSeedWeight", C.c_float), ("deltaRMin", C.c_float),
                ("compatSeedLimit", C.c_uint32)] + [(n, C.c_float) for n in (
        "good_spB_min_radius", "good_spB_weight_increase", "good_spT_max_radius",
        "good_spT_weight_increase", "good_spB_min_weight", "seed_min_weight", "spB_min_radius")]


class TrackParamsEstimationConfig(C.Structure):
    _fields_ = [("initial_sigma", C.c_float * 6), ("initial_sigma_qopt", C.c_float),
                ("initial_sigma_pt_rel", C.c_float), ("initial_inflation", C.c_float * 6)]


class Spacepoints(C.Structure):
    _fields_ = [("size", C.c_uint32), ("global_", C.c_void_p), ("radius_variance", C.c_void_p),
                ("z_variance", C.c_void_p), ("measurement_index_1", C.c_void_p)]


class Seeds(C.Structure):
    _fields_ = [("size", C.c_uint32), ("bottom", C.c_void_p), ("middle", C.c_void_p),
                ("top", C.c_void_p), ("weight", C.c_void_p)]


def _host_lib() -> C.CDLL:
    """The configuration defaults live in the host library (no device needed)."""
    from . import simulation
    lib = simulation._lib()
    if not getattr(lib, "_seed_cfg_typed", False):
        lib.ckf_seedfinder_config_default.argtypes = [C.POINTER(SeedfinderConfig)]
        lib.ckf_seedfinder_config_setup.argtypes = [C.POINTER(SeedfinderConfig)]
        lib.ckf_spacepoint_grid_config_from.argtypes = [C.POINTER(SeedfinderConfig),
                                                       C.POINTER(SpacepointGridConfig)]
        lib.ckf_seedfilter_config_default.argtypes = [C.POINTER(SeedfilterConfig)]
        lib.ckf_track_params_estimation_config_default.argtypes = [
            C.POINTER(TrackParamsEstimationConfig)]
        lib._seed_cfg_typed = True
    return lib


def seedfinder_config(setup: bool = True, **kw) -> SeedfinderConfig:
    """traccc::seedfinder_config defaults, overridden by keyword; derived
    fields recomputed afterwards when setup is True (the reference computes
    them once, in its constructor, from the defaults)."""
    c = SeedfinderConfig()
    _host_lib().ckf_seedfinder_config_default(C.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    if setup:
        _host_lib().ckf_seedfinder_config_setup(C.byref(c))
    return c


def spacepoint_grid_config(finder: SeedfinderConfig, **kw) -> SpacepointGridConfig:
    g = SpacepointGridConfig()
    _host_lib().ckf_spacepoint_grid_config_from(C.byref(finder), C.byref(g))
    for k, v in kw.items():
        setattr(g, k, v)
    return g


def seedfilter_config(**kw) -> SeedfilterConfig:
    c = SeedfilterConfig()
    _host_lib().ckf_seedfilter_config_default(C.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def track_params_estimation_config(**kw) -> TrackParamsEstimationConfig:
    c = TrackParamsEstimationConfig()
    _host_lib().ckf_track_params_estimation_config_default(C.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def _lib() -> C.CDLL:
    from .ckf import _lib as _ckf_lib
    lib = _ckf_lib()
    if getattr(lib, "_seed_typed", False):
        return lib
    vp = C.c_void_p
    pvp = C.POINTER(C.c_void_p)
    u32p = C.POINTER(C.c_uint32)
    lib.ckf_seeding_create.argtypes = [C.POINTER(SeedfinderConfig), C.POINTER(SpacepointGridConfig),
                                       C.POINTER(SeedfilterConfig), pvp]
    lib.ckf_seeding_destroy.argtypes = [vp]
    lib.ckf_seeding_set_stream.argtypes = [vp, vp]
    lib.ckf_seeding_run.argtypes = [vp, C.POINTER(Spacepoints), u32p]
    lib.ckf_seeding_run_device.argtypes = [vp, C.POINTER(Spacepoints), u32p]
    lib.ckf_seeding_result_size.argtypes = [vp, u32p]
    lib.ckf_seeding_result_download.argtypes = [vp, C.POINTER(Seeds)]
    lib.ckf_spacepoint_formation.argtypes = [vp, vp, C.c_uint32, vp, vp, u32p]
    lib.ckf_estimate_track_params.argtypes = [C.POINTER(TrackParamsEstimationConfig),
                                              C.POINTER(Spacepoints), C.POINTER(Seeds),
                                              C.POINTER(F.Measurements), C.POINTER(C.c_float * 3),
                                              vp, vp, vp]
    for name in ("ckf_seeding_create", "ckf_seeding_destroy", "ckf_seeding_set_stream",
                 "ckf_seeding_run", "ckf_seeding_run_device", "ckf_seeding_result_size",
                 "ckf_seeding_result_download", "ckf_spacepoint_formation",
                 "ckf_estimate_track_params"):
        getattr(lib, name).restype = C.c_int
    lib._seed_typed = True
    return lib


@dataclass
class SeedCollection:
    """traccc edm::seed_collection: spacepoint indices and weights."""
    bottom: np.ndarray
    middle: np.ndarray
    top: np.ndarray
    weight: np.ndarray

    def __len__(self) -> int:
        return len(self.bottom)

    def as_tuples(self):
        return list(zip(self.bottom.tolist(), self.middle.tolist(), self.top.tolist()))


def _sp_struct(xyz, var_r=None, var_z=None, midx=None, keep=None):
    xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
    arrs = [xyz]
    vr = None if var_r is None else np.ascontiguousarray(var_r, np.float32)
    vz = None if var_z is None else np.ascontiguousarray(var_z, np.float32)
    mi = None if midx is None else np.ascontiguousarray(midx, np.uint32)
    for a in (vr, vz, mi):
        if a is not None:
            arrs.append(a)
    if keep is not None:
        keep.extend(arrs)
    return Spacepoints(len(xyz), xyz.ctypes.data, None if vr is None else vr.ctypes.data,
                       None if vz is None else vz.ctypes.data, None if mi is None else mi.ctypes.data)


class seeding_algorithm:  # noqa: N801  (traccc's class name)
    """traccc's triplet seeding on the MI355X."""

    def __init__(self, finder: SeedfinderConfig | None = None,
                 grid: SpacepointGridConfig | None = None,
                 filter: SeedfilterConfig | None = None, stream: int | None = None):
        from .ckf import _check
        self.finder = finder if finder is not None else seedfinder_config()
        self.grid = grid if grid is not None else spacepoint_grid_config(self.finder)
        self.filter = filter if filter is not None else seedfilter_config()
        h = C.c_void_p()
        _check(_lib().ckf_seeding_create(C.byref(self.finder), C.byref(self.grid),
                                         C.byref(self.filter), C.byref(h)))
        self._h = h
        self._L = _lib()
        if stream is not None:
            _check(_lib().ckf_seeding_set_stream(self._h, C.c_void_p(stream)))

    def __del__(self):
        lib = getattr(self, "_L", None)
        if getattr(self, "_h", None) and lib is not None:
            lib.ckf_seeding_destroy(self._h)
            self._h = None

    def __call__(self, global_xyz, radius_variance=None, z_variance=None) -> SeedCollection:
        from .ckf import _check
        keep = []
        sp = _sp_struct(global_xyz, radius_variance, z_variance, keep=keep)
        n = C.c_uint32()
        _check(_lib().ckf_seeding_run(self._h, C.byref(sp), C.byref(n)))
        return self._download(n.value)

    def run_device(self, d_global, d_radius_variance=None, d_z_variance=None) -> SeedCollection:
        """Spacepoints as torch tensors in HBM ([n, 3] float32)."""
        from .ckf import _check
        sp = Spacepoints(d_global.shape[0], d_global.data_ptr(),
                         None if d_radius_variance is None else d_radius_variance.data_ptr(),
                         None if d_z_variance is None else d_z_variance.data_ptr(), None)
        n = C.c_uint32()
        _check(_lib().ckf_seeding_run_device(self._h, C.byref(sp), C.byref(n)))
        return self._download(n.value)

    def _download(self, n: int) -> SeedCollection:
        from .ckf import _check
        out = SeedCollection(np.zeros(n, np.uint32), np.zeros(n, np.uint32), np.zeros(n, np.uint32),
                             np.zeros(n, np.float32))
        s = Seeds(n, out.bottom.ctypes.data, out.middle.ctypes.data, out.top.ctypes.data,
                  out.weight.ctypes.data)
        _check(_lib().ckf_seeding_result_download(self._h, C.byref(s)))
        return out


def spacepoint_formation(ddet, devents, event: int = 0):
    """One spacepoint per 2D measurement of an event (device kernel):
    ([n, 3] global positions, [n] measurement indices)."""
    from .ckf import _check
    n = C.c_uint32()
    _check(_lib().ckf_spacepoint_formation(ddet._h, devents._h, event, None, None, C.byref(n)))
    g = np.zeros((max(n.value, 1), 3), np.float32)
    mi = np.zeros(max(n.value, 1), np.uint32)
    _check(_lib().ckf_spacepoint_formation(ddet._h, devents._h, event, g.ctypes.data,
                                           mi.ctypes.data, C.byref(n)))
    return g[:n.value], mi[:n.value]


class track_params_estimation:  # noqa: N801  (traccc's class name)
    """traccc's track-parameter estimation from seeds (device kernel)."""

    def __init__(self, config: TrackParamsEstimationConfig | None = None):
        self.config = config if config is not None else track_params_estimation_config()

    def __call__(self, global_xyz, measurement_index, seeds: SeedCollection, event,
                 bfield) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(surface links [n], vectors [n, 6], covariances [n, 6, 6]) of
        the seeds' bound parameters at their bottom spacepoints; `event` is
        the simulation Event the spacepoints' measurements belong to."""
        from .ckf import _check
        keep = []
        sp = _sp_struct(global_xyz, midx=measurement_index, keep=keep)
        n = len(seeds)
        s = Seeds(n, seeds.bottom.ctypes.data, seeds.middle.ctypes.data, seeds.top.ctypes.data,
                  seeds.weight.ctypes.data)
        m = event.measurements_struct()
        b = (C.c_float * 3)(*[float(v) for v in bfield])
        sl = np.zeros(max(n, 1), np.uint64)
        vec = np.zeros((max(n, 1), 6), np.float32)
        cov = np.zeros((max(n, 1), 6, 6), np.float32)
        _check(_lib().ckf_estimate_track_params(C.byref(self.config), C.byref(sp), C.byref(s),
                                                C.byref(m), C.byref(b), sl.ctypes.data,
                                                vec.ctypes.data, cov.ctypes.data))
        return sl[:n], vec[:n], cov[:n]


def helix_points(q: float, p_xyz, b_tesla: float, s_values, origin=(0.0, 0.0, 0.0)):
    """Points of the exact helix (uniform field along z, traccc units) at
    path lengths s_values from `origin` with momentum p_xyz (GeV): the
    closed form the parameter estimation is checked against."""
    px, py, pz = p_xyz
    p = math.sqrt(px * px + py * py + pz * pz)
    pt = math.hypot(px, py)
    bz = b_tesla * TESLA
    R = pt / (abs(q) * bz)           # mm
    phi0 = math.atan2(py, px)
    h = -math.copysign(1.0, q * bz)  # rotation sense
    out = []
    for s in s_values:
        st = s * pt / p              # transverse arc length
        a = st / R
        x = origin[0] + R * (math.sin(phi0 + h * a) - math.sin(phi0)) * h
        y = origin[1] - R * (math.cos(phi0 + h * a) - math.cos(phi0)) * h
        z = origin[2] + s * pz / p
        out.append((x, y, z))
    return np.asarray(out)
