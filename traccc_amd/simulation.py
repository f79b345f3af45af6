"""Synthetic inputs for the CKF: toy detectors, fields and particle-gun events.

Wraps include/traccc_amd/simulation.h.  The workload definitions (C1..C5)
follow the configurations this project is measured on: barrel-only / barrel +
disks detectors, constant 2 T or a tabulated solenoid map, muon or pion guns
with the stated pT / eta spectra, smearing, noise and dead modules.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math
from dataclasses import dataclass, field

import numpy as np

from . import _ffi as F

SIM_LIB = "libtraccc_amd_sim.so"

DET_BARREL4 = 0
DET_BARREL4_DISK4 = 1
DET_BARREL5_DISK6 = 2
DET_TELESCOPE = 3

MUON_MASS = 0.1056583755  # GeV (tests/core/particle.cpp:18-48 pins this value)
PION_MASS = 0.13957039


def _lib():
    lib = F.load_library(SIM_LIB)
    if not getattr(lib, "_tamd_typed", False):
        lib.tamd_detector_build.argtypes = [C.c_uint32, C.c_float, C.c_uint32, C.c_float,
                                            C.POINTER(F.SimDetector)]
        lib.tamd_detector_build.restype = C.c_int
        lib.tamd_detector_free.argtypes = [C.POINTER(F.SimDetector)]
        lib.tamd_solenoid_map_build.argtypes = [C.c_float] * 6 + [C.POINTER(F.BField),
                                                                  C.POINTER(F.c_float_p)]
        lib.tamd_solenoid_map_build.restype = C.c_int
        lib.tamd_solenoid_map_free.argtypes = [F.c_float_p]
        lib.tamd_solenoid_field_exact.argtypes = [C.c_float, C.c_float, C.c_float, C.c_double,
                                                  C.c_double, C.POINTER(C.c_double),
                                                  C.POINTER(C.c_double)]
        lib.tamd_cel.argtypes = [C.c_double] * 4
        lib.tamd_cel.restype = C.c_double
        lib.tamd_event_generate.argtypes = [C.POINTER(F.Detector), C.POINTER(F.BField),
                                            C.POINTER(F.EventSpec), C.POINTER(F.SimEvent)]
        lib.tamd_event_generate.restype = C.c_int
        lib.tamd_event_free.argtypes = [C.POINTER(F.SimEvent)]
        lib.ckf_finding_config_default.argtypes = [C.POINTER(F.FindingConfig)]
        lib._tamd_typed = True
    return lib


def default_config() -> F.FindingConfig:
    cfg = F.FindingConfig()
    _lib().ckf_finding_config_default(C.byref(cfg))
    return cfg


# ---------------------------------------------------------------------------
# Detector
# ---------------------------------------------------------------------------
class DetectorData:
    """Host detector description (numpy-backed ckf_surface / ckf_layer arrays)."""

    def __init__(self, surfaces: np.ndarray, layers: np.ndarray):
        self.surfaces_raw = surfaces  # bytes of ckf_surface[]
        self.layers_raw = layers
        self.n_surfaces = len(surfaces) // C.sizeof(F.Surface)
        self.n_layers = len(layers) // C.sizeof(F.Layer)
        self._struct = F.Detector(
            C.cast(self.surfaces_raw.ctypes.data, C.POINTER(F.Surface)), self.n_surfaces,
            C.cast(self.layers_raw.ctypes.data, C.POINTER(F.Layer)), self.n_layers)

    @property
    def struct(self) -> F.Detector:
        return self._struct

    def surface(self, i: int) -> F.Surface:
        return self._struct.surfaces[i]

    def layer(self, i: int) -> F.Layer:
        return self._struct.layers[i]

    @property
    def identifiers(self) -> np.ndarray:
        return np.array([self._struct.surfaces[i].identifier for i in range(self.n_surfaces)],
                        dtype=np.uint64)


def build_detector(preset: int, thickness_x0: float = 0.0, telescope_planes: int = 9,
                   telescope_spacing: float = 20.0) -> DetectorData:
    lib = _lib()
    d = F.SimDetector()
    rc = lib.tamd_detector_build(preset, thickness_x0, telescope_planes, telescope_spacing,
                                 C.byref(d))
    if rc != 0:
        raise ValueError(f"tamd_detector_build failed ({rc})")
    try:
        sb = np.frombuffer(C.string_at(d.surfaces, d.n_surfaces * C.sizeof(F.Surface)),
                           dtype=np.uint8).copy()
        lb = np.frombuffer(C.string_at(d.layers, d.n_layers * C.sizeof(F.Layer)),
                           dtype=np.uint8).copy()
    finally:
        lib.tamd_detector_free(C.byref(d))
    return DetectorData(sb, lb)


# ---------------------------------------------------------------------------
# Field
# ---------------------------------------------------------------------------
class FieldData:
    def __init__(self, struct: F.BField, table: np.ndarray | None = None):
        self.table = table
        if table is not None:
            struct.map = F.ptr(table, C.c_float)
        self._struct = struct

    @property
    def struct(self) -> F.BField:
        return self._struct


def constant_field(bz_tesla: float = 2.0, bx: float = 0.0, by: float = 0.0) -> FieldData:
    b = F.BField()
    b.kind = 0
    b.b[0], b.b[1], b.b[2] = bx, by, bz_tesla
    return FieldData(b)


def solenoid_map(radius=1200.0, length=5600.0, b_centre=2.0, r_max=1500.0, z_max=3000.0,
                 pitch=10.0) -> FieldData:
    lib = _lib()
    b = F.BField()
    t = F.c_float_p()
    rc = lib.tamd_solenoid_map_build(radius, length, b_centre, r_max, z_max, pitch, C.byref(b),
                                     C.byref(t))
    if rc != 0:
        raise ValueError("solenoid map build failed")
    n = 2 * b.n_r * b.n_z
    table = np.ctypeslib.as_array(t, shape=(n,)).copy()
    lib.tamd_solenoid_map_free(t)
    return FieldData(b, table)


def solenoid_exact(r: float, z: float, radius=1200.0, length=5600.0, b_centre=2.0):
    br, bz = C.c_double(), C.c_double()
    _lib().tamd_solenoid_field_exact(radius, length, b_centre, r, z, C.byref(br), C.byref(bz))
    return br.value, bz.value


def cel(kc, p, a, b) -> float:
    return _lib().tamd_cel(kc, p, a, b)


# ---------------------------------------------------------------------------
# Events
# ---------------------------------------------------------------------------
@dataclass
class Event:
    local_position: np.ndarray
    local_variance: np.ndarray
    dimensions: np.ndarray
    time: np.ndarray
    diameter: np.ndarray
    identifier: np.ndarray
    surface_link: np.ndarray
    subspace: np.ndarray
    cluster_index: np.ndarray
    meas_particle: np.ndarray
    seed_surface_link: np.ndarray
    seed_vector: np.ndarray
    seed_covariance: np.ndarray
    seed_particle: np.ndarray
    particle_vertex: np.ndarray
    particle_momentum: np.ndarray
    particle_charge: np.ndarray
    particle_n_meas: np.ndarray
    _keep: list = field(default_factory=list, repr=False)

    @property
    def n_meas(self) -> int:
        return len(self.dimensions)

    @property
    def n_seeds(self) -> int:
        return len(self.seed_particle)

    def measurements_struct(self) -> F.Measurements:
        m = F.Measurements()
        m.size = self.n_meas
        m.local_position = F.ptr(self.local_position, C.c_float)
        m.local_variance = F.ptr(self.local_variance, C.c_float)
        m.dimensions = F.ptr(self.dimensions, C.c_uint32)
        m.time = F.ptr(self.time, C.c_float)
        m.diameter = F.ptr(self.diameter, C.c_float)
        m.identifier = F.ptr(self.identifier, C.c_uint32)
        m.surface_link = F.ptr(self.surface_link, C.c_uint64)
        m.subspace = F.ptr(self.subspace, C.c_uint8)
        m.cluster_index = F.ptr(self.cluster_index, C.c_uint32)
        return m

    def seeds_struct(self) -> F.BoundParams:
        s = F.BoundParams()
        s.size = self.n_seeds
        s.surface_link = F.ptr(self.seed_surface_link, C.c_uint64)
        s.vector = F.ptr(self.seed_vector, C.c_float)
        s.covariance = F.ptr(self.seed_covariance, C.c_float)
        return s


@dataclass
class GunSpec:
    n_particles: int = 100
    pdg: int = 13
    mass: float = MUON_MASS
    random_charge: bool = True
    pt_distribution: int = 0  # 0 uniform, 1 log-uniform, 2 shifted exponential
    pt_a: float = 1.0
    pt_b: float = 10.0
    eta_min: float = -1.0
    eta_max: float = 1.0
    phi_min: float = -math.pi
    phi_max: float = math.pi
    vertex: tuple = (0.0, 0.0, 0.0)
    vertex_sigma: tuple = (0.1, 0.1, 50.0)
    fixed_direction: bool = False
    dir_phi: float = 0.0
    dir_theta: float = math.pi / 2
    smear: tuple = (0.05, 0.05)
    noise_fraction: float = 0.0
    dead_fraction: float = 0.0
    seed_qop_rel_sigma: float = 0.1
    seed_angle_sigma: float = 0.001
    seed_time_sigma: float = 1.0
    max_hits_per_particle: int = 64
    seed: int = 42

    def struct(self) -> F.EventSpec:
        s = F.EventSpec()
        s.n_particles = self.n_particles
        s.pdg = self.pdg
        s.mass = self.mass
        s.random_charge = int(self.random_charge)
        s.pt_distribution = self.pt_distribution
        s.pt_a, s.pt_b = self.pt_a, self.pt_b
        s.eta_min, s.eta_max = self.eta_min, self.eta_max
        s.phi_min, s.phi_max = self.phi_min, self.phi_max
        for i in range(3):
            s.vertex[i] = self.vertex[i]
            s.vertex_sigma[i] = self.vertex_sigma[i]
        s.fixed_direction = int(self.fixed_direction)
        s.dir_phi, s.dir_theta = self.dir_phi, self.dir_theta
        s.smear[0], s.smear[1] = self.smear
        s.noise_fraction = self.noise_fraction
        s.dead_fraction = self.dead_fraction
        s.seed_qop_rel_sigma = self.seed_qop_rel_sigma
        s.seed_angle_sigma = self.seed_angle_sigma
        s.seed_time_sigma = self.seed_time_sigma
        s.max_hits_per_particle = self.max_hits_per_particle
        s.seed = self.seed
        return s


def generate_event(det: DetectorData, fld: FieldData, spec: GunSpec) -> Event:
    lib = _lib()
    ev = F.SimEvent()
    sspec = spec.struct()
    rc = lib.tamd_event_generate(C.byref(det.struct), C.byref(fld.struct), C.byref(sspec),
                                 C.byref(ev))
    if rc != 0:
        raise ValueError(f"event generation failed ({rc})")
    try:
        nm, ns, npart = ev.n_meas, ev.n_seeds, ev.n_particles
        out = Event(
            local_position=F.np_from(ev.local_position, 2 * nm, np.float32, (nm, 2)),
            local_variance=F.np_from(ev.local_variance, 2 * nm, np.float32, (nm, 2)),
            dimensions=F.np_from(ev.dimensions, nm, np.uint32),
            time=F.np_from(ev.time, nm, np.float32),
            diameter=F.np_from(ev.diameter, nm, np.float32),
            identifier=F.np_from(ev.identifier, nm, np.uint32),
            surface_link=F.np_from(ev.surface_link, nm, np.uint64),
            subspace=F.np_from(ev.subspace, 2 * nm, np.uint8, (nm, 2)),
            cluster_index=F.np_from(ev.cluster_index, nm, np.uint32),
            meas_particle=F.np_from(ev.meas_particle, nm, np.int32),
            seed_surface_link=F.np_from(ev.seed_surface_link, ns, np.uint64),
            seed_vector=F.np_from(ev.seed_vector, 6 * ns, np.float32, (ns, 6)),
            seed_covariance=F.np_from(ev.seed_covariance, 36 * ns, np.float32, (ns, 6, 6)),
            seed_particle=F.np_from(ev.seed_particle, ns, np.uint32),
            particle_vertex=F.np_from(ev.particle_vertex, 3 * npart, np.float32, (npart, 3)),
            particle_momentum=F.np_from(ev.particle_momentum, 3 * npart, np.float32, (npart, 3)),
            particle_charge=F.np_from(ev.particle_charge, npart, np.float32),
            particle_n_meas=F.np_from(ev.particle_n_meas, npart, np.uint32),
        )
    finally:
        lib.tamd_event_free(C.byref(ev))
    return out


# ---------------------------------------------------------------------------
# Workload definitions
# ---------------------------------------------------------------------------
def with_seeds(ev: Event, surface_link, vector, covariance, particle) -> Event:
    """A copy of ev whose seeds are the given bound parameters (surface
    identifiers, [n, 6] vectors, [n, 6, 6] covariances) and their particles."""
    import dataclasses
    n = len(surface_link)
    return dataclasses.replace(
        ev, seed_surface_link=np.ascontiguousarray(surface_link, np.uint64),
        seed_vector=np.ascontiguousarray(np.asarray(vector, np.float32).reshape(n, 6)),
        seed_covariance=np.ascontiguousarray(np.asarray(covariance, np.float32).reshape(n, 6, 6)),
        seed_particle=np.ascontiguousarray(particle, np.uint32), _keep=[])


def workload(name: str, seed: int = 42, n_particles: int | None = None):
    """(detector, field, gun spec) of a named workload.

    C1: 4 barrel layers, no material, 2 T; 100 muons, pT U[1,10] GeV, |eta| < 1.
    C2: + 2x4 disks, 1% X0, 5% noise; 10k muons, pT log-U[0.5,100], |eta| < 2.5.
    C3: C2 detector with the solenoid map; per-event multiplicity is drawn by
        the batch driver (log-U[500, 50000]).
    C4: C2 detector; 20k pions, pT 0.3 + Exp(1) GeV, |eta| < 2.5, 10% noise.
    C5: 5 barrel + 2x6 disks, 2% X0, 3% dead, 10% noise, solenoid map;
        60k pions per event with the C4 spectrum.
    """
    if name == "C1":
        det = build_detector(DET_BARREL4, 0.0)
        fld = constant_field(2.0)
        gun = GunSpec(n_particles=n_particles or 100, seed=seed)
    elif name in ("C2", "C3"):
        det = build_detector(DET_BARREL4_DISK4, 0.01)
        fld = constant_field(2.0) if name == "C2" else solenoid_map()
        gun = GunSpec(n_particles=n_particles or 10000, pt_distribution=1, pt_a=0.5, pt_b=100.0,
                      eta_min=-2.5, eta_max=2.5, noise_fraction=0.05, seed=seed)
    elif name == "C4":
        det = build_detector(DET_BARREL4_DISK4, 0.01)
        fld = constant_field(2.0)
        gun = GunSpec(n_particles=n_particles or 20000, pdg=211, mass=PION_MASS,
                      pt_distribution=2, pt_a=0.3, pt_b=1.0, eta_min=-2.5, eta_max=2.5,
                      noise_fraction=0.10, seed=seed)
    elif name == "C5":
        det = build_detector(DET_BARREL5_DISK6, 0.02)
        fld = solenoid_map()
        gun = GunSpec(n_particles=n_particles or 60000, pdg=211, mass=PION_MASS,
                      pt_distribution=2, pt_a=0.3, pt_b=1.0, eta_min=-2.5, eta_max=2.5,
                      noise_fraction=0.10, dead_fraction=0.03, seed=seed)
    else:
        raise ValueError(name)
    return det, fld, gun


def batch_workload(name: str = "C3", n_events: int = 256, seed: int = 42,
                   n_min: int = 500, n_max: int = 50000):
    """(detector, field, [gun spec per event]) of a batched workload: per-event
    particle counts log-uniform in [n_min, n_max] (C3: 256 events, 500-50,000
    muons, solenoid map; C5: 64 events of 60,000 pions), event e seeded with
    seed + e."""
    det, fld, base = workload(name, seed=seed)
    rng = np.random.default_rng(seed)
    if name == "C5":
        counts = [60000] * n_events
    else:
        counts = np.exp(rng.uniform(math.log(n_min), math.log(n_max), n_events)).astype(int).tolist()
    guns = [dataclasses.replace(base, n_particles=int(n), seed=seed + e) for e, n in enumerate(counts)]
    return det, fld, guns


def generate_events(det: DetectorData, fld: FieldData, specs, threads: int = 8):
    """generate_event over many specs on a thread pool (the generator runs in
    C and releases the interpreter lock)."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        return list(ex.map(lambda sp: generate_event(det, fld, sp), specs))


def finding_config_for(name: str) -> F.FindingConfig:
    """finding_config used for a workload (muon / pion hypothesis, pT floor)."""
    cfg = default_config()
    if name in ("C1", "C2", "C3"):
        cfg.ptc_pdg, cfg.ptc_mass, cfg.ptc_charge = 13, MUON_MASS, -1.0
    else:
        cfg.ptc_pdg, cfg.ptc_mass, cfg.ptc_charge = 211, PION_MASS, 1.0
    cfg.min_pT = 0.2
    cfg.run_smoother = 0
    return cfg
