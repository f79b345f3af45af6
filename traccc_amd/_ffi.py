"""ctypes mirrors of the C ABI structs in include/traccc_amd/ckf.h and
include/traccc_amd/simulation.h, plus library loading.

Only plain data crosses the boundary; numpy arrays provide the host buffers.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(PKG_DIR, "_lib")

c_float_p = C.POINTER(C.c_float)
c_u32_p = C.POINTER(C.c_uint32)
c_u64_p = C.POINTER(C.c_uint64)
c_u16_p = C.POINTER(C.c_uint16)
c_u8_p = C.POINTER(C.c_uint8)
c_i32_p = C.POINTER(C.c_int32)


class PropagationConfig(C.Structure):
    _fields_ = [
        ("step_constraint", C.c_float),
        ("rk_error_tol", C.c_float),
        ("min_stepsize", C.c_float),
        ("path_limit", C.c_float),
        ("overstep_tolerance", C.c_float),
        ("min_mask_tolerance", C.c_float),
        ("max_mask_tolerance", C.c_float),
        ("mask_tolerance_scalor", C.c_float),
        ("on_surface_tolerance", C.c_float),
        ("search_window", C.c_uint32 * 2),
    ]


class FindingConfig(C.Structure):
    _fields_ = [
        ("max_num_branches_per_seed", C.c_uint32),
        ("max_num_branches_per_surface", C.c_uint32),
        ("min_track_candidates_per_track", C.c_uint32),
        ("max_track_candidates_per_track", C.c_uint32),
        ("max_num_skipping_per_cand", C.c_uint32),
        ("max_num_consecutive_skipped", C.c_uint32),
        ("max_num_tracks_per_measurement", C.c_uint32),
        ("min_measurement_voting_fraction", C.c_float),
        ("run_pkf", C.c_uint32),
        ("run_smoother", C.c_uint32),
        ("min_step_length_for_next_surface", C.c_float),
        ("max_step_counts_for_next_surface", C.c_uint32),
        ("chi2_max", C.c_float),
        ("propagation", PropagationConfig),
        ("min_p", C.c_float),
        ("min_pT", C.c_float),
        ("ptc_pdg", C.c_int32),
        ("ptc_mass", C.c_float),
        ("ptc_charge", C.c_float),
        ("duplicate_removal_minimum_length", C.c_uint32),
        ("initial_links_per_seed", C.c_uint32),
    ]


class FittingConfig(C.Structure):
    """include/traccc_amd/fitting.h ckf_fitting_config."""
    _fields_ = [
        ("n_iterations", C.c_uint32),
        ("propagation", PropagationConfig),
        ("max_step_counts", C.c_uint32),
        ("min_p", C.c_float),
        ("min_pT", C.c_float),
        ("ptc_pdg", C.c_int32),
        ("ptc_mass", C.c_float),
        ("ptc_charge", C.c_float),
        ("covariance_inflation_factor", C.c_float),
        ("surface_sequence_size_factor", C.c_uint32),
        ("min_surface_sequence_capacity", C.c_uint32),
        ("do_precise_hole_count", C.c_uint32),
    ]


class FittedTracks(C.Structure):
    """include/traccc_amd/fitting.h ckf_fitted_tracks."""
    _fields_ = [
        ("n_tracks", C.c_uint32),
        ("n_states", C.c_uint32),
        ("fit_outcome", c_u16_p),
        ("params_surface_link", c_u64_p),
        ("params_vector", c_float_p),
        ("params_covariance", c_float_p),
        ("ndf", c_float_p),
        ("chi2", c_float_p),
        ("pval", c_float_p),
        ("nholes", c_u32_p),
        ("state_offsets", c_u32_p),
        ("state_flags", c_u8_p),
        ("filtered_chi2", c_float_p),
        ("smoothed_chi2", c_float_p),
        ("backward_chi2", c_float_p),
        ("state_surface_link", c_u64_p),
        ("filtered_vector", c_float_p),
        ("filtered_covariance", c_float_p),
        ("smoothed_vector", c_float_p),
        ("smoothed_covariance", c_float_p),
        ("measurement_index", c_u32_p),
    ]


class Surface(C.Structure):
    _fields_ = [
        ("center", C.c_float * 3),
        ("axis_u", C.c_float * 3),
        ("axis_v", C.c_float * 3),
        ("normal", C.c_float * 3),
        ("bound", C.c_float * 3),
        ("thickness_x0", C.c_float),
        ("bound_kind", C.c_uint32),
        ("layer", C.c_uint32),
        ("sensitive", C.c_uint32),
        ("reserved", C.c_uint32),
        ("identifier", C.c_uint64),
    ]


class Layer(C.Structure):
    _fields_ = [
        ("kind", C.c_uint32),
        ("first_surface", C.c_uint32),
        ("n_phi", C.c_uint32),
        ("n_long", C.c_uint32),
        ("phi0", C.c_float),
        ("long0", C.c_float),
        ("long_pitch", C.c_float),
        ("env_r", C.c_float * 2),
        ("env_z", C.c_float * 2),
        ("normal", C.c_float * 3),
    ]


class Detector(C.Structure):
    _fields_ = [
        ("surfaces", C.POINTER(Surface)),
        ("n_surfaces", C.c_uint32),
        ("layers", C.POINTER(Layer)),
        ("n_layers", C.c_uint32),
    ]


class BField(C.Structure):
    _fields_ = [
        ("kind", C.c_uint32),
        ("b", C.c_float * 3),
        ("map", c_float_p),
        ("n_r", C.c_uint32),
        ("n_z", C.c_uint32),
        ("r_min", C.c_float),
        ("z_min", C.c_float),
        ("dr", C.c_float),
        ("dz", C.c_float),
    ]


class Measurements(C.Structure):
    _fields_ = [
        ("size", C.c_uint32),
        ("local_position", c_float_p),
        ("local_variance", c_float_p),
        ("dimensions", c_u32_p),
        ("time", c_float_p),
        ("diameter", c_float_p),
        ("identifier", c_u32_p),
        ("surface_link", c_u64_p),
        ("subspace", c_u8_p),
        ("cluster_index", c_u32_p),
    ]


class BoundParams(C.Structure):
    _fields_ = [
        ("size", C.c_uint32),
        ("surface_link", c_u64_p),
        ("vector", c_float_p),
        ("covariance", c_float_p),
    ]


class TrackCandidates(C.Structure):
    _fields_ = [
        ("n_tracks", C.c_uint32),
        ("n_constituents", C.c_uint32),
        ("fit_outcome", c_u16_p),
        ("params_surface_link", c_u64_p),
        ("params_vector", c_float_p),
        ("params_covariance", c_float_p),
        ("ndf", c_float_p),
        ("chi2", c_float_p),
        ("pval", c_float_p),
        ("nholes", c_u32_p),
        ("constituent_offsets", c_u32_p),
        ("constituent_type", c_u16_p),
        ("constituent_index", c_u32_p),
    ]


class RunStats(C.Structure):
    _fields_ = [
        ("n_steps", C.c_uint32),
        ("n_links", C.c_uint64),
        ("n_propagations", C.c_uint64),
        ("n_tips", C.c_uint64),
        ("n_regrows", C.c_uint32),
        ("reserved", C.c_uint32),
        ("n_scored", C.c_uint64),
    ]


# ---- simulation.h ----
class SimDetector(C.Structure):
    _fields_ = [
        ("surfaces", C.POINTER(Surface)),
        ("n_surfaces", C.c_uint32),
        ("layers", C.POINTER(Layer)),
        ("n_layers", C.c_uint32),
    ]


class EventSpec(C.Structure):
    _fields_ = [
        ("n_particles", C.c_uint32),
        ("pdg", C.c_int32),
        ("mass", C.c_float),
        ("random_charge", C.c_uint32),
        ("pt_distribution", C.c_uint32),
        ("pt_a", C.c_float),
        ("pt_b", C.c_float),
        ("eta_min", C.c_float),
        ("eta_max", C.c_float),
        ("phi_min", C.c_float),
        ("phi_max", C.c_float),
        ("vertex", C.c_float * 3),
        ("vertex_sigma", C.c_float * 3),
        ("fixed_direction", C.c_uint32),
        ("dir_phi", C.c_float),
        ("dir_theta", C.c_float),
        ("smear", C.c_float * 2),
        ("noise_fraction", C.c_float),
        ("dead_fraction", C.c_float),
        ("seed_qop_rel_sigma", C.c_float),
        ("seed_angle_sigma", C.c_float),
        ("seed_time_sigma", C.c_float),
        ("max_hits_per_particle", C.c_uint32),
        ("seed", C.c_uint64),
    ]


class SimEvent(C.Structure):
    _fields_ = [
        ("n_meas", C.c_uint32),
        ("local_position", c_float_p),
        ("local_variance", c_float_p),
        ("dimensions", c_u32_p),
        ("time", c_float_p),
        ("diameter", c_float_p),
        ("identifier", c_u32_p),
        ("surface_link", c_u64_p),
        ("subspace", c_u8_p),
        ("cluster_index", c_u32_p),
        ("meas_particle", c_i32_p),
        ("n_seeds", C.c_uint32),
        ("seed_surface_link", c_u64_p),
        ("seed_vector", c_float_p),
        ("seed_covariance", c_float_p),
        ("seed_particle", c_u32_p),
        ("n_particles", C.c_uint32),
        ("particle_vertex", c_float_p),
        ("particle_momentum", c_float_p),
        ("particle_charge", c_float_p),
        ("particle_n_meas", c_u32_p),
    ]


def ptr(arr: np.ndarray, ctype):
    """Pointer to a contiguous numpy array's data."""
    assert arr.flags["C_CONTIGUOUS"]
    return arr.ctypes.data_as(C.POINTER(ctype))


def np_from(p, n: int, dtype, shape=None):
    """Copy n elements from a C pointer into a fresh numpy array."""
    if n == 0:
        out = np.zeros(0, dtype=dtype)
    else:
        out = np.ctypeslib.as_array(p, shape=(n,)).astype(dtype, copy=True)
    if shape is not None:
        out = out.reshape(shape)
    return out


_libs: dict[str, C.CDLL] = {}


def load_library(name: str) -> C.CDLL:
    """Load a library built in-tree by traccc_amd.build; raise if missing."""
    if name in _libs:
        return _libs[name]
    path = os.path.join(LIB_DIR, name)
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is not built; run `python -m traccc_amd.build` (or __graft_entry__.build())")
    lib = C.CDLL(path)
    _libs[name] = lib
    return lib
