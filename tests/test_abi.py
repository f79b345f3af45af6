"""The C-ABI libraries load and export every function their headers declare;
configuration errors surface as the reference's invalid_argument (ValueError)
before any device call."""
import ctypes as C
import os
import re

import pytest

from traccc_amd import _ffi as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include", "traccc_amd")
DECL = re.compile(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b((?:ckf|tamd)_\w+)\s*\(", re.M)


def declared(header):
    text = open(os.path.join(INC, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(m.group(1) for m in DECL.finditer(text)))


HIP_LIB = os.path.join(F.LIB_DIR, "libtraccc_amd.so")


@pytest.mark.parametrize("header,lib", [("ckf.h", "libtraccc_amd.so"),
                                        ("ckf_debug.h", "libtraccc_amd.so"),
                                        ("scorer.h", "libtraccc_amd.so"),
                                        ("fitting.h", "libtraccc_amd.so"),
                                        ("ambiguity.h", "libtraccc_amd.so"),
                                        ("seeding.h", "libtraccc_amd.so"),
                                        ("io.h", "libtraccc_amd.so"),
                                        ("simulation.h", "libtraccc_amd_sim.so")])
def test_every_declared_symbol_is_exported(header, lib):
    path = os.path.join(F.LIB_DIR, lib)
    if not os.path.exists(path):
        pytest.skip(f"{lib} not built")
    names = declared(header)
    assert len(names) >= 2
    so = C.CDLL(path)
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


def test_headers_declare_the_operator():
    names = declared("ckf.h")
    for n in ("ckf_algorithm_create", "ckf_find_tracks", "ckf_device_detector_create",
              "ckf_device_event_create", "ckf_tracks_download", "ckf_last_error"):
        assert n in names


@pytest.mark.skipif(not os.path.exists(HIP_LIB), reason="HIP library not built")
def test_config_defaults_and_validation():
    from traccc_amd import ckf
    cfg = ckf.finding_config()
    # traccc finding_config defaults (core/include/traccc/finding/finding_config.hpp:72-146)
    assert cfg.max_num_branches_per_seed == 10
    assert cfg.max_num_branches_per_surface == 1
    assert cfg.min_track_candidates_per_track == 3
    assert cfg.max_track_candidates_per_track == 100
    assert cfg.max_num_skipping_per_cand == 3
    assert cfg.max_num_consecutive_skipped == 1
    assert cfg.max_num_tracks_per_measurement == 0
    assert abs(cfg.min_measurement_voting_fraction - 0.5) < 1e-7
    assert cfg.run_pkf == 0
    assert cfg.run_smoother == 1  # e_mbf
    assert abs(cfg.min_step_length_for_next_surface - 1.2) < 1e-6
    assert cfg.max_step_counts_for_next_surface == 100
    assert cfg.chi2_max == 30.0
    assert abs(cfg.min_p - 0.1) < 1e-7 and abs(cfg.min_pT - 0.6) < 1e-7
    assert cfg.ptc_pdg == 211
    assert cfg.duplicate_removal_minimum_length == 5
    assert cfg.initial_links_per_seed == 100
    # traccc throws std::invalid_argument for these
    bad = ckf.finding_config()
    bad.min_track_candidates_per_track = 0
    with pytest.raises(ValueError):
        ckf.combinatorial_kalman_filter_algorithm(bad)
    bad = ckf.finding_config()
    bad.min_step_length_for_next_surface = 0.05
    with pytest.raises(ValueError):
        ckf.combinatorial_kalman_filter_algorithm(bad)
    bad = ckf.finding_config()
    bad.run_smoother = 2  # Kalman smoother without a fitter
    with pytest.raises(ValueError):
        ckf.combinatorial_kalman_filter_algorithm(bad)
    bad = ckf.finding_config()
    bad.max_num_tracks_per_measurement = 2
    bad.min_measurement_voting_fraction = 1.5  # a fraction above one
    with pytest.raises(ValueError):
        ckf.combinatorial_kalman_filter_algorithm(bad)
