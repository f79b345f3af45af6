"""Progressive Kalman filter: finding_config.run_pkf with one branch per seed or
per surface (the operator's dispatch, core/src/finding/combinatorial_kalman_filter_algorithm.cpp:47-57;
core/include/traccc/finding/details/run_progressive_kalman_filter.hpp:55-258,
progressive_kalman_filter.hpp:53-209, actors/measurement_kalman_updater.hpp:94-260).

Pins on the oracle (CPU):
  * sparse telescope tracks: one entry per seed, each with one measurement per
    plane, the CKF's tracks exactly (no overlaps, one candidate per surface);
  * n identical tracks over P planes (the combinatorics setting): n entries of
    P measurements each, where the CKF with unlimited branching gives n^(P+1);
  * a seed below min_pT / 10 or a track shorter than
    min_track_candidates_per_track gives an empty entry (the reference pushes a
    default track for it); the entry count always equals the seed count;
  * the dispatch: run_pkf with more than one branch per seed and per surface
    runs the CKF.
GPU (marked gpu): the HIP walk kernel gives the oracle's entries.
"""
import math

import numpy as np
import pytest

from helpers import compare_tracks, run_finder
from traccc_amd import ckf
from traccc_amd import simulation as S

BACKENDS = ["oracle", pytest.param("gpu", marks=pytest.mark.gpu)]


def _telescope(n_planes=9):
    return S.build_detector(S.DET_TELESCOPE, 0.5 / 93.7, telescope_planes=n_planes,
                            telescope_spacing=20.0)


def _muon_cfg(**kw):
    cfg = S.default_config()
    cfg.ptc_pdg, cfg.ptc_mass, cfg.ptc_charge = 13, S.MUON_MASS, -1.0
    cfg.run_smoother = 0
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


def _gpu(backend, request):
    if backend == "gpu":
        request.getfixturevalue("gpu_available")


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("n_truth", [1, 10, 50])
def test_sparse_tracks_match_ckf(backend, n_truth, request):
    _gpu(backend, request)
    det = _telescope(9)
    fld = S.constant_field(2.0)
    gun = S.GunSpec(n_particles=n_truth, pdg=13, random_charge=True, pt_a=1.0, pt_b=10.0,
                    eta_min=-0.3, eta_max=0.3, phi_min=-0.3, phi_max=0.3,
                    vertex_sigma=(0.0, 0.0, 0.0), seed=7)
    ev = S.generate_event(det, fld, gun)
    cfg = _muon_cfg(chi2_max=200.0, min_p=0.0, min_pT=0.01)
    ref = run_finder("oracle", cfg, det, fld, ev)
    cfg.run_pkf = 1
    r = run_finder(backend, cfg, det, fld, ev)
    assert len(r["ndf"]) == ev.n_seeds == n_truth
    assert np.all(np.diff(r["offsets"]) == 9)
    assert np.array_equal(r["offsets"], ref["offsets"])
    assert np.array_equal(r["constituents"], ref["constituents"])
    # ndf = sum(dim) - 5 with 2-D measurements
    assert np.all(np.asarray(r["ndf"]) == 2 * 9 - 5)


@pytest.mark.parametrize("backend", BACKENDS)
def test_identical_tracks_one_entry_per_seed(backend, request):
    _gpu(backend, request)
    n_planes, n_truth = 9, 3
    det = _telescope(n_planes)
    fld = S.constant_field(0.0, bx=2.0)
    gun = S.GunSpec(n_particles=n_truth, pdg=13, random_charge=False, pt_a=100.0, pt_b=100.0,
                    vertex_sigma=(0.0, 0.0, 0.0), fixed_direction=True, dir_phi=0.0,
                    dir_theta=math.pi / 2, seed=11)
    ev = S.generate_event(det, fld, gun)
    cfg = _muon_cfg(max_num_branches_per_seed=1, max_num_branches_per_surface=10, chi2_max=30.0,
                    duplicate_removal_minimum_length=100, run_pkf=1)
    r = run_finder(backend, cfg, det, fld, ev)
    assert len(r["ndf"]) == n_truth
    assert np.all(np.diff(r["offsets"]) == n_planes)
    # the dispatch: more than one branch per seed and per surface runs the CKF
    cfg = _muon_cfg(max_num_branches_per_seed=2**32 - 1, max_num_branches_per_surface=10,
                    chi2_max=30.0, duplicate_removal_minimum_length=100, run_pkf=1)
    r = run_finder(backend, cfg, det, fld, ev)
    assert len(r["ndf"]) == n_truth ** (n_planes + 1)


@pytest.mark.parametrize("backend", BACKENDS)
def test_empty_entries(backend, request):
    _gpu(backend, request)
    det = _telescope(9)
    fld = S.constant_field(2.0)
    gun = S.GunSpec(n_particles=6, pdg=13, random_charge=False, pt_a=2.0, pt_b=5.0,
                    eta_min=-0.2, eta_max=0.2, phi_min=-0.2, phi_max=0.2,
                    vertex_sigma=(0.0, 0.0, 0.0), seed=3)
    ev = S.generate_event(det, fld, gun)
    # seed 1: q/p of a 1 MeV particle (below min_pT / 10); seed 4: too short
    ev.seed_vector = np.ascontiguousarray(ev.seed_vector.copy())
    ev.seed_vector[1, 4] = -1.0 / 0.001
    cfg = _muon_cfg(chi2_max=200.0, min_p=0.0, min_pT=0.1, run_pkf=1)
    r = run_finder(backend, cfg, det, fld, ev)
    lens = np.diff(r["offsets"])
    assert len(lens) == ev.n_seeds
    assert lens[1] == 0 and np.count_nonzero(lens) == ev.n_seeds - 1
    assert r["ndf"][1] == 0 and r["chi2"][1] == 0
    cfg.min_track_candidates_per_track = 10  # more states than planes
    r = run_finder(backend, cfg, det, fld, ev)
    assert len(r["ndf"]) == ev.n_seeds and np.all(np.diff(r["offsets"]) == 0)


def test_pkf_takes_overlap_hits_the_ckf_skips():
    """The progressive filter has no minimum step between surfaces, so it also
    takes the second hit of overlapping barrel modules (< 1.2 mm apart), which
    the CKF's ckf_aborter skips: its tracks are never shorter."""
    det, fld, gun = S.workload("C1", seed=42)
    ev = S.generate_event(det, fld, gun)
    cfg = S.finding_config_for("C1")
    cfg.max_num_branches_per_seed = 1
    c = run_finder("oracle", cfg, det, fld, ev)
    cfg.run_pkf = 1
    p = run_finder("oracle", cfg, det, fld, ev)
    assert len(p["ndf"]) == ev.n_seeds
    assert np.diff(p["offsets"]).sum() >= np.diff(c["offsets"]).sum()
    # every CKF track's measurements appear in the PKF entry of a seed
    pk = [set(p["constituents"][p["offsets"][i]:p["offsets"][i + 1]].tolist())
          for i in range(len(p["offsets"]) - 1)]
    contained = 0
    for i in range(len(c["offsets"]) - 1):
        t = set(c["constituents"][c["offsets"][i]:c["offsets"][i + 1]].tolist())
        contained += any(t <= s for s in pk)
    assert contained >= 0.9 * (len(c["offsets"]) - 1)


@pytest.mark.gpu
@pytest.mark.parametrize("workload,n", [("C1", 100), ("C2", 10000), ("C4", 4000)])
def test_gpu_pkf_matches_oracle(workload, n, gpu_available):
    det, fld, gun = S.workload(workload, n_particles=n, seed=42)
    ev = S.generate_event(det, fld, gun)
    cfg = S.finding_config_for(workload)
    cfg.run_pkf = 1
    print(f"[pkf] {workload} n={n}", flush=True)
    g = run_finder("gpu", cfg, det, fld, ev)
    o = run_finder("oracle", cfg, det, fld, ev)
    assert len(g["ndf"]) == len(o["ndf"]) == ev.n_seeds
    match, ng, no = compare_tracks(g["offsets"], g["constituents"], o["offsets"], o["constituents"])
    same = np.array_equal(g["offsets"], o["offsets"]) and np.array_equal(g["constituents"],
                                                                        o["constituents"])
    print(f"entries {ng}, index-set match {match:.5f}, identical {same}")
    assert match >= 0.995


@pytest.mark.gpu
def test_gpu_pkf_batched_equals_single(gpu_available):
    det, fld, _ = S.workload("C2", n_particles=10)
    cfg = S.finding_config_for("C2")
    cfg.run_pkf = 1
    events = [S.generate_event(det, fld, S.workload("C2", n_particles=n, seed=s)[2])
              for n, s in ((700, 1), (0, 2), (1200, 3), (1, 4))]
    alg = ckf.combinatorial_kalman_filter_algorithm(cfg)
    ddet = ckf.DeviceDetector(det, fld)
    batch = alg.run_device(ddet, ckf.DeviceEvents(ddet, events)).download()
    for e, ev in enumerate(events):
        single = alg.run_device(ddet, ckf.DeviceEvents(ddet, ev)).download()
        part = batch.event(e)
        assert np.array_equal(part.offsets, single.offsets)
        assert np.array_equal(part.constituents, single.constituents)
