"""CKF bookkeeping pinned by closed forms and by constructed events.

The reference's own CKF tests, restated on this project's telescope:
  * CpuCkfCombinatoricsTelescopeTests (tests/cpu/test_ckf_combinatorics_telescope.cpp:119-206):
    n identical tracks over P planes with unlimited branching give n^(P+1)
    tracks; with max_num_branches_per_seed = 500 they give n * 500.
  * CkfSparseTrackTelescopeTests (tests/cpu/test_ckf_sparse_tracks_telescope.cpp:131-290):
    #found == #truth, every track has P constituents, ndf == sum(dim) - 5.
Further constructed cases: duplicate seeds (duplicate removal keeps exactly
one), the step-0 rule on a surface without material (only chi2 == 0 accepted;
a displaced seed starts with a hole), a dead plane (one hole, nholes counts
the holes before the first measurement as traccc does).
Every case runs on the FP64 oracle and, marked gpu, on the HIP library.
"""
import math

import numpy as np
import pytest

from helpers import run_finder
from traccc_amd import simulation as S

BACKENDS = ["oracle", pytest.param("gpu", marks=pytest.mark.gpu)]
SILICON_05MM_X0 = 0.5 / 93.7  # 0.5 mm of silicon, as the reference telescope


def telescope(n_planes=9, x0=SILICON_05MM_X0):
    return S.build_detector(S.DET_TELESCOPE, x0, telescope_planes=n_planes,
                            telescope_spacing=20.0)


def muon_cfg(**kw):
    cfg = S.default_config()
    cfg.ptc_pdg, cfg.ptc_mass, cfg.ptc_charge = 13, S.MUON_MASS, -1.0
    cfg.run_smoother = 0
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


def identical_tracks_event(det, fld, n_tracks, seed=11):
    gun = S.GunSpec(n_particles=n_tracks, pdg=13, random_charge=False, pt_a=100.0, pt_b=100.0,
                    vertex_sigma=(0.0, 0.0, 0.0), fixed_direction=True, dir_phi=0.0,
                    dir_theta=math.pi / 2, seed=seed)
    return S.generate_event(det, fld, gun)


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("n_truth", [2, 3])
def test_combinatorics_closed_forms(backend, n_truth, request):
    if backend == "gpu":
        request.getfixturevalue("gpu_available")
    n_planes = 9
    det = telescope(n_planes)
    fld = S.constant_field(0.0, bx=2.0)  # field along the tracks: no bending
    ev = identical_tracks_event(det, fld, n_truth)
    assert ev.n_seeds == n_truth and ev.n_meas == n_truth * n_planes
    unlimited = muon_cfg(max_num_branches_per_seed=2**32 - 1, max_num_branches_per_surface=10,
                         chi2_max=30.0, duplicate_removal_minimum_length=100)
    limited = muon_cfg(max_num_branches_per_seed=500, max_num_branches_per_surface=10,
                       chi2_max=30.0, duplicate_removal_minimum_length=100)
    a = run_finder(backend, unlimited, det, fld, ev)
    b = run_finder(backend, limited, det, fld, ev)
    assert len(a["ndf"]) == n_truth ** (n_planes + 1)
    assert len(b["ndf"]) == n_truth * 500
    assert len(a["ndf"]) > len(b["ndf"])
    # every track has one measurement per plane
    assert np.all(np.diff(a["offsets"]) == n_planes)


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("n_truth,random_charge", [(1, False), (2, False), (4, False),
                                                    (10, False), (10, True)])
def test_sparse_tracks(backend, n_truth, random_charge, request):
    if backend == "gpu":
        request.getfixturevalue("gpu_available")
    n_planes = 9
    det = telescope(n_planes)
    fld = S.constant_field(2.0)
    cfg = muon_cfg(chi2_max=200.0, min_p=0.0, min_pT=0.01)
    n_events = 20
    for e in range(n_events):
        gun = S.GunSpec(n_particles=n_truth, pdg=13, random_charge=random_charge, pt_a=1.0,
                        pt_b=1.0, vertex_sigma=(0.0, 400.0, 400.0), fixed_direction=True,
                        dir_phi=0.0, dir_theta=math.pi / 2, seed=1000 + e)
        ev = S.generate_event(det, fld, gun)
        assert ev.n_seeds == n_truth
        r = run_finder(backend, cfg, det, fld, ev)
        assert len(r["ndf"]) == n_truth
        lens = np.diff(r["offsets"])
        assert np.all(lens == n_planes)
        # ndf == sum(dim) - 5 (tests/common/tests/kalman_fitting_test.cpp ndf_tests)
        assert np.all(np.asarray(r["ndf"]) == 2 * n_planes - 5)
        assert np.all(np.asarray(r["chi2"]) >= 0)
        pv = np.asarray(r["pval"])
        assert np.all((pv >= 0) & (pv <= 1))
        # each track is one particle's full hit list
        for i in range(n_truth):
            parts = ev.meas_particle[r["constituents"][r["offsets"][i]:r["offsets"][i + 1]]]
            assert len(set(parts.tolist())) == 1


@pytest.mark.parametrize("backend", BACKENDS)
def test_duplicate_seeds_are_removed(backend, request):
    if backend == "gpu":
        request.getfixturevalue("gpu_available")
    det = telescope(9)
    fld = S.constant_field(2.0)
    gun = S.GunSpec(n_particles=1, pdg=13, random_charge=False, pt_a=2.0, pt_b=2.0,
                    vertex_sigma=(0.0, 0.0, 0.0), fixed_direction=True, dir_phi=0.0,
                    dir_theta=math.pi / 2, seed=5)
    ev = S.generate_event(det, fld, gun)
    # the same seed twice: identical histories, equal p-values -> the later
    # link dies once the tracks are longer than duplicate_removal_minimum_length
    for name in ("seed_surface_link", "seed_vector", "seed_covariance", "seed_particle"):
        arr = getattr(ev, name)
        setattr(ev, name, np.ascontiguousarray(np.concatenate([arr, arr])))
    cfg = muon_cfg(duplicate_removal_minimum_length=5)
    r = run_finder(backend, cfg, det, fld, ev)
    assert len(r["ndf"]) == 1
    assert np.diff(r["offsets"]).tolist() == [9]
    # without duplicate removal both survive
    cfg = muon_cfg(duplicate_removal_minimum_length=100)
    r = run_finder(backend, cfg, det, fld, ev)
    assert len(r["ndf"]) == 2


@pytest.mark.parametrize("backend", BACKENDS)
def test_step_zero_rule_without_material(backend, request):
    if backend == "gpu":
        request.getfixturevalue("gpu_available")
    det = telescope(6, x0=0.0)
    fld = S.constant_field(2.0)
    gun = S.GunSpec(n_particles=2, pdg=13, random_charge=False, pt_a=5.0, pt_b=5.0,
                    vertex_sigma=(0.0, 50.0, 50.0), fixed_direction=True, dir_phi=0.0,
                    dir_theta=math.pi / 2, seed=21)
    ev = S.generate_event(det, fld, gun)
    # seed 1 is displaced by 1 um: its chi2 on its own measurement is not 0
    ev.seed_vector = ev.seed_vector.copy()
    ev.seed_vector[1, 0] += 1e-3
    cfg = muon_cfg()
    r = run_finder(backend, cfg, det, fld, ev)
    lens = np.diff(r["offsets"]).tolist()
    assert sorted(lens) == [5, 6]
    i5 = lens.index(5)
    assert int(r["nholes"][i5]) == 1          # the step-0 hole precedes the first measurement
    assert int(r["nholes"][1 - i5]) == 0


@pytest.mark.parametrize("backend", BACKENDS)
def test_dead_plane_makes_one_hole(backend, request):
    if backend == "gpu":
        request.getfixturevalue("gpu_available")
    det = telescope(8)
    fld = S.constant_field(2.0)
    gun = S.GunSpec(n_particles=3, pdg=13, random_charge=True, pt_a=3.0, pt_b=3.0,
                    vertex_sigma=(0.0, 100.0, 100.0), fixed_direction=True, dir_phi=0.0,
                    dir_theta=math.pi / 2, seed=33)
    ev = S.generate_event(det, fld, gun)
    # drop every measurement of plane 3 (identifier of layer 3)
    dead_id = det.surface(3).identifier
    keep = ev.surface_link != dead_id
    for name in ("local_position", "local_variance", "dimensions", "time", "diameter",
                 "identifier", "surface_link", "subspace", "cluster_index", "meas_particle"):
        setattr(ev, name, np.ascontiguousarray(getattr(ev, name)[keep]))
    cfg = muon_cfg()
    r = run_finder(backend, cfg, det, fld, ev)
    assert len(r["ndf"]) == 3
    assert np.all(np.diff(r["offsets"]) == 7)
    assert np.all(np.asarray(r["nholes"]) == 0)  # traccc counts holes before the first measurement
    assert np.all(np.asarray(r["ndf"]) == 2 * 7 - 5)
