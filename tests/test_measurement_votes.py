"""Measurement-vote pruning (finding_config.max_num_tracks_per_measurement > 0).

The reference runs it on its device chain only
(device/common/src/finding/combinatorial_kalman_filter_algorithm.cpp:568-634 with
gather_best_tips_per_measurement.ipp:27-180, gather_measurement_votes.ipp:19-47,
update_tip_length_buffer.ipp:20-54): every measurement keeps the K tracks with
the highest p-value of their tip's running chi2 / ndf sums, and a track
survives when at least min_measurement_voting_fraction of its measurements kept
it.  Equal p-values rank by lower track index here (the reference leaves that
to its insertion mutex), so the result is deterministic.

Pins on the oracle (CPU):
  * identical duplicate tracks (same seed twice, no duplicate removal) have equal
    p-values: K = 1 keeps the first copy only;
  * K at least the largest measurement multiplicity, or a fraction of 0, prunes
    nothing: the output equals the run without votes;
  * K = 1 with fraction 1 keeps measurement-disjoint tracks only, and every
    dropped track shares a measurement with a kept track of higher p-value;
  * the rule re-applied by brute force to the unpruned output reproduces the
    pruned output (n^(P+1) combinatorics and C4-like events).
GPU (marked gpu): the HIP library's lock-free top-K cascade gives the oracle's
tracks in the oracle's order; a batch equals its events run one by one.
"""
import math

import numpy as np
import pytest

from helpers import compare_tracks, run_finder
from traccc_amd import ckf
from traccc_amd import simulation as S

BACKENDS = ["oracle", pytest.param("gpu", marks=pytest.mark.gpu)]


def _cfg(base, **kw):
    for k, v in kw.items():
        setattr(base, k, v)
    return base


def _tracks(r):
    off, cons = r["offsets"], r["constituents"]
    return [tuple(cons[off[i]:off[i + 1]].tolist()) for i in range(len(off) - 1)]


def _brute_force_votes(r, K, frac):
    """The vote rule on an unpruned result (p-values of the output tracks)."""
    tr = _tracks(r)
    pv = np.asarray(r["pval"], dtype=np.float64)
    per_meas = {}
    for i, t in enumerate(tr):
        for m in t:
            per_meas.setdefault(m, []).append(i)
    votes = np.zeros(len(tr), dtype=np.int64)
    for m, idx in per_meas.items():
        idx.sort(key=lambda i: (-pv[i], i))
        for i in idx[:K]:
            votes[i] += 1
    keep = [i for i, t in enumerate(tr) if not (np.float32(votes[i]) / np.float32(len(t)) <
                                                np.float32(frac))]
    return [tr[i] for i in keep]


def _telescope_duplicates():
    det = S.build_detector(S.DET_TELESCOPE, 0.5 / 93.7, telescope_planes=9,
                           telescope_spacing=20.0)
    fld = S.constant_field(2.0)
    gun = S.GunSpec(n_particles=3, pdg=13, random_charge=False, pt_a=2.0, pt_b=2.0,
                    vertex_sigma=(0.0, 0.0, 0.0), fixed_direction=False, eta_min=-0.05,
                    eta_max=0.05, phi_min=-0.05, phi_max=0.05, seed=5)
    ev = S.generate_event(det, fld, gun)
    for name in ("seed_surface_link", "seed_vector", "seed_covariance", "seed_particle"):
        arr = getattr(ev, name)
        setattr(ev, name, np.ascontiguousarray(np.concatenate([arr, arr])))
    return det, fld, ev


def _muon_cfg(**kw):
    cfg = S.default_config()
    cfg.ptc_pdg, cfg.ptc_mass, cfg.ptc_charge = 13, S.MUON_MASS, -1.0
    cfg.run_smoother = 0
    return _cfg(cfg, **kw)


@pytest.mark.parametrize("backend", BACKENDS)
def test_duplicate_tracks_keep_first_copy(backend, request):
    if backend == "gpu":
        request.getfixturevalue("gpu_available")
    det, fld, ev = _telescope_duplicates()
    n = ev.n_seeds // 2
    base = run_finder(backend, _muon_cfg(duplicate_removal_minimum_length=100), det, fld, ev)
    tr = _tracks(base)
    assert len(tr) == 2 * n and tr[:n] == tr[n:]
    voted = run_finder(backend, _muon_cfg(duplicate_removal_minimum_length=100,
                                          max_num_tracks_per_measurement=1,
                                          min_measurement_voting_fraction=0.5), det, fld, ev)
    assert _tracks(voted) == tr[:n]


@pytest.mark.parametrize("backend", BACKENDS)
def test_votes_that_prune_nothing(backend, request):
    if backend == "gpu":
        request.getfixturevalue("gpu_available")
    det, fld, gun = S.workload("C4", n_particles=600, seed=21)
    ev = S.generate_event(det, fld, gun)
    cfg = S.finding_config_for("C4")
    cfg.max_num_branches_per_surface = 3
    cfg.max_num_branches_per_seed = 20
    base = run_finder(backend, cfg, det, fld, ev)
    tr = _tracks(base)
    mult = {}
    for t in tr:
        for m in t:
            mult[m] = mult.get(m, 0) + 1
    kmax = max(mult.values())
    assert kmax > 1  # the case shares measurements
    for K, frac in ((kmax, 1.0), (1, 0.0)):
        c = S.finding_config_for("C4")
        c.max_num_branches_per_surface = 3
        c.max_num_branches_per_seed = 20
        c.max_num_tracks_per_measurement = K
        c.min_measurement_voting_fraction = frac
        r = run_finder(backend, c, det, fld, ev)
        assert _tracks(r) == tr, (K, frac)


@pytest.mark.parametrize("backend", BACKENDS)
def test_full_vote_keeps_disjoint_tracks(backend, request):
    if backend == "gpu":
        request.getfixturevalue("gpu_available")
    det, fld, gun = S.workload("C4", n_particles=600, seed=22)
    ev = S.generate_event(det, fld, gun)
    cfg = S.finding_config_for("C4")
    cfg.max_num_branches_per_surface = 3
    cfg.max_num_branches_per_seed = 20
    base = run_finder(backend, cfg, det, fld, ev)
    cfg.max_num_tracks_per_measurement = 1
    cfg.min_measurement_voting_fraction = 1.0
    r = run_finder(backend, cfg, det, fld, ev)
    kept = _tracks(r)
    assert 0 < len(kept) < len(_tracks(base))
    seen = set()
    for t in kept:
        assert not (set(t) & seen)
        seen |= set(t)
    # every dropped track loses a measurement to a kept track of higher p-value
    pv = {t: p for t, p in zip(_tracks(base), base["pval"])}
    kept_set = set(kept)
    owner = {m: t for t in kept for m in t}
    for t in _tracks(base):
        if t in kept_set:
            continue
        rivals = [owner[m] for m in t if m in owner]
        assert rivals and any(pv[o] >= pv[t] for o in rivals)


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("case", ["combinatorics", "c4"])
@pytest.mark.parametrize("K,frac", [(1, 0.5), (2, 0.7), (3, 1.0)])
def test_votes_match_brute_force(backend, case, K, frac, request):
    if backend == "gpu":
        request.getfixturevalue("gpu_available")
    if case == "combinatorics":
        det = S.build_detector(S.DET_TELESCOPE, 0.5 / 93.7, telescope_planes=5,
                               telescope_spacing=20.0)
        fld = S.constant_field(0.0, bx=2.0)
        gun = S.GunSpec(n_particles=3, pdg=13, random_charge=False, pt_a=100.0, pt_b=100.0,
                        vertex_sigma=(0.0, 0.0, 0.0), fixed_direction=True, dir_phi=0.0,
                        dir_theta=math.pi / 2, seed=11)
        ev = S.generate_event(det, fld, gun)
        cfg = _muon_cfg(max_num_branches_per_seed=2**32 - 1, max_num_branches_per_surface=10,
                        chi2_max=30.0, duplicate_removal_minimum_length=100)
    else:
        det, fld, gun = S.workload("C4", n_particles=800, seed=23)
        ev = S.generate_event(det, fld, gun)
        cfg = S.finding_config_for("C4")
        cfg.max_num_branches_per_surface = 3
        cfg.max_num_branches_per_seed = 20
    base = run_finder(backend, cfg, det, fld, ev)
    cfg.max_num_tracks_per_measurement = K
    cfg.min_measurement_voting_fraction = frac
    r = run_finder(backend, cfg, det, fld, ev)
    expect = _brute_force_votes(base, K, frac)
    assert _tracks(r) == expect


@pytest.mark.gpu
@pytest.mark.parametrize("workload,n,K,frac", [("C2", 10000, 1, 0.5), ("C4", 4000, 2, 0.7)])
def test_gpu_votes_match_oracle(workload, n, K, frac, gpu_available):
    det, fld, gun = S.workload(workload, n_particles=n, seed=42)
    ev = S.generate_event(det, fld, gun)
    cfg = S.finding_config_for(workload)
    cfg.max_num_branches_per_surface = 2
    cfg.max_num_branches_per_seed = 10
    cfg.max_num_tracks_per_measurement = K
    cfg.min_measurement_voting_fraction = frac
    print(f"[votes] {workload} n={n} K={K} frac={frac}", flush=True)
    g = run_finder("gpu", cfg, det, fld, ev)
    o = run_finder("oracle", cfg, det, fld, ev)
    match, ng, no = compare_tracks(g["offsets"], g["constituents"], o["offsets"], o["constituents"])
    print(f"gpu {ng} oracle {no} tracks, index-set match {match:.5f}")
    assert match >= 0.995
    assert abs(ng - no) <= max(1, no // 1000)


@pytest.mark.gpu
def test_gpu_votes_batched_equals_single(gpu_available):
    det, fld, _ = S.workload("C4", n_particles=10)
    cfg = S.finding_config_for("C4")
    cfg.max_num_branches_per_surface = 2
    cfg.max_num_tracks_per_measurement = 1
    cfg.min_measurement_voting_fraction = 0.5
    events = [S.generate_event(det, fld, S.workload("C4", n_particles=n, seed=s)[2])
              for n, s in ((900, 1), (0, 2), (1500, 3), (40, 4))]
    alg = ckf.combinatorial_kalman_filter_algorithm(cfg)
    ddet = ckf.DeviceDetector(det, fld)
    batch = alg.run_device(ddet, ckf.DeviceEvents(ddet, events)).download()
    for e, ev in enumerate(events):
        single = alg.run_device(ddet, ckf.DeviceEvents(ddet, ev)).download()
        part = batch.event(e)
        assert np.array_equal(part.offsets, single.offsets)
        assert np.array_equal(part.constituents, single.constituents)
