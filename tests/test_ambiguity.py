"""Greedy ambiguity resolution: the oracle against the reference's own unit
tests and an independent restatement, then the HIP resolver against the
oracle (bit-exact accepted sets) on the reference's test cases, on random
track sets shaped like the reference's CUDA-vs-CPU comparison
(tests/cuda/test_ambiguity_resolution.cpp:774-957), on the one-at-a-time
path, and on fitted CKF tracks.

Reference: core/src/ambiguity_resolution/greedy_ambiguity_resolution_algorithm.cpp:19-254.
"""
import json
import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = json.load(open(os.path.join(HERE, "golden", "ambiguity_cases.json")))["cases"]


def _flatten(tracks):
    off, ids, pv = [0], [], []
    for t in tracks:
        ids += t["ids"]
        off.append(len(ids))
        pv.append(t["pval"])
    return (np.asarray(off, np.uint32), np.asarray(ids, np.uint32), np.asarray(pv, np.float32))


def _random_tracks(seed, n_tracks, lo, hi, max_id, allow_duplicate, tie_pvals=False):
    """Track sets shaped like the reference's comparison test: lengths
    uniform in [lo, hi], identifiers uniform in [0, max_id], p-values uniform
    in [0, 1) (on a coarse grid with tie_pvals)."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(lo, hi + 1, n_tracks)
    off = np.zeros(n_tracks + 1, np.uint32)
    off[1:] = np.cumsum(lens)
    ids = np.empty(int(off[-1]), np.uint32)
    for t in range(n_tracks):
        if allow_duplicate:
            ids[off[t]:off[t + 1]] = rng.integers(0, max_id + 1, lens[t])
        else:
            ids[off[t]:off[t + 1]] = rng.choice(max_id + 1, lens[t], replace=False)
    if tie_pvals:
        pv = rng.integers(0, 8, n_tracks).astype(np.float32) / 8
    else:
        pv = rng.random(n_tracks, dtype=np.float32)
    return off, ids, pv


def _shared_ids_of_kept(off, ids, keep):
    """Identifiers held by more than one kept track."""
    owners = {}
    for t in np.nonzero(keep)[0]:
        for m in set(ids[off[t]:off[t + 1]].tolist()):
            owners.setdefault(m, set()).add(int(t))
    return [m for m, s in owners.items() if len(s) > 1]


def _literal(off, ids, pv, min_meas=3, max_iter=2**32 - 1, max_shared=1):
    """The reference's loop as written there -- a sorted vector of track ids,
    the worst popped from the back, a track whose identifier becomes
    unshared re-inserted at its lower bound -- in plain Python, for inputs
    whose (rel_shared, p-value) keys are all distinct."""
    n = len(off) - 1
    acc = [t for t in range(n) if off[t + 1] - off[t] >= min_meas]
    mids = {t: ids[off[t]:off[t + 1]].tolist() for t in acc}
    tpm = {}
    for t in acc:
        for m in sorted(set(mids[t])):
            tpm.setdefault(m, []).append(t)
    nsh = {t: sum(1 for m in mids[t] if len(tpm[m]) > 1) for t in acc}
    rel = {t: np.float32(nsh[t]) / np.float32(len(mids[t])) for t in acc}

    def key(t):
        return (rel[t], -float(pv[t]))

    order = sorted(acc, key=key)
    for _ in range(min(max_iter, 10**9)):
        if not acc or max(nsh[t] for t in acc) < max_shared:
            break
        worst = order.pop()
        acc.remove(worst)
        seen = []
        for m in mids[worst]:
            if m in seen:
                continue
            seen.append(m)
            tpm[m].remove(worst)
            if len(tpm[m]) == 1:
                tid = tpm[m][0]
                nsh[tid] -= mids[tid].count(m)
                rel[tid] = np.float32(nsh[tid]) / np.float32(len(mids[tid]))
                order.remove(tid)
                k = 0
                while k < len(order) and key(order[k]) < key(tid):
                    k += 1
                order.insert(k, tid)
    keep = np.zeros(n, bool)
    keep[acc] = True
    return keep


# ---- the oracle ------------------------------------------------------------

@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_oracle_matches_the_reference_unit_tests(case):
    off, ids, pv = _flatten(case["tracks"])
    keep, _ = oracle.greedy_resolve(off, ids, pv, case["min_meas_per_track"])
    got = sorted(tuple(case["tracks"][t]["ids"]) for t in np.nonzero(keep)[0])
    assert len(got) == case["n_expected"]
    assert got == sorted(tuple(e) for e in case["expected"])


@pytest.mark.parametrize("seed", range(6))
def test_oracle_matches_the_literal_loop(seed):
    # distinct p-values keep every (rel_shared, p-value) key distinct
    off, ids, pv = _random_tracks(seed, 300, 1, 10, 250, allow_duplicate=seed % 2 == 0)
    assert len(np.unique(pv)) == len(pv)
    for max_shared, max_iter in ((1, 2**32 - 1), (2, 2**32 - 1), (1, 17)):
        keep, _ = oracle.greedy_resolve(off, ids, pv, 3, max_iter, max_shared)
        ref = _literal(off, ids, pv, 3, max_iter, max_shared)
        assert np.array_equal(keep, ref), (max_shared, max_iter)


@pytest.mark.parametrize("dup", [False, True])
def test_oracle_leaves_no_shared_measurement(dup):
    off, ids, pv = _random_tracks(7, 3000, 3, 10, 2000, dup)
    keep, removed = oracle.greedy_resolve(off, ids, pv)
    lens = np.diff(off)
    assert not _shared_ids_of_kept(off, ids, keep)
    # only tracks long enough are kept, and every removal is counted
    assert np.all(lens[keep] >= 3)
    assert removed == int(np.sum(lens >= 3)) - int(keep.sum())


def test_oracle_iteration_limit_and_threshold():
    off, ids, pv = _random_tracks(8, 500, 3, 10, 300, False)
    _, removed_all = oracle.greedy_resolve(off, ids, pv)
    keep5, removed5 = oracle.greedy_resolve(off, ids, pv, max_iterations=5)
    assert removed5 == 5 < removed_all
    keep2, _ = oracle.greedy_resolve(off, ids, pv, max_shared_meas=2)
    # with max_shared_meas = 2 a kept track may still share one measurement
    assert keep2.sum() >= oracle.greedy_resolve(off, ids, pv)[0].sum()


def test_resolver_rejects_a_zero_shared_threshold():
    from traccc_amd import _ffi as F
    if not os.path.exists(os.path.join(F.LIB_DIR, "libtraccc_amd.so")):
        pytest.skip("HIP library not built")
    from traccc_amd import ambiguity as A
    with pytest.raises(ValueError, match="max_shared_meas"):
        A.greedy_ambiguity_resolution_algorithm(A.ambiguity_resolution_config(max_shared_meas=0))
    cfg = A.ambiguity_resolution_config()
    assert (cfg.min_meas_per_track, cfg.max_iterations, cfg.max_shared_meas) == (3, 2**32 - 1, 1)


# ---- the HIP resolver --------------------------------------------------------

@pytest.fixture(scope="module")
def resolver_cls(gpu_available):
    from traccc_amd import ambiguity as A
    return A


def _gpu_vs_oracle(A, off, ids, pv, min_meas=3, max_iter=2**32 - 1, max_shared=1, id_bound=None):
    alg = A.greedy_ambiguity_resolution_algorithm(
        A.ambiguity_resolution_config(min_meas_per_track=min_meas, max_iterations=max_iter,
                                      max_shared_meas=max_shared))
    kept = alg(off, ids, pv, id_bound)
    keep, removed = oracle.greedy_resolve(off, ids, pv, min_meas, max_iter, max_shared)
    assert np.array_equal(kept, np.nonzero(keep)[0])
    st = alg.last_stats()
    assert st.n_accepted == len(kept) and st.n_removed == removed
    return alg, st


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_gpu_matches_the_reference_unit_tests(resolver_cls, case):
    off, ids, pv = _flatten(case["tracks"])
    # each reference test fills identifiers 0..100
    _gpu_vs_oracle(resolver_cls, off, ids, pv, case["min_meas_per_track"], id_bound=101)


# the reference's comparison instantiations (n_events, n_tracks, length range,
# max identifier); duplicates allowed and not
SHAPES = {"standard": (50000, 1, 10, 20000), "sparse": (5000, 3, 10, 1000000),
          "dense": (5000, 3, 10, 100), "long": (10000, 3, 500, 10000), "simple": (5, 3, 5, 10)}


@pytest.mark.gpu
@pytest.mark.parametrize("dup", [True, False])
@pytest.mark.parametrize("shape", list(SHAPES))
def test_gpu_matches_the_oracle_on_random_tracks(resolver_cls, shape, dup):
    n, lo, hi, max_id = SHAPES[shape]
    for ev in range(3):
        off, ids, pv = _random_tracks(42 + ev, n, lo, hi, max_id, dup)
        _, st = _gpu_vs_oracle(resolver_cls, off, ids, pv, id_bound=max_id + 1)
        assert not st.sequential


@pytest.mark.gpu
def test_gpu_ties_and_nan_pvalues(resolver_cls):
    off, ids, pv = _random_tracks(5, 4000, 3, 8, 1500, True, tie_pvals=True)
    pv[::37] = np.nan
    _gpu_vs_oracle(resolver_cls, off, ids, pv)


@pytest.mark.gpu
@pytest.mark.parametrize("max_shared,max_iter", [(2, 2**32 - 1), (3, 2**32 - 1), (1, 40), (2, 25)])
def test_gpu_one_at_a_time_path(resolver_cls, max_shared, max_iter):
    off, ids, pv = _random_tracks(11, 3000, 3, 10, 1500, True)
    _, st = _gpu_vs_oracle(resolver_cls, off, ids, pv, 3, max_iter, max_shared)
    assert st.sequential


@pytest.mark.gpu
def test_gpu_edge_cases(resolver_cls):
    A = resolver_cls
    alg = A.greedy_ambiguity_resolution_algorithm()
    # no tracks
    assert len(alg(np.zeros(1, np.uint32), np.zeros(0, np.uint32), np.zeros(0, np.float32))) == 0
    # every track too short
    off = np.array([0, 2, 4], np.uint32)
    assert len(alg(off, np.array([1, 2, 1, 2], np.uint32), np.array([0.5, 0.4], np.float32))) == 0
    # zero-length tracks kept with min_meas_per_track = 0
    off = np.array([0, 0, 3, 6, 6], np.uint32)
    ids = np.array([1, 2, 3, 3, 4, 5], np.uint32)
    pv = np.array([0.1, 0.2, 0.3, 0.4], np.float32)
    _gpu_vs_oracle(A, off, ids, pv, min_meas=0, id_bound=6)
    # an identifier at or above id_bound and decreasing offsets are errors
    with pytest.raises(ValueError, match="id_bound"):
        alg(np.array([0, 3], np.uint32), np.array([1, 2, 9], np.uint32),
            np.array([0.5], np.float32), id_bound=5)
    with pytest.raises(ValueError, match="decrease"):
        alg(np.array([0, 3, 2], np.uint32), np.array([1, 2, 3], np.uint32),
            np.array([0.5, 0.5], np.float32), id_bound=5)


@pytest.mark.gpu
def test_gpu_device_entry_and_repeated_calls(resolver_cls):
    import torch
    A = resolver_cls
    alg = A.greedy_ambiguity_resolution_algorithm()
    for seed, n in ((1, 20000), (2, 700)):   # the second call smaller than the first
        off, ids, pv = _random_tracks(seed, n, 3, 10, n, True)
        d_acc, na = alg.run_device(torch.from_numpy(off.astype(np.int32)).cuda(),
                                   torch.from_numpy(ids.astype(np.int32)).cuda(),
                                   torch.from_numpy(pv).cuda(), n + 1)
        keep, _ = oracle.greedy_resolve(off, ids, pv)
        got = d_acc[:n].cpu().numpy().astype(bool)
        assert np.array_equal(got, keep) and na == keep.sum()


@pytest.mark.gpu
def test_gpu_resolves_fitted_ckf_tracks(resolver_cls):
    """CKF (3 branches per surface) -> Kalman fit -> resolver, on two events
    in one batch, against the oracle on the downloaded fitted tracks with
    batch-wide measurement indices as identifiers."""
    from traccc_amd import ckf
    from traccc_amd import fitting as FT
    from traccc_amd import simulation as S
    A = resolver_cls
    events = []
    for seed in (5, 6):
        det, fld, gun = S.workload("C2", seed=seed, n_particles=1500)
        events.append(S.generate_event(det, fld, gun))
    cfg = S.finding_config_for("C2")
    cfg.max_num_branches_per_surface = 3
    ddet = ckf.DeviceDetector(det, fld)
    dev = ckf.DeviceEvents(ddet, events)
    cands = ckf.combinatorial_kalman_filter_algorithm(cfg).run_device(ddet, dev)
    fitted = FT.kalman_fitting_algorithm(FT.fitting_config(cfg)).run_device(ddet, dev, cands)
    f = fitted.download()
    n_ev_meas = [e.n_meas for e in events]
    ev_off = cands.download().event_track_offsets
    meas_off = np.concatenate([[0], np.cumsum(n_ev_meas)]).astype(np.uint64)
    ids = f.measurement_index.astype(np.uint64).copy()
    for e in range(len(events)):
        s0, s1 = f.state_offsets[ev_off[e]], f.state_offsets[ev_off[e + 1]]
        ids[s0:s1] += meas_off[e]
    alg = A.greedy_ambiguity_resolution_algorithm()
    kept = alg.run_fitted(dev, fitted)
    keep, _ = oracle.greedy_resolve(f.state_offsets, ids.astype(np.uint32), f.pval)
    assert np.array_equal(kept, np.nonzero(keep)[0])
    assert 0 < len(kept) < len(f)
    # the one-at-a-time path resolves per event
    alg2 = A.greedy_ambiguity_resolution_algorithm(A.ambiguity_resolution_config(max_shared_meas=2))
    kept2 = alg2.run_fitted(dev, fitted)
    exp2 = []
    for e in range(len(events)):
        t0, t1 = ev_off[e], ev_off[e + 1]
        o = f.state_offsets[t0:t1 + 1] - f.state_offsets[t0]
        k, _ = oracle.greedy_resolve(o, ids[f.state_offsets[t0]:f.state_offsets[t1]].astype(np.uint32),
                                     f.pval[t0:t1], max_shared_meas=2)
        exp2 += (np.nonzero(k)[0] + t0).tolist()
    assert np.array_equal(kept2, np.asarray(exp2))
