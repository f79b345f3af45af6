"""HIP CKF against the FP64 oracle on the benchmark's larger configurations.

The reference compares its host and device CKF on the same toy-detector
events as track sets (tests/cuda/test_ckf_toy_detector.cpp:217-254); the bar
here is the same comparison against the oracle, per event:

  * C5 geometry (5 barrel layers + 2x6 disks, 2 % X0, 3 % dead modules, 10 %
    noise, (r, z) solenoid map): 3,000 pions with the default finding config
    and with three branches per surface, and one full-size event of 60,000
    pions with three branches per surface.  At full size every step has
    60,000-180,000 inputs, so the find, scorer-list and scored-update work
    queues split over all eight XCD ranges (a queue of 1,024 partitions or
    more splits; the smaller tests stay one range).
  * Batch: 8 events of log-uniform 50-3,000 muons in the solenoid map (the C3
    setup at small size) in one call, each event against the oracle alone.
  * Scorer-pruned CKF at an intermediate threshold: the device logs every
    scored candidate's verdict (ckf_debug_set_verdict_log); the oracle replays
    the same verdicts -- per input the N + 4 best candidates by (chi2, index)
    are listed, the kept ones updated in order until N succeed -- and must
    find the same tracks.  C4 (3,000 pions) and C5 (full size).
"""
import ctypes as C
import time

import numpy as np
import pytest

import oracle
from helpers import compare_tracks
from traccc_amd import _ffi as F
from traccc_amd import ckf
from traccc_amd import simulation as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu(gpu_available):
    return gpu_available


def _same(a, b):
    return (np.array_equal(a.offsets, b.offsets) and
            np.array_equal(a.constituents, b.constituents))


def _report(name, g, o):
    match, ng, no = compare_tracks(g.offsets, g.constituents, o.offsets, o.constituents)
    print(f"\n{name}: gpu {ng} oracle {no} tracks, index-set match {match:.6f}, "
          f"identical order {_same(g, o)}", flush=True)
    return match


def _c5(n, seed, n_branch):
    det, fld, gun = S.workload("C5", n_particles=n, seed=seed)
    ev = S.generate_event(det, fld, gun)
    cfg = S.finding_config_for("C5")
    cfg.max_num_branches_per_surface = n_branch
    return det, fld, ev, cfg


@pytest.mark.parametrize("n_branch", [1, 3])
def test_c5_event_matches_oracle(n_branch):
    det, fld, ev, cfg = _c5(3000, 17, n_branch)
    alg = ckf.combinatorial_kalman_filter_algorithm(cfg)
    g = alg(det, fld, ev)
    o = oracle.run(cfg, det, fld, ev, 64)
    match = _report(f"C5 3,000 pions, {n_branch} branch(es) per surface", g, o)
    assert match >= 0.995
    assert abs(len(g) - len(o)) <= 0.001 * max(len(o), 1) + 1


def test_batch_matches_oracle_per_event():
    det, fld, guns = S.batch_workload("C3", n_events=8, seed=123, n_min=50, n_max=3000)
    events = S.generate_events(det, fld, guns)
    cfg = S.finding_config_for("C3")
    alg = ckf.combinatorial_kalman_filter_algorithm(cfg)
    ddet = ckf.DeviceDetector(det, fld)
    batch = alg.run_device(ddet, ckf.DeviceEvents(ddet, events)).download()
    assert len(batch.event_track_offsets) == len(events) + 1
    for e, ev in enumerate(events):
        o = oracle.run(cfg, det, fld, ev, 64)
        g = batch.event(e)
        match = _report(f"batch event {e} ({ev.n_seeds} muons)", g, o)
        assert match >= 0.995
        assert _same(g, o) or match >= 0.995


# ---------------------------------------------------------------------------
# scorer-pruned CKF replayed by the oracle
# ---------------------------------------------------------------------------
def _verdicts(alg):
    lib = F.load_library(ckf.LIB)
    lib.ckf_debug_get_verdicts.argtypes = [C.c_void_p, F.c_u64_p, C.c_uint32, F.c_u32_p]
    lib.ckf_debug_get_verdicts.restype = C.c_int
    n = C.c_uint32()
    assert lib.ckf_debug_get_verdicts(alg._h, None, 0, C.byref(n)) == 0
    keys = np.zeros(max(n.value, 1), np.uint64)
    assert lib.ckf_debug_get_verdicts(alg._h, F.ptr(keys, C.c_uint64), n.value,
                                      C.byref(n)) == 0
    return keys[:n.value]


def _set_verdict_log(alg, capacity):
    lib = F.load_library(ckf.LIB)
    lib.ckf_debug_set_verdict_log.argtypes = [C.c_void_p, C.c_uint32]
    lib.ckf_debug_set_verdict_log.restype = C.c_int
    assert lib.ckf_debug_set_verdict_log(alg._h, capacity) == 0


def _pruned_vs_replay(name, det, fld, ev, cfg, quantile, capacity):
    """Scored CKF with a scorer whose threshold prunes about `quantile` of the
    logged candidates; the oracle replays the device's verdicts."""
    from traccc_amd import scorer
    ddet = ckf.DeviceDetector(det, fld)
    dev = ckf.DeviceEvents(ddet, ev)
    alg = ckf.combinatorial_kalman_filter_algorithm(cfg)
    plain = alg.run_device(ddet, dev).download()
    _, f, _, n = alg.run_device_logged(ddet, dev, capacity=capacity)
    assert 0 < n <= capacity
    s = scorer.Scorer(seed=11)
    s.fit_normalisation(f)
    logits = s.forward(f).float().cpu().numpy()
    threshold = float(np.quantile(logits, quantile))
    alg.set_scorer(s, threshold=threshold)
    _set_verdict_log(alg, capacity)
    g = alg.run_device(ddet, dev).download()
    keys = _verdicts(alg)
    assert 0 < len(keys) <= capacity
    if alg.last_stats().n_regrows == 0:
        assert alg.last_stats().n_scored == len(keys)
    kept_bit = np.uint64(1) << np.uint64(63)
    kept = keys[(keys & kept_bit) != 0] & ~kept_bit
    frac = len(kept) / len(keys)
    t0 = time.perf_counter()
    o = oracle.run_masked(cfg, det, fld, ev, min(cfg.max_num_branches_per_surface + 1, 16), kept)
    print(f"\n{name}: {len(keys)} scored candidates, {frac:.3f} kept at threshold "
          f"{threshold:.3f}; plain {len(plain)} tracks; oracle replay {time.perf_counter() - t0:.1f} s",
          flush=True)
    assert 0.2 < frac < 0.9                       # an intermediate threshold
    assert not _same(g, plain)                    # it pruned
    return _report(name + " pruned vs oracle replay", g, o), g, o


def test_pruned_c4_matches_oracle_replay():
    det, fld, gun = S.workload("C4", n_particles=3000, seed=41)
    ev = S.generate_event(det, fld, gun)
    cfg = S.finding_config_for("C4")
    cfg.max_num_branches_per_surface = 3
    match, g, o = _pruned_vs_replay("C4 3,000 pions", det, fld, ev, cfg, 0.5, 1 << 20)
    assert match >= 0.995


@pytest.mark.timeout(600)
def test_full_size_c5_event_plain_and_pruned():
    """60,000 pions, three branches per surface: every work queue of the find
    path splits over the eight XCD ranges; plain against the oracle, pruned
    against the oracle's replay of the device's verdicts."""
    det, fld, ev, cfg = _c5(60000, 23, 3)
    alg = ckf.combinatorial_kalman_filter_algorithm(cfg)
    g = alg(det, fld, ev)
    st = alg.last_stats()
    # 60,000 step-0 inputs: >= 1,024 partitions of the plain find (32 inputs
    # per wave), the scorer list (16) and the scored update (64 at step 1)
    assert ev.n_seeds >= 32 * 1024
    o = oracle.run(cfg, det, fld, ev, 64)
    match = _report("C5 60,000 pions, 3 branches per surface (plain)", g, o)
    assert match >= 0.995
    assert st.n_links > 3 * ev.n_seeds
    match, gp, op = _pruned_vs_replay("C5 60,000 pions", det, fld, ev, cfg, 0.4, 1 << 23)
    assert match >= 0.995
