"""The branch scorer inside the CKF (GPU).

  * Branch log (training data): the logged run finds the plain run's tracks;
    one sample per measurement link; with one branch per surface the survive
    labels count exactly the tracks' measurements; feature ranges.
  * Scored finding: a scorer that keeps everything gives the plain CKF's
    tracks bit for bit (C4, three branches per surface); one that prunes
    everything gives no track; one that prunes about half gives a batch of
    events the tracks each event gets alone.
  * Gain: trained on survive / die labels the plain CKF logs on dense C4
    events (three branches per surface; a link survives when it lies on its
    seed's best found track), the pruned CKF keeps the plain CKF's
    finding efficiency within 0.5 % absolute on held-out events with markedly
    fewer propagations and links.
"""
import numpy as np
import pytest

from traccc_amd import ckf
from traccc_amd import performance as PERF
from traccc_amd import simulation as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu(gpu_available):
    return gpu_available


def _c4_cfg(n_branch=3):
    cfg = S.finding_config_for("C4")
    cfg.max_num_branches_per_surface = n_branch
    cfg.max_num_branches_per_seed = 10
    return cfg


def _events(n_events, n_particles, seed0):
    det, fld, _ = S.workload("C4", n_particles=10)
    guns = [S.workload("C4", n_particles=n_particles, seed=seed0 + e)[2] for e in range(n_events)]
    return det, fld, S.generate_events(det, fld, guns)


def test_branch_log_closed_forms():
    det, fld, evs = _events(1, 3000, 501)
    ev = evs[0]
    cfg = _c4_cfg(1)
    alg = ckf.combinatorial_kalman_filter_algorithm(cfg)
    ddet = ckf.DeviceDetector(det, fld)
    dev = ckf.DeviceEvents(ddet, ev)
    plain = alg.run_device(ddet, dev).download()
    tr, feats, labels, n = alg.run_device_logged(ddet, dev, capacity=200000)
    logged = tr.download()
    assert np.array_equal(plain.offsets, logged.offsets)
    assert np.array_equal(plain.constituents, logged.constituents)
    st = alg.last_stats()
    assert 0 < n <= st.n_links
    f = feats.cpu().numpy()
    y = labels.cpu().numpy()
    assert len(y) == n
    # one branch per surface: every measurement of a found track is one
    # logged link, marked once
    assert int(y.sum()) == len(plain.constituents)
    assert np.all((f[:, 2] >= 0) & (f[:, 2] <= cfg.chi2_max))          # chi2
    assert np.all((f[:, 7] >= 0) & (f[:, 7] < st.n_steps))             # depth
    assert np.all(f[:, 10] == 0)                                        # rank (one branch)
    assert np.all(f[:, 9] == 0)         # standing: a seed's only branch is its best
    assert np.all(f[:, 11] >= np.log(2.0) - 1e-6)                       # >= 1 candidate
    assert np.all(f[:, 6] <= cfg.max_num_skipping_per_cand)            # holes
    # three branches per surface: a branch never stands below its seed's best
    cfg3 = _c4_cfg(3)
    alg3 = ckf.combinatorial_kalman_filter_algorithm(cfg3)
    _, f3, _, n3 = alg3.run_device_logged(ddet, dev, capacity=600000)
    f3 = f3.cpu().numpy()
    assert np.all(f3[:, 9] >= 0) and np.mean(f3[:, 9] > 0) > 0.02
    assert np.all((f[:, 3] >= 0) & (f[:, 3] < 1))                      # layer fraction


def test_scored_extremes():
    from traccc_amd import scorer
    det, fld, evs = _events(1, 3000, 601)
    ev = evs[0]
    cfg = _c4_cfg(3)
    ddet = ckf.DeviceDetector(det, fld)
    dev = ckf.DeviceEvents(ddet, ev)
    plain = ckf.combinatorial_kalman_filter_algorithm(cfg).run_device(ddet, dev).download()
    s = scorer.Scorer(seed=2)
    alg = ckf.combinatorial_kalman_filter_algorithm(cfg)
    alg.set_scorer(s, threshold=-1e30)   # keeps every candidate
    keep = alg.run_device(ddet, dev).download()
    assert np.array_equal(plain.offsets, keep.offsets)
    assert np.array_equal(plain.constituents, keep.constituents)
    assert np.array_equal(plain.chi2, keep.chi2)
    alg.set_scorer(s, threshold=1e30)    # prunes every candidate
    none = alg.run_device(ddet, dev).download()
    assert len(none) == 0
    alg.set_scorer(None)
    again = alg.run_device(ddet, dev).download()
    assert np.array_equal(plain.constituents, again.constituents)


def test_scored_batched_equals_single_events():
    """The scorer decides each candidate from its own features and its seed's
    standing, so a batch of events gets, event by event, the tracks each
    event gets alone -- with a threshold (the median logit) that prunes about
    half of the candidates."""
    from traccc_amd import scorer
    det, fld, evs = _events(3, 1500, 701)
    cfg = _c4_cfg(3)
    ddet = ckf.DeviceDetector(det, fld)
    alg = ckf.combinatorial_kalman_filter_algorithm(cfg)
    plain = alg.run_device(ddet, ckf.DeviceEvents(ddet, evs)).download()
    _, f, _, n = alg.run_device_logged(ddet, ckf.DeviceEvents(ddet, evs[0]), capacity=400000)
    assert 0 < n <= 400000
    s = scorer.Scorer(seed=5)
    s.fit_normalisation(f)
    threshold = float(np.median(s.forward(f).float().cpu().numpy()))
    alg.set_scorer(s, threshold=threshold)
    batch = alg.run_device(ddet, ckf.DeviceEvents(ddet, evs)).download()
    assert alg.last_stats().n_scored > 0
    assert not np.array_equal(batch.constituents, plain.constituents)   # it pruned
    for e, ev in enumerate(evs):
        single = alg.run_device(ddet, ckf.DeviceEvents(ddet, ev)).download()
        part = batch.event(e)
        assert np.array_equal(part.offsets, single.offsets)
        assert np.array_equal(part.constituents, single.constituents)
        assert np.array_equal(part.chi2, single.chi2)


def _run(alg, ddet, events):
    tracks, props, links = [], 0, 0
    for ev in events:
        t = alg.run_device(ddet, ckf.DeviceEvents(ddet, ev)).download()
        st = alg.last_stats()
        props += st.n_propagations
        links += st.n_links
        tracks.append(t)
    eff = np.mean([PERF.finding_performance(ev, t.offsets, t.constituents).efficiency
                   for ev, t in zip(events, tracks)])
    return eff, props, links


def test_trained_scorer_prunes_at_plain_efficiency():
    import torch
    from traccc_amd import scorer
    cfg = _c4_cfg(3)
    det, fld, train_evs = _events(10, 20000, 1000)
    _, _, val_evs = _events(2, 20000, 2000)
    _, _, test_evs = _events(3, 20000, 3000)
    ddet = ckf.DeviceDetector(det, fld)
    plain = ckf.combinatorial_kalman_filter_algorithm(cfg)
    feats, labels = [], []
    for ev in train_evs:
        _, f, y, n = plain.run_device_logged(ddet, ckf.DeviceEvents(ddet, ev), capacity=600000)
        assert n <= 600000
        feats.append(f.clone())
        labels.append(y.clone())
    X = torch.cat(feats)
    Y = torch.cat(labels)
    frac = float(Y.float().mean())
    print(f"\n{X.shape[0]} samples, survive fraction {frac:.3f}")
    s = scorer.Scorer(seed=7)
    s.fit_normalisation(X)
    st = s.train(X, Y, scorer.train_config(epochs=4, batch=8192, learning_rate=2e-3))
    print(f"loss {st.first_loss:.4f} -> {st.last_loss:.4f} in {st.steps} steps, "
          f"{st.seconds:.2f} s, {st.tflops:.2f} TFLOP/s")
    assert st.last_loss < st.first_loss

    eff_plain_val, props_plain_val, _ = _run(plain, ddet, val_evs)
    pruned = ckf.combinatorial_kalman_filter_algorithm(cfg)
    chosen = None
    for thr in (2.0, 1.0, 0.0, -1.0, -2.0, -3.0, -4.0, -5.0):   # most aggressive first
        pruned.set_scorer(s, threshold=thr)
        eff, props, _ = _run(pruned, ddet, val_evs)
        print(f"threshold {thr}: efficiency {eff:.4f} (plain {eff_plain_val:.4f}), "
              f"propagations {props / props_plain_val:.3f} of plain")
        if eff >= eff_plain_val - 0.0025:
            chosen = thr
            break
    assert chosen is not None
    pruned.set_scorer(s, threshold=chosen)
    eff_p, props_p, links_p = _run(plain, ddet, test_evs)
    eff_s, props_s, links_s = _run(pruned, ddet, test_evs)
    print(f"held-out: plain eff {eff_p:.4f} props {props_p} links {links_p}; "
          f"pruned eff {eff_s:.4f} props {props_s} links {links_s}")
    assert eff_s >= eff_p - 0.005
    assert props_s < 0.85 * props_p
    assert links_s < 0.85 * links_p
