"""Benchmark: CKF tracks/s on one MI355X, plain vs MLP-pruned.

Headline (`value`): the C5 workload -- 64 events x 60,000 pions (pT 0.3 GeV
+ Exp(1 GeV), |eta| < 2.5) on 5 barrel layers + 2x6 disks with 2 % X0 per
layer, 3 % dead modules, 10 % noise hits, the (r, z) solenoid map, three
branches per surface, truth seeds -- found by the CKF pruned by the learned
branch scorer, in one operator call (traccc's
combinatorial_kalman_filter_algorithm operator()).  The scorer is trained here,
on the GPU, on survive / die labels the plain CKF logs on 32 training events;
its threshold is the most aggressive one that keeps validation efficiency
within 0.4 % absolute of plain; the timed batch is held out.  Inputs are
resident in HBM before the timed region; outputs stay in HBM.  tracks/s =
seeds processed / time per call, the median of K calls timed with HIP events
on the algorithm's stream.  The plain CKF on the same batch, the speed-up and
the efficiency loss are reported beside it, with the Kalman fit of the pruned
candidates.

Further lines: C2 (1 event x 10,000 muons, the plain CKF), C3 (256 events of
log-uniform 500-50,000 muons in the solenoid map, one call), C4 (32 events x
20,000 pions, plain vs pruned, the scorer trained on labels the plain CKF logs
on C2-configuration events), the roofline of the dominant kernel
(k_propagate, timed with HIP events inside the timed run's schedule) and the
FP32 CPU oracle timed on the host cores over C5 events.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "CKF tracks/s on 1 MI355X, plain vs MLP-pruned at <=0.5% efficiency loss"
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
MLP_FLOPS = 2 * (12 * 64 + 64 * 64 + 64)  # one scored candidate


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def host_threads() -> int:
    return len(os.sched_getaffinity(0))


def _debug_lib():
    from traccc_amd import _ffi as F
    from traccc_amd import ckf
    lib = F.load_library(ckf.LIB)
    lib.ckf_debug_set_kernel_timing.argtypes = [C.c_void_p, C.c_int]
    lib.ckf_debug_kernel_times.argtypes = [C.c_void_p] + [C.POINTER(C.c_double)] * 4 + [
        C.POINTER(C.c_uint32)]
    lib.ckf_debug_find_split.argtypes = [C.c_void_p] + [C.POINTER(C.c_double)] * 2
    lib.ckf_debug_propagate_split.argtypes = [C.c_void_p] + [C.POINTER(C.c_double)] * 2
    lib.ckf_debug_get_schedule.argtypes = [C.c_void_p, C.POINTER(C.c_uint32)]
    return lib


def kernel_times(alg, ddet, devents):
    """Device ms per call of each kernel group, from one call run step by step
    with HIP events around every kernel (ckf_debug_set_kernel_timing)."""
    lib = _debug_lib()
    lib.ckf_debug_set_kernel_timing(alg._h, 1)
    alg.run_device(ddet, devents)
    v = [C.c_double() for _ in range(8)]
    n = C.c_uint32()
    lib.ckf_debug_kernel_times(alg._h, *[C.byref(x) for x in v[:4]], C.byref(n))
    lib.ckf_debug_find_split(alg._h, C.byref(v[4]), C.byref(v[5]))
    lib.ckf_debug_propagate_split(alg._h, C.byref(v[6]), C.byref(v[7]))
    lib.ckf_debug_set_kernel_timing(alg._h, 0)
    lanes = C.c_uint32()
    lib.ckf_debug_get_schedule(alg._h, C.byref(lanes))
    keys = ("find", "dedup", "propagate", "advance", "find_kernel", "find_emit",
            "propagate_kernel", "propagate_emit")
    out = {k: x.value for k, x in zip(keys, v)}
    out["steps"] = n.value
    out["lanes"] = lanes.value
    return out


def time_calls(alg, ddet, devents, steps, warmup, torch):
    """Median and mean ms per operator call, each call timed with HIP events
    on the algorithm's stream (torch's current stream)."""
    for _ in range(warmup):
        alg.run_device(ddet, devents)
    torch.cuda.synchronize()
    ms = []
    out = None
    for _ in range(steps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        out = alg.run_device(ddet, devents)
        e1.record()
        e1.synchronize()
        ms.append(e0.elapsed_time(e1))
    torch.cuda.synchronize()
    return float(np.median(ms)), float(np.mean(ms)), out


def _efficiency(events, tracks):
    from traccc_amd import performance as PERF
    return float(np.mean([PERF.finding_performance(ev, t.offsets, t.constituents).efficiency
                          for ev, t in ((ev, tracks.event(e)) for e, ev in enumerate(events))]))


def train_scorer(a, ckf, S, torch, det, fld, label_events, cfg, tag):
    """Scorer trained on the survive / die labels the plain CKF (finding
    config `cfg`) logs on `label_events` (lists of events, one operator call
    each)."""
    from traccc_amd import scorer as SC
    ddet = ckf.DeviceDetector(det, fld)
    plain = ckf.combinatorial_kalman_filter_algorithm(cfg)
    t0 = time.perf_counter()
    feats, labels = [], []
    for evs in label_events:
        dev = ckf.DeviceEvents(ddet, evs)
        cap = 16 * sum(e.n_seeds for e in evs)
        _, f, y, n = plain.run_device_logged(ddet, dev, capacity=cap)
        if n > cap:  # rerun with room for every sample
            cap = n
            _, f, y, n = plain.run_device_logged(ddet, dev, capacity=cap)
        feats.append(f.clone())
        labels.append(y.clone())
        del dev
    X = torch.cat(feats)
    Y = torch.cat(labels)
    del feats, labels
    t_log = time.perf_counter() - t0
    sc = SC.Scorer(seed=7)
    sc.fit_normalisation(X)
    st = sc.train(X, Y, SC.train_config(epochs=a.scorer_epochs, batch=a.scorer_batch,
                                        learning_rate=a.scorer_lr))
    n_samples = int(X.shape[0])
    survive = float(Y.float().mean())
    # MFMA utilisation of inference: the scorer's forward kernel over the
    # logged samples (the CKF scores its candidates inside the find kernel)
    sc.forward(X[:1 << 20])
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    sc.forward(X)
    e1.record()
    e1.synchronize()
    infer_ms = e0.elapsed_time(e1)
    del X, Y
    torch.cuda.synchronize()
    infer_tflops = n_samples * MLP_FLOPS / (infer_ms * 1e-3) / 1e12
    log(f"[bench] {tag} scorer: {n_samples} labels (survive {survive:.3f}, logged in "
        f"{t_log:.1f} s), loss {st.first_loss:.4f} -> {st.last_loss:.4f}, {st.steps} Adam steps "
        f"in {st.seconds:.2f} s ({st.tflops:.2f} TFLOP/s); inference {n_samples} rows in "
        f"{infer_ms:.3f} ms ({infer_tflops:.1f} TFLOP/s)")
    info = {"train_labels": n_samples, "survive_fraction": round(survive, 4),
            "epochs": a.scorer_epochs, "adam_steps": int(st.steps),
            "train_seconds": round(st.seconds, 3),
            "loss": [round(st.first_loss, 5), round(st.last_loss, 5)],
            "train_mfma": {"achieved": round(st.tflops, 3), "peak": BF16_PEAK_TFLOPS,
                           "unit": "TFLOP/s", "frac": round(st.tflops / BF16_PEAK_TFLOPS, 6)},
            "inference_mfma": {"kernel": "k_mlp_infer", "rows": n_samples,
                               "ms": round(infer_ms, 4), "achieved": round(infer_tflops, 2),
                               "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                               "frac": round(infer_tflops / BF16_PEAK_TFLOPS, 6)}}
    return sc, info


def choose_threshold(ckf, ddet, plain, pruned, sc, val, max_loss=0.004):
    """The most aggressive logit threshold (0.25 grid) whose validation
    efficiency stays within max_loss of plain; None if none does."""
    dval = ckf.DeviceEvents(ddet, val)
    eff_plain = _efficiency(val, plain.run_device(ddet, dval).download())
    sweep = []
    for thr in (1.0, 0.0, -1.0, -1.5, -1.75, -2.0, -2.25, -2.5, -2.75, -3.0, -3.25, -3.5,
                -3.75, -4.0, -4.5, -5.0, -6.0):
        pruned.set_scorer(sc, threshold=thr)
        eff = _efficiency(val, pruned.run_device(ddet, dval).download())
        sweep.append({"threshold": thr, "efficiency": round(eff, 5)})
        if eff >= eff_plain - max_loss:
            return thr, sweep
    return None, sweep


def plain_vs_pruned(a, ckf, S, torch, wl, n_particles, n_test, label_events, n_val, steps,
                    warmup, stream, fit=False, label_desc="", label_cfg=None):
    """The plain CKF (three branches per surface) against the same CKF pruned
    by the learned branch scorer, on one held-out batch of `n_test` events."""
    threads = min(16, host_threads())
    cfg = S.finding_config_for(wl)
    cfg.max_num_branches_per_surface = 3
    det, fld, _ = S.workload(wl, n_particles=10)
    ddet = ckf.DeviceDetector(det, fld)

    def events(n, seed0):
        guns = [S.workload(wl, n_particles=n_particles, seed=seed0 + e)[2] for e in range(n)]
        return S.generate_events(det, fld, guns, threads=threads)

    if label_events is None:
        chunk = max(1, 640000 // n_particles)
        label_events = [events(min(chunk, a.c5_train_events - c), 100000 + c)
                        for c in range(0, a.c5_train_events, chunk)]
        label_desc = (f"{a.c5_train_events} {wl} training events of {n_particles:,} pions "
                      f"(plain CKF, 3 branches per surface)")
    sc, sinfo = train_scorer(a, ckf, S, torch, det, fld, label_events,
                             label_cfg if label_cfg is not None else cfg, wl)
    sinfo["label_events"] = label_desc
    plain = ckf.combinatorial_kalman_filter_algorithm(cfg, stream=stream)
    pruned = ckf.combinatorial_kalman_filter_algorithm(cfg, stream=stream)
    val = events(n_val, 200000)
    chosen, sweep = choose_threshold(ckf, ddet, plain, pruned, sc, val)
    # no threshold within the efficiency bound: the plain CKF runs (threshold null)
    pruned.set_scorer(sc if chosen is not None else None,
                      threshold=chosen if chosen is not None else 0.0)

    test = events(n_test, 300000)
    n_seeds = sum(e.n_seeds for e in test)
    dev = ckf.DeviceEvents(ddet, test)
    med_p, mean_p, out_p = time_calls(plain, ddet, dev, steps, warmup, torch)
    st_p = plain.last_stats()
    med_s, mean_s, out_s = time_calls(pruned, ddet, dev, steps, warmup, torch)
    st_s = pruned.last_stats()
    tr_p, tr_s = out_p.download(), out_s.download()
    eff_p, eff_s = _efficiency(test, tr_p), _efficiency(test, tr_s)
    kp = kernel_times(plain, ddet, dev)
    ks = kernel_times(pruned, ddet, dev)
    log(f"[bench] {wl} batch {len(test)} events ({n_seeds} seeds): plain median {med_p:.2f} ms "
        f"({n_seeds / med_p * 1e3:.4g} tracks/s, eff {eff_p:.4f}), pruned median {med_s:.2f} ms "
        f"({n_seeds / med_s * 1e3:.4g} tracks/s, eff {eff_s:.4f}) at threshold {chosen}: "
        f"{med_p / med_s:.3f}x; pruned kernels {ks}; plain kernels {kp}")
    line = {
        "workload": f"{wl}: {len(test)} events x {n_particles:,} pions (pT 0.3 GeV + Exp(1 GeV), "
                    f"|eta|<2.5), 3 branches per surface, one operator call",
        "seeds_per_call": n_seeds, "calls": steps,
        "plain": {"value": round(n_seeds / med_p * 1e3, 1), "unit": "tracks/s",
                  "ms_per_call_median": round(med_p, 3), "ms_per_call_mean": round(mean_p, 3),
                  "efficiency": round(eff_p, 5), "propagations": int(st_p.n_propagations),
                  "links": int(st_p.n_links), "found_tracks": len(tr_p),
                  "kernel_ms_per_call": {k: round(v, 3) for k, v in kp.items()
                                         if isinstance(v, float)}},
        "pruned": {"value": round(n_seeds / med_s * 1e3, 1), "unit": "tracks/s",
                   "ms_per_call_median": round(med_s, 3), "ms_per_call_mean": round(mean_s, 3),
                   "efficiency": round(eff_s, 5), "propagations": int(st_s.n_propagations),
                   "links": int(st_s.n_links), "found_tracks": len(tr_s), "threshold": chosen,
                   "scored_candidates": int(st_s.n_scored),
                   "kernel_ms_per_call": {k: round(v, 3) for k, v in ks.items()
                                          if isinstance(v, float)}},
        "speedup": round(med_p / med_s, 3),
        "efficiency_loss": round(eff_p - eff_s, 5),
        "threshold_sweep": sweep,
        "scorer": sinfo,
    }
    ctx = {"det": det, "fld": fld, "cfg": cfg, "test": test, "ddet": ddet, "dev": dev,
           "pruned": pruned, "plain": plain, "stats_pruned": st_s, "kernels_pruned": ks,
           "out_pruned": out_s, "n_seeds": n_seeds}
    if fit:
        from traccc_amd import fitting as FT
        fitter = FT.kalman_fitting_algorithm(FT.fitting_config(cfg))
        fitter.run_device(ddet, dev, out_s)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(5):
            fitted = fitter.run_device(ddet, dev, out_s)
        torch.cuda.synchronize()
        dt_f = (time.perf_counter() - t0) / 5
        t0 = time.perf_counter()
        for _ in range(5):
            cands = pruned.run_device(ddet, dev)
            fitted = fitter.run_device(ddet, dev, cands)
        torch.cuda.synchronize()
        dt_c = (time.perf_counter() - t0) / 5
        nt, ns = fitted.sizes()
        fd = fitted.download()
        outc = fd.fit_outcome
        ok = float(np.mean(outc == FT.OUTCOME_SUCCESS)) if len(outc) else 0.0
        log(f"[bench] {wl} fit: {nt} tracks, {ns} states in {dt_f * 1e3:.2f} ms "
            f"({nt / dt_f:.4g} fitted tracks/s, {ok:.4f} fitted); pruned CKF + fit "
            f"{dt_c * 1e3:.2f} ms per call -> {n_seeds / dt_c:.4g} tracks/s")
        line["fit"] = {"tracks": int(nt), "states": int(ns), "ms_per_call": round(dt_f * 1e3, 3),
                       "fitted_tracks_per_s": round(nt / dt_f, 1),
                       "success_fraction": round(ok, 5)}
        line["pruned_ckf_plus_fit"] = {"value": round(n_seeds / dt_c, 1), "unit": "tracks/s",
                                       "ms_per_call": round(dt_c * 1e3, 3)}
        line["ambiguity"] = ambiguity_line(wl, dev, out_s, fitted, fd, stream, torch)
    return line, ctx


def ambiguity_line(wl, dev, cands, fitted, fd, stream, torch):
    """Greedy ambiguity resolution of the fitted tracks (the step after the
    fit), timed per operator call on the host clock (the call returns with
    the accepted flags on the host); the oracle timed on a few events of the
    same tracks on one host core."""
    import oracle
    from traccc_amd import ambiguity as AMB
    res = AMB.greedy_ambiguity_resolution_algorithm(stream=stream)
    res.run_fitted(dev, fitted)
    times = []
    for _ in range(5):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        kept = res.run_fitted(dev, fitted)
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    st = res.last_stats()
    nt = len(fd)
    # the oracle on the first events' tracks (identifiers: measurement index
    # in the batch, as the resolver takes them)
    ev_off = cands.event_offsets()
    n_ev = min(4, len(ev_off) - 1)
    t1 = int(ev_off[n_ev])
    s1 = int(fd.state_offsets[t1])
    meas_off = np.zeros(len(ev_off), np.uint64)
    meas_off[1:] = np.cumsum(dev.event_n_meas)
    ids = fd.measurement_index[:s1].astype(np.uint64)
    for e in range(n_ev):
        a, b = fd.state_offsets[ev_off[e]], fd.state_offsets[ev_off[e + 1]]
        ids[a:b] += meas_off[e]
    t0 = time.perf_counter()
    keep, _ = oracle.greedy_resolve(fd.state_offsets[:t1 + 1], ids.astype(np.uint32), fd.pval[:t1])
    cpu_s = time.perf_counter() - t0
    agree = bool(np.array_equal(np.nonzero(keep)[0], kept[kept < t1]))
    log(f"[bench] {wl} ambiguity resolution: {nt} fitted tracks -> {st.n_accepted} kept "
        f"({st.n_removed} removed in {st.n_rounds} rounds) in {dt * 1e3:.2f} ms "
        f"({nt / dt:.4g} tracks/s); oracle on {n_ev} events ({t1} tracks) {cpu_s:.2f} s on 1 core "
        f"({t1 / cpu_s:.4g} tracks/s), same kept set: {agree}")
    return {"tracks": int(nt), "kept": int(st.n_accepted), "removed": int(st.n_removed),
            "rounds": int(st.n_rounds), "ms_per_call": round(dt * 1e3, 3),
            "tracks_per_s": round(nt / dt, 1),
            "timing": "median of 5 operator calls, host clock, flags returned to the host",
            "cpu_oracle": {"tracks": t1, "events": n_ev, "seconds": round(cpu_s, 3),
                           "tracks_per_s": round(t1 / cpu_s, 1), "cores": 1},
            "oracle_agrees": agree}


def propagate_roofline(ctx, kt, workload_tag):
    """Roofline of the propagation step (k_propagate and k_propagate_resume,
    the dominant kernels of the pruned call): algorithmic bytes per step /
    the step's average duration (HIP events)."""
    st = ctx["stats_pruned"]
    launches = max(kt["steps"], 1)
    lanes = kt["lanes"]
    fld = ctx["fld"]
    # per link its 32 B record and 4 B alive-prefix word read and 1 B outcome
    # flag written; per propagated branch the 112 B filtered state read and
    # the 112 B propagated state written; once per launch the field map
    # (2 x 4 B per (r, z) node) when the field is a map
    field_bytes = 0
    if fld.struct.kind == 1:
        field_bytes = 8 * fld.struct.n_r * fld.struct.n_z
    alg_bytes = st.n_links * (32 + 4 + 1) + st.n_propagations * (112 + 112) + \
        field_bytes * launches
    avg_ms = kt["propagate_kernel"] / launches
    per_launch = alg_bytes / launches
    achieved = per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    # one "launch" is the propagation step: k_propagate and, with one lane
    # per link, the k_propagate_resume that continues its deferred links
    # (the HIP events bracket both)
    kernels = [f"k_propagate<{lanes}>"] + (["k_propagate_resume"] if lanes == 1 else [])
    kernel = " + ".join(kernels)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload_tag}.json")
    if os.path.exists(tpath):
        # HBM bytes of the kernel over one pruned call of the same workload
        # (FETCH_SIZE + WRITE_SIZE, copy-calibrated), per working launch
        with open(tpath) as fh:
            tk = json.load(fh).get("calls", {}).get("pruned", {})
        if kernels[0] in tk:
            traffic = int(sum(tk[k]["hbm_bytes"] for k in kernels if k in tk) / launches)
    return {"bound": "hbm", "kernel": kernel, "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 6),
            "traffic": traffic, "avg_launch_ms": round(avg_ms, 5),
            "bytes_per_launch": int(per_launch), "launches_per_call": launches,
            "algorithmic_bytes": "37 B per link + 224 B per propagated branch + field map "
                                 f"({field_bytes} B) per launch"}


def c2_line(a, ckf, S, torch, stream):
    """C2: 1 event x 10,000 muons, plain CKF."""
    from traccc_amd import performance as PERF
    det, fld, gun = S.workload("C2", seed=42)
    ev = S.generate_event(det, fld, gun)
    cfg = S.finding_config_for("C2")
    alg = ckf.combinatorial_kalman_filter_algorithm(cfg, stream=stream)
    ddet = ckf.DeviceDetector(det, fld)
    dev = ckf.DeviceEvents(ddet, ev)
    med, mean, out = time_calls(alg, ddet, dev, max(a.steps, 20), a.warmup, torch)
    res = out.download()
    perf = PERF.finding_performance(ev, res.offsets, res.constituents)
    log(f"[bench] C2: {ev.n_seeds} seeds, median {med:.3f} ms -> "
        f"{ev.n_seeds / med * 1e3:.4g} tracks/s, efficiency {perf.efficiency:.4f}")
    return {"workload": "C2: 1 event x 10,000 muons (pT log-uniform 0.5-100 GeV, |eta|<2.5), "
                        "4 barrel + 2x4 disks, 1% X0, 5% noise, Bz = 2 T, plain CKF",
            "value": round(ev.n_seeds / med * 1e3, 1), "unit": "tracks/s",
            "ms_per_call_median": round(med, 4), "efficiency": round(perf.efficiency, 5)}


def c3_line(a, ckf, S, torch, stream):
    """C3: one operator call over 256 events (log-uniform 500-50,000 muons
    each, (r, z) solenoid map), inputs resident in HBM."""
    det, fld, guns = S.batch_workload("C3", n_events=a.batched_events, seed=42)
    events = S.generate_events(det, fld, guns, threads=min(16, host_threads()))
    n_seeds = sum(e.n_seeds for e in events)
    cfg = S.finding_config_for("C3")
    alg = ckf.combinatorial_kalman_filter_algorithm(cfg, stream=stream)
    ddet = ckf.DeviceDetector(det, fld)
    dev = ckf.DeviceEvents(ddet, events)
    med, mean, out = time_calls(alg, ddet, dev, a.batched_steps, 2, torch)
    n_tracks, _ = out.sizes()
    log(f"[bench] C3 batch: {len(events)} events, {n_seeds} seeds: median {med:.2f} ms per call "
        f"-> {n_seeds / med * 1e3:.4g} tracks/s")
    return {"workload": f"C3: {len(events)} events x log-uniform 500-50,000 muons, (r,z) solenoid "
                        f"map 10 mm, C2 detector, one operator call",
            "seeds_per_call": n_seeds, "calls": a.batched_steps,
            "ms_per_call_median": round(med, 3), "value": round(n_seeds / med * 1e3, 1),
            "unit": "tracks/s", "found_tracks": int(n_tracks)}


def c2_label_events(a, S):
    """Events of the C2 configuration (10,000 muons each) on which the plain
    CKF logs the C4 scorer's training labels, in calls of 16 events."""
    det, fld, _ = S.workload("C4", n_particles=10)
    out = []
    for c in range(0, a.c4_label_events, 16):
        guns = [S.workload("C2", seed=500000 + c + e)[2]
                for e in range(min(16, a.c4_label_events - c))]
        out.append(S.generate_events(det, fld, guns, threads=min(16, host_threads())))
    return out


def cpu_baseline_c5(a, ctx, oracle):
    """The FP32 oracle (traccc's host CKF restated) over C5 events of the
    timed batch, one event per host thread (traccc's throughput_mt pattern),
    the same finding config as the plain GPU line."""
    threads = a.cpu_threads or host_threads()
    n_ev = min(len(ctx["test"]), max(1, a.cpu_events or threads))
    evs = ctx["test"][:n_ev]
    seeds = sum(e.n_seeds for e in evs)
    secs, ntr = oracle.run_many(ctx["cfg"], ctx["det"], ctx["fld"], evs, precision=32,
                                threads=threads)
    value = seeds / secs
    log(f"[bench] CPU oracle FP32: {n_ev} C5 events ({seeds} seeds) in {secs:.2f} s on "
        f"{threads} threads -> {value:.4g} tracks/s")
    return {"value": round(value, 1), "unit": "tracks/s", "cores": threads, "kind": "port",
            "nproc": os.cpu_count(), "affinity_cores": host_threads(),
            "sample": f"{n_ev} events of the timed C5 batch ({seeds} seeds), FP32 oracle "
                      f"restating traccc's host CKF (plain, 3 branches per surface), one event "
                      f"per thread, {secs:.2f} s wall"}


def seeding_line(a, torch, stream):
    """Triplet seeding (the step before the CKF) on the spacepoints of one
    C2-configuration event (10,000 muons, spacepoints already in HBM):
    median of HIP-event-timed calls; the oracle on the same spacepoints on
    one host core; both seed lists compared."""
    import oracle
    from traccc_amd import ckf
    from traccc_amd import seeding as SD
    from traccc_amd import simulation as S
    det, fld, gun = S.workload("C2", seed=42)
    ev = S.generate_event(det, fld, gun)
    ddet = ckf.DeviceDetector(det, fld)
    dev = ckf.DeviceEvents(ddet, ev)
    xyz, midx = SD.spacepoint_formation(ddet, dev, 0)
    finder = SD.seedfinder_config(bFieldInZ=2.0 * SD.TESLA)
    grid = SD.spacepoint_grid_config(finder)
    filt = SD.seedfilter_config()
    alg = SD.seeding_algorithm(finder, grid, filt, stream=stream)
    d_xyz = torch.from_numpy(xyz).cuda()
    got = alg.run_device(d_xyz)
    times = []
    for _ in range(max(a.steps, 10)):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        got = alg.run_device(d_xyz)
        e1.record()
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1))
    ms = float(np.median(times))
    t0 = time.perf_counter()
    b, m, t, w, _ = oracle.seeding(finder, grid, filt, xyz)
    cpu_s = time.perf_counter() - t0
    same = (len(b) == len(got) and np.array_equal(b, got.bottom) and np.array_equal(m, got.middle)
            and np.array_equal(t, got.top) and np.array_equal(w.view(np.uint32), got.weight.view(np.uint32)))
    log(f"[bench] seeding: {len(xyz)} spacepoints -> {len(got)} seeds in {ms:.3f} ms "
        f"({len(xyz) / ms * 1e3:.4g} spacepoints/s); oracle {cpu_s:.2f} s on 1 core; same seeds: {same}")
    return {"workload": "spacepoints of one C2-configuration event (10,000 muons, 4 barrel + 2x4 "
                        "disk layers, 5% noise), traccc default seedfinder config at 2 T",
            "spacepoints": int(len(xyz)), "seeds": int(len(got)), "ms_per_call": round(ms, 3),
            "spacepoints_per_s": round(len(xyz) / ms * 1e3, 1),
            "timing": "median of the timed calls; the call ends with the seed count on the host",
            "cpu_oracle": {"seconds": round(cpu_s, 3), "cores": 1,
                           "spacepoints_per_s": round(len(xyz) / cpu_s, 1)},
            "oracle_agrees": bool(same)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed calls per line (median)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: every core of the affinity mask")
    ap.add_argument("--cpu-events", type=int, default=0, help="0: one C5 event per thread")
    ap.add_argument("--c5-events", type=int, default=64)
    ap.add_argument("--c5-train-events", type=int, default=32)
    ap.add_argument("--c5-val-events", type=int, default=2)
    ap.add_argument("--batched-events", type=int, default=256,
                    help="events of the C3 line (0 skips it)")
    ap.add_argument("--batched-steps", type=int, default=20)
    ap.add_argument("--c4-events", type=int, default=32, help="0 skips the C4 line")
    ap.add_argument("--c4-label-events", type=int, default=336,
                    help="C2-configuration events the C4 scorer's labels are logged on")
    ap.add_argument("--no-c2", action="store_true")
    ap.add_argument("--no-seeding", action="store_true")
    ap.add_argument("--scorer-epochs", type=int, default=3)
    ap.add_argument("--scorer-batch", type=int, default=65536,
                    help="samples per Adam step (1,024 64-sample tiles: one wave per SIMD)")
    ap.add_argument("--scorer-lr", type=float, default=4e-3)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB three times (the "
                         "reference dispatch for scaling FETCH_SIZE / WRITE_SIZE counters)")
    a = ap.parse_args()

    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the CKF shards by event, and multi-GPU runs are "
              "not part of this build (use --gpus 1)", file=sys.stderr)
        return 2

    import torch

    if a.copy_calibration > 0:
        n = a.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": 4 * n, "dispatches": 3}), flush=True)
        return 0

    from traccc_amd import ckf
    from traccc_amd import simulation as S

    torch.cuda.set_device(0)
    stream = torch.cuda.current_stream().cuda_stream

    c5, ctx = plain_vs_pruned(a, ckf, S, torch, "C5", 60000, a.c5_events, None,
                              a.c5_val_events, a.steps, a.warmup, stream, fit=True)
    roofline = propagate_roofline(ctx, ctx["kernels_pruned"], "C5_pruned")
    log(f"[bench] roofline {roofline}")

    c2 = None if a.no_c2 else c2_line(a, ckf, S, torch, stream)
    seeding = None if a.no_seeding else seeding_line(a, torch, stream)
    c3 = c3_line(a, ckf, S, torch, stream) if a.batched_events > 0 else None
    c4 = None
    if a.c4_events > 0:
        c2cfg = S.finding_config_for("C2")
        c2cfg.max_num_branches_per_surface = 3
        c4, _ = plain_vs_pruned(a, ckf, S, torch, "C4", 20000, a.c4_events, c2_label_events(a, S),
                                2, a.steps, a.warmup, stream,
                                label_desc=f"{a.c4_label_events} C2-configuration events "
                                           f"(10,000 muons each; plain CKF, muon hypothesis, "
                                           f"3 branches per surface)", label_cfg=c2cfg)

    cpu = None
    if not a.no_cpu_baseline:
        import oracle
        cpu = cpu_baseline_c5(a, ctx, oracle)

    value = c5["pruned"]["value"]
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "tracks/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": c5["pruned"]["ms_per_call_median"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "C5: 64 events x 60,000 pions, 5 barrel + 2x6 disks, 2% X0, 3% "
                               "dead modules, 10% noise, (r,z) solenoid map, truth seeds, "
                               "3 branches per surface; MLP-pruned CKF in one operator call",
                   "seeds_per_step": ctx["n_seeds"], "path": "MLP-pruned HIP CKF",
                   "plain_value": c5["plain"]["value"], "speedup_vs_plain": c5["speedup"],
                   "efficiency_loss": c5["efficiency_loss"],
                   "timing": "median of the timed calls, HIP events on the algorithm's stream"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "gpu_over_cpu": round(value / cpu["value"], 2) if cpu else None,
        "plain_over_cpu": round(c5["plain"]["value"] / cpu["value"], 2) if cpu else None,
        "c5_plain_vs_pruned": c5,
        "c2": c2,
        "batched": c3,
        "c4_plain_vs_pruned": c4,
        "seeding": seeding,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
