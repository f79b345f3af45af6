"""Benchmark: CKF tracks/s on one MI355X.

A step is one CKF operator call (traccc's combinatorial_kalman_filter_algorithm
operator()) over one event of the C2 workload: 4 barrel layers + 2x4 disks,
1% X0 per module, 5% noise, Bz = 2 T, 10,000 muons with pT log-uniform in
[0.5, 100] GeV and |eta| < 2.5, truth seeds (seed 42).  Inputs are resident in
HBM before the timed region; the output stays in HBM.  tracks/s = seeds
processed / wall time.

Extra fields: roofline of the dominant kernel (HIP-event timed), finding
efficiency, the FP32 CPU oracle timed on the host cores, the batched C3 line
(256 events per call) and the C4 line: plain CKF against the CKF pruned by the
learned branch scorer, trained here on the GPU.

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


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def kernel_times(alg, ddet, devents):
    """Per-kernel-group device ms of one timed operator call."""
    from traccc_amd import _ffi as F
    from traccc_amd import ckf
    lib = F.load_library(ckf.LIB)
    lib.ckf_debug_set_kernel_timing.argtypes = [C.c_void_p, C.c_int]
    lib.ckf_debug_kernel_times.argtypes = [C.c_void_p] + [C.POINTER(C.c_double)] * 4 + [
        C.POINTER(C.c_uint32)]
    lib.ckf_debug_set_kernel_timing(alg._h, 1)
    alg.run_device(ddet, devents)
    vals = [C.c_double() for _ in range(4)]
    n = C.c_uint32()
    lib.ckf_debug_kernel_times(alg._h, *[C.byref(v) for v in vals], C.byref(n))
    lib.ckf_debug_set_kernel_timing(alg._h, 0)
    lib.ckf_debug_get_schedule.argtypes = [C.c_void_p, C.POINTER(C.c_uint32)]
    lanes = C.c_uint32()
    lib.ckf_debug_get_schedule(alg._h, C.byref(lanes))
    return {"find": vals[0].value, "dedup": vals[1].value, "propagate": vals[2].value,
            "advance": vals[3].value, "steps": n.value, "lanes": lanes.value}


def batched_line(a, ckf, S, torch):
    """C3: one operator call over a batch of events (log-uniform 500-50,000
    muons each, (r, z) solenoid map), inputs resident in HBM."""
    det, fld, guns = S.batch_workload("C3", n_events=a.batched_events, seed=42)
    t0 = time.perf_counter()
    events = S.generate_events(det, fld, guns, threads=min(16, len(os.sched_getaffinity(0))))
    n_seeds = sum(e.n_seeds for e in events)
    log(f"[bench] C3 batch: {len(events)} events, {n_seeds} seeds, "
        f"{sum(e.n_meas for e in events)} measurements (generated in {time.perf_counter() - t0:.1f} s)")
    cfg = S.finding_config_for("C3")
    alg = ckf.combinatorial_kalman_filter_algorithm(cfg)
    ddet = ckf.DeviceDetector(det, fld)
    devents = ckf.DeviceEvents(ddet, events)
    for _ in range(2):
        alg.run_device(ddet, devents)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.batched_steps):
        out = alg.run_device(ddet, devents)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / a.batched_steps
    n_tracks, _ = out.sizes()
    log(f"[bench] C3 batch: {dt * 1e3:.2f} ms per call -> {n_seeds / dt:.4g} tracks/s")
    return {"workload": f"C3: {len(events)} events x log-uniform 500-50,000 muons, (r,z) solenoid "
                        f"map 10 mm, C2 detector, one operator call",
            "seeds_per_call": n_seeds, "calls": a.batched_steps, "ms_per_call": round(dt * 1e3, 3),
            "value": round(n_seeds / dt, 1), "unit": "tracks/s", "found_tracks": int(n_tracks)}


def _batch_efficiency(events, tracks):
    from traccc_amd import performance as PERF
    return float(np.mean([PERF.finding_performance(ev, t.offsets, t.constituents).efficiency
                          for ev, t in ((ev, tracks.event(e)) for e, ev in enumerate(events))]))


def _time_calls(alg, ddet, devents, steps, torch):
    for _ in range(2):
        alg.run_device(ddet, devents)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = alg.run_device(ddet, devents)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps, out


def pruned_line(a, ckf, S, torch, wl="C4", n_particles=20000, n_test=32, n_train=128,
                n_steps=5, fit=False):
    """A dense workload (C4: 20,000 pions per event, C2 detector; C5: 60,000
    pions, 5 barrel layers + 2x6 disks, 2% X0, 3% dead modules, solenoid
    map): the plain CKF with three branches per surface against the same CKF
    with the learned branch scorer, on one batch of events.  The scorer is
    trained here, on the GPU, on survive / die labels that the plain CKF logs
    on its own training events (a link survives when it lies on its seed's
    best found track); the threshold is the most aggressive one that keeps
    validation efficiency within 0.4% absolute of plain; the timed batch is
    held out.  With fit=True the pruned candidates then go through the Kalman
    fit + two-filter smoother, timed on its own and as the chained call."""
    from traccc_amd import scorer as SC
    threads = min(16, len(os.sched_getaffinity(0)))
    cfg = S.finding_config_for(wl)
    cfg.max_num_branches_per_surface = 3
    det, fld, _ = S.workload(wl, n_particles=10)
    ddet = ckf.DeviceDetector(det, fld)

    def events(n, seed0):
        guns = [S.workload(wl, n_particles=n_particles, seed=seed0 + e)[2] for e in range(n)]
        return S.generate_events(det, fld, guns, threads=threads)

    plain = ckf.combinatorial_kalman_filter_algorithm(cfg)
    # branch log of the plain CKF over the training events
    t0 = time.perf_counter()
    feats, labels = [], []
    chunk = max(1, 640000 // n_particles)
    for c in range(0, n_train, chunk):
        evs = events(min(chunk, n_train - c), 100000 + c)
        dev = ckf.DeviceEvents(ddet, evs)
        cap = 16 * n_particles * len(evs)
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
    tcfg = SC.train_config(epochs=a.scorer_epochs, batch=16384, learning_rate=2e-3)
    st = sc.train(X, Y, tcfg)
    n_samples = int(X.shape[0])
    survive = float(Y.float().mean())
    del X, Y
    torch.cuda.synchronize()
    log(f"[bench] {wl} scorer: {n_samples} labels from {n_train} events "
        f"(survive {survive:.3f}, {t_log:.1f} s), loss {st.first_loss:.4f} -> {st.last_loss:.4f}, "
        f"{st.steps} Adam steps in {st.seconds:.2f} s, {st.tflops:.2f} TFLOP/s")

    # threshold on validation events
    val = events(max(2, 80000 // n_particles), 200000)
    dval = ckf.DeviceEvents(ddet, val)
    eff_plain_val = _batch_efficiency(val, plain.run_device(ddet, dval).download())
    pruned = ckf.combinatorial_kalman_filter_algorithm(cfg)
    chosen, sweep = None, []
    for thr in (1.0, 0.0, -1.0, -1.5, -1.75, -2.0, -2.25, -2.5, -2.75, -3.0, -3.25, -3.5,
                -3.75, -4.0, -4.5, -5.0, -6.0):
        pruned.set_scorer(sc, threshold=thr)
        eff = _batch_efficiency(val, pruned.run_device(ddet, dval).download())
        sweep.append({"threshold": thr, "efficiency": round(eff, 5)})
        if eff >= eff_plain_val - 0.004:
            chosen = thr
            break
    del dval
    # no threshold within the efficiency bound: the plain CKF runs (threshold null)
    pruned.set_scorer(sc if chosen is not None else None,
                      threshold=chosen if chosen is not None else 0.0)

    # timed: the held-out batch, plain and pruned
    test = events(n_test, 300000)
    n_seeds = sum(e.n_seeds for e in test)
    dev = ckf.DeviceEvents(ddet, test)
    dt_p, out_p = _time_calls(plain, ddet, dev, n_steps, torch)
    st_p = plain.last_stats()
    dt_s, out_s = _time_calls(pruned, ddet, dev, n_steps, torch)
    st_s = pruned.last_stats()
    tr_p, tr_s = out_p.download(), out_s.download()
    eff_p, eff_s = _batch_efficiency(test, tr_p), _batch_efficiency(test, tr_s)
    # scorer kernels: device time over one timed call
    from traccc_amd import _ffi as F
    lib = F.load_library(ckf.LIB)
    lib.ckf_debug_set_kernel_timing.argtypes = [C.c_void_p, C.c_int]
    lib.ckf_debug_scorer_time.argtypes = [C.c_void_p, C.POINTER(C.c_double),
                                          C.POINTER(C.c_double)]
    lib.ckf_debug_kernel_times.argtypes = [C.c_void_p] + [C.POINTER(C.c_double)] * 4 + [
        C.POINTER(C.c_uint32)]
    lib.ckf_debug_set_kernel_timing(pruned._h, 1)
    pruned.run_device(ddet, dev)
    ms_list, ms_score = C.c_double(), C.c_double()
    lib.ckf_debug_scorer_time(pruned._h, C.byref(ms_list), C.byref(ms_score))
    kt = [C.c_double() for _ in range(4)]
    nst = C.c_uint32()
    lib.ckf_debug_kernel_times(pruned._h, *[C.byref(v) for v in kt], C.byref(nst))
    lib.ckf_debug_set_kernel_timing(pruned._h, 0)
    pruned_kernel_ms = {"list": round(ms_list.value, 3), "score": round(ms_score.value, 3),
                        "find_update": round(kt[0].value, 3), "dedup": round(kt[1].value, 3),
                        "propagate": round(kt[2].value, 3), "advance": round(kt[3].value, 3)}
    n_scored = pruned.last_stats().n_scored
    kp = kernel_times(plain, ddet, dev)
    plain_kernel_ms = {k: round(kp[k], 3) for k in ("find", "dedup", "propagate", "advance")}
    mlp_flops = 2 * (12 * 64 + 64 * 64 + 64)  # per scored candidate
    infer_tflops = n_scored * mlp_flops / (ms_score.value * 1e-3) / 1e12 if ms_score.value else 0.0
    log(f"[bench] {wl} batch {len(test)} events: plain {dt_p * 1e3:.2f} ms ({n_seeds / dt_p:.4g} "
        f"tracks/s, eff {eff_p:.4f}), pruned {dt_s * 1e3:.2f} ms ({n_seeds / dt_s:.4g} tracks/s, "
        f"eff {eff_s:.4f}) at threshold {chosen}; scorer {n_scored} candidates in "
        f"{ms_score.value:.3f} ms ({infer_tflops:.2f} TFLOP/s); pruned kernel ms {pruned_kernel_ms}; "
        f"plain kernel ms {plain_kernel_ms}")
    line = {
        "workload": f"{wl}: {len(test)} events x {n_particles:,} pions (pT 0.3 GeV + Exp(1 GeV), "
                    f"|eta|<2.5), 10% noise, 3 branches per surface, one operator call",
        "seeds_per_call": n_seeds, "calls": n_steps,
        "plain": {"value": round(n_seeds / dt_p, 1), "unit": "tracks/s",
                  "ms_per_call": round(dt_p * 1e3, 3), "efficiency": round(eff_p, 5),
                  "propagations": int(st_p.n_propagations), "links": int(st_p.n_links),
                  "found_tracks": len(tr_p)},
        "pruned": {"value": round(n_seeds / dt_s, 1), "unit": "tracks/s",
                   "ms_per_call": round(dt_s * 1e3, 3), "efficiency": round(eff_s, 5),
                   "propagations": int(st_s.n_propagations), "links": int(st_s.n_links),
                   "found_tracks": len(tr_s), "threshold": chosen},
        "speedup": round(dt_p / dt_s, 3),
        "efficiency_loss": round(eff_p - eff_s, 5),
        "threshold_sweep": sweep,
        "scorer": {
            "train_labels": n_samples, "train_events": n_train,
            "survive_fraction": round(survive, 4), "epochs": a.scorer_epochs,
            "adam_steps": int(st.steps), "train_seconds": round(st.seconds, 3),
            "loss": [round(st.first_loss, 5), round(st.last_loss, 5)],
            "train_mfma": {"achieved": round(st.tflops, 3), "peak": 2500.0, "unit": "TFLOP/s",
                           "frac": round(st.tflops / 2500.0, 6)},
            "inference_mfma": {"achieved": round(infer_tflops, 3), "peak": 2500.0,
                               "unit": "TFLOP/s", "frac": round(infer_tflops / 2500.0, 6),
                               "scored_candidates": int(n_scored),
                               "score_kernel_ms": round(ms_score.value, 4)},
        },
        "pruned_kernel_ms_per_call": pruned_kernel_ms,
        "plain_kernel_ms_per_call": plain_kernel_ms,
    }
    if fit:
        from traccc_amd import fitting as FT
        fitter = FT.kalman_fitting_algorithm(FT.fitting_config(cfg))
        fitter.run_device(ddet, dev, out_s)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n_steps):
            fitted = fitter.run_device(ddet, dev, out_s)
        torch.cuda.synchronize()
        dt_f = (time.perf_counter() - t0) / n_steps
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n_steps):
            cands = pruned.run_device(ddet, dev)
            fitted = fitter.run_device(ddet, dev, cands)
        torch.cuda.synchronize()
        dt_c = (time.perf_counter() - t0) / n_steps
        nt, ns = fitted.sizes()
        outc = fitted.download().fit_outcome
        ok = float(np.mean(outc == FT.OUTCOME_SUCCESS)) if len(outc) else 0.0
        log(f"[bench] {wl} fit: {nt} tracks, {ns} states in {dt_f * 1e3:.2f} ms "
            f"({nt / dt_f:.4g} fitted tracks/s, {ok:.4f} fitted); pruned CKF + fit "
            f"{dt_c * 1e3:.2f} ms per call -> {n_seeds / dt_c:.4g} tracks/s")
        line["fit"] = {"tracks": int(nt), "states": int(ns), "ms_per_call": round(dt_f * 1e3, 3),
                       "fitted_tracks_per_s": round(nt / dt_f, 1), "success_fraction": round(ok, 5)}
        line["pruned_ckf_plus_fit"] = {"value": round(n_seeds / dt_c, 1), "unit": "tracks/s",
                                       "ms_per_call": round(dt_c * 1e3, 3)}
    return line


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="C2")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-events-per-core", type=int, default=6)
    ap.add_argument("--batched-events", type=int, default=256,
                    help="events of the batched C3 line (0 skips it)")
    ap.add_argument("--batched-steps", type=int, default=5)
    ap.add_argument("--pruned-events", type=int, default=32,
                    help="events of the timed C4 plain-vs-pruned batch (0 skips the line)")
    ap.add_argument("--pruned-steps", type=int, default=5)
    ap.add_argument("--scorer-train-events", type=int, default=128)
    ap.add_argument("--scorer-epochs", type=int, default=3)
    ap.add_argument("--c5-events", type=int, default=64,
                    help="events of the C5 line (60,000 pions each; 0 skips it)")
    ap.add_argument("--c5-train-events", type=int, default=32)
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
    from traccc_amd import performance as PERF
    from traccc_amd import simulation as S

    torch.cuda.set_device(0)
    det, fld, gun = S.workload(a.workload, seed=42)
    ev = S.generate_event(det, fld, gun)
    cfg = S.finding_config_for(a.workload)
    log(f"[bench] {a.workload}: {ev.n_seeds} seeds, {ev.n_meas} measurements")

    alg = ckf.combinatorial_kalman_filter_algorithm(cfg)
    ddet = ckf.DeviceDetector(det, fld)
    devents = ckf.DeviceEvents(ddet, ev)

    for _ in range(a.warmup):
        alg.run_device(ddet, devents)
    torch.cuda.synchronize()
    outs = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        outs.append(alg.run_device(ddet, devents))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    stats = alg.last_stats()
    result = outs[-1].download()
    del outs
    n_seeds = ev.n_seeds
    value = n_seeds * a.steps / elapsed
    perf = PERF.finding_performance(ev, result.offsets, result.constituents)
    log(f"[bench] {a.steps} calls in {elapsed:.4f} s -> {value:.4g} tracks/s, "
        f"{len(result)} candidates, efficiency {perf.efficiency:.4f}")

    # roofline of the dominant kernel (HIP events on the algorithm's stream)
    kt = kernel_times(alg, ddet, devents)
    st = alg.last_stats()
    launches = max(kt["steps"], 1)
    groups = {k: kt[k] for k in ("find", "dedup", "propagate", "advance")}
    dom = max(groups, key=groups.get)
    # algorithmic bytes of k_propagate over the call: every link reads its
    # 32 B record + 4 B rank word; every propagated branch reads and writes a
    # 112 B bound state (27 floats + surface) and appends a 4 B slot index
    prop_bytes = st.n_links * 36 + st.n_propagations * (112 + 112 + 4)
    # k_find: per input its state (112 B) and previous link (32 B), per link
    # a 32 B record, per kept branch a 112 B filtered state + 8 B candidate
    find_bytes = st.n_links * (112 + 32 + 32 + 120)
    alg_bytes = {"propagate": prop_bytes, "find": find_bytes}.get(dom, prop_bytes)
    avg_ms = groups[dom] / launches
    achieved = (alg_bytes / launches) / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    # the kernels the timed group holds: the work kernel (propagation at the
    # schedule the autotuner kept) and its compaction kernel
    members = {"propagate": [f"k_propagate<{kt['lanes']}>", "k_prop_emit"],
               "find": ["k_find<0>", "k_find_emit<false>"]}.get(dom, [f"k_{dom}"])
    # measured HBM bytes per launch of those kernels: FETCH_SIZE / WRITE_SIZE
    # counter runs over this command, scaled by a 1 GiB copy (profiles/)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"hbm_traffic_{a.workload}_10k_muons.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            tk = json.load(fh).get("kernels", {})
        if all(m in tk for m in members):
            traffic = sum(tk[m]["hbm_bytes_per_launch"] for m in members)
    roofline = {"bound": "hbm", "kernel": " + ".join(members), "achieved": round(achieved, 2),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 6),
                "traffic": traffic, "avg_launch_ms": round(avg_ms, 5),
                "bytes_per_launch": int(alg_bytes / launches),
                "kernel_ms_per_call": {k: round(v, 4) for k, v in groups.items()},
                "launches_per_call": launches}
    log(f"[bench] kernel ms per call: {groups}, steps {launches}")

    batched = None
    if a.batched_events > 0:
        batched = batched_line(a, ckf, S, torch)
    pruned = None
    if a.pruned_events > 0:
        pruned = pruned_line(a, ckf, S, torch, "C4", 20000, a.pruned_events,
                             a.scorer_train_events, a.pruned_steps)
    c5 = None
    if a.c5_events > 0:
        c5 = pruned_line(a, ckf, S, torch, "C5", 60000, a.c5_events, a.c5_train_events,
                         a.pruned_steps, fit=True)

    cpu = None
    if not a.no_cpu_baseline:
        import oracle
        cores = min(16, len(os.sched_getaffinity(0)))
        n_ev = a.cpu_events_per_core * cores
        secs, ntr = oracle.run_many(cfg, det, fld, [ev] * n_ev, precision=32, threads=cores)
        cpu_value = n_seeds * n_ev / secs
        cpu = {"value": round(cpu_value, 1), "unit": "tracks/s", "cores": cores, "kind": "port",
               "sample": f"{n_ev} copies of the {a.workload} event ({n_seeds} seeds each), FP32 "
                         f"oracle restating traccc's host CKF, one event per thread, "
                         f"{secs:.2f} s wall"}
        log(f"[bench] CPU oracle FP32: {cpu_value:.4g} tracks/s on {cores} cores")

    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "tracks/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(elapsed / a.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{a.workload}: 1 event x {gun.n_particles} muons, 4 barrel + 2x4 "
                               "disks, 1% X0, 5% noise, Bz=2T, truth seeds",
                   "seeds_per_step": n_seeds, "path": "plain HIP CKF",
                   "max_num_branches_per_surface": cfg.max_num_branches_per_surface},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batched": batched,
        "mlp_pruned": pruned,
        "c5_pruned_fit": c5,
        "efficiency": round(perf.efficiency, 5),
        "found_tracks_per_s": round(len(result) * a.steps / elapsed, 1),
        "gpu_over_cpu": round(value / cpu["value"], 2) if cpu else None,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
