"""Benchmark: CKF tracks/s on one MI355X.

A step is one CKF operator call (traccc's combinatorial_kalman_filter_algorithm
operator()) over one event of the C2 workload: 4 barrel layers + 2x4 disks,
1% X0 per module, 5% noise, Bz = 2 T, 10,000 muons with pT log-uniform in
[0.5, 100] GeV and |eta| < 2.5, truth seeds (seed 42).  Inputs are resident in
HBM before the timed region; the output stays in HBM.  tracks/s = seeds
processed / wall time.

Extra fields: roofline of the dominant kernel (HIP-event timed), finding
efficiency, and the FP32 CPU oracle timed on the host cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

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
    return {"find": vals[0].value, "dedup": vals[1].value, "propagate": vals[2].value,
            "advance": vals[3].value, "steps": n.value}


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
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB three times (the "
                         "reference dispatch for scaling FETCH_SIZE / WRITE_SIZE counters)")
    a = ap.parse_args()

    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the CKF shards by event, and multi-GPU runs are "
              "not part of this build (use --gpus 1)", file=sys.stderr)
        return 2

    import numpy as np
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
    # measured HBM bytes per launch of that kernel: FETCH_SIZE / WRITE_SIZE
    # counter runs over this command, scaled by a 1 GiB copy (profiles/)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"hbm_traffic_{a.workload}_10k_muons.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get("kernels", {}).get(f"k_{dom}", {}).get(
                "hbm_bytes_per_launch")
    roofline = {"bound": "hbm", "kernel": f"k_{dom}", "achieved": round(achieved, 2),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 6),
                "traffic": traffic, "avg_launch_ms": round(avg_ms, 5),
                "bytes_per_launch": int(alg_bytes / launches),
                "kernel_ms_per_call": {k: round(v, 4) for k, v in groups.items()},
                "launches_per_call": launches}
    log(f"[bench] kernel ms per call: {groups}, steps {launches}")

    batched = None
    if a.batched_events > 0:
        batched = batched_line(a, ckf, S, torch)

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
        "efficiency": round(perf.efficiency, 5),
        "found_tracks_per_s": round(len(result) * a.steps / elapsed, 1),
        "gpu_over_cpu": round(value / cpu["value"], 2) if cpu else None,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
