"""Markdown / JSON tables from rocprofv3 result databases (rocpd SQLite).

    python -m traccc_amd.rocprof_tables kernels <results.db>
        per-kernel totals (the database's top_kernels view) with min / max
    python -m traccc_amd.rocprof_tables calls <results.db>
        kernel time per operator call of the last plain and the last
        scorer-pruned CKF call in the trace (calls end at k_build_write)
    python -m traccc_amd.rocprof_tables counters <results.db> [<results.db> ...]
        derived counters (VALUBusy, VALUUtilization, MeanOccupancyPerActiveCU,
        MfmaUtil, ...) per kernel of the last plain and the last scorer-pruned
        CKF call of each run, weighted by dispatch duration, with the kernel's
        ms per call
    python -m traccc_amd.rocprof_tables traffic <fetch.db> <write.db> <copy_fetch.db> <copy_write.db> <copy_mib>
        HBM bytes per launch from one FETCH_SIZE and one WRITE_SIZE counter
        run, each scaled by what the same counter reads for a device copy of
        copy_mib MiB (read and written once), as JSON: the mean over every
        dispatch of the run, and the total of the last plain and the last
        scorer-pruned CKF call (dispatches in order, calls end at
        k_build_write)

Host-side tooling for the committed profile summaries; nothing here touches
a GPU.
"""
from __future__ import annotations

import json
import re
import sqlite3
import sys
from collections import defaultdict


def _short(name: str) -> str:
    name = re.sub(r"\(anonymous namespace\)::", "", name)
    name = re.sub(r"^void ", "", name)
    name = re.sub(r"\(.*\)$", "", name)
    return name.replace("tamd::", "")


def kernels_table(db: str) -> str:
    con = sqlite3.connect(db)
    cur = con.cursor()
    mm = {n: (lo, hi) for n, lo, hi in
          cur.execute("select name, min(duration), max(duration) from kernels group by name")}
    out = ["| kernel | calls | total us | avg us | min us | max us | % |",
           "|---|---:|---:|---:|---:|---:|---:|"]
    for name, calls, tot, avg, pct in cur.execute(
            "select name, total_calls, total_duration, average, percentage from top_kernels"):
        lo, hi = mm.get(name, (0, 0))
        out.append(f"| `{_short(name)}` | {calls} | {tot:.1f} | {avg:.1f} | {lo / 1e3:.1f} | "
                   f"{hi / 1e3:.1f} | {pct:.2f} |")
    return "\n".join(out)


def calls_table(db: str) -> str:
    con = sqlite3.connect(db)
    rows = list(con.execute("select name, start, end from kernels order by start"))
    ends = [i for i, r in enumerate(rows) if "k_build_write" in r[0]]
    last = {}
    prev = 0
    for i in ends:
        seg = rows[prev:i + 1]
        prev = i + 1
        names = {_short(r[0]) for r in seg}
        kind = ("pruned" if "k_find_list" in names else
                "logged" if "k_find<1>" in names else "plain")
        last[kind] = seg
    out = []
    for kind in ("plain", "pruned"):
        seg = last.get(kind)
        if not seg:
            continue
        t0 = min(r[1] for r in seg if r[0].find("k_find") >= 0 or r[0].find("k_prop") >= 0)
        agg = defaultdict(lambda: [0, 0.0])
        for r in seg:
            if r[1] >= t0:
                a = agg[_short(r[0])]
                a[0] += 1
                a[1] += (r[2] - r[1]) / 1e6
        total = sum(v[1] for v in agg.values())
        out += [f"{kind} call: {(seg[-1][2] - t0) / 1e6:.2f} ms from its first step kernel to "
                f"the end of k_build_write, {total:.2f} ms in kernels", "",
                "| kernel | launches | ms per call |", "|---|---:|---:|"]
        for k, v in sorted(agg.items(), key=lambda x: -x[1][1]):
            if v[1] >= 0.01:
                out.append(f"| `{k}` | {v[0]} | {v[1]:.3f} |")
        out.append("")
    return "\n".join(out)


def _counter_means(db: str) -> dict[str, tuple[int, float]]:
    """Per kernel (full name): dispatches and the mean counter value."""
    con = sqlite3.connect(db)
    acc = defaultdict(lambda: [set(), 0.0])
    for kn, did, v in con.execute("select kernel_name, dispatch_id, value from counters_collection"):
        a = acc[_short(kn)]
        a[0].add(did)
        a[1] += v
    return {k: (len(d), s / max(len(d), 1)) for k, (d, s) in acc.items()}


def _call_sums(db: str) -> dict[str, dict[str, tuple[int, float]]]:
    """Counter totals per kernel over the last plain and the last pruned CKF
    call of the run: {kind: {kernel: (dispatches, summed value)}}."""
    con = sqlite3.connect(db)
    rows = list(con.execute("select dispatch_id, kernel_name, sum(value) from counters_collection "
                            "group by dispatch_id order by dispatch_id"))
    last: dict[str, list] = {}
    seg: list = []
    for did, kn, v in rows:
        seg.append((_short(kn), v))
        if "k_build_write" in kn:
            names = {n for n, _ in seg}
            kind = ("pruned" if "k_find_list" in names else
                    "logged" if "k_find<1>" in names else "plain")
            last[kind] = seg
            seg = []
    out: dict[str, dict[str, tuple[int, float]]] = {}
    for kind, sg in last.items():
        acc: dict[str, list] = defaultdict(lambda: [0, 0.0])
        for n, v in sg:
            acc[n][0] += 1
            acc[n][1] += v
        out[kind] = {k: (c, t) for k, (c, t) in acc.items()}
    return out


def _call_counters(db: str) -> dict[str, dict[str, dict[str, list]]]:
    """{kind: {kernel: {counter: [weighted sum, duration sum]}}} over the last
    plain and the last pruned CKF call of one run (calls end at k_build_write)."""
    con = sqlite3.connect(db)
    rows = list(con.execute("select dispatch_id, kernel_name, counter_name, value, duration "
                            "from counters_collection order by dispatch_id"))
    disp: dict[int, tuple[str, float, dict[str, float]]] = {}
    order: list[int] = []
    for did, kn, cn, v, dur in rows:
        if did not in disp:
            disp[did] = (kn, float(dur), {})
            order.append(did)
        disp[did][2][cn] = float(v)
    last: dict[str, list[int]] = {}
    seg: list[int] = []
    for did in order:
        seg.append(did)
        if "k_build_write" in disp[did][0]:
            names = {_short(disp[d][0]) for d in seg}
            kind = ("pruned" if "k_find_list" in names else
                    "logged" if "k_find<1>" in names else "plain")
            last[kind] = seg
            seg = []
    out: dict[str, dict[str, dict[str, list]]] = {}
    for kind, sg in last.items():
        acc: dict[str, dict[str, list]] = defaultdict(lambda: defaultdict(lambda: [0.0, 0.0]))
        for d in sg:
            kn, dur, cs = disp[d]
            for cn, v in cs.items():
                a = acc[_short(kn)][cn]
                a[0] += v * dur
                a[1] += dur
        out[kind] = acc
    return out


def counters_table(dbs: list[str]) -> str:
    merged: dict[str, dict[str, dict[str, list]]] = defaultdict(dict)
    names: list[str] = []
    for db in dbs:
        for kind, ks in _call_counters(db).items():
            for k, cs in ks.items():
                merged[kind].setdefault(k, {}).update(cs)
                for cn in cs:
                    if cn not in names:
                        names.append(cn)
    out = ["| call | kernel | ms | " + " | ".join(names) + " |",
           "|---|---|---:|" + "---:|" * len(names)]
    for kind in ("pruned", "plain"):
        ks = merged.get(kind, {})
        rows = []
        for k, cs in ks.items():
            if not k.startswith("k_") or k.startswith(("k_fit", "k_amb", "k_sd", "k_mlp", "k_track")):
                continue  # the CKF's own kernels only
            ms = max(v[1] for v in cs.values()) / 1e6
            vals = [f"{cs[cn][0] / cs[cn][1]:.2f}" if cn in cs and cs[cn][1] > 0 else "" for cn in names]
            rows.append((ms, k, vals))
        for ms, k, vals in sorted(rows, key=lambda r: -r[0]):
            if ms >= 0.5:
                out.append(f"| {kind} | `{k}` | {ms:.1f} | " + " | ".join(vals) + " |")
    return "\n".join(out)


def traffic(fetch_db: str, write_db: str, copy_fetch_db: str, copy_write_db: str,
            copy_mib: float) -> dict:
    kib = copy_mib * 1024.0

    def copy_scale(db: str) -> float:
        m = _counter_means(db)
        # the copy kernel is the dispatch with the largest mean
        return kib / max(v for _, v in m.values())

    fs, ws = copy_scale(copy_fetch_db), copy_scale(copy_write_db)
    f, w = _counter_means(fetch_db), _counter_means(write_db)
    kernels = {}
    for k in sorted(set(f) & set(w), key=lambda k: -f[k][0]):
        kernels[k] = {"dispatches": f[k][0], "fetch_kib_raw": round(f[k][1], 1),
                      "write_kib_raw": round(w[k][1], 1),
                      "hbm_bytes_per_launch": int(1024 * (fs * f[k][1] + ws * w[k][1]))}
    cf, cw = _call_sums(fetch_db), _call_sums(write_db)
    calls = {}
    for kind in sorted(set(cf) & set(cw)):
        ks = set(cf[kind]) & set(cw[kind])
        calls[kind] = {k: {"dispatches": cf[kind][k][0],
                           "hbm_bytes": int(1024 * (fs * cf[kind][k][1] + ws * cw[kind][k][1]))}
                       for k in sorted(ks, key=lambda k: -cf[kind][k][1])}
    return {"fetch_scale": round(fs, 4), "write_scale": round(ws, 4), "kernels": kernels,
            "calls": calls}


def main(argv: list[str]) -> int:
    if len(argv) >= 2 and argv[0] == "kernels":
        print(kernels_table(argv[1]))
    elif len(argv) >= 2 and argv[0] == "calls":
        print(calls_table(argv[1]))
    elif len(argv) >= 2 and argv[0] == "counters":
        print(counters_table(argv[1:]))
    elif len(argv) >= 6 and argv[0] == "traffic":
        print(json.dumps(traffic(*argv[1:5], float(argv[5])), indent=1))
    else:
        print(__doc__)
        return 2
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
