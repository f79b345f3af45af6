"""Comparison helpers shared by the parity tests."""
from __future__ import annotations

from collections import Counter

import numpy as np


def track_tuples(offsets, constituents):
    off = np.asarray(offsets, dtype=np.int64)
    cons = np.asarray(constituents)
    return [tuple(int(v) for v in cons[off[i]:off[i + 1]]) for i in range(len(off) - 1)]


def compare_tracks(a_offsets, a_cons, b_offsets, b_cons):
    """Multiset match rate of constituent-index tuples:
    |A and B| / max(|A|, |B|)  (traccc tests/cuda/test_ckf_toy_detector.cpp:231-254
    compare host and device tracks as sets)."""
    A = Counter(track_tuples(a_offsets, a_cons))
    B = Counter(track_tuples(b_offsets, b_cons))
    na, nb = sum(A.values()), sum(B.values())
    if max(na, nb) == 0:
        return 1.0, 0, 0
    common = sum((A & B).values())
    return common / max(na, nb), na, nb


def rel_close(a, b, rtol, atol=0.0):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b) <= atol + rtol * np.maximum(np.abs(a), np.abs(b))


def matched_scalar_agreement(a_offsets, a_cons, a_vals, b_offsets, b_cons, b_vals, rtol, atol,
                             floor=0.0):
    """Fraction of tracks present in both lists whose scalar (chi2, pval ...)
    agrees within |a - b| <= atol + rtol * max(|a|, |b|, floor)."""
    amap = {}
    for t, v in zip(track_tuples(a_offsets, a_cons), a_vals):
        amap.setdefault(t, v)
    n = ok = 0
    worst = 0.0
    for t, v in zip(track_tuples(b_offsets, b_cons), b_vals):
        if t in amap:
            n += 1
            good = bool(abs(float(amap[t]) - float(v)) <=
                        atol + rtol * max(abs(float(amap[t])), abs(float(v)), floor))
            ok += good
            d = abs(float(amap[t]) - float(v)) / max(abs(float(v)), abs(float(amap[t])), 1e-30)
            worst = max(worst, d)
    return (ok / n if n else 1.0), n, worst


def run_finder(backend, cfg, det, fld, ev):
    """Run the CKF on one event: backend "oracle" (FP64 CPU oracle) or "gpu"
    (the HIP library).  Returns a dict of the track-container fields."""
    if backend == "oracle":
        import oracle
        t = oracle.run(cfg, det, fld, ev, 64)
        return dict(offsets=t.offsets, constituents=t.constituents, ndf=t.ndf, chi2=t.chi2,
                    pval=t.pval, nholes=t.nholes, params=t.params_vector)
    if backend == "oracle32":
        import oracle
        t = oracle.run(cfg, det, fld, ev, 32)
        return dict(offsets=t.offsets, constituents=t.constituents, ndf=t.ndf, chi2=t.chi2,
                    pval=t.pval, nholes=t.nholes, params=t.params_vector)
    from traccc_amd import ckf
    alg = ckf.combinatorial_kalman_filter_algorithm(cfg)
    tc = alg(det, fld, ev)
    return dict(offsets=tc.offsets, constituents=tc.constituents, ndf=tc.ndf, chi2=tc.chi2,
                pval=tc.pval, nholes=tc.nholes, params=tc.params_vector)
