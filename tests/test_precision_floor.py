"""FP32 arithmetic floor of the CKF's chi2 against FP64.

The FP32 instantiation of the oracle (the timed CPU baseline, the precision of
traccc's default TRACCC_CUSTOM_SCALARTYPE=float build) against its FP64
instantiation: index sets are identical, chi2 sums agree within 1e-4 relative
(scale floored at 1) for the bulk and within 1e-3 for all tracks.  The GPU
parity tests use the same bands.
"""
import numpy as np

import oracle
from helpers import compare_tracks
from traccc_amd import simulation as S


def test_fp32_precision_floor():
    det, fld, _ = S.workload("C1")
    cfg = S.finding_config_for("C1")
    fr4, fr3 = [], []
    for e in range(5):
        ev = S.generate_event(det, fld, S.workload("C1", seed=42 + e)[2])
        a = oracle.run(cfg, det, fld, ev, 32)
        b = oracle.run(cfg, det, fld, ev, 64)
        match, _, _ = compare_tracks(a.offsets, a.constituents, b.offsets, b.constituents)
        assert match == 1.0
        assert np.array_equal(a.constituents, b.constituents)
        d = np.abs(a.chi2 - b.chi2) / np.maximum(np.maximum(np.abs(a.chi2), np.abs(b.chi2)), 1.0)
        fr4.append(np.mean(d <= 1e-4))
        fr3.append(np.mean(d <= 1e-3))
    assert min(fr4) >= 0.75
    assert min(fr3) == 1.0
    assert max(fr4) < 1.0  # the floor is real: FP32 does not reach 1e-4 everywhere
