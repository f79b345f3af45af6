"""Measurement ingest in traccc's CSV format (include/traccc_amd/io.h).

  * the reference's own mock measurement file (tests/io/mock_data, kept here
    as tests/golden/mock_event000000000-measurements.csv): three pixel
    measurements on one surface, local_key the raw byte 0x06, 50 um
    variances; sorting orders them by loc0 and returns the old -> new map;
  * the local_key bit rule (2 = loc0, 4 = loc1, 6 = both) on hand-written
    rows; malformed files are rejected with the line number;
  * write -> read round trip of a generated event is exact, and the CKF
    finds the same tracks from the re-read measurements (the oracle on the
    CPU; the HIP CKF against the oracle on the GPU).
"""
import os

import numpy as np
import pytest

from traccc_amd import io as IO
from traccc_amd import simulation as S

HERE = os.path.dirname(os.path.abspath(__file__))
MOCK = os.path.join(HERE, "golden", "mock_event000000000-measurements.csv")


def test_reference_mock_measurements():
    m, idx = IO.read_measurements_csv(MOCK, do_sort=False)
    assert m.n_meas == 3
    assert np.all(m.surface_link == np.uint64(1224979236083738112))
    assert np.all(m.dimensions == 2)
    assert np.array_equal(m.subspace, np.array([[0, 1]] * 3, np.uint8))
    assert np.allclose(m.local_variance, 0.0025, rtol=1e-6)
    assert np.allclose(m.local_position[:, 0], [4.2657785415649414, 2.4493119716644287,
                                                3.1442358493804932])
    assert np.array_equal(idx, [0, 1, 2])
    s, idx = IO.read_measurements_csv(MOCK, do_sort=True)
    # sorted by loc0 on the one surface: rows 1, 2, 0
    assert np.array_equal(idx, [2, 0, 1])
    assert np.array_equal(s.local_position, m.local_position[[1, 2, 0]])
    assert np.array_equal(s.identifier, [1, 2, 0])


def test_local_key_rule_and_errors(tmp_path):
    head = ("measurement_id,geometry_id,local_key,local0,local1,phi,theta,time,"
            "var_local0,var_local1,var_phi,var_theta,var_time\n")
    rows = [(b"\x02", 1.0, 9.0), (b"\x04", 2.0, 8.0), (b"\x06", 3.0, 7.0)]
    p = tmp_path / "m.csv"
    with open(p, "wb") as f:
        f.write(head.encode())
        for k, (key, a, b) in enumerate(rows):
            f.write(f"{k},5,".encode() + key + f",{a},{b},0,0,0,0.01,0.04,0,0,0\n".encode())
    m, _ = IO.read_measurements_csv(str(p), do_sort=False)
    assert list(m.dimensions) == [1, 1, 2]
    assert m.subspace[0, 0] == 0 and m.subspace[1, 0] == 1
    assert np.array_equal(m.subspace[2], [0, 1])
    assert m.local_position[0, 0] == 1.0 and m.local_position[1, 1] == 8.0
    assert m.local_variance[1, 1] == np.float32(0.04)
    # a missing column, a bad number, a key measuring nothing
    (tmp_path / "a.csv").write_text(head.replace("local1,", "") + "0,5,6,1,0,0,0,0,0,0,0,0\n")
    with pytest.raises(ValueError, match="missing column local1"):
        IO.read_measurements_csv(str(tmp_path / "a.csv"))
    (tmp_path / "b.csv").write_bytes(head.encode() + b"0,5,\x06,x1,2,0,0,0,1,1,0,0,0\n")
    with pytest.raises(ValueError, match=":2: not a number"):
        IO.read_measurements_csv(str(tmp_path / "b.csv"))
    (tmp_path / "c.csv").write_bytes(head.encode() + b"0,5,\x01,1,2,0,0,0,1,1,0,0,0\n")
    with pytest.raises(ValueError, match="neither loc0 nor loc1"):
        IO.read_measurements_csv(str(tmp_path / "c.csv"))


def test_round_trip_and_finding(tmp_path):
    import oracle
    det, fld, gun = S.workload("C2", n_particles=200, seed=5)
    ev = S.generate_event(det, fld, gun)
    p = tmp_path / "ev.csv"
    IO.write_measurements_csv(str(p), ev)
    m, idx = IO.read_measurements_csv(str(p), do_sort=True)
    # the generator's events are already in traccc order: identity map
    assert np.array_equal(idx, np.arange(ev.n_meas))
    assert np.array_equal(m.local_position, ev.local_position.reshape(-1, 2))
    assert np.array_equal(m.local_variance, ev.local_variance.reshape(-1, 2))
    assert np.array_equal(m.surface_link, ev.surface_link)
    assert np.array_equal(m.dimensions, ev.dimensions)
    # shuffled rows sort back to the same collection
    rng = np.random.default_rng(0)
    perm = rng.permutation(ev.n_meas)
    sh = IO.Measurements(m.local_position[perm].copy(), m.local_variance[perm].copy(),
                         m.dimensions[perm].copy(), m.time[perm].copy(),
                         m.identifier[perm].copy(), m.surface_link[perm].copy(),
                         m.subspace[perm].copy())
    IO.write_measurements_csv(str(tmp_path / "sh.csv"), sh)
    back, idx2 = IO.read_measurements_csv(str(tmp_path / "sh.csv"), do_sort=True)
    assert np.array_equal(back.local_position, m.local_position)
    # shuffled row r held original row perm[r]: it lands back there
    assert np.array_equal(idx2, perm)
    # the CKF (oracle) finds the same tracks from the re-read measurements
    ev2 = S.Event(**{**ev.__dict__, "local_position": back.local_position.reshape(-1).copy(),
                     "local_variance": back.local_variance.reshape(-1).copy()})
    cfg = S.finding_config_for("C2")
    a = oracle.run(cfg, det, fld, ev)
    b = oracle.run(cfg, det, fld, ev2)
    assert np.array_equal(a.offsets, b.offsets)
    assert np.array_equal(a.constituents, b.constituents)


@pytest.mark.gpu
def test_csv_measurements_through_the_hip_ckf(tmp_path, gpu_available):
    """A C2 event written to CSV with its rows shuffled, read back with the
    traccc ordering restored, runs through the HIP CKF: the same tracks as the
    FP64 oracle finds on the re-read event, and as the HIP CKF finds on the
    generator's own event (constituents mapped through the reader's
    old -> new row map)."""
    import oracle
    from traccc_amd import ckf
    det, fld, gun = S.workload("C2", n_particles=2000, seed=15)
    ev = S.generate_event(det, fld, gun)
    rng = np.random.default_rng(3)
    perm = rng.permutation(ev.n_meas)
    lp, lv = ev.local_position.reshape(-1, 2), ev.local_variance.reshape(-1, 2)
    sh = IO.Measurements(lp[perm].copy(), lv[perm].copy(), ev.dimensions[perm].copy(),
                         ev.time[perm].copy(), ev.identifier[perm].copy(),
                         ev.surface_link[perm].copy(), ev.subspace.reshape(-1, 2)[perm].copy())
    p = tmp_path / "c2.csv"
    IO.write_measurements_csv(str(p), sh)
    m, idx = IO.read_measurements_csv(str(p), do_sort=True)
    assert m.n_meas == ev.n_meas
    ev2 = S.Event(**{**ev.__dict__,
                     "local_position": np.ascontiguousarray(m.local_position, np.float32),
                     "local_variance": np.ascontiguousarray(m.local_variance, np.float32),
                     "dimensions": np.ascontiguousarray(m.dimensions, np.uint32),
                     "time": np.ascontiguousarray(m.time, np.float32),
                     "identifier": np.ascontiguousarray(m.identifier, np.uint32),
                     "surface_link": np.ascontiguousarray(m.surface_link, np.uint64),
                     "subspace": np.ascontiguousarray(m.subspace, np.uint8)})
    cfg = S.finding_config_for("C2")
    g = ckf.combinatorial_kalman_filter_algorithm(cfg)(det, fld, ev2)
    o = oracle.run(cfg, det, fld, ev2, 64)
    assert len(g) > 1500
    assert np.array_equal(g.offsets, o.offsets)
    assert np.array_equal(g.constituents, o.constituents)
    # the generator's event is already in traccc order: sorted row k of the
    # re-read event is original row perm[old] where idx[old] = k
    orig_of_sorted = np.empty(ev.n_meas, np.int64)
    orig_of_sorted[idx] = perm
    g0 = ckf.combinatorial_kalman_filter_algorithm(cfg)(det, fld, ev)
    assert np.array_equal(g0.offsets, g.offsets)
    assert np.array_equal(g0.constituents, orig_of_sorted[g.constituents])
