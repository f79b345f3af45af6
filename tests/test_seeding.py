"""Seeding chain: spacepoint formation, triplet seeding and track-parameter
estimation.  The oracle (single precision, like the reference) against the
reference's own seeding cases (tests/cpu/test_seeding.cpp:35-203) and its
helix test of the parameter estimation (tests/cpu/test_track_params_estimation.cpp:33-137),
then the HIP kernels against the oracle: identical seeds (indices, order and
weights bit for bit), spacepoints equal to numpy, parameters within 1e-5 of
the double-precision restatement, and the chain into the CKF.

Reference: core/src/seeding/*.cpp|hpp, core/include/traccc/seeding/*.hpp.
"""
import math
import os

import numpy as np
import pytest

import oracle
from traccc_amd import _ffi as F
from traccc_amd import seeding as SD

HIP_LIB = os.path.join(F.LIB_DIR, "libtraccc_amd.so")

# tests/cpu/test_seeding.cpp:48-82 (16.62 GeV muon) and :130-164 (1.85 GeV muon)
CASE1 = [(36.6706, 10.6472, 104.131), (94.2191, 29.6699, 113.628), (149.805, 47.9518, 122.979),
         (218.514, 70.3049, 134.029), (275.359, 88.668, 143.378)]
CASE2 = [(36.301, 13.1197, 106.83), (93.9366, 33.7101, 120.978), (149.192, 52.0562, 134.678),
         (218.398, 73.1025, 151.979), (275.322, 89.0663, 166.229)]


def _reference_case_configs():
    """The test's configs: the grid config copies the defaults before
    deltaRMax and maxPtScattering change (test_seeding.cpp:38-45)."""
    finder = SD.seedfinder_config()
    grid = SD.spacepoint_grid_config(finder)
    finder.deltaRMax = 100.0
    finder.maxPtScattering = 0.5
    return finder, grid, SD.seedfilter_config()


def _spacepoints(det, ev):
    """Pixel spacepoints of an event in numpy float32, rounded as the device
    kernel: centre + loc0 u + loc1 v, one per 2D measurement in order."""
    ids = det.identifiers
    keep = np.nonzero(ev.dimensions == 2)[0]
    sidx = np.searchsorted(ids, ev.surface_link[keep])
    c = np.array([[det.surface(i).center[k] for k in range(3)] for i in range(det.n_surfaces)], np.float32)
    u = np.array([[det.surface(i).axis_u[k] for k in range(3)] for i in range(det.n_surfaces)], np.float32)
    v = np.array([[det.surface(i).axis_v[k] for k in range(3)] for i in range(det.n_surfaces)], np.float32)
    lp = ev.local_position.reshape(-1, 2)[keep].astype(np.float32)
    xyz = c[sidx] + lp[:, :1] * u[sidx] + lp[:, 1:2] * v[sidx]
    return xyz.astype(np.float32), keep.astype(np.uint32)


def _event(name, n, seed, **kw):
    from traccc_amd import simulation as S
    det, fld, gun = S.workload(name, seed=seed, n_particles=n)
    for k, val in kw.items():
        setattr(gun, k, val)
    return det, fld, S.generate_event(det, fld, gun)


def _c2_configs(b_tesla=2.0):
    finder = SD.seedfinder_config(bFieldInZ=b_tesla * SD.TESLA)
    return finder, SD.spacepoint_grid_config(finder), SD.seedfilter_config()


# ---- the oracle ------------------------------------------------------------------

@pytest.mark.parametrize("case", [CASE1, CASE2])
def test_oracle_reference_seeding_cases(case):
    finder, grid, filt = _reference_case_configs()
    b, m, t, w, _ = oracle.seeding(finder, grid, filt, np.array(case, np.float32))
    # "The number of seeds should be equal to one"
    assert len(b) == 1
    assert (b[0], m[0], t[0]) == (0, 1, 2)


@pytest.mark.parametrize("q", [-1.0, 1.0])
def test_oracle_track_params_on_exact_helix(q):
    """tests/cpu/test_track_params_estimation.cpp:35-137: three points of the
    helix of a 1 GeV x 1 GeV (transverse x longitudinal) track in 2 T at
    path lengths 50, 100, 150 mm give |p| within 2e-4 GeV and the charge;
    here also phi and theta of the direction at the bottom point."""
    cfg = SD.track_params_estimation_config()
    pts = SD.helix_points(q, (1.0, 0.0, 1.0), 2.0, [50.0, 100.0, 150.0])
    vec, cov = oracle.track_params(cfg, (0.0, 0.0, 2.0 * SD.TESLA), *pts)
    p = abs(1.0 / vec[4])
    assert abs(p - math.sqrt(2.0)) < 2e-4
    assert (vec[4] < 0) == (q < 0)
    # the direction at s = 50 mm: theta 45 deg, phi turned by the arc
    R = 1.0 / (2.0 * SD.TESLA)
    phi_true = -math.copysign(1.0, q) * (50.0 / math.sqrt(2.0)) / R
    assert abs(vec[3] - math.pi / 4) < 1e-9
    assert abs(vec[2] - phi_true) < 1e-9
    # the covariance diagonal (track_params_estimation.cpp:60-87)
    deg = math.pi / 180
    assert cov[0] == pytest.approx(1.0) and cov[2] == pytest.approx(deg * deg)
    var_q = 0.01 * math.sin(vec[3]) ** 2 + (0.1 * vec[4]) ** 2 + deg * deg * (vec[4] / math.tan(vec[3])) ** 2
    assert cov[4] == pytest.approx(var_q, rel=1e-6)
    assert cov[5] == pytest.approx(100.0)


def test_oracle_seeds_on_a_simulated_event():
    """Muons through the C2 barrel with exact hits (no smearing: the
    reference's pixel spacepoints carry zero variances, so the triplet
    scattering cut leaves no room for measurement error): seeds are pure
    (three hits of one particle) and nearly every central particle gets one."""
    det, _, ev = _event("C2", 300, 11, eta_min=-1.0, eta_max=1.0, noise_fraction=0.0, pt_a=1.0,
                        smear=(0.0, 0.0))
    xyz, midx = _spacepoints(det, ev)
    finder, grid, filt = _c2_configs()
    b, m, t, w, bins = oracle.seeding(finder, grid, filt, xyz)
    part = ev.meas_particle[midx]
    pure = (part[b] == part[m]) & (part[m] == part[t])
    assert len(b) > 200 and pure.mean() > 0.9
    seeded = np.unique(part[m][pure])
    assert len(seeded) / 300 > 0.9


def test_grid_config_rejections_match():
    finder, grid, filt = _c2_configs()
    grid.minPt = 0.01                              # minimum helix radius below rMax / 2
    with pytest.raises(ValueError, match="domain_error"):
        oracle.seeding(finder, grid, filt, np.zeros((1, 3), np.float32))
    if not os.path.exists(HIP_LIB):
        pytest.skip("HIP library not built")
    with pytest.raises(ValueError, match="helix radius"):
        SD.seeding_algorithm(finder, grid, filt)


# ---- the HIP kernels -----------------------------------------------------------------

@pytest.fixture(scope="module")
def gpu(gpu_available):
    return gpu_available


def _gpu_vs_oracle(finder, grid, filt, xyz, vr=None, vz=None):
    alg = SD.seeding_algorithm(finder, grid, filt)
    got = alg(xyz, vr, vz)
    b, m, t, w, _ = oracle.seeding(finder, grid, filt, xyz, vr, vz)
    assert len(got) == len(b)
    assert np.array_equal(got.bottom, b) and np.array_equal(got.middle, m) and np.array_equal(got.top, t)
    assert np.array_equal(got.weight.view(np.uint32), w.view(np.uint32))
    return got


@pytest.mark.gpu
@pytest.mark.parametrize("case", [CASE1, CASE2])
def test_gpu_reference_seeding_cases(gpu, case):
    finder, grid, filt = _reference_case_configs()
    got = _gpu_vs_oracle(finder, grid, filt, np.array(case, np.float32))
    assert got.as_tuples() == [(0, 1, 2)]


@pytest.mark.gpu
@pytest.mark.parametrize("name,n,seed", [("C2", 2000, 3), ("C4", 8000, 5), ("C2", 10000, 7)])
def test_gpu_seeding_matches_the_oracle(gpu, name, n, seed):
    det, _, ev = _event(name, n, seed)
    xyz, _ = _spacepoints(det, ev)
    finder, grid, filt = _c2_configs()
    got = _gpu_vs_oracle(finder, grid, filt, xyz)
    assert len(got) > n // 4


@pytest.mark.gpu
def test_gpu_seeding_with_variances_and_edge_inputs(gpu):
    det, _, ev = _event("C2", 3000, 13)
    xyz, _ = _spacepoints(det, ev)
    rng = np.random.default_rng(1)
    vr = rng.uniform(0.0, 0.05, len(xyz)).astype(np.float32)
    vz = rng.uniform(0.0, 0.5, len(xyz)).astype(np.float32)
    finder, grid, filt = _c2_configs()
    _gpu_vs_oracle(finder, grid, filt, xyz, vr, vz)
    # maxSeedsPerSpM 1 and a narrow phi neighbourhood
    f2 = SD.seedfinder_config(bFieldInZ=2.0 * SD.TESLA, maxSeedsPerSpM=1)
    f2.neighbor_scope[0] = 0
    _gpu_vs_oracle(f2, SD.spacepoint_grid_config(f2), filt, xyz)
    # no spacepoint, one spacepoint, every spacepoint outside the region
    _gpu_vs_oracle(finder, grid, filt, np.zeros((0, 3), np.float32))
    _gpu_vs_oracle(finder, grid, filt, xyz[:1])
    far = xyz.copy()
    far[:, 2] += 5000.0
    assert len(_gpu_vs_oracle(finder, grid, filt, far)) == 0


@pytest.mark.gpu
def test_gpu_spacepoint_formation(gpu):
    from traccc_amd import ckf
    det, fld, ev = _event("C2", 2000, 17)
    ddet = ckf.DeviceDetector(det, fld)
    dev = ckf.DeviceEvents(ddet, [ev, ev])
    xyz, midx = SD.spacepoint_formation(ddet, dev, event=1)
    exp_xyz, exp_idx = _spacepoints(det, ev)
    assert np.array_equal(midx, exp_idx)
    assert np.array_equal(xyz.view(np.uint32), exp_xyz.view(np.uint32))


@pytest.mark.gpu
def test_gpu_track_params(gpu):
    det, fld, ev = _event("C2", 2000, 19)
    xyz, midx = _spacepoints(det, ev)
    finder, grid, filt = _c2_configs()
    seeds = SD.seeding_algorithm(finder, grid, filt)(xyz)
    tp = SD.track_params_estimation()
    b = (0.0, 0.0, 2.0 * SD.TESLA)
    sl, vec, cov = tp(xyz, midx, seeds, ev, b)
    assert len(vec) == len(seeds) > 100
    lp = ev.local_position.reshape(-1, 2)
    for i in range(0, len(seeds), max(1, len(seeds) // 200)):
        bi, mi, ti = seeds.bottom[i], seeds.middle[i], seeds.top[i]
        ov, oc = oracle.track_params(tp.config, b, xyz[bi].astype(float), xyz[mi].astype(float),
                                     xyz[ti].astype(float), lp[midx[bi]].astype(float))
        assert sl[i] == ev.surface_link[midx[bi]]
        assert np.allclose(vec[i][:2], ov[:2], rtol=0, atol=0)
        assert abs(vec[i][2] - ov[2]) < 1e-5 and abs(vec[i][3] - ov[3]) < 1e-5
        # q/p: 1e-4 relative, plus the single-precision floor of a stiff
        # track's circle through three points ~100 mm apart (a 1e-5 mm
        # sagitta error is ~2e-5 / GeV in q/p at 2 T)
        assert abs(vec[i][4] - ov[4]) <= 1e-4 * abs(ov[4]) + 2e-5
        assert np.allclose(np.diag(cov[i]), oc, rtol=1e-4)
    # the reference's helix case on the device
    pts = SD.helix_points(-1.0, (1.0, 0.0, 1.0), 2.0, [50.0, 100.0, 150.0]).astype(np.float32)
    one = SD.SeedCollection(np.array([0], np.uint32), np.array([1], np.uint32), np.array([2], np.uint32),
                            np.zeros(1, np.float32))
    sl1, v1, _ = tp(pts, np.zeros(3, np.uint32), one, ev, (0.0, 0.0, 2.0 * SD.TESLA))
    assert abs(abs(1.0 / v1[0][4]) - math.sqrt(2.0)) < 2e-4 and v1[0][4] < 0


@pytest.mark.gpu
def test_gpu_seeds_drive_the_ckf(gpu):
    """Spacepoints -> seeds -> parameters -> the CKF: the reconstructed
    seeds find most particles, as the truth seeds do.  The spacepoints carry
    the hits' 50 um resolution as variances (the formation's own are zero,
    as the reference's, which leaves smeared stiff tracks unseeded)."""
    from traccc_amd import ckf
    from traccc_amd import performance as PERF
    from traccc_amd import simulation as S
    det, fld, ev = _event("C2", 1000, 23, eta_min=-1.0, eta_max=1.0, pt_a=1.0)
    xyz, midx = _spacepoints(det, ev)
    finder, grid, filt = _c2_configs()
    var = np.full(len(xyz), 0.05 ** 2, np.float32)
    seeds = SD.seeding_algorithm(finder, grid, filt)(xyz, var, var)
    sl, vec, cov = SD.track_params_estimation()(xyz, midx, seeds, ev, (0.0, 0.0, 2.0 * SD.TESLA))
    ev2 = S.with_seeds(ev, sl, vec, cov, ev.meas_particle[midx[seeds.middle]])
    cfg = S.finding_config_for("C2")
    res = ckf.combinatorial_kalman_filter_algorithm(cfg)(det, fld, ev2)
    perf = PERF.finding_performance(ev2, res.offsets, res.constituents)
    assert perf.efficiency > 0.9
