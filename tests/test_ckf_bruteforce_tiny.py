"""Tiny hand-built events whose every CKF branch path is enumerated in numpy.

The pattern of the reference's telescope tests
(tests/cpu/test_ckf_combinatorics_telescope.cpp:119-206), with more
candidates than branches, so the best-N selection and its tie rule decide
the outcome: three straight tracks 60-120 um apart plus noise hits near them,
on four telescope planes (normal along x, 20 mm apart, 0.5 mm of silicon
each, B = 0), max_num_branches_per_surface = 2.

The enumeration below is written from the textbook: straight-line transport
between parallel planes with its exact Jacobian, Highland scattering noise
(PDG eq. 34.15) on the angle variances at each plane, the Kalman chi2 and
Joseph-form update (Fruehwirth 1987), and traccc's host CKF rules
(details/combinatorial_kalman_filter.hpp): per input keep the N lowest-chi2
candidates, ties to the lower measurement index; no candidate -> one hole
link carrying the state; a link becomes a tip when its holes exceed
max_num_skipping_per_cand or max_num_consecutive_skipped, or when the
propagation fails (here: after the last plane) from step min-1 on; tracks
keep min <= step + 1 - holes <= max.  Time is left out: it is not measured,
and no other parameter's update depends on it.
Checked: the oracle (FP64) finds exactly the enumerated track set with the
same chi2 (1e-9 relative), ndf and hole counts; the HIP library (gpu) the
same track set, chi2 within 1e-3.
"""
import math

import numpy as np
import pytest

from traccc_amd import simulation as S

X0_PLANE = 0.5 / 93.7     # 0.5 mm of silicon per plane
PLANES_X = np.array([20.0, 40.0, 60.0, 80.0])
SIGMA = 0.05              # 50 um per local coordinate
MUON_MASS = 0.1056583755


def _cfg():
    cfg = S.default_config()
    cfg.ptc_pdg, cfg.ptc_mass, cfg.ptc_charge = 13, MUON_MASS, -1.0
    cfg.run_smoother = 0
    cfg.max_num_branches_per_surface = 2
    cfg.max_num_branches_per_seed = 1000
    cfg.duplicate_removal_minimum_length = 100
    cfg.chi2_max = 30.0
    cfg.min_p = 0.0
    cfg.min_pT = 0.0
    return cfg


def _event(seed, sparse=False):
    """Three tracks along +x, loc0 = y, loc1 = z on planes x = 20..80 mm,
    smeared hits plus two noise hits per plane near the tracks; one seed per
    track on the first plane (its first hit, angles smeared by 1 mrad, q/p
    by 10 %).  sparse: the tracks 1 mm apart, track 1 without its hit on the
    third plane and track 2 without its hits on the last two (holes, a
    second consecutive hole, tips before the last plane)."""
    rng = np.random.default_rng(seed)
    n_trk = 3
    gap = 1.0 if sparse else 0.06
    y0 = rng.uniform(-0.06, 0.06) + gap * np.array([0.0, 1.0, 2.0])
    z0 = rng.uniform(-0.06, 0.06) + gap * np.array([0.0, -0.8, 1.1])
    phi = rng.normal(0.0, 0.002, n_trk)
    theta = math.pi / 2 + rng.normal(0.0, 0.002, n_trk)
    qop = -1.0 / 10.0
    pos, part, surf = [], [], []
    for k, x in enumerate(PLANES_X):
        dx = x - PLANES_X[0]
        hits = []
        for t in range(n_trk):
            y = y0[t] + dx * math.tan(phi[t])
            z = z0[t] + dx / math.tan(theta[t]) / math.cos(phi[t])
            hit = (y + rng.normal(0, SIGMA), z + rng.normal(0, SIGMA), t)
            if sparse and ((t == 1 and k == 2) or (t == 2 and k >= 2)):
                continue
            hits.append(hit)
        for _ in range(2):   # noise near the tracks
            if sparse:
                hits.append((rng.uniform(-0.1, 2.2), rng.uniform(-1.0, 1.2), -1))
            else:
                hits.append((rng.uniform(-0.1, 0.25), rng.uniform(-0.12, 0.15), -1))
        for h in hits:
            pos.append(h[:2])
            part.append(h[2])
            surf.append(k)
    nm = len(pos)
    det = S.build_detector(S.DET_TELESCOPE, X0_PLANE, telescope_planes=4,
                           telescope_spacing=20.0)
    ids = det.identifiers
    lp = np.asarray(pos, np.float32)
    vec = np.zeros((n_trk, 6), np.float32)
    for t in range(n_trk):
        first = [i for i in range(nm) if surf[i] == 0 and part[i] == t][0]
        vec[t] = [lp[first, 0], lp[first, 1], phi[t] + rng.normal(0, 1e-3),
                  theta[t] + rng.normal(0, 1e-3), qop * (1 + rng.normal(0, 0.1)), 0.0]
    cov = np.zeros((n_trk, 6, 6), np.float32)
    for t in range(n_trk):
        cov[t] = np.diag([SIGMA ** 2, SIGMA ** 2, 1e-6, 1e-6, (0.1 * qop) ** 2, 1.0])
    ev = S.Event(
        local_position=np.ascontiguousarray(lp),
        local_variance=np.full((nm, 2), SIGMA ** 2, np.float32),
        dimensions=np.full(nm, 2, np.uint32), time=np.zeros(nm, np.float32),
        diameter=np.zeros(nm, np.float32), identifier=np.arange(nm, dtype=np.uint32),
        surface_link=np.ascontiguousarray(ids[np.asarray(surf)]),
        subspace=np.tile(np.array([0, 1], np.uint8), (nm, 1)),
        cluster_index=np.arange(nm, dtype=np.uint32),
        meas_particle=np.asarray(part, np.int32),
        seed_surface_link=np.full(n_trk, ids[0], np.uint64), seed_vector=vec,
        seed_covariance=cov, seed_particle=np.arange(n_trk, dtype=np.uint32),
        particle_vertex=np.zeros((n_trk, 3), np.float32),
        particle_momentum=np.zeros((n_trk, 3), np.float32),
        particle_charge=np.full(n_trk, -1.0, np.float32),
        particle_n_meas=np.full(n_trk, 4, np.uint32))
    return det, ev, np.asarray(surf)


# ---------------------------------------------------------------------------
# the enumeration (5 parameters: loc0, loc1, phi, theta, q/p)
# ---------------------------------------------------------------------------
def _highland(x_over_x0, qop):
    p = 1.0 / abs(qop)
    beta = p / math.sqrt(p * p + MUON_MASS ** 2)
    return 0.0136 / (beta * p) * math.sqrt(x_over_x0) * (
        1 + 0.038 * math.log(x_over_x0 / beta ** 2))


def _material(x, P):
    """Scattering noise of one plane (normal along x) at the state's angles."""
    phi, th = x[2], x[3]
    cos_inc = abs(math.cos(phi) * math.sin(th))
    t0 = _highland(X0_PLANE / cos_inc, x[4])
    P = P.copy()
    P[2, 2] += t0 ** 2 / math.sin(th) ** 2
    P[3, 3] += t0 ** 2
    return P


def _transport(x, P, dx):
    """Straight line from one plane to the next (dx along the normal)."""
    y, z, phi, th, qop = x
    xn = np.array([y + dx * math.tan(phi), z + dx / math.tan(th) / math.cos(phi), phi, th, qop])
    J = np.eye(5)
    J[0, 2] = dx / math.cos(phi) ** 2
    J[1, 2] = dx * math.sin(phi) / (math.tan(th) * math.cos(phi) ** 2)
    J[1, 3] = -dx / (math.sin(th) ** 2 * math.cos(phi))
    return xn, J @ P @ J.T


def _update(x, P, m, V):
    H = np.zeros((2, 5))
    H[0, 0] = H[1, 1] = 1.0
    r = m - H @ x
    R = H @ P @ H.T + V
    Ri = np.linalg.inv(R)
    chi2 = float(r @ Ri @ r)
    K = P @ H.T @ Ri
    IKH = np.eye(5) - K @ H
    return chi2, x + K @ r, IKH @ P @ IKH.T + K @ V @ K.T


def enumerate_ckf(cfg, ev, surf):
    N = cfg.max_num_branches_per_surface
    lp = ev.local_position.astype(np.float64)
    V = np.diag([SIGMA ** 2, SIGMA ** 2]).astype(np.float32).astype(np.float64)
    n_planes = len(PLANES_X)
    tracks = []

    def expand(step, plane, x, P, hist, skip, consec, chi2s):
        # find: the candidates of this plane, best N by (chi2, index)
        cands = []
        for m in np.nonzero(surf == plane)[0]:
            c2, xf, Pf = _update(x, P, lp[m], V)
            if 0.0 <= c2 <= cfg.chi2_max:
                cands.append((c2, int(m), xf, Pf))
        cands.sort(key=lambda c: (c[0], c[1]))
        links = [(c2, m, xf, Pf, skip, 0) for c2, m, xf, Pf in cands[:N]]
        if not links:
            links = [(None, None, x, P, skip + 1, consec + 1)]
        for c2, m, xf, Pf, sk, co in links:
            h = hist + [(m, c2)]
            if sk > cfg.max_num_skipping_per_cand or co > cfg.max_num_consecutive_skipped:
                tracks.append((step, h, sk))
                continue
            if plane + 1 >= n_planes:   # no surface ahead: the propagation fails
                if step >= cfg.min_track_candidates_per_track - 1:
                    tracks.append((step, h, sk))
                continue
            Pm = _material(xf, Pf) if step == 0 else Pf   # start-surface material at step 0
            xn, Pn = _transport(xf, Pm, PLANES_X[plane + 1] - PLANES_X[plane])
            expand(step + 1, plane + 1, xn, _material(xn, Pn), h, sk, co, chi2s)

    for s in range(ev.n_seeds):
        x = ev.seed_vector[s, :5].astype(np.float64)
        P = ev.seed_covariance[s, :5, :5].astype(np.float64)
        expand(0, 0, x, P, [], 0, 0, 0.0)
    out = []
    for step, hist, sk in tracks:
        n_cands = step + 1 - sk
        if not (cfg.min_track_candidates_per_track <= n_cands <= cfg.max_track_candidates_per_track):
            continue
        meas = tuple(m for m, _ in hist if m is not None)
        chi2 = sum(c for m, c in hist if m is not None)
        first = next(i for i, (m, _) in enumerate(hist) if m is not None)
        out.append((meas, chi2, 2 * len(meas) - 5, first))   # holes before the first hit
    return out


@pytest.mark.parametrize("backend", ["oracle", pytest.param("gpu", marks=pytest.mark.gpu)])
@pytest.mark.parametrize("seed,sparse", [(1, False), (2, False), (3, False), (4, False),
                                         (5, True), (6, True), (7, True)])
def test_tiny_events_match_the_enumeration(backend, seed, sparse, request):
    if backend == "gpu":
        request.getfixturevalue("gpu_available")
    from helpers import run_finder
    cfg = _cfg()
    det, ev, surf = _event(seed, sparse)
    fld = S.constant_field(0.0)
    expect = enumerate_ckf(cfg, ev, surf)
    # more candidates than branches somewhere: the best-N cut decided something
    r = run_finder(backend, cfg, det, fld, ev)
    got = {}
    for i in range(len(r["ndf"])):
        t = tuple(int(v) for v in r["constituents"][r["offsets"][i]:r["offsets"][i + 1]])
        got.setdefault(t, []).append((float(r["chi2"][i]), float(r["ndf"][i]), int(r["nholes"][i])))
    exp = {}
    for meas, chi2, ndf, holes in expect:
        exp.setdefault(meas, []).append((chi2, ndf, holes))
    assert sorted((k, len(v)) for k, v in got.items()) == sorted((k, len(v)) for k, v in exp.items())
    tol = 1e-9 if backend == "oracle" else 1e-3
    for k, vals in exp.items():
        for (c_e, n_e, h_e), (c_g, n_g, h_g) in zip(sorted(vals), sorted(got[k])):
            assert abs(c_g - c_e) <= tol * max(1.0, abs(c_e)), (k, c_g, c_e)
            assert n_g == n_e and h_g == h_e
    if sparse:
        # holes: a track shorter than the planes, and a seed whose track ends
        # below min_track_candidates_per_track (dropped)
        assert any(len(m) < len(PLANES_X) for m, *_ in expect)
        assert len({m[0] for m, *_ in expect}) < ev.n_seeds
    else:
        assert len(expect) > ev.n_seeds   # branching happened


def test_enumeration_has_more_candidates_than_branches():
    """The events are built so that the best-N cut matters: some input sees
    more passing candidates than max_num_branches_per_surface."""
    cfg = _cfg()
    det, ev, surf = _event(1)
    lp = ev.local_position.astype(np.float64)
    V = np.diag([SIGMA ** 2, SIGMA ** 2])
    x = ev.seed_vector[0, :5].astype(np.float64)
    P = ev.seed_covariance[0, :5, :5].astype(np.float64)
    n_pass = sum(1 for m in np.nonzero(surf == 0)[0]
                 if _update(x, P, lp[m], V)[0] <= cfg.chi2_max)
    assert n_pass > cfg.max_num_branches_per_surface
