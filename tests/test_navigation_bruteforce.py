"""Brute-force pin of the oracle's navigation (CPU).

The oracle's propagation picks the next sensitive surface with its own
layer-window navigator (detray's caching navigator is not available to
compare with).  This test fixes that choice independently: the exact helix of
the start state in a uniform Bz = 2 T field is intersected with EVERY module
plane of the detector (sign changes of the plane distance along a fine path
grid, refined by bisection), and the first crossing past
min_step_length_for_next_surface that lies inside the module's exact mask is
the expected next surface -- the ckf_aborter rule of the reference
(core/include/traccc/finding/actors/ckf_aborter.hpp:37-68).  A navigator that
skips an earlier module, an overlap partner, or a module outside its search
window fails here.  Seeds are the truth-smeared seeds of the C2 detector
(4 barrel layers + 2x4 disks) and of the C5 detector (5 barrel layers + 2x6
disks, trapezoid disk modules), 200 each, followed over two surfaces.
"""
import math

import numpy as np
import pytest

import oracle
from traccc_amd import _ffi as F
from traccc_amd import simulation as S

UNIT_T = 0.000299792458   # GeV / (e T mm)


def surfaces(det):
    return np.frombuffer(det.surfaces_raw, dtype=np.dtype(F.Surface))


def free_start(sf, vec):
    """Global position and unit direction of a bound state on surface sf."""
    c = np.asarray(sf["center"], np.float64)
    u = np.asarray(sf["axis_u"], np.float64)
    v = np.asarray(sf["axis_v"], np.float64)
    pos = c + vec[0] * u + vec[1] * v
    phi, theta = float(vec[2]), float(vec[3])
    d = np.array([math.cos(phi) * math.sin(theta), math.sin(phi) * math.sin(theta),
                  math.cos(theta)])
    return pos, d


def helix(pos, d, qop, bz, s):
    """Exact helix in a uniform field (0, 0, bz) at path lengths s (array):
    dT/ds = (q/p) T x B, so the transverse direction turns by -k s."""
    k = qop * bz * UNIT_T
    c0 = complex(d[0], d[1])
    s = np.asarray(s, np.float64)
    if abs(k) < 1e-15:
        xy = complex(pos[0], pos[1]) + c0 * s
    else:
        xy = complex(pos[0], pos[1]) + c0 * (1 - np.exp(-1j * k * s)) / (1j * k)
    return np.stack([xy.real, xy.imag, pos[2] + d[2] * s], axis=-1)


def in_mask(kind, bound, l0, l1):
    if kind == 0:  # rectangle: half_u, half_v
        return abs(l0) <= bound[0] and abs(l1) <= bound[1]
    hh = bound[2]  # trapezoid: half_u at v = -hh, at v = +hh, hh
    if abs(l1) > hh:
        return False
    return abs(l0) <= bound[0] + (bound[1] - bound[0]) * (l1 + hh) / (2 * hh)


def first_crossing(sfs, start, pos, d, qop, bz, s_min, s_max=4000.0, ds=0.5):
    """Index of the first sensitive module whose plane the exact helix crosses
    inside its mask at a path length > s_min (None if none within s_max)."""
    C = np.stack(sfs["center"]).astype(np.float64)
    N = np.stack(sfs["normal"]).astype(np.float64)
    U = np.stack(sfs["axis_u"]).astype(np.float64)
    V = np.stack(sfs["axis_v"]).astype(np.float64)
    sens = sfs["sensitive"] != 0
    off = np.einsum("ij,ij->i", C, N)
    chunk = 400
    s0 = 0.0
    while s0 < s_max:
        grid = s0 + ds * np.arange(chunk + 1)
        P = helix(pos, d, qop, bz, grid)                 # [k, 3]
        G = P @ N.T - off                                # plane distances [k, n_sf]
        k_idx, m_idx = np.nonzero((G[:-1] * G[1:] <= 0) & sens[None, :])
        best = None
        if len(k_idx):
            # bisection on the exact helix, all crossings of the chunk at once
            a, b = grid[k_idx].copy(), grid[k_idx + 1].copy()
            ga = G[k_idx, m_idx].copy()
            for _ in range(60):
                mid = 0.5 * (a + b)
                gm = np.einsum("ij,ij->i", helix(pos, d, qop, bz, mid), N[m_idx]) - off[m_idx]
                same = (ga <= 0) == (gm <= 0)
                a = np.where(same, mid, a)
                ga = np.where(same, gm, ga)
                b = np.where(same, b, mid)
            s_star = 0.5 * (a + b)
            rel = helix(pos, d, qop, bz, s_star) - C[m_idx]
            l0 = np.einsum("ij,ij->i", rel, U[m_idx])
            l1 = np.einsum("ij,ij->i", rel, V[m_idx])
            for j in np.argsort(s_star):
                m = int(m_idx[j])
                if s_star[j] <= s_min:
                    continue
                if in_mask(int(sfs["bound_kind"][m]), sfs["bound"][m], l0[j], l1[j]):
                    best = (float(s_star[j]), m)
                    break
        if best is not None:
            return best[1], best[0]
        s0 += chunk * ds
    return None, None


@pytest.mark.parametrize("preset,workload", [(S.DET_BARREL4_DISK4, "C2"),
                                             (S.DET_BARREL5_DISK6, "C5")])
def test_next_surface_is_the_first_exact_helix_crossing(preset, workload):
    det = S.build_detector(preset, 0.0)     # no material: the trajectory is the helix
    fld = S.constant_field(2.0)
    # the workload's particles and truth-smeared seeds (pT >= 0.5 GeV so the
    # min_pT aborter of the propagation never ends a track here)
    _, _, gun = S.workload(workload, n_particles=400, seed=97)
    gun.noise_fraction = 0.0
    gun.dead_fraction = 0.0
    if workload == "C5":
        gun.pt_distribution, gun.pt_a, gun.pt_b = 1, 0.5, 20.0
    ev = S.generate_event(det, fld, gun)
    cfg = S.finding_config_for(workload)
    cfg.propagation.rk_error_tol = 1e-7
    cfg.max_step_counts_for_next_surface = 10000
    sfs = surfaces(det)
    start = np.searchsorted(det.identifiers, ev.seed_surface_link)
    n = min(len(start), 200)
    assert n >= 200
    agree = found = total = 0
    bad = []
    for i in range(n):
        sf0 = int(start[i])
        vec = ev.seed_vector[i].astype(np.float64)
        # two hops: from the seed surface, then from the surface reached
        for hop in range(2):
            st, end, ov, oc, path, nst = oracle.propagate(cfg, det, fld, sf0, vec,
                                                          np.eye(6) * 1e-6, False)
            pos, dd = free_start(sfs[sf0], vec)
            exp, s_exp = first_crossing(sfs, sf0, pos, dd, float(vec[4]), 2.0,
                                        cfg.min_step_length_for_next_surface)
            total += 1
            if exp is None:
                ok = st != 0          # nothing ahead: the oracle must not find a surface
            else:
                found += 1
                ok = st == 0 and int(end) == int(exp) and abs(path - s_exp) < 1e-3
            agree += ok
            if not ok:
                bad.append((i, hop, sf0, st, int(end), path, exp, s_exp))
            if st != 0 or not ok:
                break
            sf0, vec = int(end), np.asarray(ov, np.float64)
    print(f"\n{workload} detector: {agree}/{total} propagations agree with the brute-force "
          f"crossing ({found} with a next surface); disagreements: {bad[:5]}")
    assert agree == total, bad[:5]
    assert found >= 0.8 * total and total >= 1.5 * n
