"""Track-finding performance: truth matching and efficiency.

Follows traccc's definitions (performance/src/efficiency/finding_performance_writer.cpp:136-305):
a truth particle is matched by a track whose majority particle it is, with
purity >= 0.5 and completeness >= 0.5; the denominator is the selected truth
particles (pT >= 0.5 GeV, |eta| <= 3, |z_vertex| <= 500 mm, r_vertex <= 200 mm,
>= 3 measurements; performance/include/traccc/utils/truth_matching_config.hpp:14-21,
track_matching_config.hpp:14-17).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class FindingPerformance:
    n_selected: int
    n_matched: int
    n_tracks: int
    n_fake: int
    n_duplicate: int

    @property
    def efficiency(self) -> float:
        return self.n_matched / self.n_selected if self.n_selected else 0.0

    @property
    def fake_rate(self) -> float:
        return self.n_fake / self.n_tracks if self.n_tracks else 0.0

    @property
    def duplicate_rate(self) -> float:
        good = self.n_tracks - self.n_fake
        return self.n_duplicate / good if good else 0.0


def selected_particles(event, pt_min=0.5, eta_max=3.0, z_max=500.0, r_max=200.0,
                       min_meas=3) -> np.ndarray:
    p = event.particle_momentum.astype(np.float64)
    pt = np.hypot(p[:, 0], p[:, 1])
    pz = p[:, 2]
    eta = np.arcsinh(np.divide(pz, pt, out=np.zeros_like(pz), where=pt > 0))
    v = event.particle_vertex.astype(np.float64)
    rv = np.hypot(v[:, 0], v[:, 1])
    return ((pt >= pt_min) & (np.abs(eta) <= eta_max) & (np.abs(v[:, 2]) <= z_max) &
            (rv <= r_max) & (event.particle_n_meas >= min_meas))


def track_majority(event, offsets: np.ndarray, constituents: np.ndarray):
    """Per track: majority particle (-1 if none), its count, and the track length."""
    n_tracks = len(offsets) - 1
    lengths = np.diff(offsets.astype(np.int64))
    if n_tracks == 0:
        z = np.zeros(0, dtype=np.int64)
        return z - 1, z, z
    track_of = np.repeat(np.arange(n_tracks, dtype=np.int64), lengths)
    part = event.meas_particle[constituents.astype(np.int64)].astype(np.int64)
    npart = len(event.particle_charge) + 1
    key = track_of * npart + (part + 1)
    uk, cnt = np.unique(key, return_counts=True)
    tr = uk // npart
    pa = uk % npart - 1
    # noise (pa == -1) never forms a majority particle
    cnt = np.where(pa >= 0, cnt, 0)
    order = np.lexsort((-cnt, tr))
    tr_s, pa_s, cnt_s = tr[order], pa[order], cnt[order]
    first = np.ones(len(tr_s), dtype=bool)
    first[1:] = tr_s[1:] != tr_s[:-1]
    maj = np.full(n_tracks, -1, dtype=np.int64)
    mcnt = np.zeros(n_tracks, dtype=np.int64)
    maj[tr_s[first]] = np.where(cnt_s[first] > 0, pa_s[first], -1)
    mcnt[tr_s[first]] = cnt_s[first]
    return maj, mcnt, lengths


def finding_performance(event, offsets, constituents, **sel) -> FindingPerformance:
    selected = selected_particles(event, **sel)
    maj, mcnt, lengths = track_majority(event, offsets, constituents)
    n_tracks = len(lengths)
    nmeas = event.particle_n_meas.astype(np.int64)
    ok = maj >= 0
    purity = np.where(lengths > 0, mcnt / np.maximum(lengths, 1), 0.0)
    completeness = np.zeros(n_tracks)
    completeness[ok] = mcnt[ok] / np.maximum(nmeas[maj[ok]], 1)
    matched = ok & (purity >= 0.5) & (completeness >= 0.5)
    matched_particles = np.unique(maj[matched])
    n_matched_sel = int(np.count_nonzero(selected[matched_particles])) if len(matched_particles) else 0
    n_dup = int(np.count_nonzero(matched)) - len(matched_particles)
    return FindingPerformance(n_selected=int(np.count_nonzero(selected)), n_matched=n_matched_sel,
                              n_tracks=n_tracks, n_fake=int(np.count_nonzero(~matched)),
                              n_duplicate=n_dup)
