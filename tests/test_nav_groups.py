"""The navigator's grouped layer search (CPU): the same next layer and entry
distance as the plain scan of every layer, bit for bit.

The device navigator picks the next layer by taking layer groups (up to four
neighbouring layers under one (r, z) box) in increasing lower bound and
stopping once the smallest bound left exceeds the best entry distance; the
scan evaluates every unvisited layer in index order.  Both live in
traccc_amd/csrc/core/propagator.hpp; this test compiles that header for the
host with g++ and runs both on the C1 telescope, the C2 barrel + disks and the
C5 detector (53 layers), from random points inside the detector, points on
the layers' envelopes with that layer visited (the states navigation
creates), and degenerate directions (parallel to z, perpendicular to z).
"""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

from traccc_amd import simulation as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("needs g++")
    out = tmp_path_factory.mktemp("nav") / "libnavcheck.so"
    cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-ffp-contract=off",
           "-Iinclude", "-Itraccc_amd/csrc", "tests/cpp/nav_groups_check.cpp", "-o", str(out)]
    b = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stderr[-3000:]
    lib = C.CDLL(str(out))
    fp, ip, up = (np.ctypeslib.ndpointer(np.float32, flags="C"),
                  np.ctypeslib.ndpointer(np.int32, flags="C"),
                  np.ctypeslib.ndpointer(np.uint64, flags="C"))
    lib.nav_groups_check.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, fp, fp, up, C.c_float,
                                     ip, fp, ip, fp]
    lib.nav_groups_check.restype = C.c_int
    lib.nav_groups_build.argtypes = [C.c_void_p, C.c_uint32, ip, fp,
                                     np.ctypeslib.ndpointer(np.uint32, flags="C")]
    lib.nav_groups_build.restype = C.c_uint32
    return lib


def _layers(wl):
    det, _, _ = S.workload(wl, n_particles=10)
    raw = np.ascontiguousarray(det.layers_raw)
    env = np.array([[*det.layer(i).env_r, *det.layer(i).env_z] for i in range(det.n_layers)])
    return det, raw, env


def _states(det, env, rng, n):
    nl = det.n_layers
    kinds = np.array([det.layer(i).kind for i in range(nl)])
    rmax = float(np.max(np.abs(env[:, :2]))) * 1.1 + 10
    zmax = float(np.max(np.abs(env[:, 2:]))) * 1.1 + 10
    pos, dirs, vis = [], [], []
    # random points and directions, random visited sets
    m = n // 2
    r = rmax * np.sqrt(rng.random(m))
    ph = rng.uniform(-np.pi, np.pi, m)
    pos.append(np.stack([r * np.cos(ph), r * np.sin(ph), rng.uniform(-zmax, zmax, m)], 1))
    v = rng.normal(size=(m, 3))
    dirs.append(v / np.linalg.norm(v, axis=1, keepdims=True))
    vis.append(np.array([sum(1 << k for k in range(nl) if rng.random() < 0.3) for _ in range(m)],
                        np.uint64))
    # points on a layer's envelope (that layer visited), outward-going
    m2 = n - m
    li = rng.integers(0, nl, m2)
    p2 = np.zeros((m2, 3))
    d2 = rng.normal(size=(m2, 3))
    for j, l in enumerate(li):
        e = env[l]
        if kinds[l] == 2:  # plane layer: slab along its normal
            nrm = np.array(det.layer(int(l)).normal, float)
            p2[j] = nrm * rng.uniform(e[0], e[1]) + rng.normal(size=3) * 20
            d2[j] = nrm + rng.normal(size=3) * 0.3
        else:
            rr = rng.uniform(e[0], e[1])
            a = rng.uniform(-np.pi, np.pi)
            p2[j] = [rr * np.cos(a), rr * np.sin(a), rng.uniform(e[2], e[3])]
            d2[j] = [np.cos(a), np.sin(a), 0] + rng.normal(size=3) * 0.6
    d2 /= np.linalg.norm(d2, axis=1, keepdims=True)
    pos.append(p2)
    dirs.append(d2)
    vis.append(np.array([(1 << int(l)) | (1 << int(rng.integers(0, nl))) for l in li], np.uint64))
    pos, dirs, vis = np.concatenate(pos), np.concatenate(dirs), np.concatenate(vis)
    # degenerate directions: along z, and in the transverse plane
    k = n // 20
    dirs[:k] = [0.0, 0.0, 1.0]
    dirs[k:2 * k] = [0.0, 0.0, -1.0]
    dirs[2 * k:3 * k, 2] = 0.0
    dirs[2 * k:3 * k] /= np.linalg.norm(dirs[2 * k:3 * k], axis=1, keepdims=True)
    return (np.ascontiguousarray(pos, np.float32), np.ascontiguousarray(dirs, np.float32),
            np.ascontiguousarray(vis))


@pytest.mark.parametrize("wl", ["C1", "C2", "C5"])
def test_grouped_layer_search_equals_scan(harness, wl):
    det, raw, env = _layers(wl)
    rng = np.random.default_rng(11)
    n = 200000
    pos, dirs, vis = _states(det, env, rng, n)
    margin = float(S.finding_config_for(wl).propagation.max_mask_tolerance)
    b0, b1 = np.zeros(n, np.int32), np.zeros(n, np.int32)
    s0, s1 = np.zeros(n, np.float32), np.zeros(n, np.float32)
    ng = harness.nav_groups_check(raw.ctypes.data, det.n_layers, n, pos, dirs, vis, margin,
                                  b0, s0, b1, s1)
    assert 1 <= ng <= 16
    assert np.array_equal(b0, b1), np.flatnonzero(b0 != b1)[:10]
    same = (s0.view(np.uint32) == s1.view(np.uint32))
    assert same.all(), np.flatnonzero(~same)[:10]
    # the search has work to do: most states find a layer
    assert np.mean(b0 >= 0) > 0.3


@pytest.mark.parametrize("wl", ["C1", "C5"])
def test_groups_cover_every_layer_once(harness, wl):
    det, raw, env = _layers(wl)
    member = np.zeros(16 * 4, np.int32)
    box = np.zeros(16 * 4, np.float32)
    bounded = np.zeros(16, np.uint32)
    ng = harness.nav_groups_build(raw.ctypes.data, det.n_layers, member, box, bounded)
    mem = member[:ng * 4].reshape(ng, 4)
    got = np.sort(mem[mem >= 0])
    assert np.array_equal(got, np.arange(det.n_layers))
    kinds = np.array([det.layer(i).kind for i in range(det.n_layers)])
    for g in range(ng):
        ms = mem[g][mem[g] >= 0]
        if bounded[g]:
            # the box contains every member's envelope
            assert np.all(kinds[ms] != 2)
            assert box[4 * g] <= env[ms, 0].min() and box[4 * g + 1] >= env[ms, 1].max()
            assert box[4 * g + 2] <= env[ms, 2].min() and box[4 * g + 3] >= env[ms, 3].max()
        else:
            assert np.all(kinds[ms] == 2)
    if wl == "C5":
        # a disk's four rings share a group
        assert ng == 14
