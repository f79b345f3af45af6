"""Committed golden fixtures (tests/golden/, written by make_golden.py from
the FP64 oracle): the oracle reproduces them (CPU), and the device kernels
reproduce them (GPU) -- the gain update and propagation within the unit
tolerances of the parity tests, the tiny events' track candidates exactly,
the fit and the scorer within the FP32 / BF16 bounds stated below.
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle
from oracle import mlp_oracle as M
from traccc_amd import _ffi as F
from traccc_amd import fitting as FT
from traccc_amd import simulation as S

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def _check_detector(det, stored):
    c = np.array([list(det.surface(i).center) for i in range(det.n_surfaces)], np.float64)
    got = np.array([det.n_surfaces, float(np.sum(c * np.arange(1, len(c) + 1)[:, None]))])
    assert got[0] == stored[0] and abs(got[1] - stored[1]) <= 1e-9 * abs(stored[1]), \
        "geometry changed: rerun tests/golden/make_golden.py"


def _event(g, tag):
    fields = {f: g[f"{tag}_ev_{f}"] for f in
              ("local_position", "local_variance", "dimensions", "time", "diameter",
               "identifier", "surface_link", "subspace", "cluster_index", "meas_particle",
               "seed_surface_link", "seed_vector", "seed_covariance", "seed_particle",
               "particle_vertex", "particle_momentum", "particle_charge", "particle_n_meas")}
    return S.Event(**{k: np.ascontiguousarray(v) for k, v in fields.items()})


def _tiny(tag):
    g = _load("ckf_tiny_events.npz")
    wl = str(g[f"{tag}_workload"])
    det, fld, _ = S.workload(wl, n_particles=1)
    _check_detector(det, g[f"{tag}_detector"])
    cfg = S.finding_config_for(wl)
    cfg.max_num_branches_per_surface = int(g[f"{tag}_branches"])
    return g, det, fld, cfg, _event(g, tag)


# ---------------------------------------------------------------------------
# oracle
# ---------------------------------------------------------------------------
def test_oracle_kf_update_golden():
    g = _load("kf_update_cases.npz")
    for i in range(len(g["status"])):
        st, c2, v, c = oracle.kf_update(g["vec"][i].astype(np.float64), g["cov"][i].astype(np.float64),
                                        g["meas"][i], g["var"][i], int(g["dims"][i]),
                                        tuple(g["subspace"][i]))
        assert st == g["status"][i]
        assert c2 == pytest.approx(g["chi2"][i], rel=1e-12, abs=1e-14)
        assert np.allclose(v, g["out_vec"][i], rtol=1e-12, atol=1e-14)
        assert np.allclose(c, g["out_cov"][i], rtol=1e-12, atol=1e-18)


def test_oracle_propagation_golden():
    g = _load("propagation_cases.npz")
    det, fld, _ = S.workload(str(g["workload"]), n_particles=1)
    _check_detector(det, g["detector"])
    cfg = S.finding_config_for(str(g["workload"]))
    for i in range(len(g["start"])):
        st, end, v, c, path, _ = oracle.propagate(cfg, det, fld, int(g["start"][i]), g["vec"][i],
                                                  g["cov"][i], True)
        assert st == g["status"][i]
        if st == 0:
            assert end == g["end"][i]
            assert np.allclose(v, g["out_vec"][i], rtol=1e-10, atol=1e-12)
            assert path == pytest.approx(g["path"][i], rel=1e-10)


@pytest.mark.parametrize("tag", ["c1", "c4"])
def test_oracle_tiny_events_golden(tag):
    g, det, fld, cfg, ev = _tiny(tag)
    tr = oracle.run(cfg, det, fld, ev, precision=64)
    assert np.array_equal(tr.offsets, g[f"{tag}_offsets"])
    assert np.array_equal(tr.constituents, g[f"{tag}_constituents"])
    assert np.allclose(tr.chi2, g[f"{tag}_chi2"], rtol=1e-12)
    assert np.array_equal(tr.ndf, g[f"{tag}_ndf"])
    assert len(tr) >= (6 if tag == "c1" else 40)


def test_oracle_fit_golden():
    g, det, fld, cfg, ev = _tiny("c1")
    f = _load("fit_tiny.npz")
    r = oracle.fit(FT.fitting_config(cfg), det, fld, ev, f["seed_surface_link"], f["seed_vector"],
                   f["seed_covariance"], f["offsets"], f["constituents"], precision=64)
    assert np.array_equal(r.outcome, f["outcome"])
    assert np.allclose(r.chi2, f["chi2"], rtol=1e-10)
    assert np.allclose(r.params_vector, f["params_vector"], rtol=1e-10, atol=1e-12)


def test_oracle_mlp_golden():
    g = _load("mlp_cases.npz")
    p = g["params"].astype(np.float64)
    logit, _ = M.forward(p, g["X"], g["mean"], g["std"])
    assert np.allclose(logit, g["logit"], rtol=1e-12, atol=1e-12)
    loss, grad = M.loss_and_grad(p, g["X"], g["y"].astype(np.float64), g["mean"], g["std"],
                                 pos_weight=float(g["pos_weight"]))
    assert loss == pytest.approx(float(g["loss"]), rel=1e-12)
    assert np.allclose(grad, g["grad"], rtol=1e-10, atol=1e-14)


# ---------------------------------------------------------------------------
# device
# ---------------------------------------------------------------------------
def _dbg():
    from traccc_amd import ckf
    lib = F.load_library(ckf.LIB)
    lib.ckf_debug_kf_update.restype = C.c_int
    lib.ckf_debug_propagate.restype = C.c_int
    return lib


@pytest.mark.gpu
def test_device_kf_update_golden(gpu_available):
    g = _load("kf_update_cases.npz")
    n = len(g["status"])
    vec = np.ascontiguousarray(g["vec"]); cov = np.ascontiguousarray(g["cov"])
    meas = np.ascontiguousarray(g["meas"]); var = np.ascontiguousarray(g["var"])
    dims = np.ascontiguousarray(g["dims"]); subs = np.ascontiguousarray(g["subspace"])
    chi2 = np.zeros(n, np.float32); st = np.zeros(n, np.int32)
    ov = np.zeros((n, 6), np.float32); oc = np.zeros((n, 6, 6), np.float32)
    rc = _dbg().ckf_debug_kf_update(n, F.ptr(vec, C.c_float), F.ptr(cov, C.c_float),
                                    F.ptr(meas, C.c_float), F.ptr(var, C.c_float),
                                    F.ptr(dims, C.c_uint32), F.ptr(subs, C.c_uint8),
                                    F.ptr(chi2, C.c_float), F.ptr(st, C.c_int32),
                                    F.ptr(ov, C.c_float), F.ptr(oc, C.c_float))
    assert rc == 0
    assert np.array_equal(st, g["status"])
    # FP32 device vs FP64 golden: 1e-4 relative (the unit parity bar)
    assert np.all(np.abs(chi2 - g["chi2"]) <= 1e-4 * np.abs(g["chi2"]) + 1e-6)
    assert np.all(np.abs(ov - g["out_vec"]) <= 1e-4 * np.abs(g["out_vec"]) + 1e-6)


@pytest.mark.gpu
def test_device_propagation_golden(gpu_available):
    g = _load("propagation_cases.npz")
    det, fld, _ = S.workload(str(g["workload"]), n_particles=1)
    cfg = S.finding_config_for(str(g["workload"]))
    n = len(g["start"])
    st = np.zeros(n, np.int32); end = np.zeros(n, np.uint32)
    ov = np.zeros((n, 6), np.float32); oc = np.zeros((n, 6, 6), np.float32)
    path = np.zeros(n, np.float32); nst = np.zeros(n, np.uint32)
    rc = _dbg().ckf_debug_propagate(
        C.byref(cfg), C.byref(det.struct), C.byref(fld.struct), n,
        F.ptr(np.ascontiguousarray(g["start"]), C.c_uint32),
        F.ptr(np.ascontiguousarray(g["vec"], np.float32), C.c_float),
        F.ptr(np.ascontiguousarray(g["cov"], np.float32), C.c_float), 1,
        F.ptr(st, C.c_int32), F.ptr(end, C.c_uint32), F.ptr(ov, C.c_float),
        F.ptr(oc, C.c_float), F.ptr(path, C.c_float), F.ptr(nst, C.c_uint32))
    assert rc == 0
    assert np.array_equal(st, g["status"])
    ok = st == 0
    assert np.array_equal(end[ok], g["end"][ok])
    scale = np.ones(6)
    d = np.abs(ov[ok] - g["out_vec"][ok])
    d[:, 4] /= np.abs(g["out_vec"][ok][:, 4])
    assert np.all(d[:, :5] <= 1e-4 * scale[:5])


@pytest.mark.gpu
@pytest.mark.parametrize("tag", ["c1", "c4"])
def test_device_tiny_events_golden(tag, gpu_available):
    from traccc_amd import ckf
    g, det, fld, cfg, ev = _tiny(tag)
    out = ckf.combinatorial_kalman_filter_algorithm(cfg)(det, fld, ev)
    assert np.array_equal(out.offsets, g[f"{tag}_offsets"])
    assert np.array_equal(out.constituents, g[f"{tag}_constituents"])
    assert np.array_equal(out.ndf, g[f"{tag}_ndf"].astype(np.float32))


@pytest.mark.gpu
def test_device_fit_golden(gpu_available):
    from traccc_amd import ckf
    g, det, fld, cfg, ev = _tiny("c1")
    f = _load("fit_tiny.npz")
    ddet = ckf.DeviceDetector(det, fld)
    dev = ckf.DeviceEvents(ddet, ev)
    cands = ckf.combinatorial_kalman_filter_algorithm(cfg).run_device(ddet, dev)
    got = FT.kalman_fitting_algorithm(FT.fitting_config(cfg)).run_device(ddet, dev, cands).download()
    assert np.array_equal(got.fit_outcome, f["outcome"])
    ok = f["outcome"] == FT.OUTCOME_SUCCESS
    # FP32 fit vs FP64 golden (the FP32 floor of the fit, tests/test_fit.py)
    assert np.all(np.abs(got.chi2[ok] - f["chi2"][ok]) <= 1e-2 * f["chi2"][ok] + 1e-4)
    sig = np.sqrt(np.einsum("nii->ni", f["params_covariance"][ok]))
    assert np.all(np.abs(got.params_vector[ok] - f["params_vector"][ok])[:, :5] <= 0.05 * sig[:, :5])


@pytest.mark.gpu
def test_device_mlp_golden(gpu_available):
    import torch
    from traccc_amd import scorer
    g = _load("mlp_cases.npz")
    s = scorer.Scorer(seed=1)
    s.set_params(g["params"], g["mean"], g["std"])
    got = s.forward(torch.tensor(g["X"], device="cuda")).cpu().numpy()
    # BF16 operands, FP32 accumulation vs the BF16-emulating FP64 logits
    assert np.max(np.abs(got - g["logit_bf16"])) < 2e-2 * max(1.0, np.max(np.abs(g["logit_bf16"])))
