"""Python binding of the learned branch scorer (include/traccc_amd/scorer.h).

Device arrays are passed as torch tensors on the GPU (plumbing for device
memory); all arithmetic runs in the library's HIP kernels.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _ffi as F
from .ckf import LIB, CKFError, _check

N_FEATURES = 12
FEATURE_NAMES = ("pull_loc0", "pull_loc1", "chi2", "layer", "qop", "theta", "holes", "depth",
                 "chi2_per_ndf", "standing", "rank", "log_candidates")


class TrainConfig(C.Structure):
    _fields_ = [("epochs", C.c_uint32), ("batch", C.c_uint32), ("learning_rate", C.c_float),
                ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float),
                ("weight_decay", C.c_float), ("pos_weight", C.c_float),
                ("shuffle_seed", C.c_uint64)]


class TrainStats(C.Structure):
    _fields_ = [("steps", C.c_uint32), ("first_loss", C.c_float), ("last_loss", C.c_float),
                ("seconds", C.c_double), ("tflops", C.c_double)]


_bound = False


def _lib():
    global _bound
    lib = F.load_library(LIB)
    if not _bound:
        vp = C.c_void_p
        fp = C.POINTER(C.c_float)
        lib.ckf_scorer_n_params.restype = C.c_uint32
        lib.ckf_scorer_train_config_default.argtypes = [C.POINTER(TrainConfig)]
        lib.ckf_scorer_create.argtypes = [C.c_uint64, C.POINTER(vp)]
        lib.ckf_scorer_destroy.argtypes = [vp]
        lib.ckf_scorer_set_params.argtypes = [vp, fp, fp, fp]
        lib.ckf_scorer_get_params.argtypes = [vp, fp, fp, fp]
        lib.ckf_scorer_fit_normalisation.argtypes = [vp, vp, C.c_uint32, vp]
        lib.ckf_scorer_forward.argtypes = [vp, vp, C.c_uint32, vp, vp]
        lib.ckf_scorer_gradient.argtypes = [vp, vp, vp, C.c_uint32, C.c_float, fp, fp, vp]
        lib.ckf_scorer_train.argtypes = [vp, vp, vp, C.c_uint32, C.POINTER(TrainConfig),
                                         C.POINTER(TrainStats), vp]
        lib.ckf_debug_mfma_layout.argtypes = [fp, fp, fp]
        for name in ("ckf_scorer_create", "ckf_scorer_destroy", "ckf_scorer_set_params",
                     "ckf_scorer_get_params", "ckf_scorer_fit_normalisation", "ckf_scorer_forward",
                     "ckf_scorer_gradient", "ckf_scorer_train", "ckf_debug_mfma_layout"):
            getattr(lib, name).restype = C.c_int
        _bound = True
    return lib


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def n_params() -> int:
    return int(_lib().ckf_scorer_n_params())


def train_config(**kw) -> TrainConfig:
    cfg = TrainConfig()
    _lib().ckf_scorer_train_config_default(C.byref(cfg))
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


def mfma_layout(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    """D = A B for 16x16 inputs through the scorer's matrix-core fragments."""
    A = np.ascontiguousarray(A, np.float32)
    B = np.ascontiguousarray(B, np.float32)
    D = np.zeros((16, 16), np.float32)
    _check(_lib().ckf_debug_mfma_layout(_fptr(A), _fptr(B), _fptr(D)))
    return D


class Scorer:
    """ckf_scorer: the MLP's parameters and normalisation, in HBM."""

    def __init__(self, seed: int = 1):
        h = C.c_void_p()
        _check(_lib().ckf_scorer_create(seed, C.byref(h)))
        self._h = h
        self._L = _lib()

    def __del__(self):
        lib = getattr(self, "_L", None)
        if getattr(self, "_h", None) and lib is not None:
            lib.ckf_scorer_destroy(self._h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def params(self):
        p = np.zeros(n_params(), np.float32)
        m = np.zeros(N_FEATURES, np.float32)
        s = np.zeros(N_FEATURES, np.float32)
        _check(_lib().ckf_scorer_get_params(self._h, _fptr(p), _fptr(m), _fptr(s)))
        return p, m, s

    def set_params(self, params=None, mean=None, std=None):
        arrs = [None if a is None else np.ascontiguousarray(a, np.float32)
                for a in (params, mean, std)]
        ptrs = [None if a is None else _fptr(a) for a in arrs]
        _check(_lib().ckf_scorer_set_params(self._h, *ptrs))

    def fit_normalisation(self, features):
        _check(_lib().ckf_scorer_fit_normalisation(self._h, features.data_ptr(),
                                                   features.shape[0], None))

    def forward(self, features):
        import torch
        out = torch.empty(features.shape[0], dtype=torch.float32, device=features.device)
        _check(_lib().ckf_scorer_forward(self._h, features.data_ptr(), features.shape[0],
                                         out.data_ptr(), None))
        return out

    def gradient(self, features, labels, pos_weight: float = 1.0):
        g = np.zeros(n_params(), np.float32)
        loss = np.zeros(1, np.float32)
        _check(_lib().ckf_scorer_gradient(self._h, features.data_ptr(), labels.data_ptr(),
                                          features.shape[0], pos_weight, _fptr(g), _fptr(loss),
                                          None))
        return float(loss[0]), g

    def train(self, features, labels, cfg: TrainConfig | None = None) -> TrainStats:
        cfg = cfg or train_config()
        st = TrainStats()
        _check(_lib().ckf_scorer_train(self._h, features.data_ptr(), labels.data_ptr(),
                                       features.shape[0], C.byref(cfg), C.byref(st), None))
        return st
