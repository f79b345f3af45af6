"""ckf_scorer_gradient_noalloc (C ABI) against ckf_scorer_gradient (GPU).

The allocation-free entry point uses the same kernels with its per-tile
partial gradients in a fixed device array of the library instead of an
allocated buffer, so gradient and loss must equal the allocating entry
point's bit for bit -- on two calls in a row, the second with fewer samples
(the array keeps the first call's contents), and sizes above the limit are
rejected.
"""
import ctypes as C

import numpy as np
import pytest

from traccc_amd import _ffi as F
from traccc_amd import ckf
from traccc_amd import scorer as SC

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu(gpu_available):
    return gpu_available


def _lib():
    lib = F.load_library(ckf.LIB)
    vp = C.c_void_p
    lib.ckf_scorer_gradient_noalloc.argtypes = [vp, vp, vp, C.c_uint32, C.c_float, vp, vp, vp]
    lib.ckf_scorer_gradient_noalloc.restype = C.c_int
    lib.ckf_scorer_gradient.argtypes = [vp, vp, vp, C.c_uint32, C.c_float,
                                        C.POINTER(C.c_float), C.POINTER(C.c_float), vp]
    lib.ckf_scorer_gradient.restype = C.c_int
    return lib


def _samples(n, seed):
    import torch
    rng = np.random.default_rng(seed)
    f = rng.normal(size=(n, 12)).astype(np.float32)
    y = (rng.uniform(size=n) < 0.4).astype(np.uint8)
    return torch.from_numpy(f).cuda(), torch.from_numpy(y).cuda()


def _reference(lib, s, feats, labels, pos_weight):
    P = SC.n_params()
    g = np.zeros(P, np.float32)
    loss = np.zeros(1, np.float32)
    rc = lib.ckf_scorer_gradient(s.handle, feats.data_ptr(), labels.data_ptr(), feats.shape[0],
                                 pos_weight, g.ctypes.data_as(C.POINTER(C.c_float)),
                                 loss.ctypes.data_as(C.POINTER(C.c_float)), None)
    assert rc == 0
    return g, loss[0]


def _noalloc(lib, s, feats, labels, pos_weight):
    import torch
    P = SC.n_params()
    d_g = torch.full((P,), float("nan"), dtype=torch.float32, device="cuda")
    d_l = torch.full((1,), float("nan"), dtype=torch.float32, device="cuda")
    rc = lib.ckf_scorer_gradient_noalloc(s.handle, feats.data_ptr(), labels.data_ptr(),
                                         feats.shape[0], pos_weight, d_g.data_ptr(),
                                         d_l.data_ptr(), None)
    assert rc == 0, ckf._lib().ckf_last_error()
    torch.cuda.synchronize()
    return d_g.cpu().numpy(), float(d_l.cpu().numpy()[0])


def test_noalloc_gradient_equals_the_allocating_one():
    lib = _lib()
    s = SC.Scorer(seed=3)
    feats, labels = _samples(40000, 1)
    s.fit_normalisation(feats)
    # two calls in a row, the second with fewer samples (the scratch array
    # still holds the first call's partial gradients)
    for n, pw in ((40000, 1.0), (777, 2.5)):
        f, y = feats[:n].contiguous(), labels[:n].contiguous()
        g_ref, l_ref = _reference(lib, s, f, y, pw)
        g, l = _noalloc(lib, s, f, y, pw)
        assert np.all(np.isfinite(g)) and np.isfinite(l)
        assert np.array_equal(g, g_ref), float(np.max(np.abs(g - g_ref)))
        assert l == l_ref


def test_noalloc_gradient_rejects_sizes_above_its_limit():
    lib = _lib()
    s = SC.Scorer(seed=4)
    feats, labels = _samples(65537, 2)
    import torch
    d_g = torch.zeros(SC.n_params(), dtype=torch.float32, device="cuda")
    d_l = torch.zeros(1, dtype=torch.float32, device="cuda")
    rc = lib.ckf_scorer_gradient_noalloc(s.handle, feats.data_ptr(), labels.data_ptr(), 65537,
                                         1.0, d_g.data_ptr(), d_l.data_ptr(), None)
    assert rc == 1   # CKF_ERR_INVALID_ARGUMENT
    assert "65536" in ckf._lib().ckf_last_error().decode()
    # the limit itself is accepted
    rc = lib.ckf_scorer_gradient_noalloc(s.handle, feats.data_ptr(), labels.data_ptr(), 65536,
                                         1.0, d_g.data_ptr(), d_l.data_ptr(), None)
    assert rc == 0
    torch.cuda.synchronize()
    assert np.all(np.isfinite(d_g.cpu().numpy()))
