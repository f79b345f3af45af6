"""Learned branch scorer: the MLP oracle's pins (CPU) and the HIP BF16-MFMA
kernels against it (GPU).

Oracle (oracle/mlp_oracle.py) pins:
  * the FP64 gradient equals central finite differences of the loss;
  * closed forms: all-zero weights give logit = b3 and d loss / d b3 =
    mean(w (sigmoid(b3) - y)); GELU at 0 / large |z|; GELU' against finite
    differences; BF16 rounding of exact values and of ties (to even).
Device (marked gpu):
  * the 16x16x16 BF16 matrix-core fragment layout the kernels assume;
  * logits and gradients against the oracle with BF16 rounding at the
    kernels' operand points (tolerances in the tests), over whole and
    partial 64-row tiles;
  * training learns a separable problem and lowers the loss.
"""
import math

import numpy as np
import pytest

from oracle import mlp_oracle as M

RNG = np.random.default_rng(1234)


def _random_params(scale=0.3, rng=RNG):
    return rng.normal(0.0, scale, M.N_PARAMS)


def _random_data(n, rng=RNG):
    X = rng.normal(0.0, 1.0, (n, M.N_IN)) * rng.uniform(0.5, 3.0, M.N_IN) + rng.normal(0, 2, M.N_IN)
    y = (rng.uniform(size=n) < 0.6).astype(np.float64)
    return X, y


def test_oracle_gradient_matches_finite_differences():
    p = _random_params()
    X, y = _random_data(37)
    mean, std = X.mean(axis=0), X.std(axis=0)
    loss, g = M.loss_and_grad(p, X, y, mean, std, pos_weight=1.7)
    idx = RNG.choice(M.N_PARAMS, 60, replace=False)
    idx = np.concatenate([idx, [M.OFF_B3, M.OFF_W3, M.OFF_B1, M.OFF_W1, M.OFF_B2]])
    h = 1e-6
    for i in idx:
        e = np.zeros(M.N_PARAMS)
        e[i] = h
        lp, _ = M.loss_and_grad(p + e, X, y, mean, std, pos_weight=1.7)
        lm, _ = M.loss_and_grad(p - e, X, y, mean, std, pos_weight=1.7)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])


def test_oracle_closed_forms():
    p = np.zeros(M.N_PARAMS)
    p[M.OFF_B3] = 0.7
    X, y = _random_data(50)
    s, _ = M.forward(p, X, np.zeros(M.N_IN), np.ones(M.N_IN))
    assert np.allclose(s, 0.7)
    loss, g = M.loss_and_grad(p, X, y, np.zeros(M.N_IN), np.ones(M.N_IN))
    sig = 1.0 / (1.0 + math.exp(-0.7))
    assert g[M.OFF_B3] == pytest.approx(np.mean(sig - y), rel=1e-12)
    assert loss == pytest.approx(np.mean(-(y * math.log(sig) + (1 - y) * math.log(1 - sig))),
                                 rel=1e-12)
    # all weight gradients vanish with zero weights except W3 (H2 = GELU(b2) = 0)
    assert np.all(g[:M.OFF_W3] == 0)
    assert M.gelu(0.0) == 0.0
    assert M.gelu(10.0) == pytest.approx(10.0, rel=1e-12)
    assert abs(M.gelu(-10.0)) < 1e-20
    z = np.linspace(-4, 4, 41)
    fd = (M.gelu(z + 1e-6) - M.gelu(z - 1e-6)) / 2e-6
    assert np.allclose(M.gelu_grad(z), fd, atol=1e-8)


def test_bf16_rounding():
    assert M.bf16(1.0) == 1.0
    assert M.bf16(3.140625) == 3.140625  # exactly representable (8 significant bits)
    # halfway between 1 and the next BF16 (1 + 2^-7): ties to even -> 1
    assert M.bf16(1.0 + 2.0 ** -8) == 1.0
    # halfway between 1 + 2^-7 and 1 + 2^-6: ties to even -> 1 + 2^-6
    assert M.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6
    assert M.bf16(-2.5) == -2.5


# ---------------------------------------------------------------------------
# device
# ---------------------------------------------------------------------------

def _torch_dev():
    import torch
    return torch, torch.device("cuda:0")


@pytest.mark.gpu
def test_mfma_fragment_layout(gpu_available):
    from traccc_amd import scorer
    A = RNG.normal(size=(16, 16)).astype(np.float32)
    B = RNG.normal(size=(16, 16)).astype(np.float32)
    D = scorer.mfma_layout(A, B)
    ref = M.bf16(A) @ M.bf16(B)
    assert np.allclose(D, ref, rtol=1e-5, atol=1e-5)
    # a permutation-revealing case: one-hot A rows pick B rows
    A = np.zeros((16, 16), np.float32)
    for i in range(16):
        A[i, (5 * i + 3) % 16] = 1.0
    B = np.arange(256, dtype=np.float32).reshape(16, 16)
    assert np.array_equal(scorer.mfma_layout(A, B), B[[(5 * i + 3) % 16 for i in range(16)]])


def _scorer_with(p, mean, std):
    from traccc_amd import scorer
    s = scorer.Scorer(seed=3)
    s.set_params(p.astype(np.float32), mean.astype(np.float32), std.astype(np.float32))
    return s


@pytest.mark.gpu
@pytest.mark.parametrize("n", [64, 1000, 4096])
def test_forward_matches_oracle(n, gpu_available):
    torch, dev = _torch_dev()
    p = _random_params(0.25)
    X, _ = _random_data(n)
    mean, std = X.mean(axis=0), X.std(axis=0)
    s = _scorer_with(p, mean, std)
    got = s.forward(torch.tensor(X, dtype=torch.float32, device=dev)).cpu().numpy()
    ref, _ = M.forward(p.astype(np.float32).astype(np.float64), X.astype(np.float32), mean.astype(np.float32),
                       std.astype(np.float32), bf16_points=True)
    # FP32 accumulation order and the FP32 normalisation before BF16 rounding
    assert np.max(np.abs(got - ref)) < 2e-2 * max(1.0, np.max(np.abs(ref)))
    assert np.mean(np.abs(got - ref)) < 2e-3 * max(1.0, np.mean(np.abs(ref)))


@pytest.mark.gpu
@pytest.mark.parametrize("n", [64, 1000, 4096])
def test_gradient_matches_oracle(n, gpu_available):
    torch, dev = _torch_dev()
    p = _random_params(0.25)
    X, y = _random_data(n)
    mean, std = X.mean(axis=0), X.std(axis=0)
    s = _scorer_with(p, mean, std)
    loss, g = s.gradient(torch.tensor(X, dtype=torch.float32, device=dev),
                         torch.tensor(y, dtype=torch.uint8, device=dev), pos_weight=1.5)
    args = (p.astype(np.float32).astype(np.float64), X.astype(np.float32), y,
            mean.astype(np.float32), std.astype(np.float32))
    lq, gq = M.loss_and_grad(*args, pos_weight=1.5, bf16_points=True)
    l64, g64 = M.loss_and_grad(*args, pos_weight=1.5, bf16_points=False)
    assert loss == pytest.approx(lq, rel=2e-3)
    rel_q = np.linalg.norm(g - gq) / np.linalg.norm(gq)
    rel_64 = np.linalg.norm(g - g64) / np.linalg.norm(g64)
    cos = float(g @ gq / (np.linalg.norm(g) * np.linalg.norm(gq)))
    print(f"n={n}: |g - g_bf16|/|g| = {rel_q:.2e}, |g - g_fp64|/|g| = {rel_64:.2e}, cos {cos:.6f}")
    assert rel_q < 1e-2
    assert cos > 0.9999
    assert rel_64 < 5e-2
    # per block: no parameter group is off on its own
    for lo, hi in ((M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2),
                   (M.OFF_B2, M.OFF_W3), (M.OFF_W3, M.OFF_B3)):
        assert np.linalg.norm(g[lo:hi] - gq[lo:hi]) <= 3e-2 * np.linalg.norm(gq[lo:hi]) + 1e-7


@pytest.mark.gpu
def test_training_learns_separable_problem(gpu_available):
    torch, dev = _torch_dev()
    from traccc_amd import scorer
    rng = np.random.default_rng(5)
    n = 65536
    X = rng.normal(size=(n, M.N_IN)).astype(np.float32) * 2.0 + 1.0
    y = ((X[:, 0] - 1.0) * 0.8 - (X[:, 3] - 1.0) + 0.3 * (X[:, 7] - 1.0) ** 2 > 0.2).astype(np.uint8)
    Xd = torch.tensor(X, device=dev)
    yd = torch.tensor(y, device=dev)
    s = scorer.Scorer(seed=11)
    s.fit_normalisation(Xd)
    _, mean, std = s.params()
    assert np.allclose(mean, X.mean(axis=0), atol=1e-3)
    assert np.allclose(std, X.std(axis=0), rtol=1e-3)
    st = s.train(Xd, yd, scorer.train_config(epochs=12, batch=2048, learning_rate=3e-3))
    logits = s.forward(Xd).cpu().numpy()
    acc = np.mean((logits > 0) == (y == 1))
    print(f"loss {st.first_loss:.4f} -> {st.last_loss:.4f}, accuracy {acc:.4f}, "
          f"{st.steps} steps, {st.tflops:.3f} TFLOP/s")
    assert st.last_loss < st.first_loss
    assert acc > 0.97
