"""GPU: BF16 MFMA MLP kernels against the NumPy oracle (forward, gradients),
and training reduces the loss."""
import ctypes as C

import numpy as np
import pytest

from oracle import mlp_oracle as MO

pytestmark = pytest.mark.gpu


def _net(nin=7, width=256, depth=4, nout=2, seed=11):
    import nosh_amd as nosh
    return nosh.Mlp(nin, width, depth, nout, seed)


@pytest.mark.parametrize("M", [64, 1000, 4096])
def test_forward_matches_oracle(M):
    net = _net()
    rng = np.random.default_rng(1)
    X = rng.normal(0, 1, (M, 7)).astype(np.float32)
    y = net.forward(X)
    p = net.get_params()
    y_bf, _, _ = MO.forward(p, X, net.dims, bf=True)
    err = np.linalg.norm(y - y_bf) / np.linalg.norm(y_bf)
    print("forward rel err vs bf16-emulating oracle", err)
    assert err < 5e-3
    y64, _, _ = MO.forward(p, X, net.dims, bf=False)
    assert np.linalg.norm(y - y64) / np.linalg.norm(y64) < 5e-2


def test_gradients_match_oracle():
    net = _net()
    rng = np.random.default_rng(2)
    M = 3000
    X = rng.normal(0, 1, (M, 7)).astype(np.float32)
    T = rng.normal(0, 1, (M, 2)).astype(np.float32)
    loss, g = net.gradients(X, T)
    p = net.get_params()
    loss_ref, g_ref = MO.loss_and_grad(p, X, T, net.dims, bf=True)
    assert abs(loss - loss_ref) / loss_ref < 5e-3
    # Layer by layer (weights, then biases).
    dims = net.dims
    off = 0
    for a, b in zip(dims[:-1], dims[1:]):
        s = slice(off, off + a * b)
        e = np.linalg.norm(g[s] - g_ref[s]) / np.linalg.norm(g_ref[s])
        print("weights", a, "x", b, "rel err", e)
        assert e < 2e-2
        off += a * b
    for b in dims[1:]:
        s = slice(off, off + b)
        e = np.linalg.norm(g[s] - g_ref[s]) / max(np.linalg.norm(g_ref[s]), 1e-30)
        assert e < 2e-2
        off += b


def test_training_reduces_loss():
    import torch
    net = _net(seed=3)
    rng = np.random.default_rng(4)
    M = 8192
    X = rng.uniform(-1, 1, (M, 7)).astype(np.float32)
    T = np.stack([np.sin(2 * X[:, 0]) * X[:, 1], X[:, 2] * X[:, 3] - 0.5 * X[:, 4]],
                 axis=1).astype(np.float32)
    dev = torch.device("cuda:0")
    Xd = torch.from_numpy(X).to(dev)
    Td = torch.from_numpy(T).to(dev)
    losses = [net.train_step_device(Xd.data_ptr(), Td.data_ptr(), M, 1e-3) for _ in range(300)]
    print("loss", losses[0], "->", losses[-1])
    assert losses[-1] < 0.1 * losses[0]


@pytest.mark.parametrize("width,M", [(256, 1000), (256, 70000), (128, 300), (64, 129)])
def test_fused_inference_matches_forward(width, M):
    """The one-kernel inference (activations in LDS, MFMA hidden layers) gives
    forward()'s outputs: same roundings and summation order, so equal up to
    FMA contraction differences; and it agrees with the bf16-emulating
    oracle."""
    net = _net(nin=9, width=width, seed=5)
    rng = np.random.default_rng(7)
    X = rng.normal(0, 1, (M, 9)).astype(np.float32)
    y_f = net.forward(X)
    y_i = net.infer(X)
    d = np.abs(y_i - y_f).max() / max(np.abs(y_f).max(), 1e-30)
    print("fused vs layered max rel diff", d)
    assert d < 1e-5
    if M <= 1000:
        y_bf, _, _ = MO.forward(net.get_params(), X, net.dims, bf=True)
        assert np.linalg.norm(y_i - y_bf) / np.linalg.norm(y_bf) < 5e-3


def test_first_adam_step_closed_form():
    """Adam's first step (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected)
    has a closed form: m_hat = g and v_hat = g^2, so every parameter moves by
    -lr g / (|g| + eps), i.e. -lr sign(g) where |g| >> eps. The step is
    checked against the gradient the library reports for the same minibatch
    (FP32 master weights: within a few ulps of the parameter)."""
    import torch
    net = _net(seed=5)
    rng = np.random.default_rng(6)
    M = 2048
    X = rng.normal(0, 1, (M, 7)).astype(np.float32)
    T = rng.normal(0, 1, (M, 2)).astype(np.float32)
    p0 = net.get_params().astype(np.float64)
    _, g = net.gradients(X, T)
    g = g.astype(np.float64)
    dev = torch.device("cuda:0")
    Xd = torch.from_numpy(X).to(dev)
    Td = torch.from_numpy(T).to(dev)
    lr = 1e-3
    net.train_step_device(Xd.data_ptr(), Td.data_ptr(), M, lr)
    torch.cuda.synchronize()
    d = net.get_params().astype(np.float64) - p0
    want = -lr * g / (np.abs(g) + 1e-8)
    tol = 4 * np.finfo(np.float32).eps * np.maximum(np.abs(p0), lr) + 1e-3 * lr
    assert np.all(np.abs(d - want) <= tol), np.max(np.abs(d - want) - tol)
    big = np.abs(g) > 1e-5
    assert big.sum() > 0.9 * g.size
    np.testing.assert_allclose(d[big], -lr * np.sign(g[big]), rtol=0, atol=2e-3 * lr + 1e-6)
