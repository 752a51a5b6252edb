"""Pins of the MLP oracle: analytic gradient against central finite
differences of the FP64 loss; GELU derivative against its difference quotient."""
import numpy as np

from oracle import mlp_oracle as M


def test_dgelu():
    z = np.linspace(-6, 6, 1001)
    h = 1e-6
    fd = (M.gelu(z + h) - M.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - M.dgelu(z))) < 1e-8


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 65504.0], np.float32)
    r = M.bf16(x)
    # bfloat16 keeps 8 significant bits: spacing 2^-7 on [1, 2).
    assert r[0] == 1.0 and r[1] == 1.0  # halfway 1 + 2^-8 rounds to even
    assert r[2] == 1.0078125
    assert abs(r[3] - (-3.140625)) < 1e-12
    assert np.all(np.abs(r - x) <= np.abs(x) * 2 ** -8)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    dims = M.dims_of(5, 16, 3, 2)
    nparams = sum(a * b for a, b in zip(dims[:-1], dims[1:])) + sum(dims[1:])
    p = rng.normal(0, 0.4, nparams)
    X = rng.normal(0, 1, (40, 5))
    T = rng.normal(0, 1, (40, 2))
    loss, g = M.loss_and_grad(p, X, T, dims)
    h = 1e-6
    idx = rng.choice(nparams, 60, replace=False)
    for i in idx:
        e = np.zeros(nparams)
        e[i] = h
        lp, _ = M.loss_and_grad(p + e, X, T, dims)
        lm, _ = M.loss_and_grad(p - e, X, T, dims)
        assert abs((lp - lm) / (2 * h) - g[i]) < 1e-7 * max(1.0, abs(g[i]))


def test_gelu_matches_library_fixture():
    """The oracle's GELU and its derivative against PyTorch's tanh-form GELU
    and autograd derivative (tests/golden/gelu.npz, written in this
    repository by tests/golden/make_gelu.py): a dropped term or a wrong
    constant (0.044715, sqrt(2/pi)) would miss by far more than 1e-12."""
    import os
    f = np.load(os.path.join(os.path.dirname(__file__), "golden", "gelu.npz"))
    z = f["z"]
    assert np.max(np.abs(M.gelu(z) - f["gelu"])) < 1e-12
    assert np.max(np.abs(M.dgelu(z) - f["dgelu"])) < 1e-12
    # Limits: GELU(z) -> z for large z, -> 0 for very negative z, GELU(0) = 0.
    assert abs(M.gelu(np.array([0.0]))[0]) == 0.0
    assert abs(M.gelu(np.array([8.0]))[0] - 8.0) < 1e-12
