"""Writes tests/golden/gelu.npz: the tanh-form GELU used by the MLP surrogate,
0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3))), and its derivative, on
z in [-8, 8], from PyTorch's own implementation
(torch.nn.functional.gelu(approximate="tanh") and autograd, float64) -- a
routine independent of the oracle's NumPy formula, which the test checks
against these values.

Run from the repository root: python tests/golden/make_gelu.py
"""
import os

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    z = torch.linspace(-8.0, 8.0, 2001, dtype=torch.float64, requires_grad=True)
    y = torch.nn.functional.gelu(z, approximate="tanh")
    (dy,) = torch.autograd.grad(y.sum(), z)
    np.savez_compressed(os.path.join(HERE, "gelu.npz"), z=z.detach().numpy(),
                        gelu=y.detach().numpy(), dgelu=dy.numpy())


if __name__ == "__main__":
    main()
