"""GPU: nosh_nls_gibbs_energy_noalloc (no device allocation; partial sums in
the library's fixed scratch array) gives the same bits as
nosh_nls_gibbs_energy, on two calls in a row, the second on a smaller mesh,
and refuses a null argument (NOSH_ERR_INVALID)."""
import numpy as np
import pytest

from nosh_amd import meshgen as mg

pytestmark = pytest.mark.gpu


def _model(coords, cells):
    import nosh_amd as nosh
    mesh = nosh.Mesh(coords, cells)
    mvp = nosh.VectorFieldExplicitValues(mesh, mg.magnetic_potential(coords), 0.0)
    return nosh.Nls(mesh, mvp, nosh.ScalarFieldConstant(mesh, -1.0), 1.0,
                    nosh.ScalarFieldConstant(mesh, 1.0))


def test_gibbs_energy_noalloc_matches_allocating_entry_point():
    import torch
    import nosh_amd as nosh
    dev = torch.device("cuda:0")
    cases = [("kuhn 24^3 jittered", mg.kuhn_cube(24, jitter=0.2, seed=5), 41),
             ("square 16x16", mg.square_tri(16), 42)]
    out = torch.full((2,), float("nan"), dtype=torch.float64, device=dev)
    for k, (name, (c, t), seed) in enumerate(cases):
        print(f"[noalloc gibbs: {name}, call {k + 1}]", flush=True)
        model = _model(c, t)
        psi = mg.random_state(c.shape[0], seed)
        want = nosh.gibbs_energy(model, psi)
        x = torch.from_numpy(psi).to(dev)
        stream = torch.cuda.current_stream(dev)
        nosh.gibbs_energy_noalloc(model, x.data_ptr(), out[k:].data_ptr(), stream.cuda_stream)
        torch.cuda.synchronize()
        got = float(out[k].item())
        # Same partial sums in the same order: the same bits.
        assert got == want, (name, got, want)
        z = psi[0::2] + 1j * psi[1::2]
        assert want < 0.0 and abs(want) <= float(np.max(np.abs(z)) ** 4)
    with pytest.raises(ValueError):  # NOSH_ERR_INVALID
        nosh.gibbs_energy_noalloc(model, 0, out.data_ptr(), 0)
