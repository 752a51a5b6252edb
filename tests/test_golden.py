"""Golden vectors (tests/golden/*.npz, written by make_golden.py from the
oracle): the oracle must reproduce them on CPU, the HIP library on the GPU."""
import os

import numpy as np
import pytest

from conftest import rel_err
from nosh_amd import meshgen as mg

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
GEN = {"square64": lambda: mg.square_tri(64), "kuhn4": lambda: mg.kuhn_cube(4)}
VEC = ["F", "Kphi", "Jphi", "dFdg", "dFdmu", "Pphi"]


def load(name):
    return dict(np.load(os.path.join(HERE, f"{name}.npz")))


@pytest.mark.parametrize("name", list(GEN))
def test_oracle_reproduces_golden(name):
    from oracle import oracle as O
    g = load(name)
    coords, cells = GEN[name]()
    m = O.Mesh(coords, cells)
    np.testing.assert_array_equal(m.edges(), g["edges"])
    np.testing.assert_array_equal(m.cell_edges(), g["cell_edges"])
    md = O.Model(m, mg.magnetic_potential(coords))
    md.set_params(1.0, float(g["mu"]))
    psi, phi = mg.random_state(m.n, 0), mg.random_state(m.n, 1)
    md.jac_rebuild(psi)
    md.prec_rebuild(psi)
    got = dict(F=md.f(psi), Kphi=md.keo(phi), Jphi=md.jac(phi), dFdg=md.dfdp(psi, "g"),
               dFdmu=md.dfdp(psi, "mu"), Pphi=md.prec_op(phi))
    for k in VEC:
        assert rel_err(got[k], g[k]) < 1e-13, k
    assert rel_err(m.control_volumes(), g["cv"]) < 1e-14


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(GEN))
def test_library_matches_golden(name):
    import nosh_amd as nosh
    g = load(name)
    print(f"[golden {name}]", flush=True)
    coords, cells = GEN[name]()
    n = coords.shape[0]
    mu = float(g["mu"])
    mesh = nosh.Mesh(coords, cells)
    np.testing.assert_array_equal(mesh.my_edges(), g["edges"])
    np.testing.assert_array_equal(mesh.cell_edges(), g["cell_edges"])
    ln, cov = mesh.get_edge_data()
    assert rel_err(ln, g["length"]) < 1e-14 and rel_err(cov, g["covolume"]) < 1e-12
    assert rel_err(mesh.control_volumes(), g["cv"]) < 1e-12
    mvp = nosh.VectorFieldExplicitValues(mesh, mg.magnetic_potential(coords), mu)
    V = nosh.ScalarFieldConstant(mesh, -1.0)
    t = nosh.ScalarFieldConstant(mesh, 1.0)
    model = nosh.Nls(mesh, mvp, V, 1.0, t, None, "mu")
    params = {"g": 1.0, "mu": mu}
    psi, phi = mg.random_state(n, 0), mg.random_state(n, 1)
    keo = nosh.Keo(mesh, t, mvp)
    keo.set_parameters({"mu": mu})
    J = model.create_W_op()
    J.rebuild(params, psi)
    P = model.create_W_prec()
    P.rebuild(params, psi)
    got = dict(F=model.eval_f(psi, params), Kphi=keo.apply(phi), Jphi=J.apply(phi),
               dFdg=model.eval_dfdp(psi, params, 0), dFdmu=model.eval_dfdp(psi, params, 1),
               Pphi=P.apply_operator(phi))
    for k in VEC:
        assert rel_err(got[k], g[k]) < 1e-12, k
