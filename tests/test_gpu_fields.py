"""GPU: scalar-field generality against the oracle -- a constant potential
with its own parameter (V = c + p, its dF/dp column), an explicit-values
potential (V = beta * values, the "beta" column), a constant thickness other
than 1 (alpha_e = covolume/length * mean thickness, c t in F and J), and the
reference's error for an explicit-values thickness. References:
src/scalar_field_constant.cpp:43-75, src/scalar_field_explicit_values.cpp:33-57,
src/parameter_matrix_keo.cpp:202-227, src/model_evaluator_nls.cpp:102-111,630-692.
Bar: 1e-12 relative (2-norm)."""
import numpy as np
import pytest

from conftest import rel_err
from nosh_amd import meshgen as mg

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module", params=["kuhn8_jitter", "square16"])
def mesh_case(request):
    import nosh_amd as nosh
    from oracle import oracle as O
    print(f"[fields case {request.param}]", flush=True)
    if request.param == "kuhn8_jitter":
        c, t = mg.kuhn_cube(8, jitter=0.2, seed=2)
    else:
        c, t = mg.square_tri(16)
    return dict(coords=c, cells=t, gm=nosh.Mesh(c, t), om=O.Mesh(c, t),
                A=mg.magnetic_potential(c), n=c.shape[0])


def _oracle(case, V, dVdp, v_param, thickness=None):
    from oracle import oracle as O
    om = O.Model(case["om"], case["A"], thickness=thickness, V=V, dVdp=dVdp, deriv_param="mu")
    if v_param:
        O.lib().orc_set_v_param(om._h, v_param.encode())
    return om


def _check_model(model, om, params, names, n):
    psi = mg.random_state(n, 0)
    phi = mg.random_state(n, 1)
    om.set_params(params["g"], params["mu"])
    assert model.get_p_names() == names
    assert rel_err(model.eval_f(psi, params), om.f(psi)) < TOL
    J = model.create_W_op()
    J.rebuild(params, psi)
    om.jac_rebuild(psi)
    assert rel_err(J.apply(phi), om.jac(phi)) < TOL
    P = model.create_W_prec()
    P.rebuild(params, psi)
    om.prec_rebuild(psi)
    assert rel_err(P.apply_operator(phi), om.prec_op(phi)) < TOL
    for col, name in enumerate(names):
        got = model.eval_dfdp(psi, params, col)
        ref = om.dfdp(psi, name)
        assert rel_err(got, ref) < TOL, name


def test_constant_potential_with_parameter(mesh_case):
    """V = c + p: F, J and P use c + p; the p column of dF/dp is
    dK/dmu psi + c t psi (dV/dp = 1)."""
    import nosh_amd as nosh
    gm, n = mesh_case["gm"], mesh_case["n"]
    c0, p = -1.2, 0.3
    V = nosh.ScalarFieldConstant(gm, c0, "alpha", p)
    model = nosh.Nls(gm, nosh.VectorFieldExplicitValues(gm, mesh_case["A"], 0.4), V, 1.3,
                     nosh.ScalarFieldConstant(gm, 1.0))
    assert model.get_nominal_values() == {"alpha": p, "g": 1.3, "mu": 0.4}
    om = _oracle(mesh_case, np.full(n, c0 + p), np.ones(n), "alpha")
    _check_model(model, om, {"alpha": p, "g": 1.3, "mu": 0.4}, ["alpha", "g", "mu"], n)
    # Without the parameter in the map the reference's get_v gives c alone.
    om0 = _oracle(mesh_case, np.full(n, c0), np.ones(n), "alpha")
    om0.set_params(1.3, 0.4)
    psi = mg.random_state(n, 2)
    assert rel_err(model.eval_f(psi, {"g": 1.3, "mu": 0.4}), om0.f(psi)) < TOL


def test_explicit_values_potential(mesh_case):
    """V = beta * values with the scalar parameter beta (nominal 1); the beta
    column of dF/dp is dK/dmu psi + c t values psi; a map without beta is
    the reference's std::out_of_range (KeyError)."""
    import nosh_amd as nosh
    gm, n = mesh_case["gm"], mesh_case["n"]
    vals = -0.5 - mg.splitmix64_uniform(7, n)
    V = nosh.ScalarFieldExplicitValues(gm, vals)
    model = nosh.Nls(gm, nosh.VectorFieldExplicitValues(gm, mesh_case["A"], 0.2), V, 0.9,
                     nosh.ScalarFieldConstant(gm, 1.0))
    assert model.get_nominal_values() == {"beta": 1.0, "g": 0.9, "mu": 0.2}
    beta = 0.7
    om = _oracle(mesh_case, beta * vals, vals, "beta")
    _check_model(model, om, {"beta": beta, "g": 0.9, "mu": 0.2}, ["beta", "g", "mu"], n)
    with pytest.raises(KeyError):
        model.eval_f(mg.random_state(n, 3), {"g": 0.9, "mu": 0.2})


def test_thickness_not_one(mesh_case):
    """Constant thickness 1.7: alpha_e = 1.7 covolume/length and c t = 1.7 c
    in F, J, P and dF/dp (both columns)."""
    import nosh_amd as nosh
    gm, n = mesh_case["gm"], mesh_case["n"]
    t = nosh.ScalarFieldConstant(gm, 1.7)
    mvp = nosh.VectorFieldExplicitValues(gm, mesh_case["A"], 0.3)
    model = nosh.Nls(gm, mvp, nosh.ScalarFieldConstant(gm, -1.0), 1.0, t)
    om = _oracle(mesh_case, -np.ones(n), np.zeros(n), "", thickness=np.full(n, 1.7))
    _check_model(model, om, {"g": 1.0, "mu": 0.3}, ["g", "mu"], n)
    keo = nosh.Keo(gm, t, mvp)
    keo.set_parameters({"mu": 0.3})
    om.set_params(1.0, 0.3)
    phi = mg.random_state(n, 4)
    assert rel_err(keo.apply(phi), om.keo(phi)) < TOL


def test_explicit_values_thickness_is_refused_like_the_reference(mesh_case):
    """keo::build_alpha_cache_ asks the thickness for get_v with an empty
    map; an explicit-values field then throws std::out_of_range ("beta")."""
    import nosh_amd as nosh
    gm, n = mesh_case["gm"], mesh_case["n"]
    t = nosh.ScalarFieldExplicitValues(gm, np.ones(n))
    model = nosh.Nls(gm, nosh.VectorFieldExplicitValues(gm, mesh_case["A"], 0.3),
                     nosh.ScalarFieldConstant(gm, -1.0), 1.0, t)
    with pytest.raises(KeyError):
        model.eval_f(mg.random_state(n, 5), {"g": 1.0, "mu": 0.3})


def test_batch_keeps_the_potential_parameter():
    """A batch over a model whose potential has its own parameter solves with
    V = c + p (the model's nominal p), every instance checked by the oracle."""
    import nosh_amd as nosh
    from oracle import oracle as O
    c, t = mg.kuhn_cube(6)
    n = c.shape[0]
    gm = nosh.Mesh(c, t)
    A = mg.magnetic_potential(c)
    c0, p = -1.3, 0.25
    model = nosh.Nls(gm, nosh.VectorFieldExplicitValues(gm, A, 0.0),
                     nosh.ScalarFieldConstant(gm, c0, "alpha", p), 1.0,
                     nosh.ScalarFieldConstant(gm, 1.0))
    mu = np.array([0.05, 0.1, 0.15, 0.2])
    g = np.array([0.8, 1.0, 1.2, 1.4])
    psi0 = np.concatenate([np.stack([np.full(n, np.sqrt(-(c0 + p) / gb)), np.zeros(n)],
                                    1).ravel() for gb in g])
    batch = nosh.Batch([(model, 4)], mu, g)
    psi, st = batch.solve(psi0, tol=1e-10, max_iters=40, minres_rtol=1e-6, forcing=0)
    assert all(s["converged"] for s in st), st
    om = O.Model(O.Mesh(c, t), A, V=np.full(n, c0 + p))
    for k, x in enumerate(batch.split(psi)):
        om.set_params(g[k], mu[k])
        assert np.linalg.norm(om.f(x)) <= 1e-9, k


def test_mesh_and_fields_from_an_exodus_file(tmp_path):
    """A jittered cube and its fields written to an Exodus II file and read
    back (nosh_amd.io, the reference's .e input path through MOAB): the
    device mesh built from the file has the oracle's edge numbering and
    control volumes, and F, J phi and dF/dp with the file's A, V (explicit
    values) and thickness agree with the oracle to 1e-12."""
    import nosh_amd as nosh
    from nosh_amd import io as nio
    from oracle import oracle as O
    c0, t0 = mg.kuhn_cube(6, jitter=0.2, seed=8)
    n = c0.shape[0]
    V0 = -1.0 + 0.1 * np.sin(c0[:, 0])
    path = str(tmp_path / "cube.e")
    print("[exodus mesh: 6^3 jittered cube]", flush=True)
    nio.write_exodus(path, c0, t0, {"A": mg.magnetic_potential(c0), "V": V0,
                                     "thickness": np.full(n, 1.3)})
    c, t, pd = nio.read_mesh(path)
    gm, om_mesh = nosh.Mesh(c, t), O.Mesh(c, t)
    assert np.array_equal(gm.my_edges(), om_mesh.edges())
    assert rel_err(gm.control_volumes(), om_mesh.control_volumes()) < TOL
    th = float(pd["thickness"][0])
    model = nosh.Nls(gm, nosh.VectorFieldExplicitValues(gm, pd["A"], 0.0),
                     nosh.ScalarFieldExplicitValues(gm, pd["V"]), 1.1,
                     nosh.ScalarFieldConstant(gm, th))
    om = _oracle(dict(om=om_mesh, A=pd["A"]), pd["V"], pd["V"], "beta",
                 thickness=np.full(n, th))
    _check_model(model, om, {"beta": 1.0, "g": 1.1, "mu": 0.3}, ["beta", "g", "mu"], n)
