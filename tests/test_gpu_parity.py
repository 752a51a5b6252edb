"""GPU parity: the HIP library against the CPU oracle on the same seeded inputs.

Bar: integer/index outputs bit-exact; FP64 operator applies within 1e-12
relative (2-norm); Newton solutions within 1e-9 after global phase alignment.
"""
import numpy as np
import pytest

from conftest import phase_align, rel_err
from nosh_amd import meshgen as mg

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _mesh_case(name):
    if name == "square16":
        return mg.square_tri(16)
    if name == "square64":
        return mg.square_tri(64)
    if name == "kuhn6":
        return mg.kuhn_cube(6)
    if name == "kuhn8_jitter":
        return mg.kuhn_cube(8, jitter=0.2, seed=2)
    raise KeyError(name)


CASES = ["square16", "square64", "kuhn6", "kuhn8_jitter"]


@pytest.fixture(scope="module", params=CASES)
def case(request):
    import nosh_amd as nosh
    from oracle import oracle as O
    print(f"[case {request.param}]", flush=True)
    coords, cells = _mesh_case(request.param)
    om = O.Mesh(coords, cells)
    gm = nosh.Mesh(coords, cells)
    A = mg.magnetic_potential(coords)
    return dict(name=request.param, coords=coords, cells=cells, om=om, gm=gm, A=A,
                n=coords.shape[0])


def test_mesh_topology_exact(case):
    om, gm = case["om"], case["gm"]
    assert gm.num_edges == om.E
    np.testing.assert_array_equal(gm.my_edges(), om.edges())
    np.testing.assert_array_equal(gm.cell_edges(), om.cell_edges())


def test_mesh_geometry(case):
    om, gm = case["om"], case["gm"]
    ln_g, cov_g = gm.get_edge_data()
    ln_o, cov_o = om.edge_data()
    assert rel_err(ln_g, ln_o) < TOL
    assert rel_err(cov_g, cov_o) < TOL
    assert rel_err(gm.control_volumes(), om.control_volumes()) < TOL


def test_permutation_is_bijection(case):
    perm = case["gm"].permutation()
    assert np.array_equal(np.sort(perm), np.arange(case["n"]))


def _objects(case, mu, g=1.0):
    import nosh_amd as nosh
    from oracle import oracle as O
    gm = case["gm"]
    mvp = nosh.VectorFieldExplicitValues(gm, case["A"], mu)
    V = nosh.ScalarFieldConstant(gm, -1.0)
    t = nosh.ScalarFieldConstant(gm, 1.0)
    model = nosh.Nls(gm, mvp, V, g, t, None, "mu")
    om = O.Model(case["om"], case["A"])
    om.set_params(g, mu)
    return mvp, V, t, model, om


@pytest.mark.parametrize("mu", [0.0, 0.1, 0.4])
def test_keo_f_jac_dfdp(case, mu):
    import nosh_amd as nosh
    mvp, V, t, model, om = _objects(case, mu)
    n = case["n"]
    psi = mg.random_state(n, 0)
    phi = mg.random_state(n, 1)
    params = {"g": 1.0, "mu": mu}
    # Edge projection a_e.
    at_o, _ = om.edge_fields()
    np.testing.assert_allclose(mvp.edge_projections(), mu * at_o, rtol=0, atol=1e-13)
    # K phi.
    keo = nosh.Keo(case["gm"], t, mvp)
    keo.set_parameters({"mu": mu})
    assert rel_err(keo.apply(phi), om.keo(phi)) < TOL
    # F(psi).
    assert rel_err(model.eval_f(psi, params), om.f(psi)) < TOL
    # J(psi) phi, through create_W_op (shares the model's keo).
    J = model.create_W_op()
    J.rebuild(params, psi)
    om.jac_rebuild(psi)
    assert rel_err(J.apply(phi), om.jac(phi)) < TOL
    # dF/dp columns in p_names order [g, mu].
    assert model.get_p_names() == ["g", "mu"]
    for col, name in enumerate(["g", "mu"]):
        got = model.eval_dfdp(psi, params, col)
        ref = om.dfdp(psi, name)
        if np.linalg.norm(ref) == 0:
            assert np.linalg.norm(got) == 0
        else:
            assert rel_err(got, ref) < TOL, name


def test_operators_with_32bit_columns():
    """The operator kernels read the split layout's columns as 16-bit offsets
    from the row when the reordered graph's bandwidth allows (every mesh
    above), else as 32-bit indices: a 33^3 cube with shuffled vertices and no
    reordering (bandwidth ~39 k) takes the 32-bit path, and K phi, F, J phi,
    P phi and the Chebyshev apply still match the oracle."""
    import nosh_amd as nosh
    from oracle import oracle as O
    c0, t0 = mg.kuhn_cube(33)
    n = c0.shape[0]
    perm = np.random.default_rng(5).permutation(n)
    inv = np.empty(n, dtype=np.int64)
    inv[perm] = np.arange(n)
    coords, cells = c0[perm], inv[t0].astype(t0.dtype)
    gm = nosh.Mesh(coords, cells, reorder=False)
    assert gm.layout()["column_bytes"] == 4
    assert nosh.Mesh(*mg.kuhn_cube(6)).layout()["column_bytes"] == 2
    A = mg.magnetic_potential(coords)
    mu = 0.3
    mvp = nosh.VectorFieldExplicitValues(gm, A, mu)
    t = nosh.ScalarFieldConstant(gm, 1.0)
    model = nosh.Nls(gm, mvp, nosh.ScalarFieldConstant(gm, -1.0), 1.0, t, None, "mu")
    om = O.Model(O.Mesh(coords, cells), A)
    om.set_params(1.0, mu)
    psi, phi = mg.random_state(n, 0), mg.random_state(n, 1)
    params = {"g": 1.0, "mu": mu}
    keo = nosh.Keo(gm, t, mvp)
    keo.set_parameters({"mu": mu})
    assert rel_err(keo.apply(phi), om.keo(phi)) < TOL
    assert rel_err(model.eval_f(psi, params), om.f(psi)) < TOL
    J = model.create_W_op()
    J.rebuild(params, psi)
    om.jac_rebuild(psi)
    assert rel_err(J.apply(phi), om.jac(phi)) < TOL
    P = model.create_W_prec()
    P.rebuild(params, psi)
    om.prec_rebuild(psi)
    assert rel_err(P.apply_operator(phi), om.prec_op(phi)) < TOL
    P.set_chebyshev(5, 30.0, 0.0)
    assert rel_err(P.apply(phi), om.cheb(phi, 5, 30.0)) < TOL
    # The FP32 kernels on the same layout: FP32-inner Newton from psi = 1
    # reaches a state whose oracle residual is at the FP64 tolerance.
    psi0 = np.zeros(2 * n)
    psi0[0::2] = 1.0
    psi_m, st = nosh.newton_solve(model, psi0, params, tol=1e-10, max_iters=40,
                                  minres_rtol=1e-10, minres_max_iters=3000, cheb_iters=4,
                                  cheb_ratio=30.0, precision=1, forcing=1, forcing_min=1e-6,
                                  forcing_max=0.1)
    assert st["converged"], st
    assert np.linalg.norm(om.f(psi_m)) <= 1e-9


@pytest.mark.parametrize("mu", [0.1])
def test_keo_regularized(case, mu):
    mvp, V, t, model, om = _objects(case, mu)
    n = case["n"]
    psi = mg.random_state(n, 0)
    phi = mg.random_state(n, 1)
    P = model.create_W_prec()
    P.rebuild({"g": 1.0, "mu": mu}, psi)
    om.prec_rebuild(psi)
    assert rel_err(P.apply_operator(phi), om.prec_op(phi)) < TOL
    assert abs(P.upper_bound() - om.cheb_bound()) < 1e-12 * om.cheb_bound()
    for iters in (1, 2, 5):
        P.set_chebyshev(iters, 30.0, 0.0)
        assert rel_err(P.apply(phi), om.cheb(phi, iters, 30.0)) < TOL, iters


def test_multicolumn_apply_and_errors(case):
    import nosh_amd as nosh
    mvp, V, t, model, om = _objects(case, 0.1)
    n = case["n"]
    psi = mg.random_state(n, 0)
    J = model.create_W_op()
    with pytest.raises(nosh.NoshError):
        J.apply(psi)  # not rebuilt yet
    with pytest.raises(KeyError):
        J.rebuild({"mu": 0.1}, psi)  # params.at("g")
    J.rebuild({"g": 1.0, "mu": 0.1}, psi)
    om.jac_rebuild(psi)
    X = np.stack([mg.random_state(n, s) for s in (1, 2, 3)], axis=1)
    Y = J.apply(X)
    for c in range(3):
        assert rel_err(Y[:, c], om.jac(X[:, c])) < TOL
    for kw in ({"mode": nosh.TRANS}, {"alpha": 2.0}, {"beta": 1.0}):
        with pytest.raises(nosh.LogicError):
            J.apply(psi, **kw)


def test_minres_matches_oracle(case):
    import nosh_amd as nosh
    mvp, V, t, model, om = _objects(case, 0.1)
    n = case["n"]
    psi = mg.random_state(n, 0)
    params = {"g": 1.0, "mu": 0.1}
    J = model.create_W_op()
    J.rebuild(params, psi)
    P = model.create_W_prec()
    P.set_chebyshev(4, 30.0, 0.0)
    P.rebuild(params, psi)
    om.jac_rebuild(psi)
    om.prec_rebuild(psi)
    b = -om.f(psi)
    x_o, k_o, rr_o = om.minres(b, rtol=1e-8, maxit=400, cheb_iters=4, ratio=30.0)
    x_g, k_g, rr_g = nosh.minres_solve(J, P, b, rtol=1e-8, max_iters=400)
    assert abs(k_g - k_o) <= 1
    assert rel_err(x_g, x_o) < 1e-7
    # True residual of the GPU solution.
    r = b - om.jac(x_g)
    assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-5


@pytest.mark.parametrize("name", ["square64", "kuhn6"])
def test_newton_matches_oracle(name):
    import nosh_amd as nosh
    from oracle import oracle as O
    coords, cells = _mesh_case(name)
    mu = 0.1 if name.startswith("square") else 0.4
    A = mg.magnetic_potential(coords)
    n = coords.shape[0]
    psi0 = mg.random_state(n, 0)
    om = O.Model(O.Mesh(coords, cells), A)
    psi_o, it_o, nf_o, mh_o = om.newton(psi0, 1.0, mu, tol=1e-10, maxit=40,
                                        minres_rtol=1e-10, minres_maxit=1000, cheb_iters=4,
                                        ratio=30.0)
    assert it_o >= 0, "oracle Newton did not converge"
    gm = nosh.Mesh(coords, cells)
    mvp = nosh.VectorFieldExplicitValues(gm, A, mu)
    V = nosh.ScalarFieldConstant(gm, -1.0)
    t = nosh.ScalarFieldConstant(gm, 1.0)
    model = nosh.Nls(gm, mvp, V, 1.0, t)
    psi_g, st = nosh.newton_solve(model, psi0, {"g": 1.0, "mu": mu}, tol=1e-10, max_iters=40,
                                  minres_rtol=1e-10, minres_max_iters=1000, cheb_iters=4,
                                  cheb_ratio=30.0)
    print(name, "oracle", it_o, list(mh_o), "gpu", st["newton_iters"], st["minres_history"])
    assert st["converged"] and st["final_norm_f"] <= 1e-10
    assert np.linalg.norm(om.f(psi_g)) <= 1e-9
    d = phase_align(psi_g, psi_o) - psi_o
    assert np.linalg.norm(d) / np.linalg.norm(psi_o) < 1e-9
    # Mixed precision (FP32 inner MINRES, FP64 F and update).
    psi_m, st_m = nosh.newton_solve(model, psi0, {"g": 1.0, "mu": mu}, tol=1e-10,
                                    max_iters=60, minres_rtol=1e-6, minres_max_iters=1000,
                                    cheb_iters=4, cheb_ratio=30.0, precision=1, forcing=0)
    print(name, "mixed", st_m["newton_iters"], st_m["minres_history"])
    assert st_m["converged"] and st_m["final_norm_f"] <= 1e-10
    assert np.linalg.norm(om.f(psi_m)) <= 1e-9
    # From the random start, Newton wanders for some 20 steps before it
    # converges, and which nearby state it reaches depends on the rounding of
    # the inner solves (FP32 here), so the start above pins only the residual.
    # Started near the oracle's solution, mixed-precision Newton must come
    # back to it: to 1e-9 where J is well conditioned.
    rng = np.random.default_rng(7)
    start = psi_o + 1e-3 * np.linalg.norm(psi_o) / np.sqrt(2 * n) * rng.standard_normal(2 * n)
    psi_r, st_r = nosh.newton_solve(model, start, {"g": 1.0, "mu": mu}, tol=1e-10,
                                    max_iters=20, minres_rtol=1e-6, minres_max_iters=1000,
                                    cheb_iters=4, cheb_ratio=30.0, precision=1, forcing=0)
    print(name, "mixed from the perturbed solution", st_r["newton_iters"])
    assert st_r["converged"] and st_r["final_norm_f"] <= 1e-10
    d = phase_align(psi_r, psi_o) - psi_o
    assert np.linalg.norm(d) / np.linalg.norm(psi_o) < 1e-9
