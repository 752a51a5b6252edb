"""GPU: every benchmark configuration at its own size, checked by the oracle
(the reference pins its operators on full meshes too: test/keo.cpp:143-179,
test/compute_f.cpp:88-121, test/jac.cpp:127-160).

* 96^3 Kuhn cube (912,673 vertices): topology bit-exact; edge data, control
  volumes, K phi, F, J phi, dF/dp (both columns) and P phi within 1e-12 of
  the oracle at mu = 0.4; plain FP64 and FP32-inner Newton side by side from
  the seeded random start (both reach the trivial state psi = 0 here, which
  the test names) and from a perturbed vortex state (the same solution,
  up to J's soft mode there).
* 128 instances over jittered 16^3 / 24^3 / 32^3 / 48^3 cubes in one batch:
  every instance converged, each with oracle ||F|| <= 1e-9.
* the 200 k-vertex Delaunay disk: surrogate-seeded continuation past the
  40-step training point, final state checked by the oracle.
* the 160^3 cube (4.17 M vertices): 8 branches of the batched continuation,
  each final state checked by the oracle.
The 64 x 64 square is covered at full size by test_gpu_parity.py."""
import numpy as np
import pytest

from conftest import phase_align, rel_err
from nosh_amd import meshgen as mg

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _model(gm, coords, g=1.0):
    import nosh_amd as nosh
    mvp = nosh.VectorFieldExplicitValues(gm, mg.magnetic_potential(coords), 0.0)
    return nosh.Nls(gm, mvp, nosh.ScalarFieldConstant(gm, -1.0), g,
                    nosh.ScalarFieldConstant(gm, 1.0))


@pytest.fixture(scope="module")
def cube96():
    import nosh_amd as nosh
    from oracle import oracle as O
    print("[config 1: building the 96^3 cube (GPU and oracle)]", flush=True)
    c, t = mg.kuhn_cube(96)
    gm = nosh.Mesh(c, t)
    om = O.Mesh(c, t)
    return dict(coords=c, cells=t, gm=gm, om=om, n=c.shape[0])


@pytest.mark.timeout(600)
def test_config1_mesh_96(cube96):
    gm, om = cube96["gm"], cube96["om"]
    assert (cube96["n"], gm.num_edges) == (912673, 6276384)
    assert gm.num_edges == om.E
    np.testing.assert_array_equal(gm.my_edges(), om.edges())
    np.testing.assert_array_equal(gm.cell_edges(), om.cell_edges())
    ln_g, cov_g = gm.get_edge_data()
    ln_o, cov_o = om.edge_data()
    assert rel_err(ln_g, ln_o) < TOL
    assert rel_err(cov_g, cov_o) < TOL
    cv = gm.control_volumes()
    assert rel_err(cv, om.control_volumes()) < TOL
    assert abs(cv.sum() - 1000.0) < 1e-9


@pytest.mark.timeout(600)
def test_config1_operators_96(cube96):
    import nosh_amd as nosh
    from oracle import oracle as O
    c, n, gm = cube96["coords"], cube96["n"], cube96["gm"]
    mu, g = 0.4, 1.0
    A = mg.magnetic_potential(c)
    mvp = nosh.VectorFieldExplicitValues(gm, A, mu)
    t = nosh.ScalarFieldConstant(gm, 1.0)
    model = nosh.Nls(gm, mvp, nosh.ScalarFieldConstant(gm, -1.0), g, t)
    print("[config 1: oracle model at 96^3]", flush=True)
    om = O.Model(cube96["om"], A)
    om.set_params(g, mu)
    psi = mg.random_state(n, 0)
    phi = mg.random_state(n, 1)
    params = {"g": g, "mu": mu}
    keo = nosh.Keo(gm, t, mvp)
    keo.set_parameters({"mu": mu})
    assert rel_err(keo.apply(phi), om.keo(phi)) < TOL
    assert rel_err(model.eval_f(psi, params), om.f(psi)) < TOL
    J = model.create_W_op()
    J.rebuild(params, psi)
    om.jac_rebuild(psi)
    assert rel_err(J.apply(phi), om.jac(phi)) < TOL
    for col, name in enumerate(["g", "mu"]):
        assert rel_err(model.eval_dfdp(psi, params, col), om.dfdp(psi, name)) < TOL, name
    P = model.create_W_prec()
    P.rebuild(params, psi)
    om.prec_rebuild(psi)
    assert rel_err(P.apply_operator(phi), om.prec_op(phi)) < TOL


@pytest.mark.timeout(600)
def test_config1_newton_side_by_side_96(cube96):
    """Plain FP64 and FP32-inner / FP64-outer Newton-MINRES at mu = 0.4,
    g = 1. From the seeded random start both converge to the trivial state
    (psi = 0: the random phases cancel on this fine mesh); the absolute norms
    are reported and checked. Then a nontrivial solution: plain FP64 Newton
    from psi = 1 (plus 10 % seeded noise) converges to a vortex state; from
    that state perturbed by 1e-3 seeded noise both variants converge back to
    it. J has an eigenvalue near 1e-8 there, so ||F|| ~ 1e-11 fixes the
    state only to ~1e-3 along its eigenvector: the FP32-inner result is
    within 1e-9 (relative, phase-aligned) of the FP64 one, or differs from
    it along that soft mode within the residual bound. (From the far
    superconducting start itself the two variants wander to different vortex
    states.)"""
    import nosh_amd as nosh
    from oracle import oracle as O
    c, n, gm = cube96["coords"], cube96["n"], cube96["gm"]
    mu, g = 0.4, 1.0
    model = _model(gm, c)
    om = O.Model(cube96["om"], mg.magnetic_potential(c))
    om.set_params(g, mu)
    kw = dict(tol=1e-10, max_iters=60, minres_rtol=1e-10, minres_max_iters=2000, cheb_iters=6,
              cheb_ratio=300.0)
    kw_mixed = dict(kw, precision=1, forcing=1, forcing_min=1e-6, forcing_max=0.1)
    s1 = 0.1 * mg.random_state(n, 5)
    s1[0::2] += 1.0
    print("[config 1: FP64 Newton from the superconducting start]", flush=True)
    p_sc, st_sc = nosh.newton_solve(model, s1, {"g": g, "mu": mu}, **kw)
    assert st_sc["converged"] and np.linalg.norm(om.f(p_sc)) <= 1e-9
    starts = {"random": mg.random_state(n, 0),
              "perturbed vortex state": p_sc + 1e-3 * mg.random_state(n, 6)}
    for name, psi0 in starts.items():
        print(f"[config 1: Newton side by side from the {name} start]", flush=True)
        p64, st64 = nosh.newton_solve(model, psi0, {"g": g, "mu": mu}, **kw)
        p32, st32 = nosh.newton_solve(model, psi0, {"g": g, "mu": mu}, **kw_mixed)
        nf64, nf32 = np.linalg.norm(om.f(p64)), np.linalg.norm(om.f(p32))
        n64, n32 = np.linalg.norm(p64), np.linalg.norm(p32)
        d = np.linalg.norm(phase_align(p32, p64) - p64)
        print(f"  FP64: newton {st64['newton_iters']} ||F|| {nf64:.2e} ||psi|| {n64:.3e}; "
              f"FP32-inner: newton {st32['newton_iters']} ||F|| {nf32:.2e} ||psi|| {n32:.3e}; "
              f"||psi - psi_FP64|| {d:.2e}", flush=True)
        assert st64["converged"] and st32["converged"]
        assert nf64 <= 1e-9 and nf32 <= 1e-9
        if name == "random":
            # The trivial state: |psi| is round-off, and so is the difference.
            assert n64 < 1e-8 * np.sqrt(n) and n32 < 1e-8 * np.sqrt(n)
            assert d < 1e-8 * np.sqrt(n)
        else:
            assert n64 > 0.1 * np.sqrt(n)
            if d / n64 >= 1e-9:
                # Above the 1e-9 bar only along a soft mode of J: the
                # difference lies along the eigenvector of J's smallest
                # eigenvalue (gauge mode deflated) and within the linearised
                # bound (||F_64|| + ||F_32||) / lambda_min.
                lam, vec, _ = nosh.jacobian_eigenpairs(model, p64, {"g": g, "mu": mu}, nev=1,
                                                       vectors=True, tol=1e-8, max_iters=4000)
                dvec = phase_align(p32, p64) - p64
                v = vec[0] / np.linalg.norm(vec[0])
                along = abs(v @ dvec) / d
                print(f"  soft mode: lambda_min {lam[0]:.3e}, bound "
                      f"{(nf64 + nf32) / lam[0]:.2e}, share of the difference along it "
                      f"{along:.6f}", flush=True)
                assert 0.0 < lam[0] < 1e-6
                assert d <= 10.0 * (nf64 + nf32) / lam[0]
                assert along > 0.99


@pytest.mark.timeout(900)
def test_config2_all_128_instances():
    """The batched uneven workload: 128 instances round-robin over jittered
    16^3 / 24^3 / 32^3 / 48^3 cubes, mu ~ U(0,1), g ~ U(0.5,2) (seed 2),
    random starts, FP32 inner / FP64 outer, compaction: all converge and each
    instance's final state has oracle ||F|| <= 1e-9."""
    import nosh_amd as nosh
    from oracle import oracle as O
    sizes = [16, 24, 32, 48]
    print("[config 2: meshes]", flush=True)
    meshes = [mg.kuhn_cube(s, jitter=0.2, seed=2) for s in sizes]
    models = [_model(nosh.Mesh(c, t), c) for c, t in meshes]
    ninst = 128
    owner = [i % 4 for i in range(ninst)]
    counts = [owner.count(k) for k in range(4)]
    u = mg.splitmix64_uniform(2, 2 * ninst).reshape(ninst, 2)
    order = sorted(range(ninst), key=lambda i: (owner[i], i))
    mu = np.array([u[i, 0] for i in order])
    g = np.array([0.5 + 1.5 * u[i, 1] for i in order])
    psi0 = np.concatenate([mg.random_state(meshes[owner[i]][0].shape[0], 1000 + i)
                           for i in order])
    batch = nosh.Batch([(models[k], counts[k]) for k in range(4)], mu, g)
    batch.set_tuning(512, 2, 4, 16, fold=8)
    print("[config 2: batched solve of 128 instances]", flush=True)
    psi, st = batch.solve(psi0, tol=1e-10, max_iters=200, minres_rtol=1e-10,
                          minres_max_iters=4000, cheb_iters=6, cheb_ratio=300.0, forcing=1,
                          forcing_min=1e-6, forcing_max=0.1)
    print("  newton steps per instance:", sorted(s["newton_iters"] for s in st), flush=True)
    assert all(s["converged"] for s in st), [k for k, s in enumerate(st) if not s["converged"]]
    sols = batch.split(psi)
    k = 0
    for gi, (c, t) in enumerate(meshes):
        om = O.Model(O.Mesh(c, t), mg.magnetic_potential(c))
        for _ in range(counts[gi]):
            om.set_params(g[k], mu[k])
            assert np.linalg.norm(om.f(sols[k])) <= 1e-9, k
            k += 1


@pytest.mark.timeout(900)
def test_config3_disk_surrogate_continuation():
    """The 200 k-vertex disk (Delaunay, seed 3), g = 1, psi = 1 at mu = 0,
    steps of 0.01 (the 200-step schedule to mu = 2) up to mu = 0.5: the
    surrogate is trained on the first 40 converged states and then offered at
    every step; every step converges and the final state passes the oracle."""
    import nosh_amd as nosh
    from oracle import oracle as O
    print("[config 3: disk mesh]", flush=True)
    c, t = mg.disk(0.0213)
    n = c.shape[0]
    assert 195000 < n < 206000
    gm = nosh.Mesh(c, t)
    model = _model(gm, c)
    psi0 = np.zeros(2 * n)
    psi0[0::2] = 1.0
    print("[config 3: surrogate continuation, 50 steps]", flush=True)
    psi, rec, st = nosh.continuation_surrogate(
        model, psi0, {"g": 1.0}, "mu", 0.0, 0.5, 50,
        surrogate=dict(width=256, depth=4, train_after=40, train_steps=300, batch=16384,
                       lr=1e-3, newton_budget=4),
        tol=1e-10, max_iters=40, minres_rtol=1e-10, minres_max_iters=2000, cheb_iters=6,
        cheb_ratio=300.0)
    print(f"  surrogate stats {st}", flush=True)
    assert len(rec) == 51 and all(r["converged"] for r in rec)
    assert st["trained"] == 1
    om = O.Model(O.Mesh(c, t), mg.magnetic_potential(c))
    om.set_params(1.0, 0.5)
    assert np.linalg.norm(om.f(psi)) <= 1e-9


@pytest.mark.timeout(900)
def test_config4_cube160_batched_branches():
    """The 160^3 cube: 8 (mu_0 = 0, g) branches of the batched mixed-precision
    continuation over the first 3 steps of the 50-step schedule to mu = 1;
    every branch's final state passes the oracle at its own g."""
    import nosh_amd as nosh
    from oracle import oracle as O
    print("[config 4: 160^3 cube]", flush=True)
    c, t = mg.kuhn_cube(160)
    n = c.shape[0]
    assert n == 4173281
    gm = nosh.Mesh(c, t)
    model = _model(gm, c)
    nb = 8
    g = 0.6 + 0.2 * np.arange(nb)
    psi0 = np.zeros(2 * n * nb)
    for b in range(nb):
        psi0[2 * n * b:2 * n * (b + 1):2] = 1.0 / np.sqrt(g[b])
    batch = nosh.Batch([(model, nb)], np.zeros(nb), g)
    batch.set_tuning(1024, 2, 4, 16, fold=4)
    print("[config 4: batched continuation, 3 steps]", flush=True)
    psi, rec, _ = batch.continuation(psi0, 0.06, 3, tol=1e-10, max_iters=40,
                                     minres_rtol=1e-10, minres_max_iters=4000, cheb_iters=5,
                                     cheb_ratio=300.0, forcing=1, forcing_min=1e-6,
                                     forcing_max=0.1)
    assert len(rec) == 4 and all(r["converged"] == nb for r in rec)
    del batch
    print("[config 4: oracle at 160^3]", flush=True)
    om = O.Model(O.Mesh(c, t), mg.magnetic_potential(c))
    sols = np.split(psi, nb)
    for b in range(nb):
        om.set_params(g[b], 0.06)
        nf = np.linalg.norm(om.f(sols[b]))
        print(f"  branch {b}: oracle ||F|| {nf:.2e}", flush=True)
        assert nf <= 1e-9, b
