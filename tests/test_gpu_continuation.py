"""GPU: continuation driver and energy functionals, checked with the oracle."""
import numpy as np
import pytest

from nosh_amd import meshgen as mg

pytestmark = pytest.mark.gpu


def _model(coords, cells, mu=0.0, g=1.0):
    import nosh_amd as nosh
    mesh = nosh.Mesh(coords, cells)
    mvp = nosh.VectorFieldExplicitValues(mesh, mg.magnetic_potential(coords), mu)
    V = nosh.ScalarFieldConstant(mesh, -1.0)
    t = nosh.ScalarFieldConstant(mesh, 1.0)
    return mesh, nosh.Nls(mesh, mvp, V, g, t)


def test_energy_functionals():
    import nosh_amd as nosh
    from oracle import oracle as O
    coords, cells = mg.kuhn_cube(5)
    mesh, model = _model(coords, cells)
    n = coords.shape[0]
    psi, phi = mg.random_state(n, 0), mg.random_state(n, 1)
    cv = O.Mesh(coords, cells).control_volumes()
    r2 = psi[0::2] ** 2 + psi[1::2] ** 2
    e = -np.sum(cv * r2 * r2) / cv.sum()
    assert abs(nosh.gibbs_energy(model, psi) - e) < 1e-13 * abs(e)
    ip = np.sum(cv * (phi[0::2] * psi[0::2] + phi[1::2] * psi[1::2])) / cv.sum()
    assert abs(nosh.inner_product(model, phi, psi) - ip) < 1e-13


def test_continuation_in_mu_from_psi_one():
    """From psi = 1 at mu = 0 (the exact solution there) along mu to 1.0 with
    each predictor: every accepted point solves F = 0 (checked by the oracle),
    the three runs end on the same state, and the tangent predictor needs no
    more Newton steps than the trivial one."""
    import nosh_amd as nosh
    from oracle import oracle as O
    coords, cells = mg.square_tri(16)
    n = coords.shape[0]
    mesh, model = _model(coords, cells)
    psi0 = np.zeros(2 * n)
    psi0[0::2] = 1.0
    om = O.Model(O.Mesh(coords, cells), mg.magnetic_potential(coords))
    om.set_params(1.0, 1.0)
    totals, finals = {}, {}
    for predictor in (0, 1, 2):
        print(f"[continuation predictor={predictor}]", flush=True)
        psi, rec = nosh.continuation(model, psi0, {"g": 1.0}, "mu", 0.0, 1.0, 10,
                                     predictor=predictor, tol=1e-10, max_iters=20)
        assert len(rec) == 11 and all(r["converged"] for r in rec)
        assert rec[0]["newton_iters"] == 0  # psi = 1 solves F = 0 at mu = 0
        assert np.linalg.norm(om.f(psi)) <= 1e-9
        # Energy is -1 at psi = 1 and rises as the field suppresses |psi|.
        assert abs(rec[0]["gibbs"] + 1.0) < 1e-12
        assert rec[-1]["gibbs"] > rec[0]["gibbs"]
        totals[predictor] = sum(r["newton_iters"] for r in rec)
        finals[predictor] = psi
        print("predictor", predictor, [r["newton_iters"] for r in rec])
    # (Past mu ~ 0.8 a vortex enters and the runs may end on different
    # branches, so the final states are not compared.)
    assert totals[1] <= totals[0]


def test_surrogate_seeded_continuation_gains_over_tangent():
    """Continuation along mu in [0, 0.28] on the 32x32 square (a vortex
    enters near mu = 0.31 on this mesh, where Newton from any predictor needs
    dozens of steps): after training on the run's first 10 solutions, the
    surrogate guess is closer to the converged solution than the tangent
    guess on the steps it seeds, every accepted point solves F = 0 (oracle),
    and Newton needs no more steps than from the tangent."""
    import nosh_amd as nosh
    from oracle import oracle as O
    coords, cells = mg.square_tri(32)
    n = coords.shape[0]
    mesh, model = _model(coords, cells)
    psi0 = np.zeros(2 * n)
    psi0[0::2] = 1.0
    kw = dict(tol=1e-10, max_iters=20)
    steps, stop, P = 28, 0.28, 10
    psi, rec, st = nosh.continuation_surrogate(
        model, psi0, {"g": 1.0}, "mu", 0.0, stop, steps,
        surrogate=dict(train_after=P, train_steps=400, batch=4096, lr=2e-3, newton_budget=4),
        **kw)
    print("newton", [r["newton_iters"] for r in rec])
    print("seeded", st["steps_seeded"], "fallbacks", st["fallbacks"], "loss", st["final_loss"],
          "train s", st["train_seconds"])
    print("pred/tan errors", [(r["predictor_used"], round(r["pred_error"], 7),
                               round(r["tan_error"], 7)) for r in rec[P + 1:]])
    assert len(rec) == steps + 1 and all(r["converged"] for r in rec)
    assert st["trained"] == 1
    om = O.Model(O.Mesh(coords, cells), mg.magnetic_potential(coords))
    om.set_params(1.0, stop)
    assert np.linalg.norm(om.f(psi)) <= 1e-9
    # Every step after training evaluates the surrogate (its guess error is
    # recorded apart from the tangent's); Newton starts from it when its
    # residual is at most half the tangent guess's.
    probed = [r for r in rec[1:] if r["pred_error"] != r["tan_error"]]
    assert len(probed) >= steps - P - 2
    ratio = np.mean([r["pred_error"] / r["tan_error"] for r in probed])
    print("mean pred/tan error ratio", ratio)
    assert ratio < 0.9
    # Same run with the tangent predictor only, for the Newton count.
    _, rec_t = nosh.continuation(model, psi0, {"g": 1.0}, "mu", 0.0, stop, steps, predictor=1,
                                 **kw)
    n_sur = sum(r["newton_iters"] for r in rec[P + 1:])
    n_tan = sum(r["newton_iters"] for r in rec_t[P + 1:])
    print("newton after training: surrogate", n_sur, "tangent", n_tan)
    assert n_sur <= n_tan


def test_surrogate_continuation_on_disk():
    """The disk workload at small size: Delaunay disk of radius 5 (h = 0.2,
    about 2.3 k vertices, seed 3), mu from 0 to 0.25 in 25 steps from psi = 1
    with the surrogate trained on the first 10 solutions. Every step
    converges, the final state solves F = 0 (oracle), the surrogate's guesses
    are closer than the tangent's, and Newton needs no more steps than with
    the tangent predictor alone."""
    import nosh_amd as nosh
    from oracle import oracle as O
    coords, cells = mg.disk(0.2)
    n = coords.shape[0]
    mesh, model = _model(coords, cells)
    psi0 = np.zeros(2 * n)
    psi0[0::2] = 1.0
    kw = dict(tol=1e-10, max_iters=20)
    steps, stop, P = 25, 0.25, 10
    psi, rec, st = nosh.continuation_surrogate(
        model, psi0, {"g": 1.0}, "mu", 0.0, stop, steps,
        surrogate=dict(train_after=P, train_steps=400, batch=4096, lr=2e-3, newton_budget=4),
        **kw)
    print("newton", [r["newton_iters"] for r in rec], "stats", st)
    assert len(rec) == steps + 1 and all(r["converged"] for r in rec)
    om = O.Model(O.Mesh(coords, cells), mg.magnetic_potential(coords))
    om.set_params(1.0, stop)
    assert np.linalg.norm(om.f(psi)) <= 1e-9
    probed = [r for r in rec[1:] if r["pred_error"] != r["tan_error"]]
    assert len(probed) >= steps - P - 3
    ratio = np.mean([r["pred_error"] / r["tan_error"] for r in probed])
    print("mean pred/tan error ratio", ratio, "seeded", st["steps_seeded"])
    assert ratio < 0.9
    _, rec_t = nosh.continuation(model, psi0, {"g": 1.0}, "mu", 0.0, stop, steps, predictor=1,
                                 **kw)
    assert sum(r["newton_iters"] for r in rec[P + 1:]) <= sum(r["newton_iters"]
                                                             for r in rec_t[P + 1:])
