"""GPU: batched continuation (branches stepped in mu together), checked with
the oracle and against the single-instance FP64 continuation."""
import numpy as np
import pytest

from nosh_amd import meshgen as mg

pytestmark = pytest.mark.gpu

MU0 = np.array([0.0, 0.04, 0.08, 0.12])
G = np.array([1.0, 0.8, 1.25, 1.0])
KW = dict(tol=1e-10, max_iters=40, minres_rtol=1e-10, minres_max_iters=4000, forcing=1,
          forcing_min=1e-6, forcing_max=0.1)


def _setup(ncell=8):
    import nosh_amd as nosh
    c, t = mg.kuhn_cube(ncell)
    mesh = nosh.Mesh(c, t)
    mvp = nosh.VectorFieldExplicitValues(mesh, mg.magnetic_potential(c), 0.0)
    model = nosh.Nls(mesh, mvp, nosh.ScalarFieldConstant(mesh, -1.0), 1.0,
                     nosh.ScalarFieldConstant(mesh, 1.0))
    n = c.shape[0]
    psi0 = []
    for g in G:
        p = np.zeros(2 * n)
        p[0::2] = 1.0 / np.sqrt(g)
        psi0.append(p)
    batch = nosh.Batch([(model, len(G))], MU0, G)
    return c, t, model, batch, np.concatenate(psi0)


def _oracle_check(c, t, mu, psis, tol=1e-9):
    from oracle import oracle as O
    om = O.Model(O.Mesh(c, t), mg.magnetic_potential(c))
    for b, p in enumerate(psis):
        om.set_params(G[b], mu[b])
        assert np.linalg.norm(om.f(p)) <= tol, b


@pytest.mark.parametrize("predictor", ["auto", "tangent", "secant"])
def test_batch_continuation_matches_single_instance(predictor):
    """Four (mu_0, g) branches over span 0.2 in 10 steps: every branch ends on
    a solution of F = 0 (oracle) and on the state the single-instance FP64
    tangent continuation reaches (up to the global phase)."""
    import nosh_amd as nosh
    from conftest import phase_align
    c, t, model, batch, psi0 = _setup()
    span, steps = 0.2, 10
    psi, rec, st = batch.continuation(psi0, span, steps, predictor=predictor, **KW)
    print("newton", [r["newton_iters"] for r in rec], "minres", [r["minres_iters"] for r in rec])
    assert st is None
    assert len(rec) == steps + 1 and all(r["converged"] == len(G) for r in rec)
    assert max(r["max_norm_f"] for r in rec) <= 1e-10
    finals = batch.split(psi)
    _oracle_check(c, t, MU0 + span, finals)
    for b in range(len(G)):
        p, r1 = nosh.continuation(model, np.asarray(psi0[b * len(finals[b]):(b + 1) * len(finals[b])]),
                                  {"g": G[b]}, "mu", MU0[b], MU0[b] + span, steps, predictor=1,
                                  tol=1e-10, max_iters=40)
        assert all(r["converged"] for r in r1)
        d = np.linalg.norm(phase_align(finals[b], p) - p) / np.linalg.norm(p)
        print("branch", b, "rel diff vs single-instance", d)
        assert d < 1e-8


def test_batch_tangent_guess_halves_the_secant_error():
    """On smooth steps of equal size h the secant guess psi_k + (psi_k -
    psi_{k-1}) errs by h^2 psi'' + O(h^3) and the tangent guess psi_k +
    h psi' by h^2 psi'' / 2 + O(h^3): the recorded guess errors (the
    distance of each guess to the converged state, mean over the branches)
    of the tangent run are about half the secant run's, on steps 2-6 of the
    same branches (mu <= 0.24; on later steps of this run psi'' changes
    from step to step, and the ratio leaves 1/2: 0.55 and 0.6 at steps 7
    and 8, 0.96 and 0.28 at steps 9 and 10). Both runs end on the same
    solutions (oracle), up to the global phase."""
    from conftest import phase_align
    c, t, model, batch, psi0 = _setup()
    span, steps = 0.2, 10
    errs, finals = {}, {}
    for pred in ("tangent", "secant"):
        psi, rec, _ = batch.continuation(psi0, span, steps, predictor=pred, **KW)
        assert all(r["converged"] == len(G) for r in rec)
        errs[pred] = np.array([r["sec_error"] for r in rec[2:7]])
        finals[pred] = batch.split(psi)
    q = errs["tangent"] / errs["secant"]
    print("tangent / secant guess error per step", np.round(q, 3))
    assert np.all(q > 0.45) and np.all(q < 0.6)
    for a, b in zip(finals["tangent"], finals["secant"]):
        assert np.linalg.norm(phase_align(a, b) - b) <= 1e-8 * np.linalg.norm(b)
    _oracle_check(c, t, MU0 + span, finals["tangent"])


def test_batch_continuation_surrogate_seeds_and_gains():
    """Same branches over span 0.24 in 24 steps (secant predictor) with the
    surrogate trained on the first 8 steps: on the steps where it is evaluated its guesses
    are closer to the converged states than the secant's, and every branch
    ends on a solution of F = 0 (oracle)."""
    c, t, model, batch, psi0 = _setup()
    span, steps, P = 0.24, 24, 8
    psi, rec, st = batch.continuation(
        psi0, span, steps,
        surrogate=dict(train_after=P, train_steps=400, batch=4096, lr=2e-3, newton_budget=4),
        predictor="secant", **KW)
    print("stats", st)
    print("newton", [r["newton_iters"] for r in rec])
    print("pred/sec", [(r["seeded"], round(r["pred_error"], 7), round(r["sec_error"], 7))
                       for r in rec[P + 2:]])
    assert len(rec) == steps + 1 and all(r["converged"] == len(G) for r in rec)
    assert st["trained"] == 1
    # Steps where the surrogate was evaluated (its guess error is recorded
    # apart from the secant's); branches start from it only when its residual
    # is at most half the secant guess's.
    probed = [r for r in rec if r["pred_error"] != r["sec_error"]]
    assert len(probed) >= 3
    ratio = np.mean([r["pred_error"] / r["sec_error"] for r in probed])
    print("probed steps", len(probed), "seeded branch-steps", sum(r["seeded"] for r in rec),
          "mean pred/secant error ratio", ratio)
    assert ratio < 0.9
    _oracle_check(c, t, MU0 + span, batch.split(psi))


def test_surrogate_that_never_wins_leaves_the_run_unchanged():
    """A surrogate whose guesses are never accepted (accept_ratio 1e-12) is
    trained and probed, and every branch still takes its short try from the
    predictor's guess: the same Newton and MINRES counts at every step and
    the same final states, bit for bit, as the run without a surrogate."""
    c, t, model, batch, psi0 = _setup()
    span, steps, P = 0.24, 16, 6
    runs = {}
    for label, sur in (("plain", None),
                       ("probed", dict(train_after=P, train_steps=100, batch=4096, lr=2e-3,
                                       newton_budget=4, accept_ratio=1e-12))):
        print(f"[surrogate never wins: {label}]", flush=True)
        psi, rec, st = batch.continuation(psi0, span, steps, surrogate=sur, predictor="auto", **KW)
        assert all(r["converged"] == len(G) for r in rec)
        runs[label] = (psi, [(r["newton_iters"], r["minres_iters"]) for r in rec], rec, st)
    assert runs["probed"][3]["trained"] == 1 and runs["probed"][3]["steps_seeded"] == 0
    assert any(r["pred_error"] != r["sec_error"] for r in runs["probed"][2])  # it was probed
    assert runs["probed"][1] == runs["plain"][1]
    assert np.array_equal(runs["probed"][0], runs["plain"][0])


def test_batch_continuation_halves_steps_for_failing_branches():
    """A large step (0 -> 0.1 in mu) with a 3-step Newton limit: the step
    fails from every guess in the lockstep tries and in the single-instance
    FP64 Newton fallback, so the branches are split into halved sub-steps
    (the FP64 tangent continuation, batched sub-steps after it), and every
    branch ends on a solution of F = 0 (oracle)."""
    import nosh_amd as nosh
    import torch
    # Device memory freed with NaN in it: the continuation's buffers must not read
    # anything they did not write (the first secant once computed
    # psi_k + 0 * (psi_k - stale) = NaN in reused memory).
    junk = torch.full((64 << 20,), float("nan"), dtype=torch.float64, device="cuda")
    del junk
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    c, t, model, _, psi0 = _setup()
    mu0 = np.zeros(len(G))  # psi_0 = 1/sqrt(g) is exact at mu = 0
    batch = nosh.Batch([(model, len(G))], mu0, G)
    kw = dict(tol=1e-10, max_iters=3, minres_rtol=1e-12, minres_max_iters=4000, forcing=0)
    span, steps = 0.1, 1
    psi, rec, _ = batch.continuation(psi0, span, steps, predictor="secant", **kw)
    print([(r["newton_iters"], r["halvings"], r["converged"]) for r in rec])
    assert len(rec) == steps + 1 and all(r["converged"] == len(G) for r in rec)
    assert max(r["halvings"] for r in rec) >= 1
    from oracle import oracle as O
    om = O.Model(O.Mesh(c, t), mg.magnetic_potential(c))
    for b, p in enumerate(batch.split(psi)):
        om.set_params(G[b], span)
        assert np.linalg.norm(om.f(p)) <= 1e-9, b
