"""GPU: the batched solver (uneven meshes, per-instance mu/g, FP32 inner
MINRES) — every instance's solution checked with the oracle."""
import numpy as np
import pytest

from nosh_amd import meshgen as mg

pytestmark = pytest.mark.gpu


def _model(coords, cells):
    import nosh_amd as nosh
    mesh = nosh.Mesh(coords, cells)
    mvp = nosh.VectorFieldExplicitValues(mesh, mg.magnetic_potential(coords), 0.0)
    V = nosh.ScalarFieldConstant(mesh, -1.0)
    t = nosh.ScalarFieldConstant(mesh, 1.0)
    return nosh.Nls(mesh, mvp, V, 1.0, t)


@pytest.fixture(scope="module")
def setup():
    import nosh_amd as nosh
    meshes = [mg.kuhn_cube(6, jitter=0.2, seed=2), mg.square_tri(16), mg.kuhn_cube(4)]
    counts = [3, 5, 2]
    models = [_model(c, t) for c, t in meshes]
    u = mg.splitmix64_uniform(2, 2 * sum(counts)).reshape(-1, 2)
    mu = u[:, 0]
    g = 0.5 + 1.5 * u[:, 1]
    # Starts near the superconducting branch (psi = 1 plus 10 % seeded noise):
    # Newton from fully random phases on these coarse meshes is chaotic, and
    # these tests are about the batching, not Newton's global behaviour.
    psi0 = []
    k = 0
    for (c, t), cnt in zip(meshes, counts):
        for _ in range(cnt):
            p = 0.1 * mg.random_state(c.shape[0], 100 + k)
            p[0::2] += 1.0
            psi0.append(p)
            k += 1
    return meshes, counts, models, mu, g, np.concatenate(psi0)


def _check_with_oracle(meshes, counts, mu, g, sols, tol=1e-9):
    from oracle import oracle as O
    k = 0
    for (c, t), cnt in zip(meshes, counts):
        om = O.Model(O.Mesh(c, t), mg.magnetic_potential(c))
        for _ in range(cnt):
            om.set_params(g[k], mu[k])
            assert np.linalg.norm(om.f(sols[k])) <= tol, k
            k += 1


@pytest.mark.parametrize("threads,fold", [(128, 1), (256, 1), (256, 4)])
def test_batched_graph_and_persistent_agree(setup, threads, fold):
    import nosh_amd as nosh
    meshes, counts, models, mu, g, psi0 = setup
    batch = nosh.Batch(list(zip(models, counts)), mu, g)
    # Random starts: Newton needs 20-50 steps; with inner MINRES to 1e-6 the
    # batched counts match the single-instance solver's.
    opts = dict(tol=1e-10, max_iters=80, minres_rtol=1e-6, minres_max_iters=4000, cheb_iters=4,
                cheb_ratio=100.0, forcing=0)
    out = {}
    for mode in (0, 1):
        print(f"[batched threads={threads} fold={fold} mode={mode}]", flush=True)
        batch.set_tuning(threads, mode, 4, 16, fold=fold)
        psi, st = batch.solve(psi0, **opts)
        assert all(s["converged"] for s in st), st
        assert max(s["final_norm_f"] for s in st) <= 1e-10
        _check_with_oracle(meshes, counts, mu, g, batch.split(psi))
        out[mode] = psi
    # Same phases, same deterministic reductions: bitwise identical.
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("threads,fold", [(128, 1), (256, 4)])
def test_two_lanes_per_thread_match_one(setup, threads, fold):
    """Mode 2 (every thread of the gathering kernels takes 2 lanes of a row)
    against mode 0 (one lane per thread) on groups of 3 (padded to 4 lanes),
    5 (padded to 8) and 2 instances: the same arithmetic per row and lane and
    the same partial sums, so bitwise identical; a batch with a 1-lane group
    refuses mode 2 and stays in mode 0."""
    import nosh_amd as nosh
    meshes, _, models, mu, g, _ = setup
    counts = [3, 5, 2]
    u = mg.splitmix64_uniform(11, 2 * sum(counts)).reshape(-1, 2)
    mu, g = 0.3 * u[:, 0], 0.6 + 0.8 * u[:, 1]
    psi0 = []
    for (c, t), cnt in zip(meshes, counts):
        for q in range(cnt):
            p = 0.1 * mg.random_state(c.shape[0], 700 + len(psi0))
            p[0::2] += 1.0
            psi0.append(p)
    psi0 = np.concatenate(psi0)
    batch = nosh.Batch(list(zip(models, counts)), mu, g)
    opts = dict(tol=1e-10, max_iters=60, minres_rtol=1e-6, minres_max_iters=4000, cheb_iters=4,
                cheb_ratio=100.0, forcing=0)
    out = {}
    for mode in (0, 2):
        print(f"[batched threads={threads} fold={fold} mode={mode}]", flush=True)
        batch.set_tuning(threads, mode, 4, 16, fold=fold)
        psi, st = batch.solve(psi0, **opts)
        assert all(s["converged"] for s in st), st
        out[mode] = (psi, [(s["newton_iters"], s["minres_iters"]) for s in st])
    assert out[0][1] == out[2][1]
    assert np.array_equal(out[0][0], out[2][0])
    _check_with_oracle(meshes, counts, mu, g, batch.split(out[2][0]))
    b1 = nosh.Batch(list(zip(models, [3, 5, 1])), mu[:9], g[:9])
    with pytest.raises(nosh.NoshError):
        b1.set_tuning(256, 2, 4, 16)


def test_two_lanes_per_thread_match_one_wide_groups(setup):
    """Mode 2 against mode 0 on groups of 12 and 9 instances (16 lanes: 8
    threads per row, two quads that each share the slot records by DPP, the
    threads of lanes 12-15 and 9-15 with no instance). Instances finish their
    MINRES solves at different iterations, so quads also run with threads
    whose lanes are all done. Bitwise identical results and counts."""
    import nosh_amd as nosh
    meshes, _, models, _, _, _ = setup
    sub = [(meshes[1], models[1], 12), (meshes[2], models[2], 9)]
    counts = [k for _, _, k in sub]
    u = mg.splitmix64_uniform(23, 2 * sum(counts)).reshape(-1, 2)
    mu, g = 0.3 * u[:, 0], 0.6 + 0.8 * u[:, 1]
    psi0 = []
    for (c, t), _, cnt in sub:
        for q in range(cnt):
            p = 0.1 * mg.random_state(c.shape[0], 900 + len(psi0))
            p[0::2] += 1.0
            psi0.append(p)
    psi0 = np.concatenate(psi0)
    batch = nosh.Batch([(m, k) for _, m, k in sub], mu, g)
    opts = dict(tol=1e-10, max_iters=60, minres_rtol=1e-6, minres_max_iters=4000, cheb_iters=4,
                cheb_ratio=100.0, forcing=1, forcing_min=1e-6, forcing_max=0.1)
    out = {}
    for mode in (0, 2):
        print(f"[batched wide groups mode={mode}]", flush=True)
        batch.set_tuning(256, mode, 4, 16, fold=2)
        psi, st = batch.solve(psi0, **opts)
        assert all(s["converged"] for s in st), st
        out[mode] = (psi, [(s["newton_iters"], s["minres_iters"]) for s in st])
    assert out[0][1] == out[2][1]
    assert len({m for _, m in out[2][1]}) > 1  # the instances did not all stop together
    assert np.array_equal(out[0][0], out[2][0])
    _check_with_oracle([m for m, _, _ in sub], counts, mu, g, batch.split(out[2][0]))


def test_batched_autotune(setup, tmp_path, monkeypatch):
    import nosh_amd as nosh
    monkeypatch.setenv("NOSH_AUTOTUNE_CACHE", str(tmp_path / "tune.tsv"))
    meshes, counts, models, mu, g, psi0 = setup
    batch = nosh.Batch(list(zip(models, counts)), mu, g)
    t, m, f = batch.autotune(psi0, cheb_iters=4, cheb_ratio=100.0)
    assert t in (128, 256, 512, 1024) and m in (0, 1, 2) and f in (1, 2, 4, 8)
    assert (tmp_path / "tune.tsv").read_text().strip().endswith(f"{t} {f} {m}")
    assert batch.autotune(psi0, cheb_iters=4, cheb_ratio=100.0) == (t, m, f)  # from the cache
    psi, st = batch.solve(psi0, tol=1e-10, max_iters=80, minres_rtol=1e-6, cheb_iters=4,
                          cheb_ratio=100.0, forcing=0)
    assert all(s["converged"] for s in st)
    _check_with_oracle(meshes, counts, mu, g, batch.split(psi))


def test_persistent_mode_large_grid_matches_graph():
    """Persistent mode on a batch whose row tiles outnumber its grid (one or
    two workgroups per CU, each striding over the tiles): bitwise the graph
    mode's result. With the barrier wait limit forced to 0 from the host, the
    barriers time out, every workgroup leaves the kernel and the solve
    reports an error instead of hanging; the batch then solves normally."""
    import os
    import nosh_amd as nosh
    c, t = mg.kuhn_cube(32)
    m = _model(c, t)
    n = c.shape[0]
    mu = np.linspace(0.02, 0.16, 8)
    batch = nosh.Batch([(m, 8)], mu, np.ones(8))
    psi0 = []
    for b in range(8):
        p = 0.1 * mg.random_state(n, 300 + b)
        p[0::2] += 1.0
        psi0.append(p)
    psi0 = np.concatenate(psi0)
    opts = dict(tol=1e-10, max_iters=60, minres_rtol=1e-6, minres_max_iters=2000, cheb_iters=4,
                cheb_ratio=100.0, forcing=0)
    out = {}
    # 512 threads: the persistent kernel's blocks are twice the size of the
    # finalize kernels' (whose sums of the > 256 block partials must not
    # depend on it); mode 2 runs the two-lane kernels.
    for bt, modes in ((512, (0, 1, 2)), (256, (0, 1))):
        for mode in modes:
            print(f"[persistent large grid: mode {mode}, {bt} threads, 35937 rows x 8 lanes]",
                  flush=True)
            batch.set_tuning(bt, mode, 4, 16, fold=1)
            psi, st = batch.solve(psi0, **opts)
            assert all(s["converged"] for s in st), st
            out[bt, mode] = psi
        for mode in modes[1:]:
            assert np.array_equal(out[bt, 0], out[bt, mode]), (bt, mode)
    out[0] = out[256, 0]
    print("[persistent large grid: forced barrier time-out]", flush=True)
    os.environ["NOSH_BARRIER_TIMEOUT_S"] = "0"
    try:
        with pytest.raises(nosh.NoshError):
            batch.solve(psi0, **opts)
    finally:
        del os.environ["NOSH_BARRIER_TIMEOUT_S"]
    psi, st = batch.solve(psi0, **opts)
    assert all(s["converged"] for s in st)
    assert np.array_equal(psi, out[0])


def test_full_group_matches_padded_group():
    """A full group (8 instances, 8 lanes: slot metadata shared by lane
    shuffles) and a padded one (the first 7 of them, 8 lanes: every lane
    loads its own) take the same Newton steps and reach the same solutions
    for the common instances, up to the global phase (the two code paths may
    contract FMAs differently, so not bitwise), and every solution passes
    the oracle."""
    import nosh_amd as nosh
    c, t = mg.kuhn_cube(6, jitter=0.2, seed=2)
    m = _model(c, t)
    n = c.shape[0]
    u = mg.splitmix64_uniform(5, 16).reshape(8, 2)
    # Below the first vortex entry, where Newton from psi ~ 1 converges quickly.
    mu, g = 0.25 * u[:, 0], 0.8 + 0.4 * u[:, 1]
    psi0 = []
    for k in range(8):
        p = 0.1 * mg.random_state(n, 300 + k)
        p[0::2] += 1.0
        psi0.append(p)
    opts = dict(tol=1e-10, max_iters=40, minres_rtol=1e-6, minres_max_iters=4000, cheb_iters=4,
                cheb_ratio=100.0, forcing=0)
    b8 = nosh.Batch([(m, 8)], mu, g)
    b7 = nosh.Batch([(m, 7)], mu[:7], g[:7])
    for b in (b8, b7):
        b.set_tuning(256, 0, 4, 16, fold=2)
    p8, s8 = b8.solve(np.concatenate(psi0), **opts)
    p7, s7 = b7.solve(np.concatenate(psi0[:7]), **opts)
    print("full group", [(s["newton_iters"], s["converged"]) for s in s8])
    print("padded group", [(s["newton_iters"], s["converged"]) for s in s7])
    assert all(s["converged"] for s in s8) and all(s["converged"] for s in s7)
    assert [s["newton_iters"] for s in s8[:7]] == [s["newton_iters"] for s in s7]
    # Rounding differences move the solution along the gauge direction i psi
    # (J's null space), which F does not see: compare up to the global phase.
    from conftest import phase_align
    for a_, b_ in zip(b8.split(p8)[:7], b7.split(p7)):
        assert np.abs(phase_align(a_, b_) - b_).max() <= 1e-9
    _check_with_oracle([(c, t)], [8], mu, g, b8.split(p8))


def test_shared_mu_slot_values_match_per_lane_rebuild():
    """When every instance of a full group shares mu, the row products read
    the off-diagonal values precomputed once per slot instead of rebuilding
    them per lane; both are the same FP32 operations, so instances 0-6 of a
    group sharing mu (precomputed values) and of one whose instance 7 has
    another mu (per-lane rebuild) take identical steps to identical bits."""
    import nosh_amd as nosh
    c, t = mg.kuhn_cube(6, jitter=0.2, seed=2)
    m = _model(c, t)
    n = c.shape[0]
    u = mg.splitmix64_uniform(7, 8)
    g = 0.8 + 0.4 * u
    psi0 = []
    for k in range(8):
        p = 0.1 * mg.random_state(n, 500 + k)
        p[0::2] += 1.0
        psi0.append(p)
    psi0 = np.concatenate(psi0)
    opts = dict(tol=1e-10, max_iters=40, minres_rtol=1e-6, minres_max_iters=4000, cheb_iters=4,
                cheb_ratio=100.0, forcing=0)
    mu_shared = np.full(8, 0.2)
    mu_mixed = mu_shared.copy()
    mu_mixed[7] = 0.21
    sols, stats = [], []
    for mu in (mu_shared, mu_mixed):
        b = nosh.Batch([(m, 8)], mu, g)
        b.set_tuning(256, 0, 4, 16, fold=2)
        p, st = b.solve(psi0, **opts)
        assert all(s["converged"] for s in st)
        sols.append(b.split(p))
        stats.append(st)
    for k in range(7):
        assert stats[0][k]["newton_iters"] == stats[1][k]["newton_iters"]
        assert stats[0][k]["minres_iters"] == stats[1][k]["minres_iters"]
        assert np.array_equal(sols[0][k], sols[1][k]), k
    _check_with_oracle([(c, t)], [8], mu_shared, g, sols[0])
    _check_with_oracle([(c, t)], [8], mu_mixed, g, sols[1])


def test_cheb_step_timing_leaves_the_batch_usable(setup):
    """The measurement entry point times the batched Chebyshev step with every
    instance active, reports its algorithmic bytes (12 per padded slot + 44
    per row and lane at Chebyshev degree 4), and a solve after it still
    converges (oracle)."""
    import nosh_amd as nosh
    meshes, counts, models, mu, g, psi0 = setup
    batch = nosh.Batch(list(zip(models, counts)), mu, g)
    opts = dict(tol=1e-10, max_iters=80, minres_rtol=1e-6, cheb_iters=4, cheb_ratio=100.0,
                forcing=0)
    batch.solve(psi0, **opts)
    us, by = batch.time_cheb_step(20)
    lanes = [1 << max(0, (c - 1).bit_length()) for c in counts]
    expect = sum(12 * m.mesh.layout()["nnz_pad"] + 44 * m.mesh.num_vertices * bp
                 for m, bp in zip(models, lanes))
    print("kb_cheb_step", us, "us", by, "bytes")
    assert us > 0 and by == expect
    psi, st = batch.solve(psi0, **opts)
    assert all(s["converged"] for s in st)
    _check_with_oracle(meshes, counts, mu, g, batch.split(psi))


def test_random_start_batch_compacts_and_converges(capfd, monkeypatch):
    """24 random-start instances on two small meshes: Newton from random
    phases is chaotic (tens of steps for some instances), so the batch
    compacts the last quarter into a batch of their own; every instance ends
    on an oracle solution and its statistics add up over the phases."""
    import nosh_amd as nosh
    monkeypatch.setenv("NOSH_PROGRESS", "1")
    meshes = [mg.kuhn_cube(4, jitter=0.2, seed=2), mg.square_tri(12)]
    counts = [12, 12]
    models = [_model(c, t) for c, t in meshes]
    u = mg.splitmix64_uniform(7, 2 * sum(counts)).reshape(-1, 2)
    mu, g = u[:, 0], 0.5 + 1.5 * u[:, 1]
    psi0 = np.concatenate([mg.random_state(c.shape[0], 300 + k)
                           for (c, t), cnt in zip(meshes, counts) for k in range(cnt)])
    batch = nosh.Batch(list(zip(models, counts)), mu, g)
    psi, st = batch.solve(psi0, tol=1e-10, max_iters=120, minres_rtol=1e-6, cheb_iters=4,
                          cheb_ratio=100.0, forcing=0)
    err = capfd.readouterr().err
    print([s["newton_iters"] for s in st])
    assert "compacted 24 ->" in err
    assert all(s["converged"] for s in st), st
    assert all(s["newton_iters"] >= 1 and s["minres_iters"] >= s["newton_iters"] for s in st)
    _check_with_oracle(meshes, counts, mu, g, batch.split(psi))
