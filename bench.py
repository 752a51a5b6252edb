"""Benchmark: Newton-MINRES solves/sec at ||F||_2 <= 1e-10 on the 96^3 Kuhn
cube, mixed-precision batched surrogate-seeded continuation against plain
FP64 GPU Newton-MINRES; and the Jacobian SpMV's HBM GB/s.

Mesh: cube [-5,5]^3, 96^3 cells split into 6 Kuhn tetrahedra each (912,673
vertices, 5,308,416 tets, 6,276,384 edges); A = mu*0.5*(-y, x, 0); V = -1;
thickness 1.

Headline workload (``value``; one step = one continuation step of every
branch): mu-continuation from mu = 0 to 1 in ``--steps`` equal steps (default
50) of 8 branches (mu_0 = 0, g_b = 0.6 + 0.2 b,
psi_0 = 1/sqrt(g_b), exact at mu = 0), each step solved to ||F|| <= 1e-10.
Combined method: the branches batched in one instance-interleaved state (one
launch per phase, hipGraph-captured MINRES iterations, autotuned block size),
FP32 inner MINRES with FP64 Newton (Eisenstat-Walker forcing), the tangent
predictor (one lockstep solve of J dpsi/dmu = -dF/dmu for all branches) on the
first step and after a step that a branch did not pass from its first guess,
the secant otherwise (``--cont-predictor`` tangent / secant for one of them
throughout), lockstep retries at branch transitions, and the BF16 MFMA MLP surrogate trained online on the run's own
solutions seeding Newton where its guess has the smaller residual (the
predictor guess otherwise). ``value`` = 8 x steps / wall time of the run
(training and inference included). Plain baseline: the single-instance FP64
continuation (MINRES rtol 1e-10, tangent predictor, Chebyshev-Jacobi(6, 300),
its fastest setting) of every branch; the ratio is value / its solves/sec.

Side by side (the 96^3 random-start solve): plain FP64 and FP32-inner
Newton-MINRES from psi_0 seeded random (|psi| ~ U(0.5,1), arg ~ U(0, 2 pi),
seed 0) at mu = 0.4, g = 1, with ||F|| and ||psi - psi_FP64||.

Roofline: the FP64 Jacobian apply J(psi) phi with cold caches (a 1 GiB read
before each launch), its duration from the start/end events the HIP runtime
stamps on each dispatch on the launch stream; algorithmic bytes per launch
40E + 68n (complex CSR storing both directions: 16 B value + 4 B column per
off-diagonal entry, 4 B row pointer and 8 B diagonal per row, phi / y / psi
16 B and c*t 8 B per vertex) against 8 TB/s, beside the library's own
streaming-kernel ceiling.

CPU baseline: the oracle (a C++ restatement of the reference operators and
solver, OpenMP) on a bounded sample on this host: the cost of F, the
Jacobian rebuild and MINRES iterations is measured and combined with the
plain continuation's own per-step Newton / MINRES counts.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Newton-MINRES solves/sec at ||F||<=1e-10 vs plain FP64 GPU; SpMV HBM GB/s"
# Iteration modes of the batched solver (nosh_batch_set_tuning).
MODE_NAMES = ("hipGraph", "persistent", "hipGraph, 2 lanes per thread")


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=None,
                   help="timed steps: continuation steps (default 50; every branch solved "
                        "once per step), or batch solves for --workload batched (default 1)")
    p.add_argument("--warmup", type=int, default=None,
                   help="untimed steps before the timed run (default 2 continuation steps, "
                        "1 batch solve)")
    p.add_argument("--ncell", type=int, default=96)
    p.add_argument("--mu", type=float, default=0.4, help="mu of the random-start solves")
    p.add_argument("--solve-reps", type=int, default=2,
                   help="random-start solves timed per variant")
    p.add_argument("--cheb", type=int, default=6, help="Chebyshev degree, plain FP64 solver")
    p.add_argument("--cont-cheb", type=int, default=5, help="Chebyshev degree, batched solver")
    p.add_argument("--ratio", type=float, default=300.0)
    p.add_argument("--cont-ratio", type=float, default=300.0,
                   help="Chebyshev interval ratio, batched solver")
    p.add_argument("--cont-forcing-max", type=float, default=0.1)
    p.add_argument("--minres-maxit", type=int, default=2000)
    p.add_argument("--spmv-reps", type=int, default=50)
    p.add_argument("--variants", default="plain,mixed")
    p.add_argument("--roofline-only", action="store_true",
                   help="only the roofline measurements (for profilers)")
    p.add_argument("--no-roofline-160", dest="roofline_160", action="store_false",
                   help="skip the 160^3 J.phi timing")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--verbose", action="store_true")
    p.add_argument("--cont-branches", type=int, default=8)
    p.add_argument("--cont-span", type=float, default=1.0)
    p.add_argument("--cont-train-after", type=int, default=10)
    p.add_argument("--cont-retrain", type=int, default=0,
                   help="Adam steps per continuation step after the first training")
    p.add_argument("--cont-batch", type=int, default=8192, help="surrogate minibatch rows")
    p.add_argument("--cont-train-steps", type=int, default=150,
                   help="Adam steps of the surrogate's first training")
    p.add_argument("--cont-plain-branches", type=int, default=8,
                   help="branches run with the plain FP64 continuation for the ratio")
    p.add_argument("--no-ablation", action="store_true",
                   help="skip the single-instance ablation rows")
    p.add_argument("--no-persistent", action="store_true",
                   help="skip the persistent-vs-graph MINRES comparison")
    p.add_argument("--no-extras", action="store_true",
                   help="only the timed continuation (no second step size, no surrogate-free "
                        "run, no ablation; for profilers)")
    p.add_argument("--cont-tuning", default="",
                   help="block_threads,fold[,mode] for the batched continuation (skips the "
                        "autotuner; mode 0 by default)")
    p.add_argument("--no-surrogate", action="store_true")
    p.add_argument("--cont-predictor", default="auto", choices=["auto", "tangent", "secant"],
                   help="predictor of the batched continuation (auto: the tangent on the first "
                        "step and after a step a branch did not pass from its first guess, the "
                        "secant otherwise)")
    p.add_argument("--cont-smooth", type=int, default=4,
                   help="Jacobi sweeps on the residual of each surrogate guess")
    p.add_argument("--cont-accept", type=float, default=0.5,
                   help="surrogate guess used when ||F|| <= this times the secant guess's")
    p.add_argument("--workload", default="cont96",
                   choices=["cont96", "batched", "cont", "disk", "stability"],
                   help="cont96: the headline (continuation, side-by-side solves, roofline, "
                        "CPU baseline); cont: the continuation part only; batched: 128 uneven "
                        "random-start instances on jittered 16^3/24^3/32^3/48^3 cubes; disk: "
                        "surrogate continuation on the 200 k-vertex disk; stability: "
                        "eigenvalues of J and arclength steps on the cube")
    p.add_argument("--instances", type=int, default=128)
    p.add_argument("--plain-sample", type=int, default=8,
                   help="batched: instances solved one by one with plain FP64 for the ratio")
    a = p.parse_args()
    if a.steps is None:
        a.steps = {"batched": 1, "disk": 200, "stability": 12}.get(a.workload, 50)
    if a.warmup is None:
        a.warmup = 1 if a.workload == "batched" else 2
    return a


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def _profiled_traffic(ncell):
    """HBM bytes per J.phi launch from the committed rocprofv3 PMC summary of
    this workload (FETCH_SIZE + WRITE_SIZE, calibrated), if there is one."""
    path = os.path.join(ROOT, "profiles", f"jphi_fp64_kuhn{ncell}.md")
    if not os.path.exists(path):
        return None, None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(float(line.split(":")[1])), os.path.relpath(path, ROOT)
    return None, None


def cpu_baseline(coords, cells, A, mu, psi0, phi, counts, cheb, ratio, what, model=None):
    """The oracle on this host's cores: one full FP64 Newton-MINRES solve at
    the benchmark's size (the first continuation step of branch g = 1: mu =
    `mu`, from psi = 1, ||F|| <= 1e-10, MINRES rtol 1e-10, Chebyshev(cheb)),
    timed end to end; ``value`` = 1 / its seconds. The GPU's plain FP64 solve
    of the same problem is timed beside it. Also reported: medians of 20 of
    F, K phi and J phi, and the per-solve cost extrapolated from per-operation
    medians and the plain GPU continuation's own per-solve counts."""
    import nosh_amd as nosh
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.perf_counter()
    om = O.Model(O.Mesh(coords, cells), A)
    om.set_params(1.0, mu)
    setup = time.perf_counter() - t0

    def med(fn, reps=5):
        ts = []
        for _ in range(reps):
            s = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - s)
        return float(np.median(ts))

    t_f = med(lambda: om.f(psi0), 20)
    t_k = med(lambda: om.keo(phi), 20)

    def rebuild():
        om.jac_rebuild(psi0)
        om.prec_rebuild(psi0)

    t_rb = med(rebuild, 3)
    t_j = med(lambda: om.jac(phi), 20)
    threads = O.set_threads(0)
    O.set_threads(1)
    t_j1 = med(lambda: om.jac(phi), 20)
    O.set_threads(threads)
    b = -om.f(psi0)
    k_sample = 20
    s = time.perf_counter()
    _, k, _ = om.minres(b, rtol=1e-30, maxit=k_sample, cheb_iters=cheb, ratio=ratio)
    t_it = (time.perf_counter() - s) / max(k, 1)
    steps = counts["newton_iters"]
    its = counts["minres_iters"]
    t_extrap = (steps + 1) * t_f + steps * t_rb + its * t_it
    # The full solve, CPU and GPU.
    n = coords.shape[0]
    start = np.zeros(2 * n)
    start[0::2] = 1.0
    s = time.perf_counter()
    pc, it_c, nf_c, mh_c = om.newton(start, 1.0, mu, tol=1e-10, maxit=40, minres_rtol=1e-10,
                                     minres_maxit=4000, cheb_iters=cheb, ratio=ratio)
    t_cpu = time.perf_counter() - s
    gpu = None
    if model is not None:
        kw = dict(tol=1e-10, max_iters=40, minres_rtol=1e-10, minres_max_iters=4000,
                  cheb_iters=cheb, cheb_ratio=ratio)
        nosh.newton_solve(model, start, {"g": 1.0, "mu": mu}, **kw)
        s = time.perf_counter()
        pg, sg = nosh.newton_solve(model, start, {"g": 1.0, "mu": mu}, **kw)
        gpu = {"seconds": time.perf_counter() - s, "newton": sg["newton_iters"],
               "minres": sg["minres_iters_total"],
               "rel_diff_vs_cpu": float(np.linalg.norm(_phase_align(pg, pc) - pc) /
                                        np.linalg.norm(pc))}
    sample = (f"oracle on {threads} threads ({_cpu_model()}, nproc {os.cpu_count()}, "
              f"OMP_NUM_THREADS {os.environ.get('OMP_NUM_THREADS', 'unset')}): one full FP64 "
              f"Newton-MINRES solve of the 96^3 cube at mu = {mu}, g = 1 from psi = 1 "
              f"(Chebyshev({cheb}) in place of MueLu AMG), {int(it_c)} Newton steps, "
              f"{int(np.sum(mh_c))} MINRES iterations, {t_cpu:.1f} s; setup {setup:.1f} s")
    say("cpu baseline: " + sample + (f"; GPU plain FP64 {gpu['seconds']:.3f} s" if gpu else ""))
    return {"value": 1.0 / t_cpu if it_c >= 0 else 0.0, "unit": "solves/s", "cores": threads,
            "kind": "port", "sample": sample, "seconds_per_solve": t_cpu,
            "cpu_newton": int(it_c), "cpu_minres": int(np.sum(mh_c)),
            "cpu_final_norm_f": float(nf_c[-1]) if len(nf_c) else None,
            "gpu_plain_same_solve": gpu,
            "jphi_ms": t_j * 1e3, "jphi_ms_1thread": t_j1 * 1e3, "f_ms": t_f * 1e3,
            "keo_ms": t_k * 1e3, "rebuild_ms": t_rb * 1e3, "minres_iter_ms": t_it * 1e3,
            "extrapolated_continuation_solve_s": t_extrap,
            "extrapolation_note": f"{steps + 1:.2f} F + {steps:.2f} rebuilds + {its:.1f} MINRES "
                                  f"iterations, the {what}'s mean per-solve counts",
            "cpu_model": _cpu_model(), "nproc": os.cpu_count()}


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def config0_solves(args):
    """The small plain case, 64x64 square of right triangles (mu = 0.1, g = 1,
    seeded SplitMix64 psi_0): one full FP64 Newton-MINRES solve to
    ||F|| <= 1e-10 on the GPU and with the oracle on the host, same settings."""
    import nosh_amd as nosh
    from nosh_amd import meshgen as mg
    from oracle import oracle as O
    c, t = mg.square_tri(64)
    n = c.shape[0]
    psi0 = mg.random_state(n, 0)
    mesh = nosh.Mesh(c, t)
    model = nosh.Nls(mesh, nosh.VectorFieldExplicitValues(mesh, mg.magnetic_potential(c), 0.0),
                     nosh.ScalarFieldConstant(mesh, -1.0), 1.0,
                     nosh.ScalarFieldConstant(mesh, 1.0))
    kw = dict(tol=1e-10, max_iters=30, minres_rtol=1e-10, minres_max_iters=1000,
              cheb_iters=args.cheb, cheb_ratio=args.ratio)
    nosh.newton_solve(model, psi0, {"g": 1.0, "mu": 0.1}, **kw)  # warm-up
    t0 = time.perf_counter()
    pg, sg = nosh.newton_solve(model, psi0, {"g": 1.0, "mu": 0.1}, **kw)
    t_gpu = time.perf_counter() - t0
    om = O.Model(O.Mesh(c, t), mg.magnetic_potential(c))
    t0 = time.perf_counter()
    pc, it, nf, mh = om.newton(psi0, 1.0, 0.1, tol=1e-10, maxit=30, minres_rtol=1e-10,
                               minres_maxit=1000, cheb_iters=args.cheb, ratio=args.ratio)
    t_cpu = time.perf_counter() - t0
    d = float(np.linalg.norm(_phase_align(pg, pc) - pc) / np.linalg.norm(pc))
    return {"mesh": f"square 64x64 (n = {n})", "gpu_seconds": t_gpu, "cpu_seconds": t_cpu,
            "cpu_threads": O.set_threads(0), "gpu_newton": sg["newton_iters"],
            "gpu_minres": sg["minres_iters_total"], "gpu_final_norm_f": sg["final_norm_f"],
            "cpu_newton": int(it), "cpu_minres": int(np.sum(mh)),
            "cpu_final_norm_f": float(nf[-1]), "rel_diff_gpu_vs_cpu_solution": d}


def _phase_align(psi, ref):
    z = psi[0::2] + 1j * psi[1::2]
    zr = ref[0::2] + 1j * ref[1::2]
    w = z * np.exp(-1j * np.angle(np.vdot(zr, z)))
    out = np.empty_like(psi)
    out[0::2], out[1::2] = w.real, w.imag
    return out


def build_model(ncell):
    import nosh_amd as nosh
    from nosh_amd import meshgen as mg
    coords, cells = mg.kuhn_cube(ncell)
    t0 = time.perf_counter()
    mesh = nosh.Mesh(coords, cells)
    t_mesh = time.perf_counter() - t0
    A = mg.magnetic_potential(coords)
    mvp = nosh.VectorFieldExplicitValues(mesh, A, 0.0)
    model = nosh.Nls(mesh, mvp, nosh.ScalarFieldConstant(mesh, -1.0), 1.0,
                     nosh.ScalarFieldConstant(mesh, 1.0))
    return coords, cells, A, mesh, model, t_mesh


def _plain_continuation(model, n, psi0, g, mu0, span, steps, branches, args, **extra):
    """Single-instance continuation of the given branches (FP64 Newton-MINRES
    rtol 1e-10 and the tangent predictor unless `extra` changes the Newton
    options); returns (solves/s, per-branch records)."""
    import nosh_amd as nosh
    t_all, recs = 0.0, []
    for b in branches:
        kw = dict(tol=1e-10, max_iters=40, minres_rtol=1e-10, minres_max_iters=args.minres_maxit,
                  cheb_iters=args.cheb, cheb_ratio=args.ratio)
        kw.update(extra)
        t1 = time.perf_counter()
        _, r1 = nosh.continuation(model, psi0[2 * n * b:2 * n * (b + 1)], {"g": g[b]}, "mu",
                                  mu0[b], mu0[b] + span, steps, predictor=1, **kw)
        dt = time.perf_counter() - t1
        t_all += dt
        ok = len(r1) == steps + 1 and all(r["converged"] for r in r1)
        recs.append({"branch": b, "seconds": round(dt, 3), "converged": ok,
                     "newton_iters": sum(r["newton_iters"] for r in r1[1:]),
                     "minres_iters": sum(r["minres_iters"] for r in r1[1:]),
                     "halvings": sum(r["halvings"] for r in r1[1:]),
                     "per_step_seconds": [round(r["seconds"], 4) for r in r1[1:]]})
        say(f"continuation: single-instance {extra or 'plain FP64'} branch {b}: {dt:.2f}s, "
            f"newton {recs[-1]['newton_iters']}, converged {ok}")
    ok = all(r["converged"] for r in recs)
    return (len(branches) * steps / t_all if ok and t_all > 0 else 0.0), recs


def _batched_run(batch, psi0, span, steps, kw, sur, predictor):
    """One batched continuation; returns (seconds, records, surrogate stats, psi)."""
    t0 = time.perf_counter()
    psi, rec, sst = batch.continuation(psi0, span, steps, surrogate=sur, predictor=predictor,
                                       **kw)
    return time.perf_counter() - t0, rec, sst, psi


def _run_summary(nb, steps, dt, rec):
    conv = len(rec) == steps + 1 and all(r["converged"] == nb for r in rec)
    per = [r["seconds"] for r in rec[1:]]
    return {"seconds": round(dt, 3), "solves_per_s": (nb * steps / dt) if conv else 0.0,
            "converged": conv, "steps": steps,
            "lockstep_newton_iters": sum(r["newton_iters"] for r in rec),
            "lockstep_minres_iters": sum(r["minres_iters"] for r in rec),
            "max_norm_f": max(r["max_norm_f"] for r in rec),
            "largest_step_share": round(max(per) / max(sum(per), 1e-12), 4),
            "per_step_seconds": [round(r["seconds"], 4) for r in rec],
            "per_step_newton": [r["newton_iters"] for r in rec]}


def continuation_phase(args, model, n):
    """mu-continuation of 8 (mu_0 = 0, g) branches on one mesh.

    Timed (the headline ``value``): the batched, mixed-precision,
    surrogate-seeded driver over args.steps steps of mu (after args.warmup
    untimed steps). Beside it, untimed: the same run at the other step size
    of the workload (50 steps when the headline has 20, else 20); the plain
    FP64 single-instance tangent continuation of every branch at the headline
    step size (the ratio's denominator) and of two branches at the other; an
    ablation at the headline step size that adds one change at a time (FP64
    with Eisenstat-Walker forcing; FP32 inner MINRES; batching; surrogate)."""
    import nosh_amd as nosh
    nb, span, steps = args.cont_branches, args.cont_span, args.steps
    g = 0.6 + 0.2 * np.arange(nb)
    mu0 = np.zeros(nb)
    psi0 = np.zeros(2 * n * nb)
    for b in range(nb):
        psi0[2 * n * b:2 * n * (b + 1):2] = 1.0 / np.sqrt(g[b])
    kw = dict(tol=1e-10, max_iters=40, minres_rtol=1e-10, minres_max_iters=4000,
              cheb_iters=args.cont_cheb, cheb_ratio=args.cont_ratio, forcing=1,
              forcing_min=1e-6, forcing_max=args.cont_forcing_max)
    batch = nosh.Batch([(model, nb)], mu0, g)
    t0 = time.perf_counter()
    if args.cont_tuning:
        v = [int(x) for x in args.cont_tuning.split(",")]
        bt, fold, mode = v[0], v[1], (v[2] if len(v) > 2 else 0)
        batch.set_tuning(bt, mode, 4, 16, fold=fold)
        tuned = (bt, mode, fold)
    else:
        tuned = batch.autotune(psi0, use_cache=False, **kw)
    t_tune = time.perf_counter() - t0
    say(f"continuation: autotune {t_tune:.2f}s -> {tuned}")
    sur = None if args.no_surrogate else dict(
        width=256, depth=4, train_after=args.cont_train_after, train_steps=args.cont_train_steps,
        batch=args.cont_batch, lr=1e-3, newton_budget=6, retrain_steps=args.cont_retrain,
        smooth_sweeps=args.cont_smooth, accept_ratio=args.cont_accept)
    if args.warmup > 0:
        t0 = time.perf_counter()
        batch.continuation(psi0, span * args.warmup / steps, args.warmup,
                           predictor=args.cont_predictor, **kw)
        say(f"continuation: warmup {args.warmup} steps {time.perf_counter() - t0:.2f}s")
    # ---- timed region
    t_comb, rec, sst, psi = _batched_run(batch, psi0, span, steps, kw, sur, args.cont_predictor)
    # ----
    head = _run_summary(nb, steps, t_comb, rec)
    comb_rate = head["solves_per_s"]
    say(f"continuation: combined {t_comb:.2f}s for {nb}x{steps} solves -> {comb_rate:.3f}/s, "
        f"converged {head['converged']}, largest step {head['largest_step_share']:.1%}, "
        f"surrogate {sst}")
    # Roofline of the timed region's dominant kernel (the batched Chebyshev
    # step), every branch active, at the final states, timed with HIP events
    # on the library stream.
    us_cs, bytes_cs = batch.time_cheb_step(50)
    kname = "kb_cheb_step_v<2>" if tuned[1] == 2 else "kb_cheb_step"
    cheb_roof = {"bound": "hbm", "kernel": f"{kname} (batched FP32 Chebyshev step, "
                                           f"{nb} lanes, all active)",
                 "achieved": round(bytes_cs / us_cs / 1e3, 1), "peak": 8000.0, "unit": "GB/s",
                 "frac": round(bytes_cs / us_cs / 1e3 / 8000.0, 4),
                 "us_per_launch": round(us_cs, 2), "bytes_per_launch": int(bytes_cs),
                 "layout_note": "12 B per padded slot (packed column, alpha, a~; shared by the "
                                "lanes) + 44 B per row and lane (z_in, r, z_{j-2} read and "
                                "overwritten by z_j, P block; FP32), step j = 2"}
    say(f"continuation: {kname} {us_cs:.1f} us, {bytes_cs / 1e6:.1f} MB -> "
        f"{bytes_cs / us_cs / 1e3:.0f} GB/s")
    finals = batch.split(psi)
    # Error of the mixed-precision states against FP64: FP64 Newton from each
    # final state (same branch) and the distance to where it lands.
    fp64_err = []
    for b in sorted({0, nb - 1}):
        p, s64 = nosh.newton_solve(model, finals[b], {"g": g[b], "mu": mu0[b] + span}, tol=1e-12,
                                   max_iters=10, minres_rtol=1e-12,
                                   minres_max_iters=args.minres_maxit, cheb_iters=args.cheb,
                                   cheb_ratio=args.ratio)
        fp64_err.append(float(np.linalg.norm(_phase_align(finals[b], p) - p) /
                              np.linalg.norm(p)))
    say(f"continuation: mixed vs FP64-polished final states, rel. err. max {max(fp64_err):.2e}")
    # Persistent (cooperative) MINRES against the tuned hipGraph mode: one
    # batched Newton solve of the 8 branches at mu_0 + span from their final
    # states perturbed by 1e-4, same block size and fold (bitwise equal
    # results expected).
    pers = None
    if not args.no_persistent:
        from nosh_amd import meshgen as mg
        bp = nosh.Batch([(model, nb)], mu0 + span, g)
        p_start = psi + 1e-4 * np.concatenate([mg.random_state(n, 900 + b) for b in range(nb)])
        kw_p = dict(kw, forcing=1)
        res = {}
        for mode in sorted({0, 1, tuned[1]}):
            bp.set_tuning(tuned[0], mode, 4, 16, fold=tuned[2])
            bp.solve(p_start, **kw_p)  # warm-up (graph capture, first launch)
            t1 = time.perf_counter()
            p_m, st_m = bp.solve(p_start, **kw_p)
            res[mode] = (time.perf_counter() - t1, p_m, st_m)
        t_g, p_g, st_g = res[0]
        t_p, p_p, st_p = res[1]
        pers = {"block_threads": tuned[0], "fold": tuned[2],
                "graph_seconds": round(t_g, 4), "persistent_seconds": round(t_p, 4),
                "tuned_mode": MODE_NAMES[tuned[1]],
                "tuned_mode_seconds": round(res[tuned[1]][0], 4),
                "bitwise_equal_persistent_vs_graph": bool(np.array_equal(p_g, p_p)),
                "newton_iters": max(s["newton_iters"] for s in st_g),
                "minres_iters": max(s["minres_iters"] for s in st_g),
                "converged": all(s["converged"] for s in st_p),
                "note": "one batched Newton solve of the 8 branches at the end of the span; "
                        "graph = hipGraph mode 0 (one lane per thread), persistent = mode 1, "
                        "same block size and fold, so the same arithmetic"}
        say(f"continuation: persistent MINRES {t_p:.3f}s vs hipGraph {t_g:.3f}s "
            f"({MODE_NAMES[tuned[1]]} {res[tuned[1]][0]:.3f}s), bitwise equal "
            f"{pers['bitwise_equal_persistent_vs_graph']}")
        del bp
    # The batched run without the surrogate (the ablation's batching row, and
    # the surrogate's net gain in this same process).
    other_steps = 50 if steps != 50 else 20
    no_sur, other, alt_pred = {"solves_per_s": None}, None, None
    if not args.no_extras:
        t_ns, rec_ns, _, _ = _batched_run(batch, psi0, span, steps, kw, None,
                                          args.cont_predictor)
        no_sur = _run_summary(nb, steps, t_ns, rec_ns)
        say(f"continuation: batched without surrogate {t_ns:.2f}s -> "
            f"{no_sur['solves_per_s']:.3f}/s")
        # The other step size of the workload.
        t_o, rec_o, sst_o, _ = _batched_run(batch, psi0, span, other_steps, kw, sur,
                                            args.cont_predictor)
        other = _run_summary(nb, other_steps, t_o, rec_o)
        other["surrogate"] = sst_o
        say(f"continuation: combined at {other_steps} steps {t_o:.2f}s -> "
            f"{other['solves_per_s']:.3f}/s")
        # The headline run with the other predictors (same process, same
        # surrogate settings).
        alt_pred = {}
        for alt in ("auto", "tangent", "secant"):
            if alt == args.cont_predictor:
                continue
            t_a, rec_a, _, _ = _batched_run(batch, psi0, span, steps, kw, sur, alt)
            alt_pred[alt] = _run_summary(nb, steps, t_a, rec_a)
            say(f"continuation: {alt} predictor {t_a:.2f}s -> "
                f"{alt_pred[alt]['solves_per_s']:.3f}/s")
    del batch
    # Plain FP64 single-instance continuation of every branch.
    branches = list(range(min(args.cont_plain_branches, nb)))
    plain_rate, plain_rec = _plain_continuation(model, n, psi0, g, mu0, span, steps, branches,
                                                args) if branches else (None, [])
    two = sorted({0, nb - 1})
    if other is not None:
        plain_o_rate, plain_o_rec = _plain_continuation(model, n, psi0, g, mu0, span,
                                                        other_steps, two, args)
        other["plain_fp64_solves_per_s"] = plain_o_rate
        other["plain_sample"] = plain_o_rec
        other["speedup_vs_plain_fp64"] = (other["solves_per_s"] / plain_o_rate
                                          if plain_o_rate else None)
    # Ablation at the headline step size (single-instance rows on two branches).
    abl = {"plain FP64 (tangent predictor, MINRES rtol 1e-10)": plain_rate}
    if not args.no_ablation and not args.no_extras:
        r_ew, _ = _plain_continuation(model, n, psi0, g, mu0, span, steps, two, args, forcing=1,
                                      forcing_min=1e-6, forcing_max=args.cont_forcing_max)
        abl["+ Eisenstat-Walker forcing (FP64 inner)"] = r_ew
        r_32, _ = _plain_continuation(model, n, psi0, g, mu0, span, steps, two, args, forcing=1,
                                      forcing_min=1e-6, forcing_max=args.cont_forcing_max,
                                      precision=1)
        abl["+ FP32 inner MINRES"] = r_32
    abl[f"+ batching (8 branches in lockstep, {args.cont_predictor} predictor, Chebyshev "
        f"{args.cont_cheb})"] = no_sur["solves_per_s"]
    abl["+ surrogate (the headline)"] = comb_rate
    seeded = [r for r in rec if r["seeded"]]
    out = dict(head)
    out.update({
        "branches": nb, "span": span, "g": [round(v, 3) for v in g],
        "autotuned": {"block_threads": tuned[0], "fold": tuned[2],
                      "mode": MODE_NAMES[tuned[1]], "seconds": round(t_tune, 2)},
        "rel_err_vs_fp64_polished": {"branches": sorted({0, nb - 1}), "values": fp64_err},
        "surrogate": sst,
        "seeded_branch_steps": sum(r["seeded"] for r in rec),
        "fallback_branch_steps": sum(r["fallbacks"] for r in rec),
        "mean_pred_error_seeded_steps": (float(np.mean([r["pred_error"] for r in seeded]))
                                         if seeded else None),
        "mean_secant_error_seeded_steps": (float(np.mean([r["sec_error"] for r in seeded]))
                                           if seeded else None),
        "without_surrogate": no_sur,
        "surrogate_net_gain": (comb_rate / no_sur["solves_per_s"]
                               if no_sur.get("solves_per_s") else None),
        "dominant_kernel_roofline": cheb_roof,
        "persistent_vs_graph": pers,
        "plain_fp64_solves_per_s": plain_rate, "plain_sample": plain_rec,
        "speedup_vs_plain_fp64": comb_rate / plain_rate if plain_rate else None,
        f"at_{other_steps}_steps": other,
        "predictor": args.cont_predictor,
        "other_predictors": alt_pred,
        "ablation_solves_per_s": abl,
    })
    return out


def _event_time(stream, fn, reps, warm=3):
    """Average seconds per call of fn() over reps calls, HIP events on stream."""
    import torch
    with torch.cuda.stream(stream):
        for _ in range(warm):
            fn()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
    e1.synchronize()
    return e0.elapsed_time(e1) / 1e3 / reps


def stream_ceiling():
    """Attainable HBM bandwidth from the library's own streaming kernels
    (read, write, copy over 2 GiB buffers; non-temporal and ordinary
    accesses; 8 workgroups of 256 threads per CU): the best GB/s of each."""
    import nosh_amd as nosh
    nbytes = 2 << 30
    out = {}
    for kind in ("read", "write", "copy"):
        best = 0.0
        for nt in (True, False):
            t = nosh.bench_stream(kind, nbytes, reps=10, blocks_per_cu=8, nontemporal=nt)
            moved = (2 if kind == "copy" else 1) * nbytes
            best = max(best, moved / t / 1e9)
        out[kind] = round(best, 1)
    say(f"stream ceiling (own kernels, 2 GiB): read {out['read']:.0f}, write {out['write']:.0f}, "
        f"copy {out['copy']:.0f} GB/s")
    return out


def jphi_timings(args, mesh, model, params, psi0, phi, stream, dev, fp32=True):
    """FP64 (and FP32) J(psi) phi on the library's launch stream: average
    kernel durations, stamped by the runtime on each kernel dispatch (HIP
    events; launch gaps excluded), warm (back-to-back) and cold (a 1 GiB read
    before each launch), against the canonical bytes 40E + 68n (FP32:
    24E + 36n)."""
    import torch
    E, n = mesh.num_edges, mesh.num_vertices
    J = model.create_W_op()
    J.rebuild(params, psi0)
    x_ref = torch.from_numpy(phi).to(dev)
    x = torch.empty_like(x_ref)
    y = torch.empty_like(x_ref)
    s = stream.cuda_stream
    mesh.to_internal(x_ref.data_ptr(), x.data_ptr(), s)
    res = {"n": n, "E": E, "bytes64": 40 * E + 68 * n, "bytes32": 24 * E + 36 * n,
           "warm64": J.bench_apply(x.data_ptr(), y.data_ptr(), reps=args.spmv_reps, stream=s),
           "cold64": J.bench_apply(x.data_ptr(), y.data_ptr(), cold=True, reps=20, stream=s)}
    if fp32:
        x32 = x.float()
        y32 = torch.empty_like(x32)
        res["warm32"] = J.bench_apply(x32.data_ptr(), y32.data_ptr(), fp32=True,
                                      reps=args.spmv_reps, stream=s)
        res["cold32"] = J.bench_apply(x32.data_ptr(), y32.data_ptr(), fp32=True, cold=True,
                                      reps=20, stream=s)
    torch.cuda.synchronize()
    return res, J, x, y


def jphi_roofline(args, mesh, model, params, psi0, phi, stream, dev, A):
    """The roofline of the Jacobian SpMV. Primary figure: the FP64 J(psi) phi
    at 96^3 with cold caches (its 313 MB working set otherwise partly sits in
    the 256 MB Infinity Cache), canonical bytes / average launch time against
    8 TB/s and against the attainable ceiling of the library's streaming
    kernels. Beside it: the warm (back-to-back) figure, FP32 J phi, K phi,
    F(psi), and the 160^3 J phi (1.44 GB, no cache effects)."""
    import torch
    import nosh_amd as nosh
    E, n = mesh.num_edges, mesh.num_vertices
    t, J, x, y = jphi_timings(args, mesh, model, params, psi0, phi, stream, dev)
    bytes_canon = t["bytes64"]
    gbs = lambda b, s: b / s / 1e9  # noqa: E731
    ceiling = stream_ceiling()
    # Copy-kernel reference of the same size class (torch's copy_).
    cbuf = torch.empty(bytes_canon // 8, dtype=torch.float64, device=dev)
    cdst = torch.empty_like(cbuf)
    t_copy = _event_time(stream, lambda: cdst.copy_(cbuf), 20)
    copy_gbs = 2 * cbuf.numel() * 8 / t_copy / 1e9
    del cbuf, cdst
    layout = mesh.layout()
    # Split layout: per upper slot a column and a 16 B value, per lower slot a
    # column and the 4 B position of the mirrored value (which comes from
    # L2), columns as 16-bit offsets from the row (or 32-bit indices when the
    # bandwidth does not allow), two slice pointers per slice, 56 B per
    # vertex.
    cb = layout["column_bytes"]
    bytes_layout = ((16 + cb) * layout["nnz_upper"] + (4 + cb) * layout["nnz_lower"] +
                    8 * (layout["nslices"] + 1) + 56 * n)
    # K.phi and F(psi) (FP64, internal order), canonical bytes 40E + 44n and
    # 40E + 52n.
    others = {}
    keo = nosh.Keo(mesh, nosh.ScalarFieldConstant(mesh, 1.0),
                   nosh.VectorFieldExplicitValues(mesh, A, 0.0))
    keo.set_parameters(params)
    fk = lambda: keo.apply_internal(x.data_ptr(), y.data_ptr(), stream.cuda_stream)  # noqa
    ff = lambda: model.eval_f_internal(x.data_ptr(), params, y.data_ptr(),  # noqa: E731
                                       stream.cuda_stream)
    # Warm only, HIP events around back-to-back launches (launch gaps
    # included).
    for name, fn, b_ in (("keo_apply (K.phi FP64)", fk, 40 * E + 44 * n),
                         ("compute_f (F(psi) FP64)", ff, 40 * E + 52 * n)):
        tw = _event_time(stream, fn, args.spmv_reps)
        others[name] = {"us_per_launch": round(tw * 1e6, 2), "bytes_per_launch": b_,
                        "achieved": round(gbs(b_, tw), 1),
                        "frac": round(gbs(b_, tw) / HBM_PEAK_GBS, 4)}
    others["k_block_apply<float> (J.phi FP32, blocks converted once per rebuild)"] = {
        "us_per_launch": round(t["warm32"] * 1e6, 2),
        "cold_us_per_launch": round(t["cold32"] * 1e6, 2), "bytes_per_launch": t["bytes32"],
        "achieved": round(gbs(t["bytes32"], t["warm32"]), 1),
        "cold_achieved": round(gbs(t["bytes32"], t["cold32"]), 1),
        "frac": round(gbs(t["bytes32"], t["warm32"]) / HBM_PEAK_GBS, 4),
        "cold_frac": round(gbs(t["bytes32"], t["cold32"]) / HBM_PEAK_GBS, 4)}
    del J, keo
    # The 160^3 cube: 1.44 GB per J launch, no cache effects.
    k160 = None
    if args.roofline_160:
        from nosh_amd import meshgen as mg
        c, cl = mg.kuhn_cube(160)
        m160 = nosh.Mesh(c, cl)
        md160 = nosh.Nls(m160, nosh.VectorFieldExplicitValues(m160, mg.magnetic_potential(c), 0.0),
                         nosh.ScalarFieldConstant(m160, -1.0), 1.0,
                         nosh.ScalarFieldConstant(m160, 1.0))
        n1 = c.shape[0]
        t1, J1, _, _ = jphi_timings(args, m160, md160, params, mg.random_state(n1, 0),
                                    mg.random_state(n1, 1), stream, dev, fp32=False)
        k160 = {"n_vertices": n1, "n_edges": t1["E"], "bytes_per_launch": t1["bytes64"],
                "us_per_launch": round(t1["warm64"] * 1e6, 2),
                "cold_us_per_launch": round(t1["cold64"] * 1e6, 2),
                "achieved": round(gbs(t1["bytes64"], t1["cold64"]), 1),
                "frac": round(gbs(t1["bytes64"], t1["cold64"]) / HBM_PEAK_GBS, 4),
                "frac_of_attainable_read": round(gbs(t1["bytes64"], t1["cold64"]) /
                                                 ceiling["read"], 4)}
        say(f"J.phi 160^3: warm {t1['warm64']*1e6:.1f} us, cold {t1['cold64']*1e6:.1f} us -> "
            f"{k160['achieved']:.0f} GB/s")
        del J1, md160, m160
    traffic, traffic_src = _profiled_traffic(args.ncell)
    ach_cold = gbs(bytes_canon, t["cold64"])
    ach_warm = gbs(bytes_canon, t["warm64"])
    say(f"J.phi 96^3 FP64: cold {t['cold64']*1e6:.1f} us -> {ach_cold:.0f} GB/s "
        f"({ach_cold/HBM_PEAK_GBS:.1%}); warm {t['warm64']*1e6:.1f} us -> {ach_warm:.0f} GB/s; "
        f"FP32 cold {t['cold32']*1e6:.1f} us, warm {t['warm32']*1e6:.1f} us; torch copy "
        f"{copy_gbs:.0f} GB/s")
    return {"bound": "hbm", "achieved": round(ach_cold, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach_cold / HBM_PEAK_GBS, 4), "traffic": traffic,
            "traffic_source": traffic_src,
            "kernel": "k_block_apply<double> (J.phi FP64, 96^3, cold caches)",
            "timing": "average kernel duration from start/end events that the HIP runtime "
                      "stamps on each dispatch (hipExtLaunchKernelGGL) on the launch stream; "
                      "cold: a 1 GiB read before each launch",
            "us_per_launch": round(t["cold64"] * 1e6, 2), "bytes_per_launch": bytes_canon,
            "attainable": {"unit": "GB/s", "own_stream_kernels": ceiling,
                           "torch_copy": round(copy_gbs, 1)},
            "frac_of_attainable_read": round(ach_cold / ceiling["read"], 4),
            "warm": {"us_per_launch": round(t["warm64"] * 1e6, 2), "achieved": round(ach_warm, 1),
                     "frac": round(ach_warm / HBM_PEAK_GBS, 4),
                     "note": "back-to-back launches: part of the 313 MB working set is served "
                             "by the 256 MB Infinity Cache"},
            "kuhn160": k160,
            "layout_bytes_per_launch": bytes_layout,
            "layout_note": f"split sliced ELLPACK (64-row slices): upper slots (column above "
                           f"the row) {cb} B column + 16 B value, lower slots {cb} B column + 4 B "
                           f"position of the mirrored upper slot (conjugate value, gathered "
                           f"from L2), 8 B per slice, 24 B 2x2 block + 16 B x + 16 B y per "
                           f"vertex; columns as {'16-bit offsets from the row' if cb == 2 else '32-bit indices'}",
            "nnz_upper": layout["nnz_upper"], "nnz_lower": layout["nnz_lower"],
            "column_bytes": cb,
            "other_kernels": others}


def surrogate_mfma(rows, stream, dev, nin=9, width=256, depth=4, reps=5):
    """Fused MLP inference (the surrogate's per-step cost) on `rows` feature
    rows: TFLOP/s of its dense layers against the bf16 MFMA peak."""
    import torch
    import nosh_amd as nosh
    net = nosh.Mlp(nin, width, depth, 2, 1)
    X = torch.randn(rows, nin, dtype=torch.float32, device=dev)
    Y = torch.empty(rows, 2, dtype=torch.float32, device=dev)
    with torch.cuda.stream(stream):
        net.infer_device(X.data_ptr(), rows, Y.data_ptr(), stream.cuda_stream)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            net.infer_device(X.data_ptr(), rows, Y.data_ptr(), stream.cuda_stream)
        e1.record(stream)
    e1.synchronize()
    t = e0.elapsed_time(e1) / 1e3 / reps
    flop_mfma = 2.0 * rows * (depth - 1) * width * width
    flop_all = 2.0 * rows * (nin * width + (depth - 1) * width * width + width * 2)
    ach = flop_mfma / t / 1e12
    say(f"surrogate inference: {rows} rows in {t*1e3:.2f} ms -> {ach:.1f} TFLOP/s on MFMA")
    return {"bound": "mfma", "kernel": "k_mlp_infer<8> (fused BF16 MLP inference)",
            "rows": rows, "ms_per_launch": round(t * 1e3, 3), "achieved": round(ach, 1),
            "peak": 2500.0, "unit": "TFLOP/s", "frac": round(ach / 2500.0, 4),
            "all_layers_TFLOPs": round(flop_all / t / 1e12, 1)}


def single_solves(args, model, params, psi0, stream, dev):
    """The 96^3 solve side by side, plain FP64 and FP32-inner / FP64-outer
    Newton-MINRES, from two starts: the seeded random state (|psi| ~
    U(0.5,1), arg ~ U(0, 2 pi); on this fine mesh the random phases cancel
    and both variants converge to the trivial state psi = 0, which the
    record names), and a nontrivial vortex state perturbed by 1e-3, which
    both take back to that state. Reported per variant: ||F||, ||psi|| and
    the phase-aligned ||psi - psi_FP64|| (absolute, and relative when
    psi_FP64 is nontrivial)."""
    import torch
    import nosh_amd as nosh
    from nosh_amd import meshgen as mg
    n = psi0.size // 2
    s1 = 0.1 * mg.random_state(n, 5)
    s1[0::2] += 1.0
    # A nontrivial solution (plain FP64 Newton from the superconducting
    # start lands on a vortex state), perturbed by 1e-3 seeded noise: from
    # there both variants converge back to it. (From the far start itself
    # they wander to different vortex states.)
    t0 = time.perf_counter()
    p_sc, st_sc = nosh.newton_solve(model, s1, params, tol=1e-10, max_iters=60, minres_rtol=1e-10,
                                    minres_max_iters=args.minres_maxit, cheb_iters=args.cheb,
                                    cheb_ratio=args.ratio)
    say(f"side by side: FP64 from the superconducting start {time.perf_counter() - t0:.2f}s, "
        f"newton {st_sc['newton_iters']}, ||psi|| {np.linalg.norm(p_sc):.3e}")
    out = {"vortex_state_from_superconducting_start": {
        "newton_iters": st_sc["newton_iters"], "converged": bool(st_sc["converged"]),
        "norm_psi": float(np.linalg.norm(p_sc))}}
    for start, p0 in (("random", psi0),
                      ("perturbed vortex state", p_sc + 1e-3 * mg.random_state(n, 6))):
        psi_dev0 = torch.from_numpy(p0).to(dev)
        psi = torch.empty_like(psi_dev0)
        results, sols = {}, {}
        for var in [v for v in args.variants.split(",") if v]:
            kw = dict(tol=1e-10, max_iters=40, minres_rtol=1e-10,
                      minres_max_iters=args.minres_maxit, cheb_iters=args.cheb,
                      cheb_ratio=args.ratio)
            if var == "mixed":
                kw.update(precision=1, forcing=1, forcing_min=1e-6, forcing_max=0.1)
            opts = nosh.default_newton_options(**kw)
            psi.copy_(psi_dev0)
            torch.cuda.synchronize()
            nosh.newton_solve_device(model, psi.data_ptr(), params, opts, stream.cuda_stream)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            all_stats = []
            for _ in range(args.solve_reps):
                with torch.cuda.stream(stream):
                    psi.copy_(psi_dev0)
                all_stats.append(nosh.newton_solve_device(model, psi.data_ptr(), params, opts,
                                                          stream.cuda_stream))
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            ok = all(s["converged"] and s["final_norm_f"] <= 1e-10 for s in all_stats)
            sols[var] = psi.cpu().numpy()
            results[var] = {"solves_per_s": args.solve_reps / dt,
                            "s_per_solve": dt / args.solve_reps, "converged": ok,
                            "norm_f": all_stats[-1]["final_norm_f"],
                            "norm_psi": float(np.linalg.norm(sols[var])),
                            "newton_iters": all_stats[-1]["newton_iters"],
                            "minres_iters": all_stats[-1]["minres_iters_total"]}
        if "plain" in sols:
            ref = sols["plain"]
            nref = float(np.linalg.norm(ref))
            trivial = nref < 1e-6 * np.sqrt(n)
            for var, x in sols.items():
                d = float(np.linalg.norm(_phase_align(x, ref) - ref))
                results[var]["norm_diff_vs_fp64"] = d
                results[var]["rel_diff_vs_fp64"] = None if trivial else d / nref
            results["state"] = ("trivial (psi = 0: both variants converge to the normal state)"
                                if trivial else "nontrivial")
        say(start, results)
        out[start] = results
    return out


def headline(args):
    import torch
    from nosh_amd import meshgen as mg
    dev = torch.device("cuda:0")
    torch.cuda.init()
    coords, cells, A, mesh, model, t_mesh = build_model(args.ncell)
    n, E = mesh.num_vertices, mesh.num_edges
    say(f"mesh {args.ncell}^3: n={n} E={E} build {t_mesh:.2f}s")
    params = {"g": 1.0, "mu": args.mu}
    psi0 = mg.random_state(n, 0)
    phi = mg.random_state(n, 1)
    stream = torch.cuda.Stream(device=dev)
    roofline = jphi_roofline(args, mesh, model, params, psi0, phi, stream, dev, A)
    if args.roofline_only:
        print(json.dumps({"roofline": roofline}))
        return
    solves = single_solves(args, model, params, psi0, stream, dev)
    mfma = surrogate_mfma(n * args.cont_branches, stream, dev)
    cont = continuation_phase(args, model, n)
    cpu = None
    if not args.no_cpu and cont["plain_sample"]:
        ps = cont["plain_sample"][0]
        cpu = cpu_baseline(coords, cells, A, args.cont_span / args.steps, psi0, phi,
                           {"newton_iters": ps["newton_iters"] / args.steps,
                            "minres_iters": ps["minres_iters"] / args.steps},
                           args.cheb, args.ratio, "plain FP64 continuation", model=model)
    c0 = None if args.no_cpu else config0_solves(args)
    if c0:
        say(f"config 0 solve: GPU {c0['gpu_seconds']:.3f}s, CPU {c0['cpu_seconds']:.3f}s, "
            f"rel. diff {c0['rel_diff_gpu_vs_cpu_solution']:.2e}")
    value = cont["solves_per_s"] if cont["converged"] else 0.0
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(cont["seconds"] / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64/f32 (FP32 inner MINRES, FP64 Newton); bf16/f32 MFMA surrogate",
        "data": "synthetic (structured Kuhn mesh; branch start states 1/sqrt(g); seeded "
                "SplitMix64 psi_0 for the side-by-side solves)",
        "config": {"workload": f"Kuhn-tet cube {args.ncell}^3: mu-continuation 0 -> "
                               f"{args.cont_span} in {args.steps} steps of "
                               f"{args.cont_branches} (mu_0 = 0, g) branches, batched",
                   "n_vertices": n, "n_edges": E, "n_tets": int(cells.shape[0]),
                   "V": -1.0, "branches_g": cont["g"],
                   "combined": f"batched FP32-inner/FP64 Newton, secant + surrogate predictor, "
                               f"Chebyshev-Jacobi({args.cont_cheb}, {args.cont_ratio}), "
                               f"Eisenstat-Walker forcing <= {args.cont_forcing_max}",
                   "plain": f"FP64 Newton-MINRES rtol 1e-10, tangent predictor, "
                            f"Chebyshev-Jacobi({args.cheb}, {args.ratio})"},
        "speedup_vs_plain_fp64": (round(cont["speedup_vs_plain_fp64"], 3)
                                  if cont["speedup_vs_plain_fp64"] else None),
        "continuation": cont,
        "side_by_side_solves": solves,
        "roofline": roofline,
        "roofline_timed_region": cont["dominant_kernel_roofline"],
        "surrogate_mfma": mfma,
        "mesh_build_s": round(t_mesh, 3),
        "cpu_baseline": cpu,
        "small_case_solve": c0,
    }
    print(json.dumps(out))


def disk_workload(args):
    """The disk workload: Delaunay disk of radius 5 (h = 0.0213, about 200 k
    vertices, seed 3), V = -1, g = 1, mu from 0 to 2 in 200 steps from
    psi = 1. The surrogate-seeded continuation (MLP trained on the first 40
    solutions, tangent fallback) against the plain tangent continuation, both
    with FP64 Newton-MINRES: Newton iterations and seconds per step."""
    import torch
    import nosh_amd as nosh
    from nosh_amd import meshgen as mg
    torch.cuda.init()
    coords, cells = mg.disk(0.0213)
    n = coords.shape[0]
    mesh = nosh.Mesh(coords, cells)
    mvp = nosh.VectorFieldExplicitValues(mesh, mg.magnetic_potential(coords), 0.0)
    model = nosh.Nls(mesh, mvp, nosh.ScalarFieldConstant(mesh, -1.0), 1.0,
                     nosh.ScalarFieldConstant(mesh, 1.0))
    say(f"disk workload: n = {n}, {cells.shape[0]} triangles")
    psi0 = np.zeros(2 * n)
    psi0[0::2] = 1.0
    steps = args.steps
    kw = dict(tol=1e-10, max_iters=40, minres_rtol=1e-10, minres_max_iters=args.minres_maxit,
              cheb_iters=args.cheb, cheb_ratio=args.ratio)
    t0 = time.perf_counter()
    _, rec_s, st = nosh.continuation_surrogate(
        model, psi0, {"g": 1.0}, "mu", 0.0, 2.0, steps,
        surrogate=dict(width=256, depth=4, train_after=40, train_steps=300, batch=16384,
                       lr=1e-3, newton_budget=4), **kw)
    t_s = time.perf_counter() - t0
    say(f"disk: surrogate run {t_s:.1f}s, {st}")
    t0 = time.perf_counter()
    _, rec_p = nosh.continuation(model, psi0, {"g": 1.0}, "mu", 0.0, 2.0, steps, predictor=1,
                                 **kw)
    t_p = time.perf_counter() - t0
    say(f"disk: plain run {t_p:.1f}s")
    after = slice(41, None)
    out = {"workload": f"Delaunay disk r = 5, n = {n}: mu 0 -> 2 in {steps} steps",
           "surrogate": {"seconds": t_s, "steps_done": len(rec_s) - 1, "stats": st,
                         "newton_total": sum(r["newton_iters"] for r in rec_s[1:]),
                         "newton_after_training": sum(r["newton_iters"] for r in rec_s[after]),
                         "seconds_after_training": sum(r["seconds"] for r in rec_s[after]),
                         "seeded_steps": sum(r["predictor_used"] == 3 for r in rec_s)},
           "plain": {"seconds": t_p, "steps_done": len(rec_p) - 1,
                     "newton_total": sum(r["newton_iters"] for r in rec_p[1:]),
                     "newton_after_training": sum(r["newton_iters"] for r in rec_p[after]),
                     "seconds_after_training": sum(r["seconds"] for r in rec_p[after])},
           "per_step_newton": {"surrogate": [r["newton_iters"] for r in rec_s],
                               "plain": [r["newton_iters"] for r in rec_p]},
           "per_step_seconds": {"surrogate": [round(r["seconds"], 4) for r in rec_s],
                                "plain": [round(r["seconds"], 4) for r in rec_p]}}
    print(json.dumps(out))


def cont_workload(args):
    import torch
    torch.cuda.init()
    coords, cells, A, mesh, model, t_mesh = build_model(args.ncell)
    say(f"cont workload: mesh {args.ncell}^3, n={mesh.num_vertices}")
    print(json.dumps(continuation_phase(args, model, mesh.num_vertices)))


def stability_workload(args):
    """Stability eigenvalues and arc-length continuation on the Kuhn cube
    (default 96^3, g = 1): the Meissner state at mu = 0.2 by natural
    continuation; the 4 smallest eigenvalues of J with the phase mode
    deflated (LOBPCG, Chebyshev preconditioner); then pseudo-arclength steps
    from mu = 0.2 with those eigenvalues at every accepted point, against the
    natural-parameter continuation's time per step (FP64 Newton-MINRES, with
    Eisenstat-Walker forcing for the natural steps, where it is faster; the
    arclength corrector solves exactly, which is faster and keeps it on the
    branch at folds)."""
    import torch
    import nosh_amd as nosh
    torch.cuda.init()
    coords, cells, A, mesh, model, t_mesh = build_model(args.ncell)
    n = mesh.num_vertices
    say(f"stability workload: mesh {args.ncell}^3, n={n}")
    kw_arc = dict(tol=1e-10, max_iters=20, minres_rtol=1e-10, minres_max_iters=args.minres_maxit,
                  cheb_iters=args.cheb, cheb_ratio=args.ratio)
    kw = dict(kw_arc, forcing=1, forcing_min=1e-6, forcing_max=0.1)
    psi0 = np.zeros(2 * n)
    psi0[0::2] = 1.0
    t0 = time.perf_counter()
    psi, rec_n = nosh.continuation(model, psi0, {"g": 1.0}, "mu", 0.0, 0.2, 10, predictor=1, **kw)
    t_nat = time.perf_counter() - t0
    assert all(r["converged"] for r in rec_n)
    say(f"natural continuation 0 -> 0.2 in 10 steps: {t_nat:.2f}s")
    eig = {}
    for prec in (1, 0):
        t0 = time.perf_counter()
        ev, _, st = nosh.jacobian_eigenpairs(model, psi, {"g": 1.0, "mu": 0.2}, nev=4, tol=1e-8,
                                             use_prec=prec, cheb_iters=args.cheb,
                                             cheb_ratio=args.ratio, max_iters=2000)
        st["wall"] = time.perf_counter() - t0
        eig["prec" if prec else "noprec"] = {"evals": ev.tolist(), **st}
        say(f"eigenpairs (prec={prec}): {ev}, {st}")
    t0 = time.perf_counter()
    _, rec_a = nosh.arclength(model, psi, {"g": 1.0, "mu": 0.2}, "mu", start=0.2,
                              max_steps=args.steps, initial_step=0.02, max_step=0.05,
                              eigenvalues=4, newton_kw=kw_arc)
    t_arc = time.perf_counter() - t0
    say(f"arclength {len(rec_a) - 1} steps: {t_arc:.2f}s")
    t0 = time.perf_counter()
    _, rec_b = nosh.arclength(model, psi, {"g": 1.0, "mu": 0.2}, "mu", start=0.2,
                              max_steps=args.steps, initial_step=0.02, max_step=0.05,
                              eigenvalues=0, newton_kw=kw_arc)
    t_arc0 = time.perf_counter() - t0
    out = {"workload": f"Kuhn-tet cube {args.ncell}^3, g = 1: stability eigenvalues at the "
                       f"Meissner state mu = 0.2 and {args.steps} arclength steps from there",
           "n_vertices": n,
           "natural_continuation": {"steps": 10, "seconds": t_nat,
                                    "newton": sum(r["newton_iters"] for r in rec_n[1:]),
                                    "minres": sum(r["minres_iters"] for r in rec_n[1:])},
           "eigenpairs": eig,
           "arclength": {"steps": len(rec_a) - 1, "seconds_with_eigs": t_arc,
                         "seconds_without_eigs": t_arc0,
                         "records": [{k: (round(v, 6) if isinstance(v, float) else v)
                                      for k, v in r.items()} for r in rec_a]}}
    print(json.dumps(out))


def batched_workload(args):
    """128 instances drawn round-robin over jittered Kuhn cubes 16^3, 24^3, 32^3,
    48^3 (interior vertices moved by U(-0.2h, 0.2h), seed 2); mu ~ U(0,1),
    g ~ U(0.5,2) from the SplitMix64 stream of seed 2 (two draws per
    instance); psi_0 of instance i from seed 1000 + i; V = -1. One step = the
    whole batch solved to ||F|| <= 1e-10 (FP32 inner MINRES, FP64 outer)."""
    print("batched workload: importing torch", file=sys.stderr, flush=True)
    import torch
    import nosh_amd as nosh
    from nosh_amd import meshgen as mg
    print("batched workload: building meshes", file=sys.stderr, flush=True)
    sizes = [16, 24, 32, 48]
    meshes = [mg.kuhn_cube(s, jitter=0.2, seed=2) for s in sizes]
    models = []
    for c, t in meshes:
        mesh = nosh.Mesh(c, t)
        mvp = nosh.VectorFieldExplicitValues(mesh, mg.magnetic_potential(c), 0.0)
        models.append(nosh.Nls(mesh, mvp, nosh.ScalarFieldConstant(mesh, -1.0), 1.0,
                               nosh.ScalarFieldConstant(mesh, 1.0)))
    ninst = args.instances
    owner = [i % 4 for i in range(ninst)]
    counts = [owner.count(k) for k in range(4)]
    u = mg.splitmix64_uniform(2, 2 * ninst).reshape(ninst, 2)
    # Group-major order: instance i sits in group owner[i].
    order = sorted(range(ninst), key=lambda i: (owner[i], i))
    mu = np.array([u[i, 0] for i in order])
    g = np.array([0.5 + 1.5 * u[i, 1] for i in order])
    psi0 = np.concatenate([mg.random_state(meshes[owner[i]][0].shape[0], 1000 + i)
                           for i in order])
    batch = nosh.Batch([(models[k], counts[k]) for k in range(4)], mu, g)
    # Newton from random phases is chaotic on these meshes (5 to about 120
    # steps per instance); the limit leaves room for the slowest.
    kw = dict(tol=1e-10, max_iters=200, minres_rtol=1e-10, minres_max_iters=4000,
              cheb_iters=args.cheb, cheb_ratio=args.ratio, forcing=1, forcing_min=1e-6,
              forcing_max=0.1)

    t0 = time.perf_counter()
    tuned = batch.autotune(psi0, use_cache=False, **kw)
    t_tune = time.perf_counter() - t0
    say(f"autotune {t_tune:.2f}s -> block_threads {tuned[0]}, mode {tuned[1]}, fold {tuned[2]}")
    dev = torch.device("cuda:0")
    x0 = torch.from_numpy(psi0).to(dev)
    x = torch.empty_like(x0)
    for w in range(args.warmup):
        x.copy_(x0)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        st = batch.solve_device(x.data_ptr(), **kw)
        say(f"warmup {w}: {time.perf_counter() - t1:.2f}s, converged "
            f"{sum(s['converged'] for s in st)}/{ninst}, newton max "
            f"{max(s['newton_iters'] for s in st)}")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        x.copy_(x0)
        st = batch.solve_device(x.data_ptr(), **kw)
        say(f"step {k}: {time.perf_counter() - t0:.2f}s cumulative")
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    conv = sum(s["converged"] for s in st)
    # Plain FP64 single-instance solves of a sample, same options otherwise.
    sample = [order.index(i) for i in range(min(args.plain_sample, ninst))]
    t_plain, worst = 0.0, 0.0
    for j in sample:
        k = owner[order[j]]
        p0 = batch.split(psi0)[j]
        t1 = time.perf_counter()
        p, s = nosh.newton_solve(models[k], p0, {"g": g[j], "mu": mu[j]}, tol=1e-10,
                                 max_iters=200, minres_rtol=1e-10, minres_max_iters=4000,
                                 cheb_iters=args.cheb, cheb_ratio=args.ratio)
        t_plain += time.perf_counter() - t1
        worst = max(worst, s["final_norm_f"])
        say(f"plain instance {j} (mesh {sizes[k]}^3): {time.perf_counter() - t1:.2f}s, "
            f"newton {s['newton_iters']}, minres {s['minres_iters_total']}")
    plain_rate = len(sample) / t_plain
    # Only converged instances count as solves.
    rate = conv / dt
    out = {"metric": "batched Newton-MINRES solves/sec at ||F||<=1e-10 (128 uneven instances)",
           "value": round(rate, 3), "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 2), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None,
           "dtype": "f64/f32 (FP32 inner MINRES, FP64 Newton)",
           "data": "synthetic (jittered Kuhn meshes, seeded psi_0, mu, g)",
           "config": {"workload": "batched uneven solves", "instances": ninst,
                      "meshes": [f"{s}^3 jittered" for s in sizes], "counts": counts,
                      "autotuned": {"block_threads": tuned[0], "fold": tuned[2],
                                    "mode": MODE_NAMES[tuned[1]]},
                      "autotune_s": round(t_tune, 2)},
           "converged": conv,
           "newton_iters_max": max(s["newton_iters"] for s in st),
           "minres_iters_max": max(s["minres_iters"] for s in st),
           "plain_fp64_sequential_solves_per_s": round(plain_rate, 3),
           "plain_sample": len(sample), "plain_worst_norm_f": worst,
           "speedup_vs_plain_fp64": round(rate / plain_rate, 2)}
    print(json.dumps(out))




def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is design-only in this build.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batched":
        batched_workload(args)
    elif args.workload == "cont":
        cont_workload(args)
    elif args.workload == "disk":
        disk_workload(args)
    elif args.workload == "stability":
        stability_workload(args)
    else:
        headline(args)


if __name__ == "__main__":
    main()
