"""Benchmark: Newton-MINRES solves/sec at ||F||_2 <= 1e-10 on the 96^3 Kuhn
cube, mixed-precision batched surrogate-seeded continuation against plain
FP64 GPU Newton-MINRES; and the Jacobian SpMV's HBM GB/s.

Mesh: cube [-5,5]^3, 96^3 cells split into 6 Kuhn tetrahedra each (912,673
vertices, 5,308,416 tets, 6,276,384 edges); A = mu*0.5*(-y, x, 0); V = -1;
thickness 1.

Headline workload (``value``; one step = one continuation step of every
branch): mu-continuation from mu = 0 to 1 in 50 equal steps of 8 branches
(mu_0 = 0, g_b = 0.6 + 0.2 b, psi_0 = 1/sqrt(g_b), exact at mu = 0), each step
solved to ||F|| <= 1e-10. Combined method: the branches batched in one
instance-interleaved state (one launch per phase, hipGraph-captured MINRES
iterations, autotuned block size), FP32 inner MINRES with FP64 Newton
(Eisenstat-Walker forcing), secant predictor, and the BF16 MFMA MLP surrogate
trained online on the run's own solutions seeding Newton where its guess has
the smaller residual (tangent-free fallback to the secant). ``value`` = 8 x
50 / wall time of the run (training and inference included). Plain
baseline: the single-instance FP64 continuation (MINRES rtol 1e-10, tangent
predictor, Chebyshev-Jacobi(6, 300), its fastest setting) of branch 0; the
ratio is value / its solves/sec.

Side by side (the 96^3 random-start solve): plain FP64 and FP32-inner
Newton-MINRES from psi_0 seeded random (|psi| ~ U(0.5,1), arg ~ U(0, 2 pi),
seed 0) at mu = 0.4, g = 1, with ||F|| and ||psi - psi_FP64||.

Roofline: the FP64 Jacobian apply J(psi) phi, timed with HIP events on the
stream it is launched on; algorithmic bytes per launch 40E + 68n (complex CSR
storing both directions: 16 B value + 4 B column per off-diagonal entry, 4 B
row pointer and 8 B diagonal per row, phi / y / psi 16 B and c*t 8 B per
vertex) against 8 TB/s.

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
                   help="only the J.phi timing loop and a copy kernel (for profilers)")
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
    p.add_argument("--cont-plain-branches", type=int, default=1,
                   help="branches run with the plain FP64 continuation for the ratio")
    p.add_argument("--cont-tuning", default="",
                   help="block_threads,fold[,mode] for the batched continuation (skips the "
                        "autotuner; mode 0 by default)")
    p.add_argument("--no-surrogate", action="store_true")
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


def cpu_baseline(coords, cells, A, mu, psi0, phi, counts, cheb, ratio, what):
    """Time the oracle on this host; extrapolate one solve from the GPU
    solver's per-solve counts (Newton steps and MINRES iterations)."""
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
    t_solve = (steps + 1) * t_f + steps * t_rb + its * t_it
    sample = (f"oracle on {threads} threads ({_cpu_model()}, nproc {os.cpu_count()}, "
              f"OMP_NUM_THREADS {os.environ.get('OMP_NUM_THREADS', 'unset')}), 96^3 cube, "
              f"medians of 20: F {t_f*1e3:.1f} ms, K.phi {t_k*1e3:.1f} ms, J.phi "
              f"{t_j*1e3:.1f} ms (1 thread: {t_j1*1e3:.1f} ms), rebuild {t_rb*1e3:.1f} ms, MINRES "
              f"iteration {t_it*1e3:.1f} ms (mean of {k} with Chebyshev {cheb}); one solve = "
              f"{steps + 1:.2f} F + {steps:.2f} rebuilds + {its:.1f} MINRES iterations, the "
              f"{what}'s per-solve counts; setup {setup:.1f} s")
    return {"value": 1.0 / t_solve, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": sample, "jphi_ms": t_j * 1e3, "jphi_ms_1thread": t_j1 * 1e3,
            "f_ms": t_f * 1e3, "keo_ms": t_k * 1e3, "seconds_per_solve": t_solve,
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


def continuation_phase(args, model, n):
    """mu-continuation of several (mu_0, g) branches on one mesh: the batched,
    mixed-precision, surrogate-seeded driver (timed over args.steps steps
    after args.warmup untimed ones) against the plain FP64 single-instance
    tangent continuation of a sample of the branches."""
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
        batch=args.cont_batch, lr=1e-3, newton_budget=6, retrain_steps=args.cont_retrain)
    if args.warmup > 0:
        t0 = time.perf_counter()
        batch.continuation(psi0, span * args.warmup / steps, args.warmup, **kw)
        say(f"continuation: warmup {args.warmup} steps {time.perf_counter() - t0:.2f}s")
    t0 = time.perf_counter()
    psi, rec, sst = batch.continuation(psi0, span, steps, surrogate=sur, **kw)
    t_comb = time.perf_counter() - t0
    conv = len(rec) == steps + 1 and all(r["converged"] == nb for r in rec)
    comb_rate = nb * steps / t_comb
    say(f"continuation: combined {t_comb:.2f}s for {nb}x{steps} solves -> {comb_rate:.3f}/s, "
        f"converged {conv}, surrogate {sst}")
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
    t_plain, plain_steps, plain_rec = 0.0, 0, []
    for b in range(min(args.cont_plain_branches, nb)):
        t1 = time.perf_counter()
        p, r1 = nosh.continuation(model, psi0[2 * n * b:2 * n * (b + 1)], {"g": g[b]}, "mu",
                                  mu0[b], mu0[b] + span, steps, predictor=1, tol=1e-10,
                                  max_iters=40, minres_rtol=1e-10,
                                  minres_max_iters=args.minres_maxit, cheb_iters=args.cheb,
                                  cheb_ratio=args.ratio)
        dt = time.perf_counter() - t1
        t_plain += dt
        plain_steps += steps
        ok = len(r1) == steps + 1 and all(r["converged"] for r in r1)
        plain_rec.append({"branch": b, "seconds": round(dt, 3), "converged": ok,
                          "newton_iters": sum(r["newton_iters"] for r in r1[1:]),
                          "minres_iters": sum(r["minres_iters"] for r in r1[1:]),
                          "halvings": sum(r["halvings"] for r in r1[1:]),
                          "final_gibbs": r1[-1]["gibbs"]})
        say(f"continuation: plain branch {b}: {dt:.2f}s, {plain_rec[-1]}")
    plain_rate = plain_steps / t_plain if t_plain > 0 else None
    seeded = [r for r in rec if r["seeded"]]
    return {
        "solves_per_s": comb_rate, "seconds": t_comb, "converged": conv,
        "branches": nb, "steps": steps, "span": span, "g": [round(v, 3) for v in g],
        "autotuned": {"block_threads": tuned[0], "fold": tuned[2],
                      "mode": MODE_NAMES[tuned[1]], "seconds": round(t_tune, 2)},
        "lockstep_newton_iters": sum(r["newton_iters"] for r in rec),
        "lockstep_minres_iters": sum(r["minres_iters"] for r in rec),
        "max_norm_f": max(r["max_norm_f"] for r in rec),
        "rel_err_vs_fp64_polished": {"branches": sorted({0, nb - 1}), "values": fp64_err},
        "surrogate": sst,
        "seeded_branch_steps": sum(r["seeded"] for r in rec),
        "fallback_branch_steps": sum(r["fallbacks"] for r in rec),
        "mean_pred_error_seeded_steps": (float(np.mean([r["pred_error"] for r in seeded]))
                                         if seeded else None),
        "mean_secant_error_seeded_steps": (float(np.mean([r["sec_error"] for r in seeded]))
                                           if seeded else None),
        "per_step_seconds": [round(r["seconds"], 4) for r in rec],
        "per_step_newton": [r["newton_iters"] for r in rec],
        "dominant_kernel_roofline": cheb_roof,
        "plain_fp64_solves_per_s": plain_rate, "plain_sample": plain_rec,
        "speedup_vs_plain_fp64": comb_rate / plain_rate if plain_rate else None,
    }


def jphi_roofline(args, mesh, model, params, psi0, phi, stream, dev, A):
    """FP64 J(psi) phi timed on the library's launch stream (warm and cold)."""
    import torch
    E, n = mesh.num_edges, mesh.num_vertices
    J = model.create_W_op()
    J.rebuild(params, psi0)
    x_ref = torch.from_numpy(phi).to(dev)
    x = torch.empty_like(x_ref)
    y = torch.empty_like(x_ref)
    mesh.to_internal(x_ref.data_ptr(), x.data_ptr(), stream.cuda_stream)
    with torch.cuda.stream(stream):
        for _ in range(5):
            J.apply_internal(x.data_ptr(), y.data_ptr(), stream.cuda_stream)
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for _ in range(args.spmv_reps):
            J.apply_internal(x.data_ptr(), y.data_ptr(), stream.cuda_stream)
        ev1.record(stream)
    ev1.synchronize()
    t_j = ev0.elapsed_time(ev1) / 1e3 / args.spmv_reps
    bytes_canon = 40 * E + 68 * n
    ach = bytes_canon / t_j / 1e9
    # Cold-cache variant: reading 1 GiB between launches replaces the 256 MB
    # Infinity Cache and the L2s with clean lines, so every launch streams its
    # operands from HBM (a write-based flush would leave dirty lines whose
    # write-back the next launch would pay for).
    flush = torch.ones(1 << 27, dtype=torch.float64, device=dev)
    cold = []
    with torch.cuda.stream(stream):
        for _ in range(10):
            fsum = flush.sum()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            J.apply_internal(x.data_ptr(), y.data_ptr(), stream.cuda_stream)
            e1.record(stream)
            cold.append((e0, e1))
    torch.cuda.synchronize()
    t_cold = float(np.median([a.elapsed_time(b) for a, b in cold])) / 1e3
    del flush, fsum
    layout = mesh.layout()
    bytes_layout = 20 * layout["nnz_pad"] + 4 * (layout["nslices"] + 1) + 56 * n
    # Copy-kernel reference of the same size class (stream bandwidth).
    cbuf = torch.empty(bytes_canon // 8, dtype=torch.float64, device=dev)
    cdst = torch.empty_like(cbuf)
    with torch.cuda.stream(stream):
        for _ in range(3):
            cdst.copy_(cbuf)
        ev0.record(stream)
        for _ in range(20):
            cdst.copy_(cbuf)
        ev1.record(stream)
    ev1.synchronize()
    t_copy = ev0.elapsed_time(ev1) / 1e3 / 20
    copy_gbs = 2 * cbuf.numel() * 8 / t_copy / 1e9
    del cbuf, cdst
    # FP32 J apply (inner solver's operator).
    x32 = x.float()
    y32 = torch.empty_like(x32)
    with torch.cuda.stream(stream):
        J.apply_internal_fp32(x32.data_ptr(), y32.data_ptr(), stream.cuda_stream)
        ev0.record(stream)
        for _ in range(args.spmv_reps):
            J.apply_internal_fp32(x32.data_ptr(), y32.data_ptr(), stream.cuda_stream)
        ev1.record(stream)
    ev1.synchronize()
    t_j32 = ev0.elapsed_time(ev1) / 1e3 / args.spmv_reps  # includes the block conversion
    # K.phi and F(psi) (FP64, internal order), against the canonical bytes
    # 40E + 44n and 40E + 52n.
    import nosh_amd as nosh
    others = {}
    keo = nosh.Keo(mesh, nosh.ScalarFieldConstant(mesh, 1.0),
                   nosh.VectorFieldExplicitValues(mesh, A, 0.0))
    keo.set_parameters(params)

    def timed(fn):
        with torch.cuda.stream(stream):
            for _ in range(3):
                fn()
            ev0.record(stream)
            for _ in range(args.spmv_reps):
                fn()
            ev1.record(stream)
        ev1.synchronize()
        return ev0.elapsed_time(ev1) / 1e3 / args.spmv_reps

    t_k = timed(lambda: keo.apply_internal(x.data_ptr(), y.data_ptr(), stream.cuda_stream))
    t_f = timed(lambda: model.eval_f_internal(x.data_ptr(), params, y.data_ptr(),
                                              stream.cuda_stream))
    for name, t_, b_ in (("keo_apply (K.phi FP64)", t_k, 40 * E + 44 * n),
                         ("compute_f (F(psi) FP64)", t_f, 40 * E + 52 * n),
                         ("k_block_apply<float> (J.phi FP32, with the block conversion)", t_j32,
                          24 * E + 36 * n)):
        others[name] = {"us_per_launch": round(t_ * 1e6, 2), "bytes_per_launch": b_,
                        "achieved": round(b_ / t_ / 1e9, 1),
                        "frac": round(b_ / t_ / 1e9 / HBM_PEAK_GBS, 4)}
    say("K.phi {:.1f} us ({:.0f} GB/s), F {:.1f} us ({:.0f} GB/s)".format(
        t_k * 1e6, (40 * E + 44 * n) / t_k / 1e9, t_f * 1e6, (40 * E + 52 * n) / t_f / 1e9))
    traffic, traffic_src = _profiled_traffic(args.ncell)
    say(f"J.phi {t_j*1e6:.1f} us -> {ach:.0f} GB/s ({ach/HBM_PEAK_GBS:.1%}); copy "
        f"{copy_gbs:.0f} GB/s; fp32 J.phi {t_j32*1e6:.1f} us")
    return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic,
            "traffic_source": traffic_src,
            "kernel": "k_block_apply<double> (J.phi FP64)",
            "us_per_launch": round(t_j * 1e6, 2), "bytes_per_launch": bytes_canon,
            "cold_us_per_launch": round(t_cold * 1e6, 2),
            "cold_achieved": round(bytes_canon / t_cold / 1e9, 1),
            "cold_frac": round(bytes_canon / t_cold / 1e9 / HBM_PEAK_GBS, 4),
            "copy_kernel_GBs": round(copy_gbs, 1),
            "layout_bytes_per_launch": bytes_layout,
            "layout_note": "sliced ELLPACK (64-row slices) of the off-diagonal entries: "
                           "16 B value + 4 B column per padded slot, 4 B per slice, "
                           "24 B 2x2 block + 16 B x + 16 B y per vertex",
            "nnz_pad": layout["nnz_pad"], "jphi_fp32_us": round(t_j32 * 1e6, 2),
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
    """The random-start solve, plain FP64 and FP32-inner, side by side."""
    import torch
    import nosh_amd as nosh
    psi_dev0 = torch.from_numpy(psi0).to(dev)
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
                        "final_norm_f": all_stats[-1]["final_norm_f"],
                        "newton_iters": all_stats[-1]["newton_iters"],
                        "minres_iters": all_stats[-1]["minres_iters_total"]}
        say(var, results[var])
    if "plain" in sols:
        ref = sols["plain"]
        for var, s in sols.items():
            d = np.linalg.norm(_phase_align(s, ref) - ref)
            results[var]["rel_diff_vs_fp64"] = float(d / max(np.linalg.norm(ref), 1e-300))
    return results


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
        cpu = cpu_baseline(coords, cells, A, 0.5 * args.cont_span, psi0, phi,
                           {"newton_iters": ps["newton_iters"] / args.steps,
                            "minres_iters": ps["minres_iters"] / args.steps},
                           args.cheb, args.ratio, "plain FP64 continuation")
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
                "SplitMix64 psi_0 for the random-start solves)",
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
        "random_start_solves": solves,
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
    kw = dict(tol=1e-10, max_iters=80, minres_rtol=1e-10, minres_max_iters=4000,
              cheb_iters=args.cheb, cheb_ratio=args.ratio, forcing=1, forcing_min=1e-6,
              forcing_max=0.1)
    def say(*a):
        print(*a, file=sys.stderr, flush=True)

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
                                 max_iters=80, minres_rtol=1e-10, minres_max_iters=4000,
                                 cheb_iters=args.cheb, cheb_ratio=args.ratio)
        t_plain += time.perf_counter() - t1
        worst = max(worst, s["final_norm_f"])
        say(f"plain instance {j} (mesh {sizes[k]}^3): {time.perf_counter() - t1:.2f}s, "
            f"newton {s['newton_iters']}, minres {s['minres_iters_total']}")
    plain_rate = len(sample) / t_plain
    rate = ninst / dt
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
