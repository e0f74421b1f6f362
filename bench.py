"""Benchmark: timesteps/s of the 1M-node 3D RBF-FD lid-driven cavity.

Workload (one GPU): 100^3 lattice on [0,1]^3, interior jitter +-0.3h (seeded),
edges/corners removed -> 998,816 nodes; PHS r^3 + degree-3 polynomial (20
monomials); stencils of 40 nodes, 60/55/50 on the face layer and the two
layers under it; Re = 400 (rho = 1, mu = 0.0025), lid u = 1 on z = 1,
dt = 5e-4, AB2; pressure-Poisson relative residual 1e-9 (--tol).

Forms timed over the same number of steps from the same warmed-up state: the
plain FP64 path (FP64 Jacobi-scaled BiCGSTAB, previous pressure as the guess,
as the reference's solveWithGuess(rhs, p_old)), the mixed form (FP32-operator
inner BiCGSTAB inside FP64 defect correction), the multigrid form (FP32-
operator inner GMRES(30), the reference's other Krylov option, preconditioned
by an FP32 smoothed-aggregation multigrid V-cycle; previous pressure as the
guess) and the headline form (the multigrid form with the learned guess;
--krylov bicgstab runs these two with BiCGSTAB).  The learned guess is
trained first, untimed, on a 63,536-node cube (40^3) with the same physics:
2000 solver steps in 10 episodes from rest with seeded random lid speeds in
[0.5, 1.5].  `value` is the headline
form's timesteps/s; the FP64 rate, the ratio, the residuals and the velocity
and pressure differences between the two are reported beside it.  The
roofline entry times the PPE SpMV with HIP events on the library's stream,
over the launches that did the work (launches of a system that converged
inside a chunk of iterations return at once and are counted apart).

Prints one JSON line.  --gpus N > 1 exits: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "timesteps/s, 1M-node 3D RBF-FD NS: learned+mixed vs plain FP64, same residual"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def lid_bc(cloud):
    nv = cloud.nv
    u_dir = np.ones(nv, dtype=np.uint8)
    p_dir = np.zeros(nv, dtype=np.uint8)
    u0 = np.zeros(nv)
    lid = (cloud.boundary == 1) & (np.abs(cloud.xyz[:, 2] - 1.0) < 1e-12)
    u0[lid] = 1.0
    return u_dir, p_dir, u0


def measured_traffic(n: int, kernel_index: int = 0,
                     summary: str = "ppe_spmv_traffic_cavity3d_1M.md"):
    """HBM bytes per launch measured with rocprofv3 FETCH_SIZE and WRITE_SIZE runs on
    the 100^3 workload, read from a committed summary: the `kernel_index`-th
    `traffic_bytes_per_launch` line (PPE SpMV summary: 0 the headline form's
    FP32-operator GMRES product, 1 the FP64 form's product; multigrid summary: 0 the
    level-0 post-smoothing product); None for other workloads."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", summary)
    if n != 100 or not os.path.exists(path):
        return None
    vals = [float(line.split("=")[1]) for line in open(path)
            if line.startswith("traffic_bytes_per_launch")]
    return vals[kernel_index] if len(vals) > kernel_index else None


def run_form(M, ctx, cloud, prm, precision, steps, warmup, label, guess=None, precond=0,
             tune=True, krylov=0):
    u_dir, p_dir, u0 = lid_bc(cloud)
    t = time.perf_counter()
    ctx.fs_create(prm["rho"], prm["mu"], prm["dt"], 2, prm["tol"], prm["n_iter"], u_dir, u_dir,
                  u_dir, p_dir, precision=precision, inner_tol=prm["inner_tol"],
                  check_every=prm["check_every"], precond=precond, krylov=krylov)
    t_create = time.perf_counter() - t
    hier = ctx.precond_info()
    log(f"[{label}] fs_create {t_create:.2f}s, hierarchy {hier}")
    if guess is None:
        ctx.guess_mode(M.MP_GUESS_PREVIOUS)
    else:
        ctx.guess_mode(M.MP_GUESS_LEARNED)
        ctx.guess_params(guess)
        ctx.guess_train(False)
    nv = cloud.nv
    ctx.set_field(M.MP_U, u0)
    ctx.set_field(M.MP_V, np.zeros(nv))
    ctx.set_field(M.MP_W, np.zeros(nv))
    ctx.set_field(M.MP_P, np.zeros(ctx.n_sys))
    for it in range(warmup):
        ctx.step(1, it)
        st = ctx.stats()
        log(f"[{label}] warmup step {it}: iters {st.iterations} rel_res {st.rel_res:.2e}")
    if tune:
        interval, tune_ms = ctx.autotune()
        log(f"[{label}] poll interval {interval} (ms per solve {tune_ms})")
    else:
        interval = prm["check_every"]
    it0 = warmup
    iters = []
    res = []
    ctx.sync()
    t0 = time.perf_counter()
    for k in range(steps):
        ctx.step(1, it0 + k)
        st = ctx.stats()
        iters.append(st.iterations)
        res.append(st.rel_res)
    ctx.sync()
    t1 = time.perf_counter()
    fields = [ctx.get_field(w) for w in (M.MP_U, M.MP_V, M.MP_W, M.MP_P)]
    div = ctx.divergence()
    # per-kernel HIP-event timing launches the Krylov chunks one by one (events
    # cannot sit inside the replayed graphs): two further steps of the same run
    ctx.kernel_timing(True)
    ctx.step(2, it0 + steps)
    spmv = ctx.kernel_timing_get(0)
    pred = ctx.kernel_timing_get(1)
    smooth = ctx.kernel_timing_get(5) if precond == M.MP_PRECOND_AMG else None
    ctx.kernel_timing(False)
    log(f"[{label}] {steps} steps in {t1 - t0:.3f}s, iters/step {np.mean(iters):.1f}, "
        f"max rel_res {max(res):.2e}, spmv {spmv}")
    return dict(seconds=t1 - t0, iters=iters, rel_res=res, spmv=spmv, predictor=pred, smooth=smooth,
                fields=fields, divergence=div.tolist(), hierarchy=hier, create_s=t_create,
                poll_interval=interval)


def smooth_roofline(t, copy_gbps, n_lattice: int = 0):
    """Roofline entry of the multigrid's level-0 post-smoothing product (the
    V-cycle's largest kernel): algorithmic bytes per launch over its mean
    HIP-event duration, launches that did the work; `traffic` is its measured
    HBM bytes per launch on the 100^3 workload (None on others)."""
    if not t or t[1] <= 0:
        return None
    n, ms, bytes_l, all_l = t
    ach = bytes_l / (ms * 1e-3) / 1e9
    return {"kernel": "k_amg_op<0,2,double,double> (level-0 post-smoothing, BF16 values)",
            "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": ach / HBM_PEAK_GBS, "frac_of_copy": ach / copy_gbps, "bytes_per_launch": bytes_l,
            "mean_ms": ms, "launches": n, "launches_returned_early": all_l - n,
            "traffic": measured_traffic(n_lattice, 0, "amg_level0_traffic_cavity3d_1M.md")}


def stream_copy_gbps(mib: int = 1024, reps: int = 10) -> float:
    """HBM rate of a plain elementwise kernel (torch.neg: reads and writes
    `mib` MiB once each), timed with events on torch's stream it runs on;
    median of `reps` launches.  A practical streaming ceiling beside the
    published peak, and the calibration kernel of HBM byte counters (it moves
    exactly 2 x `mib` MiB)."""
    import torch
    n = (mib << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.neg(src, out=dst)
    times = []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        torch.neg(src, out=dst)
        e1.record()
        e1.synchronize()
        times.append(e0.elapsed_time(e1))
    del src, dst
    torch.cuda.empty_cache()
    ms = sorted(times)[len(times) // 2]
    return 2.0 * (mib << 20) / (ms * 1e-3) / 1e9


def run_batched_cavities(M, steps, seq_steps, tol, precond, krylov=0, nprob=64):
    """64 independent 2D cavities (N log-uniform on [2k, 40k], Re ~ U[50, 400])
    stepped together in one context: one launch per phase serves every cavity,
    each pressure system (FP32-operator BiCGSTAB in FP64 defect correction,
    Jacobi or one multigrid hierarchy per cavity) stops on its own.  Beside
    it, the same cavities one context each, one after another, for seq_steps
    steps, with the same preconditioner."""
    probs = M.cavity_batch(nprob)
    clouds = [p[0] for p in probs]
    ctx = M.Context(0)
    t = time.perf_counter()
    cat = ctx.upload_batch(clouds)
    ctx.weights(3, 3, precision=M.MP_PREC_MIXED, download=False)
    u_dir, p_dir, u0 = M.lid_conditions(cat)
    ctx.fs_create(1.0, 0.01, 1e-4, 2, tol, 20000, u_dir, u_dir, None, p_dir,
                  precision=M.MP_PREC_MIXED, inner_tol=1e-5, precond=precond, krylov=krylov)
    ctx.batch_params([1.0] * nprob, [1.0 / p[1] for p in probs], [p[2] for p in probs])
    ctx.set_field(M.MP_U, u0)
    ctx.set_field(M.MP_V, np.zeros(cat.nv))
    ctx.set_field(M.MP_P, np.zeros(ctx.n_sys))
    t_setup = time.perf_counter() - t
    ctx.step(2, 0)
    interval, tune_ms = ctx.autotune()
    log(f"[batch] {nprob} cavities, {cat.nv} nodes, {cat.nnz} stencil entries; set-up "
        f"{t_setup:.2f}s; poll interval {interval} (ms per solve {tune_ms})")
    ctx.sync()
    t0 = time.perf_counter()
    worst = 0.0
    it_sum = 0
    for k in range(steps):
        ctx.step(1, 2 + k)
        st = ctx.stats()
        worst = max(worst, st.rel_res)
        it_sum += st.iterations
    ctx.sync()
    t_batch = time.perf_counter() - t0
    per_prob = [ctx.batch_stats(q) for q in range(nprob)]
    div = ctx.divergence()
    ctx.close()
    log(f"[batch] {steps} steps in {t_batch:.2f}s; slowest system {it_sum / steps:.0f} "
        f"iterations/step; max rel_res {worst:.2e}")
    # the same cavities one after another, one context each
    pname = ("GMRES(30)" if krylov == M.MP_KRYLOV_GMRES else "BiCGSTAB") + ", " + \
        ("multigrid V-cycle per cavity" if precond == M.MP_PRECOND_AMG else "Jacobi")
    t_seq = 0.0
    for q, (cloud, re, dt) in enumerate(probs):
        with M.Context(0) as c1:
            c1.upload(cloud)
            c1.weights(3, 3, precision=M.MP_PREC_MIXED, download=False)
            ud, pd, uz = M.lid_conditions(cloud)
            c1.fs_create(1.0, 1.0 / re, dt, 2, tol, 20000, ud, ud, None, pd,
                         precision=M.MP_PREC_MIXED, inner_tol=1e-5, precond=precond,
                         krylov=krylov)
            c1.set_field(M.MP_U, uz)
            c1.set_field(M.MP_V, np.zeros(cloud.nv))
            c1.set_field(M.MP_P, np.zeros(c1.n_sys))
            c1.step(2, 0)
            c1.sync()
            t0 = time.perf_counter()
            c1.step(seq_steps, 2)
            c1.sync()
            t_seq += time.perf_counter() - t0
    log(f"[batch] one context per cavity, sequential: {seq_steps} steps in {t_seq:.2f}s")
    return {
        "workload": (f"{nprob} 2D lid-driven cavities, N log-uniform on [2000, 40000] "
                     f"({cat.nv} nodes total), Re ~ U[50, 400], jitter +-0.25h, k=30, "
                     f"FP32-operator Krylov ({pname}) in FP64 defect correction, "
                     f"rel. residual {tol:g}"),
        "steps": steps,
        "cavity_steps_per_s": nprob * steps / t_batch,
        "batch_timesteps_per_s": steps / t_batch,
        "sequential_cavity_steps_per_s": nprob * seq_steps / t_seq,
        "speedup_vs_sequential": (nprob * steps / t_batch) / (nprob * seq_steps / t_seq),
        "poll_interval": interval,
        "max_rel_res": worst,
        "iterations_per_step_slowest": it_sum / steps,
        "iterations_last_step": {"min": min(s.iterations for s in per_prob),
                                 "max": max(s.iterations for s in per_prob)},
        "continuity_max_all": div[4],
        "setup_s": t_setup,
    }


def cpu_baseline(cloud, w, prm, iters_per_step, max_iter_sample=10, repeats=3):
    """Oracle FP64 step on the same 1M-node cloud, on 1 thread and on all the
    box's cores (OpenMP rows, at most 16): predictor/RHS and corrector timed in
    full, BiCGSTAB (Jacobi) timed over a few iterations and scaled to the
    iterations per step the FP64 GPU path needed; median of `repeats` runs."""
    import oracle as O
    u_dir, p_dir, u0 = lid_bc(cloud)
    nv = cloud.nv
    fs = O.FractionalStep(cloud, w, u_dir, u_dir, u_dir, p_dir, prm["rho"], prm["mu"], prm["dt"],
                          2, solver_kind=0, precond=1, tol=1e-30, maxit=max_iter_sample)
    cores = max(1, min(16, os.cpu_count() or 1))
    out = {}
    for nthr in (1, cores):
        O.set_threads(nthr)
        per_steps = []
        for _ in range(repeats):
            u, v, ww = u0.copy(), np.zeros(nv), np.zeros(nv)
            p = np.zeros(fs.n_sys)
            t0 = time.perf_counter()
            uh, b = fs.predict(0, u, v, ww, p)
            t1 = time.perf_counter()
            x, _, it, _ = fs.solve(b, p)
            t2 = time.perf_counter()
            fs.correct(u, v, ww, uh, x)
            t3 = time.perf_counter()
            t_iter = (t2 - t1) / max(it, 1)
            per_steps.append(((t1 - t0) + (t3 - t2)) + iters_per_step * t_iter)
        out[nthr] = (float(np.median(per_steps)), t1 - t0, t3 - t2, t_iter, it)
    O.set_threads(cores)
    ps, tp, tc, ti, it = out[cores]
    ps1, tp1, tc1, ti1, _ = out[1]
    return {"value": 1.0 / ps, "unit": "timesteps/s", "cores": cores, "kind": "port",
            "single_thread_value": 1.0 / ps1,
            "sample": (f"oracle FP64 step on the same {nv}-node cloud, median of {repeats}: "
                       f"predictor+rhs and corrector timed in full ({tp:.3f}s + {tc:.3f}s on "
                       f"{cores} threads, {tp1:.2f}s + {tc1:.2f}s on 1), Jacobi BiCGSTAB "
                       f"{it} iterations ({ti * 1e3:.1f} ms/iter on {cores} threads, "
                       f"{ti1 * 1e3:.1f} on 1) scaled to {iters_per_step:.0f} iterations/step "
                       f"(the FP64 GPU path's count)")}


def cloud_setup_timing(M, ctx, cloud, repeats=3):
    """Stencil selection and RCM order of the cloud's nodes: GPU (host arrays
    in and out, transfers included) against the host code on all cores;
    median of `repeats`, and whether both give the same CSR and order."""
    nv, dim = cloud.xyz.shape
    lib = M.host_lib()
    xyz = np.ascontiguousarray(cloud.xyz).ravel()
    ks = np.ascontiguousarray(cloud.ksize, dtype=np.int32)
    rp_h = np.zeros(nv + 1, dtype=np.int32)
    col_h = np.zeros(int(ks.sum()), dtype=np.int32)
    ord_h = np.zeros(nv, dtype=np.int32)
    ctx.knn_stencils(cloud.xyz[:64], np.zeros(64, dtype=np.uint8), np.full(64, 8, dtype=np.int32))
    t = {"knn_gpu": [], "knn_host": [], "rcm_gpu": [], "rcm_host": []}
    for _ in range(repeats):
        t0 = time.perf_counter()
        rp_g, col_g = ctx.knn_stencils(cloud.xyz, cloud.boundary, ks)
        t1 = time.perf_counter()
        lib.mph_knn(nv, dim, xyz, cloud.boundary, ks, rp_h, col_h, 0)
        t2 = time.perf_counter()
        ord_g = ctx.rcm_order(rp_g, col_g)
        t3 = time.perf_counter()
        lib.mph_rcm(nv, rp_h, col_h, ord_h)
        t4 = time.perf_counter()
        for key, dt in (("knn_gpu", t1 - t0), ("knn_host", t2 - t1), ("rcm_gpu", t3 - t2),
                        ("rcm_host", t4 - t3)):
            t[key].append(dt)
    out = {k + "_s": float(np.median(v)) for k, v in t.items()}
    out["host_threads"] = os.cpu_count()
    out["identical"] = bool(np.array_equal(rp_g, rp_h) and np.array_equal(col_g, col_h)
                            and np.array_equal(ord_g, ord_h))
    out["note"] = "GPU times include host-device transfers of the inputs and the CSR/order"
    return out


def cpu_weights_rate(cloud, rows, threads):
    """Oracle FP64 stencil weights on the sampled rows: stencils per second."""
    import oracle as O
    t0 = time.perf_counter()
    O.stencil_weights(cloud, 3, 3, rows=rows, threads=threads)
    return len(rows) / (time.perf_counter() - t0)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=100, help="lattice points per axis")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tol", type=float, default=1e-9)
    ap.add_argument("--train-steps", type=int, default=2000,
                    help="solver steps of guess training (10 episodes)")
    ap.add_argument("--batch-steps", type=int, default=200,
                    help="steps of the 64-cavity batch (0: skip it)")
    ap.add_argument("--batch-precond", choices=["multigrid", "jacobi"], default="multigrid",
                    help="preconditioner of the 64-cavity batch and its sequential comparison")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch every kernel instead of replaying hipGraphs (same results; "
                         "for profilers)")
    ap.add_argument("--krylov", choices=["gmres", "bicgstab"], default="gmres",
                    help="Krylov method of the multigrid and headline forms and of the batch "
                         "(the FP64 and mixed Jacobi forms use BiCGSTAB, the reference's)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default poll interval and multigrid layouts (for counter "
                         "runs, under which the tuning's timed solves take minutes)")
    ap.add_argument("--forms", default="fp64,mixed,multigrid,headline",
                    help="comma list of forms to time besides fp64: mixed (FP32-inner, "
                         "Jacobi), multigrid (FP32-inner, multigrid, previous pressure as "
                         "the guess), headline (FP32-inner, multigrid, learned guess)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported",
              file=sys.stderr)
        sys.exit(2)
    import memphys_amd as M

    if args.no_graphs:
        M.Context.default_graphs = False
    # torch's own HIP runtime initialises only before the library's does
    copy_gbps = stream_copy_gbps()
    log(f"streaming ceiling (torch.neg, 1 GiB each way): {copy_gbps:.0f} GB/s")
    prm = dict(rho=1.0, mu=0.0025, dt=5e-4, tol=args.tol, n_iter=20000, inner_tol=1e-5,
               check_every=8)
    t = time.perf_counter()
    small = M.box_cloud(dim=3, n=40, jitter=0.3, seed=101, k_interior=40, k_layers=(60, 55, 50))
    episodes = 10
    guess = M.train_guess(small, prm["rho"], prm["mu"], prm["dt"], 1e-8, episodes,
                          max(1, args.train_steps // episodes), seed=2024, log=log)
    t_train = time.perf_counter() - t
    log(f"guess training: {t_train:.1f}s on {small.nv} nodes")
    t = time.perf_counter()
    cloud = M.box_cloud(dim=3, n=args.n, jitter=0.3, seed=7, k_interior=40,
                        k_layers=(60, 55, 50))
    log(f"cloud nv={cloud.nv} nnz={cloud.nnz} setup {time.perf_counter() - t:.2f}s")
    ctx = M.Context(0)
    setup = cloud_setup_timing(M, ctx, cloud)
    log(f"cloud set-up: {setup}")
    ctx.upload(cloud)
    ctx.kernel_timing(True)
    t = time.perf_counter()
    (gx, gy, gz, lap), _ = ctx.weights(3, 3, download=not args.no_cpu)
    t_w = time.perf_counter() - t
    wk = ctx.kernel_timing_get(4)
    ctx.kernel_timing(False)
    log(f"weights: {t_w:.3f}s wall incl. download, kernel {wk[1]:.2f} ms")
    # FP32 LU + FP64 refinement of the same weights (config: batched assembly,
    # FP64 vs FP32+FP64-refined); the FP64 weights stay in use for the steps
    ctx.kernel_timing(True)
    (mx, my, mz, ml), _ = ctx.weights(3, 3, precision=M.MP_PREC_MIXED, download=not args.no_cpu)
    wkm = ctx.kernel_timing_get(4)
    ctx.kernel_timing(False)
    weights_check = {"fp64_ms": wk[1], "fp32_refined_ms": wkm[1]}
    if not args.no_cpu:
        import oracle as O
        rows = np.random.default_rng(3).choice(cloud.nv, 3000, replace=False)
        ox, oy, oz, ol, _ = O.stencil_weights(cloud, 3, 3, rows=rows, threads=0)
        worst64 = worst32 = 0.0
        for r in rows:
            s_ = slice(cloud.row_ptr[r], cloud.row_ptr[r + 1])
            for a, b, o in ((gx, mx, ox), (gy, my, oy), (gz, mz, oz), (lap, ml, ol)):
                den = np.linalg.norm(o[s_])
                worst64 = max(worst64, np.linalg.norm(a[s_] - o[s_]) / den)
                worst32 = max(worst32, np.linalg.norm(b[s_] - o[s_]) / den)
        cores = max(1, min(16, os.cpu_count() or 1))
        weights_check.update({"sample_stencils": 3000, "fp64_worst_rel_vs_oracle": worst64,
                              "fp32_refined_worst_rel_vs_oracle": worst32,
                              "gpu_fp64_stencils_per_s": cloud.nv / (wk[1] * 1e-3),
                              "cpu_oracle_stencils_per_s": {
                                  "1": cpu_weights_rate(cloud, rows, 1),
                                  str(cores): cpu_weights_rate(cloud, rows, cores)}})
        del mx, my, mz, ml
    ctx.weights(3, 3, download=False)  # FP64 weights back for the time steps
    log(f"weights check: {weights_check}")

    forms = args.forms.split(",")
    tune = not args.no_autotune
    kry = M.MP_KRYLOV_GMRES if args.krylov == "gmres" else M.MP_KRYLOV_BICGSTAB
    fp64 = run_form(M, ctx, cloud, prm, M.MP_PREC_FP64, args.steps, args.warmup, "fp64",
                    tune=tune)
    extra = {}
    if "mixed" in forms:
        extra["mixed_jacobi"] = run_form(M, ctx, cloud, prm, M.MP_PREC_MIXED, args.steps,
                                         args.warmup, "mixed", tune=tune)
    if "multigrid" in forms:  # the headline form without the learned guess
        extra["mixed_multigrid"] = run_form(M, ctx, cloud, prm, M.MP_PREC_MIXED, args.steps,
                                            args.warmup, "mixed+multigrid",
                                            precond=M.MP_PRECOND_AMG, tune=tune, krylov=kry)
    if "headline" in forms:
        head = run_form(M, ctx, cloud, prm, M.MP_PREC_MIXED, args.steps, args.warmup,
                        "mixed+multigrid+learned", guess=guess, precond=M.MP_PRECOND_AMG,
                        tune=tune, krylov=kry)
    else:
        head = fp64
    vdiff = max(np.linalg.norm(a - b) / max(np.linalg.norm(a), 1e-300)
                for a, b in zip(fp64["fields"][:3], head["fields"][:3]))
    pdiff = np.linalg.norm(fp64["fields"][3] - head["fields"][3]) / np.linalg.norm(fp64["fields"][3])

    sps_head = args.steps / head["seconds"]
    sps_fp64 = args.steps / fp64["seconds"]
    n_l, ms, bytes_l, all_l = head["spmv"]
    achieved = bytes_l / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    n64, ms64, b64, all64 = fp64["spmv"]
    ach64 = b64 / (ms64 * 1e-3) / 1e9 if ms64 > 0 else 0.0
    out = {
        "metric": METRIC,
        "value": sps_head,
        "unit": "timesteps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * head["seconds"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": (f"3D lid-driven cavity Re=400, {cloud.nv} jittered nodes "
                         f"({args.n}^3 lattice, +-0.3h), PHS r^3 + deg-3 poly, k=40/50/55/60, "
                         f"dt=5e-4, AB2, PPE rel. residual {args.tol:g}"),
            "nv": cloud.nv, "nnz": cloud.nnz,
            "form": (f"FP32-operator inner {'GMRES(30)' if kry else 'BiCGSTAB'} with FP32 "
                     "smoothed-aggregation multigrid V-cycle, in FP64 defect correction, + "
                     "learned guess"),
            "guess_training": f"{args.train_steps} solver steps on {small.nv} nodes, "
                              f"lid speeds U[0.5,1.5], {t_train:.1f}s",
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(args.n, 0) if kry else None,
            "kernel": ("k_gm_spmv<float,float,float> (PPE SpMV of GMRES, FP32 operator, FP32 "
                       "gathers of the V-cycle output)" if kry else
                       "k_spmv_alpha/k_spmv_omega<float,double,float> (PPE SpMV, FP32 "
                       "operator, FP32 gathers of the V-cycle output)"),
            "bytes_per_launch": bytes_l, "mean_ms": ms, "launches": n_l,
            "launches_returned_early": all_l - n_l,
            "copy_kernel_GBps": copy_gbps, "frac_of_copy": achieved / copy_gbps,
        },
        "fp64_timesteps_per_s": sps_fp64,
        "speedup_vs_fp64": sps_head / sps_fp64,
        "ppe_iters_per_step": {"fp64": float(np.mean(fp64["iters"])),
                               "headline": float(np.mean(head["iters"]))},
        "multigrid": {"levels": head["hierarchy"]["levels"], "rows": head["hierarchy"]["rows"],
                      "nnz": head["hierarchy"]["nnz"],
                      "host_setup_s": head["hierarchy"]["setup_s"]},
        "poll_interval": {"fp64": fp64["poll_interval"], "headline": head["poll_interval"]},
        "other_forms": {k: {"timesteps_per_s": args.steps / f["seconds"],
                            "ppe_iters_per_step": float(np.mean(f["iters"])),
                            "max_ppe_rel_res": max(f["rel_res"])} for k, f in extra.items()},
        "max_ppe_rel_res": {"fp64": max(fp64["rel_res"]), "headline": max(head["rel_res"])},
        "headline_vs_fp64_rel_diff": {"velocity": vdiff, "pressure": pdiff},
        "fp64_spmv_roofline": {"achieved": ach64, "frac": ach64 / HBM_PEAK_GBS,
                               "frac_of_copy": ach64 / copy_gbps,
                               "bytes_per_launch": b64, "mean_ms": ms64, "launches": n64,
                               "launches_returned_early": all64 - n64,
                               "traffic": measured_traffic(args.n, 1)},
        "amg_level0_roofline": smooth_roofline(head.get("smooth"), copy_gbps, args.n),
        "cloud_setup": setup,
        "weights_kernel_ms": wk[1],
        "weights_fp64_vs_fp32_refined": weights_check,
        "divergence_fp64": fp64["divergence"],
    }
    if args.batch_steps > 0:
        ctx.close()
        bp = M.MP_PRECOND_AMG if args.batch_precond == "multigrid" else M.MP_PRECOND_JACOBI
        out["batched_cavities"] = run_batched_cavities(M, args.batch_steps, 5, args.tol, bp, kry)
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(cloud, (gx, gy, gz, lap), prm,
                                               float(np.mean(fp64["iters"])))
        except Exception as e:  # the baseline is reported, never the product
            out["cpu_baseline"] = {"error": str(e)}
    if args.batch_steps <= 0:
        ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
