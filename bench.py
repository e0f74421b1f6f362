"""Benchmark: timesteps/s of the 1M-node 3D RBF-FD lid-driven cavity.

Workload (one GPU): 100^3 lattice on [0,1]^3, interior jitter +-0.3h (seeded),
edges/corners removed -> 998,816 nodes; PHS r^3 + degree-3 polynomial (20
monomials); stencils of 40 nodes, 60/55/50 on the face layer and the two
layers under it; Re = 400 (rho = 1, mu = 0.0025), lid u = 1 on z = 1,
dt = 5e-4, AB2, 500 steps from rest; pressure-Poisson relative residual 1e-9
(--tol).

`value` is the headline form's timesteps/s over the run's steps W..W+K-1
(defaults 5 and 495: the whole 500-step run after 5 warm-up steps): FP32-
operator inner GMRES(30) preconditioned by an FP32 smoothed-aggregation
multigrid V-cycle, in FP64 defect correction, with the learned initial guess
correcting the projection onto the last 20 solutions (the network trained
first, untimed, on a 63,536-node cube with the same physics: 2000 solver steps
in 10 episodes from rest with seeded lid speeds in [0.5, 1.5]).

Beside it, over the same steps of the same run: the multigrid form without a
guess (previous pressure), with the order-4 extrapolation in time, with the
projection alone, and the same solver in FP64 started from the previous
pressure (FP64 GMRES(30) + the multigrid V-cycle):
`speedup_vs_fp64_same_solver`.
The plain FP64 path (FP64 Jacobi-scaled BiCGSTAB, the reference's
solveWithGuess(rhs, p_old)) and its FP32-operator variant are too slow for 500
steps: they run an early window (steps 5-24) and a late window (steps 482-499,
restarted exactly from the FP64 multigrid run's state at step 480), and
`speedup_vs_fp64` compares the headline's own steps over the same windows.

Also reported: the fields after 3 steps against the FP64 oracle solved to
1e-12 (`vs_oracle_1M`) of the plain FP64 form (at the bench's tolerance and
at a tenth of it), the FP64 multigrid form and the headline; whole oracle steps
on the host cores as the CPU baseline (1M nodes, and the 2D 10k-node cavity
with the reference's ILUT(1e-4, 10) BiCGSTAB); the stencil weights against
the oracle on every stencil, with the kernel's FP64 rate; the 64-cavity batch.
The roofline entry times the PPE SpMV with HIP events on the library's stream,
over the launches that did the work.

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
             tune=True, krylov=0, start=None, save_at=None, timing_steps=2):
    """One form from rest (or from `start`, a saved state): `warmup` untimed
    steps, the autotuner, then `steps` steps each timed on the host clock (a
    step returns after the stream is synchronised).  guess: None (previous
    pressure), "extrapolate", or the learned network's parameters.  save_at:
    a step index whose end state (fields, AB2 history, step counter) is
    returned for an exact restart.  Two further steps time the kernels with
    HIP events (the replayed graphs cannot hold events); when the guess already
    meets the tolerance in those steps (no Krylov launch), up to eight more
    steps run until the solver's kernels have been timed."""
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
    elif isinstance(guess, str) and guess == "project":
        ctx.guess_mode(M.MP_GUESS_PROJECT)
    elif isinstance(guess, str):
        ctx.guess_mode(M.MP_GUESS_EXTRAPOLATE)
    else:
        ctx.guess_mode(M.MP_GUESS_LEARNED)
        ctx.guess_params(guess)
        ctx.guess_train(False)
    nv = cloud.nv
    fields = (M.MP_U, M.MP_V, M.MP_W, M.MP_P, M.MP_SU, M.MP_SV, M.MP_SW)
    if start is None:
        ctx.set_field(M.MP_U, u0)
        ctx.set_field(M.MP_V, np.zeros(nv))
        ctx.set_field(M.MP_W, np.zeros(nv))
        ctx.set_field(M.MP_P, np.zeros(ctx.n_sys))
        it = 0
    else:
        for f in fields:
            ctx.set_field(f, start[f])
        it = start["it"]
    saved = None
    for _ in range(warmup):
        ctx.step(1, it)
        st = ctx.stats()
        log(f"[{label}] warmup step {it}: iters {st.iterations} rel_res {st.rel_res:.2e}")
        it += 1
    if tune:
        interval, tune_ms = ctx.autotune()
        log(f"[{label}] poll interval {interval} (ms per solve {tune_ms})")
    else:
        interval = prm["check_every"]
    first = it
    iters, res, dts, rej, skipped, cycles, g_res = [], [], [], 0, 0, [], []
    ctx.sync()
    t0 = time.perf_counter()
    tp = t0
    for _ in range(steps):
        ctx.step(1, it)
        st = ctx.stats()
        tn = time.perf_counter()
        dts.append(tn - tp)
        tp = tn
        iters.append(st.iterations)
        res.append(st.rel_res)
        cycles.append(st.outer_cycles)
        g_res.append(st.guess_rel_res)
        rej += st.guess_rejected
        skipped += st.guess_skipped
        if save_at is not None and it == save_at:
            saved = {f: ctx.get_field(f) for f in fields}
            saved["it"] = it + 1
            tp = time.perf_counter()  # the download is not part of a step
        it += 1
    ctx.sync()
    t1 = time.perf_counter()
    out_fields = [ctx.get_field(w) for w in (M.MP_U, M.MP_V, M.MP_W, M.MP_P)]
    div = ctx.divergence()
    ctx.kernel_timing(True)
    if timing_steps > 0:
        ctx.step(timing_steps, it)
        extra = 0
        while ctx.kernel_timing_get(0)[0] < 4 and extra < 8:
            ctx.step(1, it + timing_steps + extra)
            extra += 1
    spmv = ctx.kernel_timing_get(0)
    pred = ctx.kernel_timing_get(1)
    rhs = ctx.kernel_timing_get(2)
    corr = ctx.kernel_timing_get(3)
    smooth = ctx.kernel_timing_get(5) if precond == M.MP_PRECOND_AMG else None
    ctx.kernel_timing(False)
    log(f"[{label}] steps {first}-{first + steps - 1}: {sum(dts):.3f}s, iters/step "
        f"{np.mean(iters):.1f}, max rel_res {max(res):.2e}, spmv {spmv}")
    return dict(seconds=float(sum(dts)), wall=t1 - t0, first=first, step_s=dts, iters=iters,
                rel_res=res, spmv=spmv, predictor=pred, rhs=rhs, corrector=corr, smooth=smooth, fields=out_fields,
                divergence=div.tolist(), hierarchy=hier, create_s=t_create,
                poll_interval=interval, saved=saved, guess_rejected=rej, guess_skipped=skipped,
                cycles=cycles, guess_res=g_res)


def step_kernels(form, copy_gbps):
    """Predictor (with its packing of u, v, w), right-hand side and corrector of
    the form's timing steps: mean ms per launch (HIP events) and the
    algorithmic bytes per launch over that time, against the 8 TB/s peak."""
    out = {}
    for name in ("predictor", "rhs", "corrector"):
        t = form.get(name)
        if not t or t[0] == 0:
            continue
        gbps = t[2] / (t[1] * 1e-3) / 1e9
        out[name] = {"mean_ms": t[1], "bytes_per_launch": t[2], "achieved_GBps": gbps,
                     "frac": gbps / 8000.0, "frac_of_copy": gbps / copy_gbps}
    return out


def window_seconds(form, lo, hi):
    """Host-clock seconds a full-run form spent on steps lo..hi."""
    f = form["first"]
    return float(sum(form["step_s"][lo - f:hi - f + 1]))


def form_summary(f):
    g = np.asarray(f["guess_res"], dtype=float)
    g = g[np.isfinite(g)]
    return {"timesteps_per_s": len(f["step_s"]) / f["seconds"], "steps": [f["first"],
            f["first"] + len(f["step_s"]) - 1], "ppe_iters_per_step": float(np.mean(f["iters"])),
            "outer_cycles_per_step": float(np.mean(f["cycles"])),
            "guess_rel_res_median": float(np.median(g)) if g.size else None,
            "max_ppe_rel_res": max(f["rel_res"]), "poll_interval": f["poll_interval"]}


def vs_oracle(M, O, ctx, cloud, w_oracle, prm, forms, steps=3, threads=0):
    """Fields after `steps` steps from rest of each GPU form (label -> run_form
    keyword arguments) against the FP64 oracle (Jacobi BiCGSTAB solved to
    1e-12) on the same cloud and weights: relative 2-norm error per field."""
    u_dir, p_dir, u0 = lid_bc(cloud)
    nv = cloud.nv
    O.set_threads(threads if threads > 0 else max(1, min(16, os.cpu_count() or 1)))
    fs = O.FractionalStep(cloud, w_oracle, u_dir, u_dir, u_dir, p_dir, prm["rho"], prm["mu"],
                          prm["dt"], 2, solver_kind=0, precond=1, tol=1e-12, maxit=20000)
    U = [u0.copy(), np.zeros(nv), np.zeros(nv)]
    p = np.zeros(fs.n_sys)
    t = time.perf_counter()
    its = [fs.step(it, U[0], U[1], U[2], p) for it in range(steps)]
    t_or = time.perf_counter() - t
    log(f"[vs oracle] oracle {steps} steps in {t_or:.1f}s, iterations {its}")
    out = {"steps": steps, "oracle": f"FP64 Jacobi BiCGSTAB to rel. residual 1e-12, iterations "
                                     f"{its}, {t_or:.1f}s"}
    for label, kw in forms.items():
        kw = dict(kw)
        p_form = dict(prm, tol=kw.pop("tol", prm["tol"]))
        r = run_form(M, ctx, cloud, p_form, steps=steps, warmup=0, label=f"vs oracle {label}",
                     tune=False, timing_steps=0, **kw)
        errs = {}
        for name, g, o in zip("uvwp", r["fields"], U + [p]):
            den = np.linalg.norm(o)
            errs[name] = float(np.linalg.norm(g - o) / den) if den > 0 else float(np.linalg.norm(g))
        out[label] = errs
        log(f"[vs oracle] {label}: {errs}")
    return out


def cpu_baseline_1m(O, cloud, w, prm, steps=2, threads=0):
    """Whole oracle FP64 steps (Jacobi BiCGSTAB to the bench's tolerance) from
    rest on the 1M-node cloud, on the host cores (OpenMP, at most 16)."""
    u_dir, p_dir, u0 = lid_bc(cloud)
    nv = cloud.nv
    cores = threads if threads > 0 else max(1, min(16, os.cpu_count() or 1))
    O.set_threads(cores)
    fs = O.FractionalStep(cloud, w, u_dir, u_dir, u_dir, p_dir, prm["rho"], prm["mu"], prm["dt"],
                          2, solver_kind=0, precond=1, tol=prm["tol"], maxit=20000)
    U = [u0.copy(), np.zeros(nv), np.zeros(nv)]
    p = np.zeros(fs.n_sys)
    t = time.perf_counter()
    its = [fs.step(it, U[0], U[1], U[2], p) for it in range(steps)]
    sec = time.perf_counter() - t
    return {"value": steps / sec, "unit": "timesteps/s", "cores": cores, "kind": "port",
            "sample": (f"{steps} whole oracle FP64 steps from rest on the same {nv}-node cloud "
                       f"(Jacobi BiCGSTAB to rel. residual {prm['tol']:g}, {its} iterations), "
                       f"{sec:.1f}s on {cores} threads"),
            "cpu_model": cpu_model()}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cavity2d_c1(M, O, steps=20, warmup=5, with_cpu=True):
    """The 2D 10k-node cavity (100^2 lattice, jitter +-0.25h, k = 30, Re = 100,
    BiCGSTAB to 1e-10; dt = 2e-4, inside the explicit scheme's stability limit):
    the oracle with the reference's ILUT(1e-4, fill 10) BiCGSTAB on 1 thread and
    on all host cores, and the GPU's plain FP64 path and headline solver."""
    cloud = M.box_cloud(dim=2, n=100, jitter=0.25, seed=20240601, k_interior=30)
    nv = cloud.nv
    rho, mu, dt, tol = 1.0, 0.01, 2e-4, 1e-10
    u_dir, p_dir, u0 = M.lid_conditions(cloud)
    out = {"workload": f"2D lid-driven cavity Re=100, {nv} nodes, k=30, dt={dt:g}, BiCGSTAB "
                       f"to rel. residual {tol:g}, steps {warmup}-{warmup + steps - 1}"}
    w = O.stencil_weights(cloud, 3, 3)[:4] if with_cpu else None
    if with_cpu:
        for thr in (1, max(1, min(16, os.cpu_count() or 1))):
            O.set_threads(thr)
            t = time.perf_counter()
            fs = O.FractionalStep(cloud, w, u_dir, u_dir, None, p_dir, rho, mu, dt, 2,
                                  solver_kind=0, precond=2, tol=tol, maxit=20000)
            t_setup = time.perf_counter() - t
            U = [u0.copy(), np.zeros(nv)]
            p = np.zeros(fs.n_sys)
            for it in range(warmup):
                fs.step(it, U[0], U[1], None, p)
            t = time.perf_counter()
            its = [fs.step(warmup + k, U[0], U[1], None, p) for k in range(steps)]
            sec = time.perf_counter() - t
            out[f"cpu_oracle_ilut_bicgstab_{thr}_threads"] = {
                "timesteps_per_s": steps / sec, "ppe_iters_per_step": float(np.mean(its)),
                "operator_and_ilut_setup_s": t_setup}
    for label, kw in (("gpu_fp64_jacobi_bicgstab", dict(precision=M.MP_PREC_FP64)),
                      ("gpu_mixed_multigrid_gmres", dict(precision=M.MP_PREC_MIXED,
                                                         precond=M.MP_PRECOND_AMG,
                                                         krylov=M.MP_KRYLOV_GMRES))):
        with M.Context(0) as ctx:
            ctx.upload(cloud)
            ctx.weights(3, 3, download=False)
            ctx.fs_create(rho, mu, dt, 2, tol, 20000, u_dir, u_dir, None, p_dir,
                          inner_tol=1e-5, **kw)
            ctx.set_field(M.MP_U, u0)
            ctx.set_field(M.MP_V, np.zeros(nv))
            ctx.set_field(M.MP_P, np.zeros(ctx.n_sys))
            ctx.step(warmup, 0)
            ctx.sync()
            its = []
            t = time.perf_counter()
            for k in range(steps):
                ctx.step(1, warmup + k)
                its.append(ctx.stats().iterations)
            sec = time.perf_counter() - t
            out[label] = {"timesteps_per_s": steps / sec, "ppe_iters_per_step": float(np.mean(its))}
    log(f"[2D cavity] {out}")
    return out


def smooth_roofline(t, copy_gbps, n_lattice: int = 0, gmres: bool = True):
    """Roofline entry of the multigrid's level-0 post-smoothing product (the
    V-cycle's largest kernel): algorithmic bytes per launch over its mean
    HIP-event duration, launches that did the work; `traffic` is its measured
    HBM bytes per launch on the 100^3 workload (None on others)."""
    if not t or t[1] <= 0:
        return None
    n, ms, bytes_l, all_l = t
    ach = bytes_l / (ms * 1e-3) / 1e9
    name = ("k_amg_op<0,2,float,float> (level-0 post-smoothing, BF16 values; reads the FP32 "
            "GMRES basis vector, writes the FP32 direction)" if gmres else
            "k_amg_op<0,2,double,double> (level-0 post-smoothing, BF16 values)")
    return {"kernel": name,
            "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": ach / HBM_PEAK_GBS, "frac_of_copy": ach / copy_gbps, "bytes_per_launch": bytes_l,
            "mean_ms": ms, "launches": n, "launches_returned_early": all_l - n,
            "traffic": measured_traffic(n_lattice, 0, "amg_level0_traffic_cavity3d_1M.md")
            if gmres else None}


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
    each pressure system (FP32-operator Krylov in FP64 defect correction,
    Jacobi or one multigrid hierarchy per cavity, started from the pressure
    extrapolated through the last four steps) stops on its own.  Beside it,
    the same cavities one context each, one after another, with the same
    solver, timed over seq_steps steps after 20 untimed ones (past the start
    from rest, with the guess's history full)."""
    probs = M.cavity_batch(nprob)
    clouds = [p[0] for p in probs]
    ctx = M.Context(0)
    t = time.perf_counter()
    cat = ctx.upload_batch(clouds)
    ctx.weights(3, 3, precision=M.MP_PREC_MIXED, download=False)
    u_dir, p_dir, u0 = M.lid_conditions(cat)
    ctx.fs_create(1.0, 0.01, 1e-4, 2, tol, 20000, u_dir, u_dir, None, p_dir,
                  precision=M.MP_PREC_MIXED, inner_tol=1e-5, precond=precond, krylov=krylov)
    ctx.guess_mode(M.MP_GUESS_EXTRAPOLATE)
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
            c1.guess_mode(M.MP_GUESS_EXTRAPOLATE)
            c1.set_field(M.MP_U, uz)
            c1.set_field(M.MP_V, np.zeros(cloud.nv))
            c1.set_field(M.MP_P, np.zeros(c1.n_sys))
            c1.step(20, 0)
            c1.sync()
            t0 = time.perf_counter()
            c1.step(seq_steps, 20)
            c1.sync()
            t_seq += time.perf_counter() - t0
    log(f"[batch] one context per cavity, sequential: steps 20-{19 + seq_steps} in "
        f"{t_seq:.2f}s")
    return {
        "workload": (f"{nprob} 2D lid-driven cavities, N log-uniform on [2000, 40000] "
                     f"({cat.nv} nodes total), Re ~ U[50, 400], jitter +-0.25h, k=30, "
                     f"FP32-operator Krylov ({pname}) in FP64 defect correction, "
                     f"guess extrapolated through the last 4 pressures, rel. residual {tol:g}"),
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


def weights_flops(cloud, m=20):
    """Algorithmic FP64 flops of the weight assembly: per stencil of k points
    (M = k + m) the LU (2/3 M^3), the forward and back substitution of d + 1
    right-hand sides (2 (d+1) M^2) and the entries of A and the right-hand
    sides (~10 flops per pair, 15 per right-hand-side entry)."""
    k = np.asarray(cloud.ksize, dtype=np.float64)
    Mv = k + m
    d = cloud.dim
    return float(np.sum((2.0 / 3.0) * Mv ** 3 + 2.0 * (d + 1) * Mv ** 2 + 10.0 * Mv ** 2 +
                        15.0 * (d + 1) * Mv))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=495,
                    help="timed steps of the headline run (with --warmup: the config's 500)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=100, help="lattice points per axis")
    ap.add_argument("--no-cpu", action="store_true", help="skip every host-side oracle leg")
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
                    help="Krylov method of the multigrid forms and of the batch (the plain FP64 "
                         "and mixed Jacobi forms use BiCGSTAB, the reference's)")
    ap.add_argument("--only-weights", action="store_true",
                    help="build the 1M-node cloud, time the stencil weights (FP64, then FP32 + "
                         "FP64 refinement), print them as one JSON line and stop (for profilers)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default poll interval and multigrid layouts (for counter "
                         "runs, under which the tuning's timed solves take minutes)")
    ap.add_argument("--forms", default="fp64,mixed,multigrid,extrapolate,project,fp64_multigrid,headline",
                    help="comma list of forms: fp64 (plain FP64 Jacobi BiCGSTAB, two windows), "
                         "mixed (FP32-inner Jacobi BiCGSTAB, two windows), multigrid, extrapolate, "
                         "project, fp64_multigrid, headline (full runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported",
              file=sys.stderr)
        sys.exit(2)
    import memphys_amd as M
    O = None
    if not args.no_cpu:
        import oracle as O

    if args.no_graphs:
        M.Context.default_graphs = False
    # torch's own HIP runtime initialises only before the library's does
    copy_gbps = stream_copy_gbps()
    log(f"streaming ceiling (torch.neg, 1 GiB each way): {copy_gbps:.0f} GB/s")
    prm = dict(rho=1.0, mu=0.0025, dt=5e-4, tol=args.tol, n_iter=20000, inner_tol=1e-5,
               check_every=8)
    forms = args.forms.split(",")
    tune = not args.no_autotune
    kry = M.MP_KRYLOV_GMRES if args.krylov == "gmres" else M.MP_KRYLOV_BICGSTAB
    kname = "GMRES(30)" if kry else "BiCGSTAB"
    guess = None
    t_train = 0.0
    small = M.box_cloud(dim=3, n=40, jitter=0.3, seed=101, k_interior=40, k_layers=(60, 55, 50))
    if "headline" in forms and not args.only_weights:
        t = time.perf_counter()
        episodes = 10
        guess = M.train_guess(small, prm["rho"], prm["mu"], prm["dt"], 1e-8, episodes,
                              max(1, args.train_steps // episodes), seed=2024, log=log,
                              precision=M.MP_PREC_MIXED, precond=M.MP_PRECOND_AMG,
                              krylov=M.MP_KRYLOV_GMRES, inner_tol=1e-5)
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

    # ---- stencil weights: FP64 (register kernel) and FP32 + FP64 refinement
    ctx.kernel_timing(True)
    (gx, gy, gz, lap), _ = ctx.weights(3, 3, download=not args.no_cpu)
    wk = ctx.kernel_timing_get(4)
    ctx.kernel_timing(False)
    ctx.kernel_timing(True)
    (mx, my, mz, ml), _ = ctx.weights(3, 3, precision=M.MP_PREC_MIXED, download=not args.no_cpu)
    wkm = ctx.kernel_timing_get(4)
    ctx.kernel_timing(False)
    flops = weights_flops(cloud)
    tflops = flops / (wk[1] * 1e-3) / 1e12
    weights_check = {"fp64_ms": wk[1], "fp32_refined_ms": wkm[1],
                     "gpu_fp64_stencils_per_s": cloud.nv / (wk[1] * 1e-3),
                     # the four FP64 weight arrays written (8 B per stencil entry each)
                     "fp64_write_GBps": 4 * 8 * cloud.nnz / (wk[1] * 1e-3) / 1e9,
                     "fp64_roofline": {"bound": "fp64", "achieved": tflops, "peak": 78.6,
                                       "unit": "TFLOP/s", "frac": tflops / 78.6,
                                       "flops_per_launch": flops}}
    w_oracle = None
    if O is not None:
        cores = max(1, min(16, os.cpu_count() or 1))
        O.set_threads(cores)
        t = time.perf_counter()
        ox, oy, oz, ol, _ = O.stencil_weights(cloud, 3, 3, threads=cores)
        t_or = time.perf_counter() - t
        w_oracle = (ox, oy, oz, ol)
        rp = cloud.row_ptr
        worst32 = np.zeros(cloud.nv)
        for b_, o in ((mx, ox), (my, oy), (mz, oz), (ml, ol)):
            num = np.add.reduceat((b_ - o) ** 2, rp[:-1])
            den = np.add.reduceat(o ** 2, rp[:-1])
            worst32 = np.maximum(worst32, np.sqrt(num / den))
        weights_check.update({
            "compared_stencils": cloud.nv,
            "fp64_bit_identical_to_oracle": bool(all(np.array_equal(a_, o) for a_, o in
                                                     zip((gx, gy, gz, lap), w_oracle))),
            "fp32_refined_worst_rel_vs_oracle": float(worst32.max()),
            "cpu_oracle_stencils_per_s": {str(cores): cloud.nv / t_or},
            "cpu_oracle_stencils_per_s_1_thread":
                cpu_weights_rate(cloud, np.random.default_rng(3).choice(cloud.nv, 3000,
                                                                        replace=False), 1)})
        del mx, my, mz, ml
    log(f"weights: {weights_check}")
    if args.only_weights:
        print(json.dumps({"workload": f"stencil weights, {cloud.nv} nodes (PHS r^3 + deg-3 "
                                      f"poly, k=40/50/55/60)", "weights": weights_check}))
        ctx.close()
        return
    ctx.weights(3, 3, download=False)  # FP64 weights back for the time steps

    # ---- full runs: W warm-up steps + K timed steps each
    full = {}
    late_lo = args.warmup + args.steps - 18
    save_at = late_lo - 3  # the state before the late window's two warm-up steps
    mg = dict(precision=M.MP_PREC_MIXED, precond=M.MP_PRECOND_AMG, krylov=kry)
    if "fp64_multigrid" in forms:
        full["fp64_multigrid"] = run_form(
            M, ctx, cloud, prm, M.MP_PREC_FP64, args.steps, args.warmup, "fp64+multigrid",
            precond=M.MP_PRECOND_AMG, krylov=kry, tune=tune, save_at=save_at)
    if "multigrid" in forms:
        full["mixed_multigrid"] = run_form(M, ctx, cloud, prm, M.MP_PREC_MIXED, args.steps,
                                           args.warmup, "mixed+multigrid", precond=M.MP_PRECOND_AMG,
                                           tune=tune, krylov=kry)
    if "project" in forms:
        full["mixed_multigrid_project"] = run_form(
            M, ctx, cloud, prm, M.MP_PREC_MIXED, args.steps, args.warmup,
            "mixed+multigrid+projection", guess="project", precond=M.MP_PRECOND_AMG,
            tune=tune, krylov=kry)
    if "extrapolate" in forms:
        full["mixed_multigrid_extrapolate"] = run_form(
            M, ctx, cloud, prm, M.MP_PREC_MIXED, args.steps, args.warmup,
            "mixed+multigrid+extrapolation", guess="extrapolate", precond=M.MP_PRECOND_AMG,
            tune=tune, krylov=kry)
    if "headline" in forms:
        head = run_form(M, ctx, cloud, prm, M.MP_PREC_MIXED, args.steps, args.warmup,
                        "mixed+multigrid+learned", guess=guess, precond=M.MP_PRECOND_AMG,
                        tune=tune, krylov=kry)
    else:
        head = next(iter(full.values()))
    # ---- windows of the Jacobi forms: early (steps W..W+19), late (the last 18
    # steps, restarted from the FP64 multigrid run's state three steps before)
    state = full["fp64_multigrid"]["saved"] if "fp64_multigrid" in full else None
    windows = {}
    for key, label, prec in (("fp64", "fp64 jacobi", M.MP_PREC_FP64),
                             ("mixed", "mixed jacobi", M.MP_PREC_MIXED)):
        if key not in forms:
            continue
        e = run_form(M, ctx, cloud, prm, prec, 20, args.warmup, label + " early", tune=tune)
        late = None
        if state is not None:
            late = run_form(M, ctx, cloud, prm, prec, 18, 2, label + " late", tune=tune,
                            start=state)
        windows[key] = (e, late)
    fp64 = windows["fp64"][0] if "fp64" in windows else head

    # ---- accuracy against the FP64 oracle after 3 steps
    accuracy = None
    if O is not None:
        # (the plain FP64 form also at a tenth of the bench's tolerance: the
        # error a tighter solve of the timed comparison's slower side leaves)
        accuracy = vs_oracle(M, O, ctx, cloud, w_oracle, prm, {
            "fp64_jacobi_bicgstab": dict(precision=M.MP_PREC_FP64),
            f"fp64_jacobi_bicgstab_tol_{prm['tol'] / 10:g}": dict(precision=M.MP_PREC_FP64,
                                                                  tol=prm["tol"] / 10),
            "fp64_multigrid_gmres": dict(precision=M.MP_PREC_FP64, precond=M.MP_PRECOND_AMG,
                                         krylov=kry),
            "headline": dict(precision=M.MP_PREC_MIXED, precond=M.MP_PRECOND_AMG, krylov=kry,
                             guess=guess)})

    # ---- ratios
    sps_head = args.steps / head["seconds"]
    ratios = {}
    if "fp64_multigrid" in full:
        fm = full["fp64_multigrid"]
        ratios["speedup_vs_fp64_same_solver"] = fm["seconds"] / head["seconds"]
    win = {}
    for key, (e, late) in windows.items():
        lo_e, hi_e = e["first"], e["first"] + len(e["step_s"]) - 1
        t_form = e["seconds"]
        t_head = window_seconds(head, lo_e, hi_e)
        desc = f"steps {lo_e}-{hi_e}"
        if late is not None:
            lo_l, hi_l = late["first"], late["first"] + len(late["step_s"]) - 1
            t_form += late["seconds"]
            t_head += window_seconds(head, lo_l, hi_l)
            desc += f" and {lo_l}-{hi_l}"
        win[key] = {"steps": desc, "timesteps_per_s": (len(e["step_s"]) + (len(late["step_s"])
                    if late else 0)) / t_form,
                    "ppe_iters_per_step_early": float(np.mean(e["iters"])),
                    "ppe_iters_per_step_late": float(np.mean(late["iters"])) if late else None,
                    "headline_speedup_same_steps": t_form / t_head}
    if "fp64" in win:
        ratios["speedup_vs_fp64"] = win["fp64"]["headline_speedup_same_steps"]
    fields_diff = None
    if "fp64_multigrid" in full and head is not full.get("fp64_multigrid"):
        a_, b_ = full["fp64_multigrid"]["fields"], head["fields"]
        fields_diff = {n: float(np.linalg.norm(x - y) / max(np.linalg.norm(x), 1e-300))
                       for n, x, y in zip("uvwp", a_, b_)}

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
                         f"dt=5e-4, AB2, steps {head['first']}-"
                         f"{head['first'] + args.steps - 1} of a run from rest, PPE rel. "
                         f"residual {args.tol:g}"),
            "nv": cloud.nv, "nnz": cloud.nnz,
            "form": (f"FP32-operator inner {kname} with FP32 smoothed-aggregation multigrid "
                     "V-cycle, in FP64 defect correction, + learned guess correcting the "
                     "projection onto the last 20 solutions"),
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
        **ratios,
        "headline": form_summary(head) | {
            "guess": {"steps": len(head["step_s"]),
                      "network_evaluated": len(head["step_s"]) - head["guess_skipped"],
                      "rejected_for_projection": head["guess_rejected"],
                      "accepted": len(head["step_s"]) - head["guess_skipped"] -
                                  head["guess_rejected"],
                      "rule": "a guess whose FP64 residual is non-finite or above 10x that "
                              "of the projection it corrects is replaced by the projection; "
                              "after each rejection the network is skipped for 1, 2, 4 ... 64 "
                              "steps"}},
        "full_run_forms": {k: form_summary(f) for k, f in full.items()},
        "jacobi_forms_windows": win,
        "fields_headline_vs_fp64_multigrid_rel_diff": fields_diff,
        "vs_oracle_1M": accuracy,
        "multigrid": {"levels": head["hierarchy"]["levels"], "rows": head["hierarchy"]["rows"],
                      "nnz": head["hierarchy"]["nnz"],
                      "host_setup_s": head["hierarchy"]["setup_s"]},
        "fp64_spmv_roofline": {"achieved": ach64, "frac": ach64 / HBM_PEAK_GBS,
                               "frac_of_copy": ach64 / copy_gbps,
                               "bytes_per_launch": b64, "mean_ms": ms64, "launches": n64,
                               "launches_returned_early": all64 - n64,
                               "traffic": measured_traffic(args.n, 1)},
        "amg_level0_roofline": smooth_roofline(head.get("smooth"), copy_gbps, args.n, bool(kry)),
        "step_kernels": step_kernels(head, copy_gbps),
        "cloud_setup": setup,
        "weights_kernel_ms": wk[1],
        "weights_fp64_vs_fp32_refined": weights_check,
        "divergence_headline": head["divergence"],
    }
    if args.batch_steps > 0:
        ctx.close()
        bp = M.MP_PRECOND_AMG if args.batch_precond == "multigrid" else M.MP_PRECOND_JACOBI
        out["batched_cavities"] = run_batched_cavities(M, args.batch_steps, 5, args.tol, bp, kry)
    else:
        ctx.close()
    out["cavity2d_10k"] = cavity2d_c1(M, O, with_cpu=O is not None)
    if O is not None:
        try:
            out["cpu_baseline"] = cpu_baseline_1m(O, cloud, w_oracle, prm)
        except Exception as e:  # the baseline is reported, never the product
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
