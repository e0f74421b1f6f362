"""Text summaries of rocprofv3 result databases (rocpd SQLite).

    python -m memphys_amd.profiling stats DB [--top N]
        per kernel and grid size: dispatches, total ms, mean/min/max us, and the
        mean over the launches that did the work (at least half the median
        duration: Krylov launches of a converged system return at once)
    python -m memphys_amd.profiling iteration DB [--first K] [--marker K] [--product K]
        one preconditioned Krylov iteration of the largest system: the
        stretches between consecutive launches of `first` (default
        k_update_p<double>, BiCGSTAB's; k_gm_scale for GMRES) followed by a
        `marker` launch (default k_amg_op, the V-cycle), with a working
        `product` launch (default k_spmv_alpha; k_gm_spmv for GMRES); their
        median wall and busy time, and the kernel list of the iteration of
        median wall time
    python -m memphys_amd.profiling steps DB [--marker K] [--last N]
        per timestep of the largest system (stretches between launches of
        `marker`, default predictor_kernel): wall, busy and idle time, launches
        and kernel time per category (predictor, right-hand side, corrector,
        learned guess, Krylov, the multigrid's level-0 smoothing and the rest)
        and the idle stretches of the last step
    python -m memphys_amd.profiling pmc DB [--top N]
        per kernel, grid size and counter: dispatches and mean value per dispatch
    python -m memphys_amd.profiling traffic FETCH_DB WRITE_DB --kernel SUBSTR
                                    --copy-kernel SUBSTR --copy-bytes B [--grid G]
        HBM bytes per dispatch of a kernel: FETCH_SIZE and WRITE_SIZE (KiB per
        dispatch) each divided by what the same counter reads, per byte, for a
        plain copy kernel of known size (B bytes read and B bytes written)

FETCH_SIZE and WRITE_SIZE come from separate runs (one counter group each).
"""
from __future__ import annotations

import argparse
import sqlite3
import sys


def _short(name: str) -> str:
    name = name.replace("(anonymous namespace)::", "").replace("void ", "")
    depth, out = 0, []
    for ch in name:  # drop the argument list, keep template arguments
        if ch == "(" and depth == 0 and out:
            break
        depth += ch == "<"
        depth -= ch == ">"
        out.append(ch)
    return "".join(out).strip()


def _working(durs: list) -> list:
    """Launches that did the work: duration at least half the median launch's
    (Krylov kernels launched after their system converged return at once).
    The same rule as the library's HIP-event kernel timer."""
    srt = sorted(durs)
    cut = 0.5 * srt[len(srt) // 2]
    return [d for d in durs if d >= cut]


def stats(db: str, top: int) -> str:
    cur = sqlite3.connect(db).cursor()
    per = {}
    for name, grid, d in cur.execute("select name, grid_x, end-start from kernels"):
        per.setdefault((name, grid), []).append(d)
    rows = sorted(per.items(), key=lambda kv: -sum(kv[1]))
    tot = sum(sum(v) for _, v in rows) or 1
    lines = [f"{'kernel':56s} {'grid':>9s} {'calls':>7s} {'working':>7s} {'total_ms':>9s} "
             f"{'mean_us':>8s} {'work_us':>8s} {'min_us':>7s} {'max_us':>8s} {'pct':>6s}"]
    for (name, grid), v in rows[:top]:
        w = _working(v)
        lines.append(f"{_short(name)[:56]:56s} {grid:9d} {len(v):7d} {len(w):7d} "
                     f"{sum(v) / 1e6:9.2f} {sum(v) / len(v) / 1e3:8.1f} "
                     f"{sum(w) / len(w) / 1e3:8.1f} {min(v) / 1e3:7.1f} {max(v) / 1e3:8.1f} "
                     f"{100 * sum(v) / tot:6.2f}")
    lines.append(f"total kernel time {tot / 1e6:.2f} ms over {sum(len(v) for _, v in rows)} "
                 f"dispatches; working = launches of at least half the median duration of "
                 f"their kernel and grid, work_us their mean")
    return "\n".join(lines)


def iteration(db: str, first: str = "k_update_p<double>", marker: str = "k_amg_op",
              product: str = "k_spmv_alpha") -> str:
    cur = sqlite3.connect(db).cursor()
    rows = cur.execute("select name, start, end, grid_x from kernels order by start").fetchall()
    grid = max((r[3] for r in rows if first in r[0]), default=None)
    if grid is None:
        raise SystemExit(f"no launch of '{first}' in {db}")
    heads = [i for i, r in enumerate(rows[:-1])
             if first in r[0] and r[3] == grid and marker in rows[i + 1][0]]
    prod = [r[2] - r[1] for r in rows if product in r[0] and r[3] >= grid]
    if not heads or not prod:
        raise SystemExit(f"no preconditioned iteration at grid {grid} in {db}")
    cut = min(_working(prod))
    its = []
    for a, b in zip(heads[:-1], heads[1:]):
        seg = rows[a:b]
        if b - a > 80 or not any(product in r[0] and r[3] >= grid and r[2] - r[1] >= cut
                                 for r in seg):
            continue  # a solve's last iteration, or a converged system's chunk
        its.append((rows[b][1] - rows[a][1], sum(r[2] - r[1] for r in seg), a, b))
    if not its:
        raise SystemExit("no complete iteration found")
    its.sort()
    wall, _, a, b = its[len(its) // 2]
    busy = sorted(x[1] for x in its)[len(its) // 2]
    lines = [f"iterations {len(its)}, median wall {wall / 1e3:.1f} us, "
             f"median busy {busy / 1e3:.1f} us (grid of {first}: {grid})"]
    t0 = rows[a][1]
    for name, s, e, g in rows[a:b]:
        lines.append(f"{_short(name)[:46]:46s} grid {g:9d} start {(s - t0) / 1e3:8.1f} us "
                     f"dur {(e - s) / 1e3:7.1f} us")
    return "\n".join(lines)


_CATEGORIES = (  # (label, name substrings), first match wins
    ("predictor", ("predictor_kernel",)),
    ("ppe rhs", ("ppe_rhs_kernel",)),
    ("corrector + boundary", ("corrector_kernel", "boundary_fix_kernel")),
    ("learned guess", ("gnn", "k_guess")),
    ("gmres", ("k_gm_",)),
    ("bicgstab", ("k_update_", "k_spmv_alpha", "k_spmv_omega", "k_finish_", "k_residual_init",
                  "k_sc_reset")),
)


def steps(db: str, marker: str = "predictor_kernel", last: int = 0) -> str:
    """Per timestep of the largest system: the stretches between consecutive
    launches of `marker` at its largest grid.  Wall time, busy time (union of
    kernel intervals), launches, and kernel time per category; multigrid
    launches split into the finest level's two smoothing products and the rest
    (restriction, prolongation, coarse levels).  Then the idle stretches of
    the last step: the largest, and a histogram by length."""
    cur = sqlite3.connect(db).cursor()
    rows = cur.execute("select name, start, end, grid_x from kernels order by start").fetchall()
    grid = max((r[3] for r in rows if marker in r[0]), default=None)
    if grid is None:
        raise SystemExit(f"no launch of '{marker}' in {db}")
    heads = [i for i, r in enumerate(rows) if marker in r[0] and r[3] == grid]
    wins = list(zip(heads[:-1], heads[1:]))
    if last:
        wins = wins[-last:]
    cats = [c for c, _ in _CATEGORIES] + ["mg l0 smooth", "multigrid other", "other"]
    lines = [f"timesteps {len(wins)} (marker {marker}, grid {grid}); times in ms"]
    lines.append(f"{'step':>4s} {'wall':>7s} {'busy':>7s} {'idle':>6s} {'launch':>6s} " +
                 " ".join(f"{c[:12]:>12s}" for c in cats))
    other_names = {}
    for n, (a, b) in enumerate(wins):
        seg = rows[a:b]
        per = dict.fromkeys(cats, 0.0)
        busy, cur_s, cur_e = 0, None, None
        for name, s, e, g in seg:
            if cur_e is None or s > cur_e:
                if cur_e is not None:
                    busy += cur_e - cur_s
                cur_s, cur_e = s, e
            else:
                cur_e = max(cur_e, e)
            for c, keys in _CATEGORIES:
                if any(k in name for k in keys):
                    per[c] += e - s
                    break
            else:
                if "k_amg" in name:
                    # the finest level's two smoothing products (SELL layout,
                    # one lane per row: their grid is the marker's)
                    l0 = ("k_amg_op<0, 2" in name or "k_amg_op<0, 3" in name) and g >= grid
                    per["mg l0 smooth" if l0 else "multigrid other"] += e - s
                else:
                    per["other"] += e - s
                    k = _short(name)
                    other_names[k] = other_names.get(k, 0) + e - s
        busy += cur_e - cur_s
        wall = rows[b][1] - seg[0][1]
        lines.append(f"{n:4d} {wall / 1e6:7.2f} {busy / 1e6:7.2f} {(wall - busy) / 1e6:6.2f} "
                     f"{len(seg):6d} " + " ".join(f"{per[c] / 1e6:12.3f}" for c in cats))
    top = sorted(other_names.items(), key=lambda kv: -kv[1])[:8]
    lines.append("other: " + ", ".join(f"{k} {v / 1e6:.3f}" for k, v in top))
    # idle stretches of the last step: where the device waited for the host
    a, b = wins[-1]
    gaps = []
    for k in range(a, b):
        end = max(r[2] for r in rows[a:k + 1])
        g = rows[k + 1][1] - end
        if g > 0:
            gaps.append((g, k))
    gaps.sort(reverse=True)
    lines.append(f"largest idle stretches of step {len(wins) - 1} (total "
                 f"{sum(g for g, _ in gaps) / 1e6:.3f} ms over {len(gaps)} stretches):")
    for g, k in gaps[:12]:
        lines.append(f"  {g / 1e3:8.1f} us after {_short(rows[k][0])[:40]} "
                     f"(grid {rows[k][3]}) before {_short(rows[k + 1][0])[:40]}")
    hist = {}
    for g, _ in gaps:
        key = "<2us" if g < 2e3 else "2-10us" if g < 1e4 else "10-50us" if g < 5e4 else ">=50us"
        hist[key] = hist.get(key, [0, 0])
        hist[key][0] += 1
        hist[key][1] += g
    lines.append("  by size: " + ", ".join(f"{k}: {v[0]} stretches {v[1] / 1e6:.3f} ms"
                                          for k, v in sorted(hist.items())))
    return "\n".join(lines)


def _pmc_rows(db: str):
    cur = sqlite3.connect(db).cursor()
    return cur.execute(
        "select kernel_name, grid_size, counter_name, count(*), avg(value), avg(duration) "
        "from counters_collection group by kernel_name, grid_size, counter_name "
        "order by count(*) * avg(duration) desc").fetchall()


def pmc(db: str, top: int) -> str:
    lines = [f"{'kernel':64s} {'grid':>10s} {'counter':>12s} {'calls':>7s} {'mean':>14s} "
             f"{'mean_us':>8s}"]
    for r in _pmc_rows(db)[:top]:
        lines.append(f"{_short(r[0])[:64]:64s} {r[1]:10d} {r[2]:>12s} {r[3]:7d} {r[4]:14.1f} "
                     f"{r[5] / 1e3:8.1f}")
    return "\n".join(lines)


def _counter(db: str, counter: str, pattern: str, grid: int | None):
    """Mean counter value of the largest launch shape of the kernel, over the
    dispatches that did the work: duration at least half the median dispatch's
    (Krylov launches of a converged system return at once)."""
    best = None
    for name, g, cname, n, mean, dur in _pmc_rows(db):
        if cname != counter or pattern not in name or (grid is not None and g != grid):
            continue
        if best is None or g > best[1]:  # the largest launch of the kernel
            best = (name, g)
    if best is None:
        raise SystemExit(f"no dispatch of '{pattern}' with {counter} in {db}")
    cur = sqlite3.connect(db).cursor()
    rows = cur.execute("select value, duration from counters_collection where kernel_name = ? "
                       "and grid_size = ? and counter_name = ?",
                       (best[0], best[1], counter)).fetchall()
    cut = min(_working([r[1] for r in rows]))
    work = [r for r in rows if r[1] >= cut]
    mean = sum(r[0] for r in work) / len(work)
    dur = sum(r[1] for r in work) / len(work)
    return best[0], best[1], len(work), mean, dur, len(rows)


def traffic(fetch_db: str, write_db: str, kernel: str, copy_kernel: str, copy_bytes: int,
            grid: int | None) -> str:
    cf = _counter(fetch_db, "FETCH_SIZE", copy_kernel, None)
    cw = _counter(write_db, "WRITE_SIZE", copy_kernel, None)
    kf = _counter(fetch_db, "FETCH_SIZE", kernel, grid)
    kw = _counter(write_db, "WRITE_SIZE", kernel, grid)
    f_scale = cf[3] * 1024.0 / copy_bytes
    w_scale = cw[3] * 1024.0 / copy_bytes
    fetch_b = kf[3] * 1024.0 / f_scale
    write_b = kw[3] * 1024.0 / w_scale
    return "\n".join([
        f"copy kernel: {_short(cf[0])} grid {cf[1]}: FETCH_SIZE {cf[3]:.1f} KiB, "
        f"WRITE_SIZE {cw[3]:.1f} KiB for {copy_bytes} bytes each way",
        f"scale factors (counter bytes per true byte): fetch {f_scale:.4f}, write {w_scale:.4f}",
        f"kernel: {_short(kf[0])} grid {kf[1]}: {kf[2]} of {kf[5]} dispatches did the work "
        f"(the rest returned at once); over those: FETCH_SIZE {kf[3]:.1f} KiB, "
        f"WRITE_SIZE {kw[3]:.1f} KiB per dispatch, mean duration under the counters "
        f"{kf[4] / 1e3:.1f} us",
        f"fetch_bytes_per_launch = {fetch_b:.0f}",
        f"write_bytes_per_launch = {write_b:.0f}",
        f"traffic_bytes_per_launch = {fetch_b + write_b:.0f}",
    ])


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m memphys_amd.profiling")
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("stats")
    a.add_argument("db")
    a.add_argument("--top", type=int, default=40)
    it = sub.add_parser("iteration")
    it.add_argument("db")
    it.add_argument("--first", default="k_update_p<double>")
    it.add_argument("--marker", default="k_amg_op")
    it.add_argument("--product", default="k_spmv_alpha")
    st = sub.add_parser("steps")
    st.add_argument("db")
    st.add_argument("--marker", default="predictor_kernel")
    st.add_argument("--last", type=int, default=0)
    b = sub.add_parser("pmc")
    b.add_argument("db")
    b.add_argument("--top", type=int, default=40)
    c = sub.add_parser("traffic")
    c.add_argument("fetch_db")
    c.add_argument("write_db")
    c.add_argument("--kernel", required=True)
    c.add_argument("--copy-kernel", required=True)
    c.add_argument("--copy-bytes", type=int, required=True)
    c.add_argument("--grid", type=int, default=None)
    args = ap.parse_args(argv)
    if args.cmd == "stats":
        print(stats(args.db, args.top))
    elif args.cmd == "iteration":
        print(iteration(args.db, args.first, args.marker, args.product))
    elif args.cmd == "steps":
        print(steps(args.db, args.marker, args.last))
    elif args.cmd == "pmc":
        print(pmc(args.db, args.top))
    else:
        print(traffic(args.fetch_db, args.write_db, args.kernel, args.copy_kernel,
                      args.copy_bytes, args.grid))


if __name__ == "__main__":
    sys.exit(main())
