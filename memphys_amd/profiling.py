"""Text summaries of rocprofv3 result databases (rocpd SQLite).

    python -m memphys_amd.profiling stats DB [--top N] [--busy-us T]
        per kernel and grid size: dispatches, total ms, mean/min/max us, and the
        mean over launches longer than T us
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


def stats(db: str, top: int, busy_us: float = 3.0) -> str:
    """busy: launches longer than busy_us (solver kernels launched after their
    system converged return at entry in ~1-2 us and are left out there)."""
    cur = sqlite3.connect(db).cursor()
    rows = cur.execute(
        "select name, grid_x, count(*), sum(end-start), avg(end-start), min(end-start), "
        "max(end-start), sum(case when end-start > ? then 1 else 0 end), "
        "avg(case when end-start > ? then end-start end) from kernels group by name, grid_x "
        "order by sum(end-start) desc", (busy_us * 1e3, busy_us * 1e3)).fetchall()
    tot = sum(r[3] for r in rows) or 1
    lines = [f"{'kernel':56s} {'grid':>9s} {'calls':>7s} {'busy':>7s} {'total_ms':>9s} "
             f"{'mean_us':>8s} {'busy_us':>8s} {'min_us':>7s} {'max_us':>8s} {'pct':>6s}"]
    for r in rows[:top]:
        busy_mean = (r[8] or 0.0) / 1e3
        lines.append(f"{_short(r[0])[:56]:56s} {r[1]:9d} {r[2]:7d} {r[7]:7d} {r[3] / 1e6:9.2f} "
                     f"{r[4] / 1e3:8.1f} {busy_mean:8.1f} {r[5] / 1e3:7.1f} {r[6] / 1e3:8.1f} "
                     f"{100 * r[3] / tot:6.2f}")
    lines.append(f"total kernel time {tot / 1e6:.2f} ms over {sum(r[2] for r in rows)} dispatches; "
                 f"busy = launches longer than {busy_us:g} us")
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
    best = None
    for name, g, cname, n, mean, dur in _pmc_rows(db):
        if cname != counter or pattern not in name or (grid is not None and g != grid):
            continue
        if best is None or g > best[1]:  # the largest launch of the kernel
            best = (name, g, n, mean, dur)
    if best is None:
        raise SystemExit(f"no dispatch of '{pattern}' with {counter} in {db}")
    return best


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
        f"kernel: {_short(kf[0])} grid {kf[1]} ({kf[2]} dispatches): FETCH_SIZE {kf[3]:.1f} KiB, "
        f"WRITE_SIZE {kw[3]:.1f} KiB per dispatch",
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
    a.add_argument("--busy-us", type=float, default=3.0)
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
        print(stats(args.db, args.top, args.busy_us))
    elif args.cmd == "pmc":
        print(pmc(args.db, args.top))
    else:
        print(traffic(args.fetch_db, args.write_db, args.kernel, args.copy_kernel,
                      args.copy_bytes, args.grid))


if __name__ == "__main__":
    sys.exit(main())
