"""Text summaries of rocprofv3 result databases (rocpd SQLite).

    python -m astroemperor_amd.profiling RESULTS.db [--counters | --gaps] [--by-grid] [--top N]

Without options: one row per kernel (dispatches, total and mean device
time, VGPRs, LDS) from a ``--kernel-trace`` run (per kernel and grid size
with --by-grid).  With --gaps: dispatches in
time order grouped by (previous kernel, kernel), with the mean duration and
the mean idle time between the previous kernel's end and this one's start:
the launch sequence of a captured sweep and what each boundary costs.  With
--counters: the mean value per dispatch of every collected counter, per
kernel (or per kernel and grid size with --by-grid, which separates launches
of one kernel over different batch sizes), from a ``--pmc`` run (FETCH_SIZE /
WRITE_SIZE are in KiB).
"""
from __future__ import annotations

import argparse
import sqlite3


def kernel_table(db, top=12, by_grid=False):
    c = sqlite3.connect(db)
    group = "name, grid_x" if by_grid else "name"
    rows = c.execute(
        "select name, count(*), sum(duration), avg(duration), avg(vgpr_count), avg(lds_size), avg(grid_x), "
        f"avg(workgroup_x) from kernels group by {group} order by sum(duration) desc limit ?", (top,)).fetchall()
    total = c.execute("select sum(duration) from kernels").fetchone()[0] or 1
    out = ["| kernel | dispatches | total ms | mean us | share | VGPRs | LDS B | grid | WG |",
           "|---|---|---|---|---|---|---|---|---|"]
    for name, n, tot, avg, vg, lds, grid, wg in rows:
        out.append(f"| `{name[:80]}` | {n} | {tot / 1e6:.3f} | {avg / 1e3:.2f} | {100 * tot / total:.1f}% | "
                   f"{vg:.0f} | {lds:.0f} | {grid:.0f} | {wg:.0f} |")
    return "\n".join(out)


def gap_table(db, top=12, min_count=50):
    c = sqlite3.connect(db)
    cols = [r[1] for r in c.execute("pragma table_info(kernels)").fetchall()]
    t0, t1 = ("start", "end") if "start" in cols else ("start_time", "end_time")
    rows = c.execute(f'select name, "{t0}", "{t1}" from kernels order by "{t0}"').fetchall()
    stats = {}
    prev = None
    for name, a, b in rows:
        if prev is not None:
            key = (prev[0], name)
            st = stats.setdefault(key, [0, 0.0, 0.0])
            st[0] += 1
            st[1] += b - a
            st[2] += a - prev[2]
        prev = (name, a, b)
    items = sorted(((k, v) for k, v in stats.items() if v[0] >= min_count), key=lambda kv: -kv[1][1])[:top]
    out = ["| after | kernel | dispatches | mean us | mean gap before, us |", "|---|---|---|---|---|"]
    for (pn, n), (cnt, dur, gap) in items:
        out.append(f"| `{pn[:40]}` | `{n[:40]}` | {cnt} | {dur / cnt / 1e3:.2f} | {gap / cnt / 1e3:.2f} |")
    return "\n".join(out)


def counter_table(db, top=12, by_grid=False):
    c = sqlite3.connect(db)
    if by_grid:
        rows = c.execute(
            "select kernel_name, grid_size, counter_name, count(*), avg(value), min(value), max(value) "
            "from counters_collection group by kernel_name, grid_size, counter_name order by avg(value) desc "
            "limit ?", (top,)).fetchall()
        out = ["| kernel | grid | counter | dispatches | mean | min | max |", "|---|---|---|---|---|---|---|"]
        for name, grid, cn, n, avg, mn, mx in rows:
            out.append(f"| `{name[:80]}` | {grid} | {cn} | {n} | {avg:.1f} | {mn:.1f} | {mx:.1f} |")
        return "\n".join(out)
    rows = c.execute(
        "select kernel_name, counter_name, count(*), avg(value), min(value), max(value) from counters_collection "
        "group by kernel_name, counter_name order by avg(value) desc limit ?", (top,)).fetchall()
    out = ["| kernel | counter | dispatches | mean | min | max |", "|---|---|---|---|---|---|"]
    for name, cn, n, avg, mn, mx in rows:
        out.append(f"| `{name[:80]}` | {cn} | {n} | {avg:.1f} | {mn:.1f} | {mx:.1f} |")
    return "\n".join(out)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("db")
    ap.add_argument("--counters", action="store_true")
    ap.add_argument("--gaps", action="store_true")
    ap.add_argument("--by-grid", action="store_true",
                    help="one row per kernel and grid size (launches of one kernel over different batch sizes)")
    ap.add_argument("--top", type=int, default=12)
    a = ap.parse_args()
    if a.counters:
        print(counter_table(a.db, a.top, a.by_grid))
    elif a.gaps:
        print(gap_table(a.db, a.top))
    else:
        print(kernel_table(a.db, a.top, a.by_grid))


if __name__ == "__main__":
    main()
