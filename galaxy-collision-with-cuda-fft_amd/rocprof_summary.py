"""Summarize rocprofv3 result databases (SQLite, ROCm 7) into short text tables.

    python rocprof_summary.py trace  <bench_results.db>                 > summary.txt
    python rocprof_summary.py pmc    <fetch.db> <write.db> [--json out] > traffic.txt

``trace``: per kernel and grid size, dispatch count and average/min/max duration. With
``--kernel NAME --last N`` it adds, for the last N dispatches of that kernel (the timed
steps), the mean duration, the mean gap since the previous dispatch ended, and their sum: the
span of a HIP event pair that the kernel's own launch records (hipExtLaunchKernelGGL), whose
start marker completes when the previous kernel does.
``pmc``: per kernel and grid size, average FETCH_SIZE and WRITE_SIZE (KiB per dispatch),
scaled by the same counters of a plain device copy of known size found in the same run
(copy_f32_kernel moving a 256 MiB buffer, bench.py --copy-probe): factor = expected bytes /
counted bytes for that copy. The scaled figures estimate HBM bytes per dispatch.
"""

from __future__ import annotations

import argparse
import json
import re
import sqlite3
import sys
from collections import defaultdict

COPY_BYTES = 256 * 1024 * 1024  # bench.py --copy-probe buffer


def short(name: str) -> str:
    """Kernel name without namespaces, template arguments and parameter lists."""
    n = name.replace("(anonymous namespace)::", "")
    n = re.sub(r"\(.*\)$", "", n)
    n = re.sub(r"^void ", "", n)
    return n.strip()


def _timed_dispatches(db, kernel: str, last: int) -> str:
    rows = db.execute("select name, start, end from kernels order by start").fetchall()
    durs, gaps = [], []
    for (_, _, e0), (n1, s1, e1) in zip(rows, rows[1:]):
        if short(n1) == kernel:
            durs.append(e1 - s1)
            gaps.append(max(s1 - e0, 0))
    durs, gaps = durs[-last:], gaps[-last:]
    if not durs:
        return f"# {kernel}: no dispatches"
    d = sum(durs) / len(durs) / 1e3
    g = sum(gaps) / len(gaps) / 1e3
    # device occupancy from the first of those dispatches to the end of the trace: the union
    # of kernel intervals against the span (the rest is time with no kernel running)
    starts = [s1 for (n1, s1, _) in rows if short(n1) == kernel][-last:]
    t0, t1 = starts[0], rows[-1][2]
    busy, cur_s, cur_e, kernels = 0, None, None, 0
    for _, s1, e1 in rows:
        if e1 <= t0:
            continue
        s1 = max(s1, t0)
        kernels += 1
        if cur_e is None or s1 > cur_e:
            if cur_e is not None:
                busy += cur_e - cur_s
            cur_s, cur_e = s1, e1
        else:
            cur_e = max(cur_e, e1)
    if cur_e is not None:
        busy += cur_e - cur_s
    span = t1 - t0
    return (f"# {kernel}, last {len(durs)} dispatches: duration {d:.2f} us, gap since the "
            f"previous kernel {g:.2f} us, duration + gap {d + g:.2f} us\n"
            f"# from the first of them to the end: {span / 1e3:.1f} us, {kernels} kernels, "
            f"some kernel running {busy / 1e3:.1f} us ({100.0 * busy / span:.1f}%), none "
            f"{(span - busy) / 1e3:.1f} us")


def trace(db_path: str, kernel: str | None = None, last: int = 0) -> str:
    db = sqlite3.connect(db_path)
    rows = db.execute("select name, grid_x, grid_y, grid_z, duration from kernels").fetchall()
    agg = defaultdict(list)
    for name, gx, gy, gz, dur in rows:
        agg[(short(name), gx * gy * gz)].append(dur)
    total = sum(sum(v) for v in agg.values())
    out = [f"{'kernel':44s} {'grid':>10s} {'calls':>6s} {'avg_us':>9s} {'min_us':>9s} "
           f"{'max_us':>9s} {'total_ms':>9s} {'%':>6s}"]
    for (name, grid), d in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
        out.append(f"{name[:44]:44s} {grid:10d} {len(d):6d} {sum(d) / len(d) / 1e3:9.2f} "
                   f"{min(d) / 1e3:9.2f} {max(d) / 1e3:9.2f} {sum(d) / 1e6:9.3f} "
                   f"{100.0 * sum(d) / total:6.2f}")
    if kernel:
        out.insert(0, _timed_dispatches(db, kernel, last or 1 << 30))
    return "\n".join(out)


def _counter(db_path: str, counter: str):
    db = sqlite3.connect(db_path)
    rows = db.execute(
        "select kernel_name, grid_size, value from counters_collection where counter_name = ?",
        (counter,)).fetchall()
    agg = defaultdict(list)
    for name, grid, val in rows:
        agg[(short(name), grid)].append(float(val))
    return agg


def _copy_factor(agg) -> tuple[float, str]:
    """Expected/counted bytes for the probe copy (the copy kernel moving COPY_BYTES)."""
    cands = [(k, v) for k, v in agg.items()
             if k[0] == "copy_f32_kernel" and k[1] * 4 == COPY_BYTES]
    if not cands:
        return 1.0, "no copy probe found: unscaled"
    (name, grid), vals = cands[0]
    kib = sum(vals) / len(vals)
    return (COPY_BYTES / 1024.0) / kib if kib > 0 else 1.0, f"{name} grid {grid}: {kib:.0f} KiB"


def pmc(fetch_db: str, write_db: str, json_out: str | None) -> str:
    fetch = _counter(fetch_db, "FETCH_SIZE")
    write = _counter(write_db, "WRITE_SIZE")
    f_fac, f_src = _copy_factor(fetch)
    w_fac, w_src = _copy_factor(write)
    out = [f"FETCH_SIZE scale {f_fac:.4f} from {f_src}; WRITE_SIZE scale {w_fac:.4f} from {w_src}",
           f"{'kernel':44s} {'grid':>10s} {'fetch_KiB':>11s} {'write_KiB':>11s} "
           f"{'scaled_MB':>10s}"]
    table = {}
    for key in sorted(set(fetch) | set(write), key=lambda k: -(sum(fetch.get(k, [0])) / max(len(fetch.get(k, [1])), 1))):
        fv = fetch.get(key, [])
        wv = write.get(key, [])
        fk = sum(fv) / len(fv) if fv else float("nan")
        wk = sum(wv) / len(wv) if wv else float("nan")
        scaled = ((fk * f_fac if fv else 0.0) + (wk * w_fac if wv else 0.0)) * 1024.0
        table[f"{key[0]}@{key[1]}"] = scaled
        out.append(f"{key[0][:44]:44s} {key[1]:10d} {fk:11.0f} {wk:11.0f} {scaled / 1e6:10.2f}")
    if json_out:
        with open(json_out, "w") as fh:
            json.dump({"fetch_scale": f_fac, "write_scale": w_fac,
                       "hbm_bytes_per_dispatch": table}, fh, indent=0, sort_keys=True)
    return "\n".join(out)


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("db")
    t.add_argument("--kernel", default=None)
    t.add_argument("--last", type=int, default=0)
    p = sub.add_parser("pmc")
    p.add_argument("fetch_db")
    p.add_argument("write_db")
    p.add_argument("--json", default=None)
    a = ap.parse_args()
    print(trace(a.db, a.kernel, a.last) if a.cmd == "trace"
          else pmc(a.fetch_db, a.write_db, a.json))
    return 0


if __name__ == "__main__":
    sys.exit(main())
