"""Summaries of rocprofv3 CSV output (kernel stats and PMC counters) and the
HBM-traffic figure the bench reports.

    python -m cudec_amd.profiling stats RUN_kernel_stats.csv --command "..."
    python -m cudec_amd.profiling traffic --fetch F.csv --write W.csv \\
        --kernel tuned_batch --copy neg_kernel --copy-bytes N -o traffic.json

Traffic per launch of a kernel = FETCH_SIZE and WRITE_SIZE (KiB per
dispatch, averaged over the kernel's dispatches), each scaled by what the
same counter reads for a plain elementwise copy of known size (bench.py
--copy-probe-mib: torch.neg over int32, N bytes read and N bytes written):
    scale_fetch = N / (1024 * FETCH_KiB(copy)), scale_write likewise.
"""
from __future__ import annotations

import argparse
import csv
import json
import sys
from collections import defaultdict


def short_name(name: str, limit: int = 60) -> str:
    """`void ns::(anonymous namespace)::tuned_batch<0>(args...)` ->
    `tuned_batch<0>`; long template arguments are cut to `limit` chars."""
    s = name.replace("(anonymous namespace)::", "")
    if s.startswith("void "):
        s = s[5:]
    depth = 0
    cut = len(s)
    last_sep = 0
    i = 0
    while i < len(s):
        ch = s[i]
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            cut = i
            break
        elif ch == ":" and depth == 0 and s[i:i + 2] == "::":
            last_sep = i + 2
            i += 1
        i += 1
    s = s[last_sep:cut]
    return s if len(s) <= limit else s[:limit - 3] + "..."


def read_kernel_stats(path: str):
    with open(path, newline="") as f:
        rows = list(csv.DictReader(f))
    return [{"kernel": short_name(r["Name"]), "calls": int(r["Calls"]),
             "avg_ns": float(r["AverageNs"]), "min_ns": float(r["MinNs"]),
             "max_ns": float(r["MaxNs"]), "pct": float(r["Percentage"])} for r in rows]


def stats_markdown(path: str, command: str, title: str) -> str:
    rows = read_kernel_stats(path)
    out = [f"# {title}", "", "Command (repo root, MI355X):", "", f"    {command}", "",
           "| kernel | calls | average ns | min ns | max ns | % of GPU time |",
           "|---|---|---|---|---|---|"]
    for r in rows:
        out.append(f"| `{r['kernel']}` | {r['calls']} | {r['avg_ns']:.0f} | {r['min_ns']:.0f} | "
                   f"{r['max_ns']:.0f} | {r['pct']:.2f} |")
    return "\n".join(out) + "\n"


def read_counters(path: str):
    """{(full kernel name, counter): [value per dispatch]}; kernels are
    matched by a substring of their full name (template arguments
    included), so a functor name inside them can select a kernel."""
    acc = defaultdict(list)
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            acc[(r["Kernel_Name"], r["Counter_Name"])].append(float(r["Counter_Value"]))
    return acc


def _mean_for(acc, kernel: str, counter: str) -> tuple[float, int]:
    vals = [v for (k, c), vs in acc.items() if kernel in k and c == counter for v in vs]
    if not vals:
        raise SystemExit(f"no {counter} dispatches of a kernel matching {kernel!r}")
    return sum(vals) / len(vals), len(vals)


def traffic(fetch_csv: str, write_csv: str, kernel: str, copy: str, copy_bytes: int) -> dict:
    fa, wa = read_counters(fetch_csv), read_counters(write_csv)
    kf, nkf = _mean_for(fa, kernel, "FETCH_SIZE")
    kw, nkw = _mean_for(wa, kernel, "WRITE_SIZE")
    cf_, ncf = _mean_for(fa, copy, "FETCH_SIZE")
    cw, ncw = _mean_for(wa, copy, "WRITE_SIZE")
    sf = copy_bytes / (cf_ * 1024.0)
    sw = copy_bytes / (cw * 1024.0)
    fetch = kf * 1024.0 * sf
    write = kw * 1024.0 * sw
    return {"kernel": kernel, "bytes_per_launch": round(fetch + write),
            "fetch_bytes": round(fetch), "write_bytes": round(write),
            "fetch_kib_raw": kf, "write_kib_raw": kw, "dispatches": [nkf, nkw],
            "copy_kernel": copy, "copy_bytes": copy_bytes, "copy_fetch_kib": cf_,
            "copy_write_kib": cw, "copy_dispatches": [ncf, ncw],
            "scale_fetch": sf, "scale_write": sw}


def counters_markdown(path: str, kernel: str, command: str, title: str,
                      per_unit: int = 0, unit_name: str = "") -> str:
    """Average per dispatch of every counter collected for kernels whose
    full name contains `kernel`, and, with per_unit, the count per unit."""
    acc = read_counters(path)
    names = sorted({c for (k, c) in acc if kernel in k})
    rows = []
    for c in names:
        mean, n = _mean_for(acc, kernel, c)
        extra = f" | {mean / per_unit:.2f}" if per_unit else ""
        rows.append(f"| {c} | {n} | {mean:,.0f}{extra} |")
    head = "| counter | dispatches | mean per dispatch |" + (f" per {unit_name} |" if per_unit else "")
    sep = "|---|---|---|" + ("---|" if per_unit else "")
    return "\n".join([f"# {title}", "", "Command (repo root, MI355X):", "", f"    {command}", "",
                      head, sep] + rows) + "\n"


def traffic_markdown(t: dict, commands: list[str], source: str, algo_read: int,
                     algo_write: int, title: str) -> str:
    """The committed summary of a traffic measurement. It records the
    sha256 of the kernel source measured, so a reader (bench.py) can tell
    whether the figure still describes the current kernel."""
    import hashlib
    with open(source, "rb") as f:
        digest = hashlib.sha256(f.read()).hexdigest()
    cmd = "\n".join(f"    {c}" for c in commands)
    return f"""# {title}

Commands (repo root, MI355X), one counter per pass:

{cmd}

Calibration: the copy probe (`torch.neg` over 1 GiB of int32, so
{t['copy_bytes']:,} B read and written per dispatch, {t['copy_dispatches'][0]} dispatches) reads
FETCH_SIZE = {t['copy_fetch_kib']:.0f} KiB and WRITE_SIZE = {t['copy_write_kib']:.0f} KiB per dispatch. The scale
factors are therefore {t['scale_fetch']:.4f} (fetch) and {t['scale_write']:.4f} (write).

| per launch of `{t['kernel']}` ({t['dispatches'][0]} dispatches) | raw KiB | scaled bytes | algorithmic bytes |
|---|---|---|---|
| FETCH_SIZE | {t['fetch_kib_raw']:.0f} | {t['fetch_bytes']:,} | {algo_read:,} (compressed) |
| WRITE_SIZE | {t['write_kib_raw']:.0f} | {t['write_bytes']:,} | {algo_write:,} (decoded) |
| total | | {t['bytes_per_launch']:,} | {algo_read + algo_write:,} |

Machine-readable lines (bench.py reports the figure as roofline.traffic
only while the kernel source still has this hash):

    bytes per launch: {t['bytes_per_launch']}
    tuned.hip sha256: {digest}
"""


def main(argv=None) -> int:
    p = argparse.ArgumentParser(prog="python -m cudec_amd.profiling")
    sub = p.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    s.add_argument("--command", required=True)
    s.add_argument("--title", default="rocprofv3 kernel stats")
    s.add_argument("-o", "--out", default=None)
    c = sub.add_parser("counters")
    c.add_argument("csv")
    c.add_argument("--kernel", required=True)
    c.add_argument("--command", required=True)
    c.add_argument("--title", default="rocprofv3 counters")
    c.add_argument("--per-unit", type=int, default=0)
    c.add_argument("--unit-name", default="")
    c.add_argument("-o", "--out", default=None)
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--kernel", required=True)
    t.add_argument("--copy", required=True)
    t.add_argument("--copy-bytes", type=int, required=True)
    t.add_argument("-o", "--out", required=True)
    t.add_argument("--md", default=None, help="also write the committed markdown summary here")
    t.add_argument("--source", default="cudec_amd/csrc/tuned.hip")
    t.add_argument("--algo-read", type=int, default=0)
    t.add_argument("--algo-write", type=int, default=0)
    t.add_argument("--title", default="HBM traffic (rocprofv3 PMC)")
    t.add_argument("--command", action="append", default=[])
    a = p.parse_args(argv)
    if a.cmd == "counters":
        md = counters_markdown(a.csv, a.kernel, a.command, a.title, a.per_unit, a.unit_name)
        if a.out:
            with open(a.out, "w") as f:
                f.write(md)
        else:
            sys.stdout.write(md)
    elif a.cmd == "stats":
        md = stats_markdown(a.csv, a.command, a.title)
        if a.out:
            with open(a.out, "w") as f:
                f.write(md)
        else:
            sys.stdout.write(md)
    else:
        res = traffic(a.fetch, a.write, a.kernel, a.copy, a.copy_bytes)
        with open(a.out, "w") as f:
            json.dump(res, f, indent=1)
        if a.md:
            with open(a.md, "w") as f:
                f.write(traffic_markdown(res, a.command, a.source, a.algo_read, a.algo_write,
                                         a.title))
        print(json.dumps(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
