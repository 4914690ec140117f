"""Summaries of rocprofv3 result databases (kernel trace and PMC passes).

    python -m vca_amd.rocprof_summary trace  TRACE.db [ROWS]
    python -m vca_amd.rocprof_summary pmc    FETCH.db WRITE.db KERNELS COPY_BYTES OUT.json KEY VARIANT
    python -m vca_amd.rocprof_summary counters PMC.db [KERNEL_SUBSTRING]

`trace` prints the per-kernel table (calls, average and total duration),
the ROWS (default 14) kernels of largest total time.
`counters` prints, per kernel (those whose name holds KERNEL_SUBSTRING),
the dispatch count and the average per dispatch of every counter of a PMC
pass.
`pmc` reads a FETCH_SIZE and a WRITE_SIZE pass (KiB per dispatch), finds
the calibration copy kernel of COPY_BYTES bytes each way (the largest
__amd_rocclr_copyBuffer / elementwise copy dispatch) to turn the counters
into bytes, and stores under KEY in OUT.json the HBM traffic of one
application: the sum over the KERNELS entries (separated by ';', each a
substring of a kernel name with an optional '*N', N launches of it per
application, default 1) of N times the per-dispatch traffic of the
most-dispatched kernel matching it, with the launch shape VARIANT the run
picked, so that a later run can tell whether the counters are its own.
"""
from __future__ import annotations

import json
import re
import sqlite3
import sys


def _short(name: str) -> str:
    name = re.sub(r"\(vca::\w+\)$", "", name)
    return name if len(name) < 80 else name[:77] + "..."


def trace_table(db: str, top: int = 14) -> str:
    c = sqlite3.connect(db)
    rows = c.execute("select name, total_calls, average, total_duration, percentage "
                     "from top_kernels order by total_duration desc limit ?", (top,)).fetchall()
    out = ["| kernel | calls | avg us | total us | % |", "|---|---:|---:|---:|---:|"]
    for r in rows:
        out.append(f"| `{_short(r[0])}` | {r[1]} | {r[2]:.2f} | {r[3]:.1f} | {r[4]:.1f} |")
    return "\n".join(out)


def _per_kernel(db: str, counter: str) -> dict:
    c = sqlite3.connect(db)
    res = {}
    for name, n, avg, mx in c.execute(
            "select kernel_name, count(*), avg(value), max(value) from counters_collection "
            "where counter_name=? group by kernel_name", (counter,)):
        res[name] = (n, avg, mx)
    return res


def pmc_traffic(fetch_db: str, write_db: str, kernels: str, copy_bytes: int) -> dict:
    fetch = _per_kernel(fetch_db, "FETCH_SIZE")
    write = _per_kernel(write_db, "WRITE_SIZE")
    # calibration: the copy dispatch with the largest counts
    copies = [k for k in fetch if "copy" in k.lower()]
    if not copies:
        raise SystemExit("no copy kernel in the FETCH_SIZE pass")
    ck = max(copies, key=lambda k: fetch[k][2])
    cf_kib = fetch[ck][2]
    cw_kib = write.get(ck, (0, 0, 0))[2]
    f_scale = copy_bytes / (cf_kib * 1024.0)
    w_scale = copy_bytes / (cw_kib * 1024.0) if cw_kib else None
    names, f_kib, w_kib, disp, per_app = [], 0.0, 0.0, [], []
    for ent in kernels.split(";"):
        sub, _, n = ent.rpartition("*") if "*" in ent else (ent, "", "1")
        cand = [k for k in fetch if sub in k]
        if not cand:
            raise SystemExit(f"no kernel matching {sub!r}")
        kn = max(cand, key=lambda k: fetch[k][0])
        names.append(kn)
        disp.append(fetch[kn][0])
        per_app.append(int(n))
        f_kib += int(n) * fetch[kn][1]
        w_kib += int(n) * write[kn][1]
    return {
        "kernels": names,
        "launches_per_application": per_app,
        "dispatches": disp,
        "fetch_kib_per_dispatch": f_kib,
        "write_kib_per_dispatch": w_kib,
        "calibration_kernel": ck,
        "calibration_bytes_each_way": copy_bytes,
        "fetch_scale": f_scale,
        "write_scale": w_scale,
        "traffic_bytes": f_kib * 1024.0 * f_scale + w_kib * 1024.0 * (w_scale or 1.0),
    }


def counter_table(db: str, sub: str = "") -> str:
    c = sqlite3.connect(db)
    rows = c.execute("select kernel_name, counter_name, count(*), avg(value) "
                     "from counters_collection group by kernel_name, counter_name "
                     "order by kernel_name, counter_name").fetchall()
    out = ["| kernel | counter | dispatches | average per dispatch |", "|---|---|---:|---:|"]
    for name, cn, n, avg in rows:
        if sub in name:
            out.append(f"| `{_short(name)}` | {cn} | {n} | {avg:.6g} |")
    return "\n".join(out)


def main(argv):
    if len(argv) >= 2 and argv[0] == "trace":
        print(trace_table(argv[1], int(argv[2]) if len(argv) > 2 else 14))
    elif len(argv) >= 2 and argv[0] == "counters":
        print(counter_table(argv[1], argv[2] if len(argv) > 2 else ""))
    elif len(argv) >= 8 and argv[0] == "pmc":
        d = pmc_traffic(argv[1], argv[2], argv[3], int(argv[4]))
        d["variant"] = int(argv[7])
        try:
            with open(argv[5]) as f:
                store = json.load(f)
        except FileNotFoundError:
            store = {}
        store[argv[6]] = d
        with open(argv[5], "w") as f:
            json.dump(store, f, indent=1)
        print(json.dumps(d, indent=1))
    else:
        print(__doc__)
        return 2
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
