"""Summaries of rocprofv3 output for the profiles kept with the library.

    python -m uavs3e_amd.rocprof_summary stats DIR OUT.md --title T --cmd C
        kernel statistics of a `rocprofv3 --kernel-trace --stats` run
        (CSV output, or the rocpd database when no CSV was written)

    python -m uavs3e_amd.rocprof_summary traffic FETCH_DIR WRITE_DIR OUT.json
            --kernel NAME --copy-kernel NAME --copy-bytes N --cmd-fetch C --cmd-write C
        HBM bytes per launch of one kernel from a `--pmc FETCH_SIZE` run and a
        `--pmc WRITE_SIZE` run, each scaled by what the same counter reads for a
        plain copy kernel that moves N bytes each way in the same run
"""
import argparse
import csv
import glob
import json
import os
import sqlite3
import sys


def _files(d, suffix):
    return sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True))


def kernel_stats(d):
    """[(name, calls, total_us, avg_us, pct)] sorted by total time."""
    rows = []
    for f in _files(d, "kernel_stats.csv"):
        with open(f) as fh:
            for r in csv.DictReader(fh):
                rows.append((r["Name"], int(r["Calls"]), float(r["TotalDurationNs"]) / 1e3,
                             float(r["AverageNs"]) / 1e3, float(r["Percentage"])))
    if not rows:
        for f in _files(d, "results.db"):
            c = sqlite3.connect(f)
            for n, calls, tot, avg, pct in c.execute(
                    "select name, total_calls, total_duration, average, percentage from top_kernels"):
                # the rocpd view already reports microseconds
                rows.append((n, int(calls), float(tot), float(avg), float(pct)))
    rows.sort(key=lambda r: -r[2])
    return rows


def counter_per_dispatch(d, counter):
    """{kernel name: [value per dispatch]} for one counter (values summed over
    the instances the CSV lists per dispatch)."""
    per = {}
    for f in _files(d, "counter_collection.csv"):
        with open(f) as fh:
            for r in csv.DictReader(fh):
                if r["Counter_Name"] != counter:
                    continue
                key = (r["Kernel_Name"], r.get("Dispatch_Id") or r.get("Correlation_Id"))
                per[key] = per.get(key, 0.0) + float(r["Counter_Value"])
    out = {}
    for (name, _), v in per.items():
        out.setdefault(name, []).append(v)
    return out


def _match(table, sub):
    hits = [n for n in table if sub in n]
    if not hits:
        raise SystemExit("no kernel matching %r in %s" % (sub, sorted(table)[:20]))
    return max(hits, key=lambda n: len(table[n]))


def cmd_stats(a):
    rows = kernel_stats(a.dir)
    if not rows:
        raise SystemExit("no kernel statistics under " + a.dir)
    with open(a.out, "w") as f:
        f.write("# %s\n\nCommand (run from the repository root on one MI355X):\n\n```\n%s\n```\n\n" % (a.title, a.cmd))
        f.write("rocprofv3 kernel statistics (durations in microseconds):\n\n"
                "| kernel | calls | total us | average us | % |\n|---|---|---|---|---|\n")
        for n, calls, tot, avg, pct in rows[:a.top]:
            f.write("| `%s` | %d | %.1f | %.2f | %.2f |\n" % (n, calls, tot, avg, pct))
    print(open(a.out).read())


def cmd_traffic(a):
    res = {"kernel": a.kernel, "copy_kernel": a.copy_kernel, "copy_bytes_each_way": a.copy_bytes,
           "commands": {"FETCH_SIZE": a.cmd_fetch, "WRITE_SIZE": a.cmd_write}}
    total = 0.0
    for counter, d in (("FETCH_SIZE", a.fetch_dir), ("WRITE_SIZE", a.write_dir)):
        t = counter_per_dispatch(d, counter)
        k = _match(t, a.kernel)
        c = _match(t, a.copy_kernel)
        kv = sum(t[k]) / len(t[k])
        cv = max(t[c])
        factor = a.copy_bytes / (cv * 1024.0)   # counters read KiB
        res[counter] = {"kernel_kib_per_dispatch": kv, "dispatches": len(t[k]),
                        "copy_kib_per_dispatch": cv, "scale": factor, "bytes_per_launch": kv * 1024.0 * factor}
        total += kv * 1024.0 * factor
    res["bytes_per_launch"] = total
    with open(a.out, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res, indent=1))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="what", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    s.add_argument("out")
    s.add_argument("--title", required=True)
    s.add_argument("--cmd", required=True)
    s.add_argument("--top", type=int, default=20)
    t = sub.add_parser("traffic")
    t.add_argument("fetch_dir")
    t.add_argument("write_dir")
    t.add_argument("out")
    t.add_argument("--kernel", required=True)
    t.add_argument("--copy-kernel", required=True)
    t.add_argument("--copy-bytes", type=float, required=True)
    t.add_argument("--cmd-fetch", required=True)
    t.add_argument("--cmd-write", required=True)
    a = ap.parse_args(argv)
    (cmd_stats if a.what == "stats" else cmd_traffic)(a)


if __name__ == "__main__":
    sys.exit(main())
