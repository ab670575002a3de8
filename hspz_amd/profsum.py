"""Summarises rocprofv3 output databases into the short text kept in
profiles/: per-kernel call counts and durations, and per-dispatch counter
averages for selected kernels.

Usage: python -m hspz_amd.profsum DB [DB ...] [--kernel SUBSTR ...]
"""
import argparse
import collections
import sqlite3


def kernel_stats(db):
    con = sqlite3.connect(db)
    rows = con.execute("select name, total_calls, total_duration, average, percentage "
                       "from top_kernels").fetchall()
    return rows


def counters(db, substr):
    con = sqlite3.connect(db)
    rows = con.execute("select dispatch_id, kernel_name, counter_name, value from counters_collection "
                       "where kernel_name like ?", ("%" + substr + "%",)).fetchall()
    agg, disp = collections.defaultdict(float), set()
    for d, k, c, v in rows:
        agg[c] += v
        disp.add(d)
    n = max(len(disp), 1)
    return len(disp), {c: v / n for c, v in agg.items()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("dbs", nargs="+")
    ap.add_argument("--kernel", action="append", default=[])
    a = ap.parse_args()
    for db in a.dbs:
        print("## %s" % db)
        try:
            ks = kernel_stats(db)
        except sqlite3.Error:
            ks = []
        if ks:
            print("%-60s %8s %14s %12s %7s" % ("kernel", "calls", "total_us", "avg_us", "pct"))
            for name, calls, tot, avg, pct in ks[:20]:
                print("%-60s %8d %14.1f %12.2f %7.2f" % (name.split("(")[0][:60], calls, tot,
                                                          avg, pct))
        for k in a.kernel:
            n, c = counters(db, k)
            if n:
                print("counters for %s (%d dispatches, per-dispatch mean): %s" % (
                    k, n, ", ".join("%s=%.1f" % kv for kv in sorted(c.items()))))
        print()


if __name__ == "__main__":
    main()
