"""Per-kernel summary of hardware counters in a rocprofv3 database (rocpd, SQLite).

    python3 -m adscf_amd.rocpd_counters RUN_results.db [REGEX]

prints, for every kernel whose name matches REGEX, the number of dispatches and the
median and maximum of each collected counter over its dispatches (counters summed over hardware
instances per dispatch, as rocprofv3 reports *_sum counters).  Used to turn
`rocprofv3 --pmc FETCH_SIZE ...` and `--pmc WRITE_SIZE ...` runs of bench.py into the
traffic figures of the line passes.
"""
from __future__ import annotations

import collections
import re
import sqlite3
import statistics
import sys


def summarize(db: str, pattern: str = ".") -> dict:
    con = sqlite3.connect(f"file:{db}?mode=ro", uri=True)
    rows = con.execute("select name, dispatch_id, counter_name, counter_value from pmc_events").fetchall()
    per = collections.defaultdict(lambda: collections.defaultdict(lambda: collections.defaultdict(float)))
    rx = re.compile(pattern)
    for name, did, cname, val in rows:
        if rx.search(name):
            per[name][cname][did] += val
    out = {}
    for name, counters in per.items():
        out[name] = {c: (len(v), statistics.median(v.values()), max(v.values())) for c, v in counters.items()}
    return out


def main(argv):
    if len(argv) < 2:
        print(__doc__)
        return 2
    res = summarize(argv[1], argv[2] if len(argv) > 2 else ".")
    for name in sorted(res):
        for cname, (n, med, mx) in sorted(res[name].items()):
            print(f"{cname}\tdispatches={n}\tmedian={med:.1f}\tmax={mx:.1f}\t{name[:100]}")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
