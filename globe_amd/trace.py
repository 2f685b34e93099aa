"""GPU busy fraction of a rocprofv3 kernel trace, leg by leg.

``bench.py --mark`` brackets each serving replay and each AI-REML fit with a 16-byte device
copy (kernel ``k_copy16``). Given the directory a ``rocprofv3 --kernel-trace -f csv`` run
wrote, this reports for every pair of markers the span between them, the time at least
one kernel was running (overlapping kernels merged), the idle time split into short
(< 20 us, launch-to-launch) and long gaps, and the kernels that took the most time.

    python -m globe_amd.trace <trace dir> [--top 8]
"""
import argparse
import csv
import glob
import os
from collections import defaultdict

MARKER = "k_copy16"
MARKER_MAX_NS = 200_000  # a marker copies 16 bytes; a longer k_copy16 is real work


def short_name(name):
    k = name.replace("void ", "").replace("(anonymous namespace)::", "").replace("globe::", "")
    return k.split("<")[0].split("(")[0]


def load(trace_dir):
    rows = []
    for f in glob.glob(os.path.join(trace_dir, "**", "*kernel_trace.csv"), recursive=True):
        with open(f) as fh:
            for r in csv.DictReader(fh):
                rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"]))
    rows.sort()
    return rows


def segments(rows, short_gap_ns=20_000):
    marks = [i for i, r in enumerate(rows) if MARKER in r[2] and r[1] - r[0] < MARKER_MAX_NS]
    out = []
    for p in range(0, len(marks) - 1, 2):
        a, b = marks[p], marks[p + 1]
        seg = rows[a + 1:b]
        t0, t1 = rows[a][1], rows[b][0]
        busy = 0
        short = 0
        cs = ce = None
        for s, e, _ in seg:
            if ce is None or s > ce:
                if ce is not None:
                    busy += ce - cs
                    if s - ce < short_gap_ns:
                        short += s - ce
                cs, ce = s, e
            else:
                ce = max(ce, e)
        if ce is not None:
            busy += ce - cs
        per = defaultdict(lambda: [0, 0])
        for s, e, n in seg:
            k = short_name(n)
            per[k][0] += 1
            per[k][1] += e - s
        span = t1 - t0
        out.append({"kernels": len(seg), "span_s": span / 1e9, "busy_s": busy / 1e9,
                    "busy_frac": busy / span if span else 0.0, "short_gaps_s": short / 1e9,
                    "long_gaps_s": (span - busy - short) / 1e9,
                    "by_kernel": sorted(((k, c, t / 1e9) for k, (c, t) in per.items()),
                                        key=lambda x: -x[2])})
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("trace_dir")
    ap.add_argument("--top", type=int, default=8)
    a = ap.parse_args(argv)
    segs = segments(load(a.trace_dir))
    print(f"{len(segs)} marked legs")
    for i, s in enumerate(segs):
        print(f"leg {i}: {s['kernels']} kernels, span {s['span_s']:.3f} s, busy {s['busy_s']:.3f} s "
              f"= {100 * s['busy_frac']:.1f}%, gaps < 20 us {1e3 * s['short_gaps_s']:.1f} ms, "
              f"longer gaps {1e3 * s['long_gaps_s']:.1f} ms")
        for k, c, t in s["by_kernel"][:a.top]:
            print(f"    {1e3 * t:10.2f} ms {c:7d}  {k}")


if __name__ == "__main__":
    main()
