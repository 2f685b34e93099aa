"""Reports from rocprofv3 csv output.

busy: ``bench.py --mark`` brackets each serving replay and each AI-REML fit with a 16-byte
device copy (kernel ``k_copy16``). Given the directory a ``rocprofv3 --kernel-trace -f csv``
run wrote, this reports for every pair of markers the span between them, the time at
least one kernel was running (overlapping kernels merged), the idle time split into short
(< 20 us, launch-to-launch) and long gaps, and the kernels that took the most time.

    python -m globe_amd.trace <trace dir> [--top 8] [--gaps 5]

With --gaps, the long gaps of each leg are also grouped by the kernels that end before
and start after them (a device-to-host copy before a gap is a host round trip).

counters: per-kernel tables of the contraction kernels from ``rocprofv3 --pmc`` runs of
``bench.py --calib`` (one directory per pass): clock, MFMA busy share, issued int8 ops,
LDS instructions, bank-conflict and LDS-busy shares, waits on LDS issue, and HBM bytes
from FETCH_SIZE / WRITE_SIZE scaled by what they read for the calibration copies (three
plain 1 GiB copies, ``k_copy16``).

    python -m globe_amd.trace counters --fetch D1 --write D2 --mfma D3 [--json OUT]

kernels: launches and mean duration per kernel of a kernel trace, over the launches at
least --min-us long (the contraction launches of one problem size among the tuner's and
other legs' launches).

    python -m globe_amd.trace kernels <trace dir> --prefix k_geno_wide --min-us 4000
"""
import argparse
import csv
import glob
import json
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
        prev = None
        pairs = defaultdict(lambda: [0, 0])
        for s, e, n in seg:
            if ce is None or s > ce:
                if ce is not None:
                    busy += ce - cs
                    if s - ce < short_gap_ns:
                        short += s - ce
                    else:
                        g = pairs[(short_name(prev), short_name(n))]
                        g[0] += 1
                        g[1] += s - ce
                cs, ce = s, e
            else:
                ce = max(ce, e)
            prev = n
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
                                        key=lambda x: -x[2]),
                    "long_gaps_by_pair": sorted(((a, b, c, t / 1e9) for (a, b), (c, t) in pairs.items()),
                                                key=lambda x: -x[3])})
    return out


XCDS = 8            # GRBM_GUI_ACTIVE is summed over the XCDs
CUS = 256
SIMDS = 4 * CUS
CALIB_BYTES = 1 << 30


def load_counters(pmc_dir):
    """{dispatch: (kernel name, duration ns, {counter: value})} of a --pmc run."""
    out = {}
    for f in glob.glob(os.path.join(pmc_dir, "**", "*counter_collection.csv"), recursive=True):
        with open(f) as fh:
            for r in csv.DictReader(fh):
                d = int(r["Dispatch_Id"])
                if d not in out:
                    out[d] = (r["Kernel_Name"], int(r["End_Timestamp"]) - int(r["Start_Timestamp"]), {})
                out[d][2][r["Counter_Name"]] = float(r["Counter_Value"])
    return out


def template_name(name):
    k = name.replace("void ", "").replace("(anonymous namespace)::", "").replace("globe::", "")
    return k.split("(")[0]


def calib_scale(disp, counter):
    """Bytes per counter unit (KiB) from the 1 GiB calibration copies."""
    vals = [c[counter] for n, _, c in disp.values() if "k_copy16" in n and counter in c and c[counter] > 2e5]
    if not vals:
        return None
    return CALIB_BYTES / (sum(vals) / len(vals) * 1024.0)


def per_kernel(disp, prefix):
    agg = defaultdict(lambda: defaultdict(list))
    for n, dur, c in disp.values():
        k = template_name(n)
        if not k.startswith(prefix):
            continue
        agg[k]["dur"].append(dur)
        for cn, v in c.items():
            agg[k][cn].append(v)
    return agg


def mean(x):
    return sum(x) / len(x) if x else float("nan")


def counters_report(a):
    fetch, write, mf = load_counters(a.fetch), load_counters(a.write), load_counters(a.mfma)
    fs, ws = calib_scale(fetch, "FETCH_SIZE"), calib_scale(write, "WRITE_SIZE")
    print(f"calibration: FETCH_SIZE x{fs:.4f}, WRITE_SIZE x{ws:.4f} (bytes per KiB unit / 1024)"
          if fs and ws else "calibration copies not found")
    report = {"fetch_scale": fs, "write_scale": ws}
    for prefix in ("k_geno_wide", "k_geno_thin"):
        fk, wk, mk = per_kernel(fetch, prefix), per_kernel(write, prefix), per_kernel(mf, prefix)
        names = sorted(set(fk) | set(mk), key=lambda k: -len(mk.get(k, fk.get(k, {})).get("dur", [])))
        print(f"\n{prefix}: launches, ms (counter run), clock GHz, MFMA busy, issued int8 TOP/s, "
              "LDS instr / launch, bank-conflict cycles / CU-cycle, LDS-busy / CU-cycle, "
              "LDS-issue wait / wave-cycle, fetch GB, write GB")
        rows = {}
        for k in names:
            m = mk.get(k, {})
            row = {"launches": len(m.get("dur", fk.get(k, {}).get("dur", [])))}
            if m:
                dur = mean(m["dur"]) * 1e-9
                gui = mean(m.get("GRBM_GUI_ACTIVE", [])) / XCDS
                row.update(ms=dur * 1e3, clock=gui / dur / 1e9)
                if "SQ_VALU_MFMA_BUSY_CYCLES" in m:
                    row["mfma_busy"] = mean(m["SQ_VALU_MFMA_BUSY_CYCLES"]) / (gui * SIMDS)
                if "SQ_INSTS_VALU_MFMA_MOPS_I8" in m:
                    row["int8_tops"] = mean(m["SQ_INSTS_VALU_MFMA_MOPS_I8"]) * 512 / dur / 1e12
                if "SQ_INSTS_LDS" in m:
                    row["lds_insts"] = mean(m["SQ_INSTS_LDS"])
                if "SQ_LDS_BANK_CONFLICT" in m:
                    row["bank_conflict"] = mean(m["SQ_LDS_BANK_CONFLICT"]) / (gui * CUS)
                if "SQ_LDS_IDX_ACTIVE" in m:
                    row["lds_busy"] = mean(m["SQ_LDS_IDX_ACTIVE"]) / (gui * CUS)
                if "SQ_WAIT_INST_LDS" in m and "SQ_WAVE_CYCLES" in m:
                    row["lds_wait"] = mean(m["SQ_WAIT_INST_LDS"]) / mean(m["SQ_WAVE_CYCLES"])
            if fs and k in fk and "FETCH_SIZE" in fk[k]:
                row["fetch"] = mean(fk[k]["FETCH_SIZE"]) * 1024 * fs
            if ws and k in wk and "WRITE_SIZE" in wk[k]:
                row["write"] = mean(wk[k]["WRITE_SIZE"]) * 1024 * ws
            rows[k] = row

            def f(key, fmt):
                return format(row[key], fmt) if key in row else "-"
            print(f"  {k:34s} {row['launches']:4d} {f('ms', '8.3f')} {f('clock', '5.2f')} {f('mfma_busy', '6.1%')} "
                  f"{f('int8_tops', '7.0f')} {f('lds_insts', '10.3e')} {f('bank_conflict', '6.3f')} "
                  f"{f('lds_busy', '6.3f')} {f('lds_wait', '6.3f')} "
                  f"{format(row['fetch'] / 1e9, '7.2f') if 'fetch' in row else '-'} "
                  f"{format(row['write'] / 1e9, '6.2f') if 'write' in row else '-'}")
        # tuned variants: every candidate is launched the same number of times while the
        # tuner times them, so within one digit width (wide: the first template argument;
        # thin: one width) the variants launched more often than the least launched one
        # are those the tuner kept (one per pass and shape)
        groups = defaultdict(list)
        for k in rows:
            groups[k.split("<")[1].split(",")[0] if prefix == "k_geno_wide" and "<" in k else ""].append(k)
        tuned = []
        for w in sorted(groups):
            low = min(rows[k]["launches"] for k in groups[w])
            kept = [k for k in groups[w] if rows[k]["launches"] > low]
            tuned += kept if kept else groups[w]
        n = sum(rows[k]["launches"] for k in tuned if "fetch" in rows[k] and "write" in rows[k])
        if n:
            fb = sum(rows[k]["fetch"] * rows[k]["launches"] for k in tuned if "fetch" in rows[k]) / n
            wb = sum(rows[k]["write"] * rows[k]["launches"] for k in tuned if "write" in rows[k]) / n
            print(f"  tuned ({', '.join(tuned)}), launch-weighted: fetch {fb / 1e9:.2f} GB, write {wb / 1e9:.2f} GB")
            role = "wide" if prefix == "k_geno_wide" else "thin"
            report[role] = {"kernel": f"{prefix} (tuned variants: " + ", ".join(tuned) + ")", "fetch_bytes": fb,
                            "write_kernel": f"{prefix} (tuned variants)", "write_bytes": wb, "launches": n,
                            "by_name": rows}
    if a.json:
        with open(a.json, "w") as fh:
            json.dump(report, fh, indent=1)


def kernels_report(a):
    per = defaultdict(list)
    for s0, e0, n in load(a.trace_dir):
        k = template_name(n)
        if k.startswith(a.prefix) and (e0 - s0) >= a.min_us * 1e3:
            per[k].append(e0 - s0)
    tot = sum(len(v) for v in per.values())
    print(f"{a.prefix} launches >= {a.min_us:g} us: {tot}")
    for k, v in sorted(per.items(), key=lambda kv: -len(kv[1])):
        print(f"  {k:40s} {len(v):5d} launches, mean {mean(v) / 1e6:8.3f} ms, min {min(v) / 1e6:8.3f}, "
              f"max {max(v) / 1e6:8.3f}")
    if tot:
        print(f"  all: mean {sum(sum(v) for v in per.values()) / tot / 1e6:.3f} ms")


def main(argv=None):
    import sys
    argv = sys.argv[1:] if argv is None else argv
    if argv and argv[0] == "kernels":
        ap = argparse.ArgumentParser(prog="python -m globe_amd.trace kernels")
        ap.add_argument("trace_dir")
        ap.add_argument("--prefix", required=True)
        ap.add_argument("--min-us", type=float, default=0.0)
        kernels_report(ap.parse_args(argv[1:]))
        return
    if argv and argv[0] == "counters":
        ap = argparse.ArgumentParser(prog="python -m globe_amd.trace counters")
        ap.add_argument("--fetch", required=True)
        ap.add_argument("--write", required=True)
        ap.add_argument("--mfma", required=True)
        ap.add_argument("--json", help="write the per-launch traffic (bench.py reads roofline.traffic from it)")
        counters_report(ap.parse_args(argv[1:]))
        return
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("trace_dir")
    ap.add_argument("--top", type=int, default=8)
    ap.add_argument("--gaps", type=int, default=0, help="also list the N largest long-gap kernel pairs")
    a = ap.parse_args(argv)
    segs = segments(load(a.trace_dir))
    print(f"{len(segs)} marked legs")
    for i, s in enumerate(segs):
        print(f"leg {i}: {s['kernels']} kernels, span {s['span_s']:.3f} s, busy {s['busy_s']:.3f} s "
              f"= {100 * s['busy_frac']:.1f}%, gaps < 20 us {1e3 * s['short_gaps_s']:.1f} ms, "
              f"longer gaps {1e3 * s['long_gaps_s']:.1f} ms")
        for k, c, t in s["by_kernel"][:a.top]:
            print(f"    {1e3 * t:10.2f} ms {c:7d}  {k}")
        if a.gaps:
            print("  long gaps by the kernels either side:")
            for k0, k1, c, t in s["long_gaps_by_pair"][:a.gaps]:
                print(f"    {1e3 * t:10.2f} ms {c:7d}  {k0} -> {k1}")


if __name__ == "__main__":
    main()
