"""Summarise a rocprofv3 run of bench.py over its timed windows only.

    python -m grus_framework_amd.rocprof_summary --regions R.json \\
        (--trace X_kernel_trace.csv | --counters X_counter_collection.csv) \\
        [--traffic-out traffic.json] [--title T]

bench.py --regions-out R.json records, for each timed loop, the host clock
(CLOCK_MONOTONIC and CLOCK_BOOTTIME, ns) at its begin and end and the units
the loop processed (traversals, PageRank iterations).  rocprofv3 stamps
every dispatch on the host clock; a dispatch belongs to a window when it
starts inside it.  The clock whose windows capture dispatches is picked
automatically.

--trace     per kernel: calls, total and average duration, time per unit.
--counters  per kernel: calls, counter sum per unit and average per
            dispatch (FETCH_SIZE / WRITE_SIZE are KiB).  A window named
            copy_1GiB (bench.py --copy-calibration: one 1 GiB device-to-
            device copy) calibrates the counter: factor = 2^30 bytes /
            (counter KiB * 1024), so bytes = KiB * 1024 * factor.  With
            --traffic-out the scaled bytes per unit of every other window
            are merged into that JSON file as <window>.fetch_bytes /
            .write_bytes, with the settings the window ran with (its
            "config" in R.json, when bench.py recorded one).
Output is markdown on stdout.
"""
import argparse
import csv
import json
import os
from collections import defaultdict

CLOCKS = ("monotonic", "boottime")


def load_regions(path):
    with open(path) as f:
        return json.load(f)


def load_rows(path):
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def pick_clock(regions, starts):
    """The clock under which the most dispatches fall into some window."""
    best, best_n = CLOCKS[0], -1
    for c in CLOCKS:
        n = 0
        for r in regions:
            b, e = r["begin"][c], r["end"][c]
            n += sum(1 for s in starts if b <= s <= e)
        if n > best_n:
            best, best_n = c, n
    return best


def in_window(r, clock, start):
    return r["begin"][clock] <= start <= r["end"][clock]


def summarise_trace(regions, rows, clock):
    out = []
    for r in regions:
        agg = defaultdict(lambda: [0, 0])
        for row in rows:
            s = int(row["Start_Timestamp"])
            if in_window(r, clock, s):
                a = agg[row["Kernel_Name"]]
                a[0] += 1
                a[1] += int(row["End_Timestamp"]) - s
        units = r["units"]
        total = sum(v[1] for v in agg.values())
        out.append("### %s  (%d units, %d dispatches, %.4f ms kernel time per unit)\n" %
                   (r["name"], units, sum(v[0] for v in agg.values()), total / units / 1e6))
        out.append("| kernel | calls | calls/unit | total ms | avg us | ms/unit | share |")
        out.append("|---|---:|---:|---:|---:|---:|---:|")
        for k, (n, ns) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
            out.append("| `%s` | %d | %.2f | %.3f | %.2f | %.4f | %.1f%% |" %
                       (short(k), n, n / units, ns / 1e6, ns / n / 1e3, ns / units / 1e6,
                        100.0 * ns / max(total, 1)))
        out.append("")
    return "\n".join(out)


def summarise_counters(regions, rows, clock, traffic_out):
    counters = sorted(set(row["Counter_Name"] for row in rows))
    if len(counters) != 1:
        raise SystemExit("one counter per run expected, got %s" % counters)
    name = counters[0]
    out = []
    factor = None
    per_region = []
    for r in regions:
        agg = defaultdict(lambda: [0, 0.0])
        for row in rows:
            if in_window(r, clock, int(row["Start_Timestamp"])):
                a = agg[row["Kernel_Name"]]
                a[0] += 1
                a[1] += float(row["Counter_Value"])
        per_region.append((r, agg))
        if r["name"] == "copy_1GiB":
            kib = sum(v[1] for v in agg.values())
            factor = (1 << 30) / (kib * 1024.0) if kib > 0 else None
    out.append("Counter %s (KiB per dispatch as rocprofv3 reports it)." % name)
    if factor:
        out.append("Calibration: a 1 GiB device-to-device copy counts %.0f KiB, so bytes = KiB x "
                   "1024 x %.4f.\n" % ((1 << 30) / factor / 1024.0, factor))
    merged = {}
    if traffic_out and os.path.exists(traffic_out):
        with open(traffic_out) as f:
            merged = json.load(f)
    field = {"FETCH_SIZE": "fetch", "WRITE_SIZE": "write"}.get(name, name.lower())
    for r, agg in per_region:
        if r["name"] == "copy_1GiB":
            continue
        units = r["units"]
        kib = sum(v[1] for v in agg.values())
        line = "### %s  (%d units, %.0f KiB per unit" % (r["name"], units, kib / units)
        if factor:
            line += ", %.1f MB per unit scaled" % (kib / units * 1024 * factor / 1e6)
        out.append(line + ")\n")
        out.append("| kernel | calls | KiB/unit | avg KiB/dispatch |")
        out.append("|---|---:|---:|---:|")
        for k, (n, v) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
            out.append("| `%s` | %d | %.0f | %.1f |" % (short(k), n, v / units, v / n))
        out.append("")
        if factor:
            e = merged.setdefault(r["name"], {})
            e[field + "_bytes"] = int(round(kib / units * 1024 * factor))
            e[field + "_factor"] = round(factor, 4)
            if "config" in r:
                e["config"] = r["config"]
    if traffic_out and factor:
        with open(traffic_out, "w") as f:
            json.dump(merged, f, indent=1, sort_keys=True)
            f.write("\n")
    return "\n".join(out)


def short(k):
    """Kernel name without namespaces, arguments and long template lists."""
    for ns in ("mgg::", "(anonymous namespace)::", "ROCPRIM_400200_NS::", "detail::"):
        k = k.replace(ns, "")
    if k.startswith("void "):
        k = k[5:]
    p = k.find("(")
    k = k[:p] if p > 0 else k
    return k if len(k) <= 72 else k[:69] + "..."


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--regions", required=True)
    ap.add_argument("--trace")
    ap.add_argument("--counters")
    ap.add_argument("--traffic-out", default="")
    ap.add_argument("--title", default="")
    a = ap.parse_args()
    if bool(a.trace) == bool(a.counters):
        raise SystemExit("give exactly one of --trace / --counters")
    regions = [r for r in load_regions(a.regions) if "end" in r]
    rows = load_rows(a.trace or a.counters)
    clock = pick_clock(regions, [int(r["Start_Timestamp"]) for r in rows])
    if a.title:
        print("## %s\n" % a.title)
    print("Dispatches kept: those starting inside a timed window (host clock %s).\n" % clock)
    if a.trace:
        print(summarise_trace(regions, rows, clock))
    else:
        print(summarise_counters(regions, rows, clock, a.traffic_out))


if __name__ == "__main__":
    main()
