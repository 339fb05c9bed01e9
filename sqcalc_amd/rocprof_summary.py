"""Short text summaries of rocprofv3 CSV output for the S(q) kernels.

  python -m sqcalc_amd.rocprof_summary stats STATS_CSV [--trace TRACE_CSV] --command "..."
      per-kernel calls / total / average duration from a
      ``rocprofv3 --kernel-trace --stats --output-format csv`` run
      (the ``*_kernel_stats.csv`` file), and with --trace the same
      dispatches grouped by launch shape.

  python -m sqcalc_amd.rocprof_summary counters CSV [CSV ...] --command "..."
      per-kernel averages of every counter or derived metric in one or more
      ``rocprofv3 --pmc`` runs (``*_counter_collection.csv``), each kernel
      in its most frequent launch shape, with VALU and LDS instructions per
      wave beside them.

  python -m sqcalc_amd.rocprof_summary traffic --fetch FETCH_CSV --write WRITE_CSV
         --copy-bytes B --json out.json --command "..." > out.md
      HBM bytes per dispatch from two ``rocprofv3 --pmc FETCH_SIZE`` and
      ``--pmc WRITE_SIZE`` runs (``*_counter_collection.csv``, KiB per
      dispatch).  Each counter is scaled by the factor that makes it read
      B bytes for the calibration copy kernel (torch's elementwise ``neg``
      over B bytes in and B bytes out, launched by ``bench.py
      --calibrate-copy``), and the JSON maps each stage name to the scaled
      fetch + write bytes per launch.
"""

import argparse
import collections
import csv
import json
import re
import sys

# kernel template name -> stage name used by the plan's stage timers
STAGE_OF = {
    "k_prep": "atom_sort", "k_scan": "atom_sort", "k_scatter": "atom_sort",
    "k_binsort": "atom_sort", "k_spread_z": "spread_z", "k_fft_y": "dft_y",
    "k_fft_x_bin": "dft_x_bin", "k_finalize": "finalize", "k_adj_x": "adj_x",
    "k_adj_y": "adj_y", "k_adj_z_gather": "adj_z_gather", "k_adj_sum": "adj_sum",
}
COPY_PATTERN = "neg"


def short_name(name):
    """'void sqc::k_spread_z<float, 480>(sqc::Geo, ...)' -> 'k_spread_z<float, 480>'."""
    n = name.replace("(anonymous namespace)::", "")
    n = n.split("(")[0].replace("void ", "").replace("sqc::", "").strip()
    return n


def base_name(name):
    return re.split(r"[<(]", short_name(name))[0]


def is_copy(name):
    return COPY_PATTERN in name and "elementwise" in name


def read_stats(path):
    rows = []
    with open(path) as fh:
        for r in csv.DictReader(fh):
            rows.append((r["Name"], int(r["Calls"]), float(r["TotalDurationNs"]),
                         float(r["AverageNs"]), float(r["Percentage"])))
    return rows


def read_counter(path, counter):
    """Mean counter value per dispatch and the number of dispatches, per
    kernel name, over the kernel's most frequent launch shape (grid and LDS
    size): the tuner's trial launches of other tiles are left out."""
    acc = collections.defaultdict(lambda: collections.defaultdict(list))
    with open(path) as fh:
        for r in csv.DictReader(fh):
            if r["Counter_Name"] != counter:
                continue
            shape = (r.get("Grid_Size"), r.get("LDS_Block_Size"))
            acc[r["Kernel_Name"]][shape].append(float(r["Counter_Value"]))
    out = {}
    for k, shapes in acc.items():
        v = max(shapes.values(), key=len)
        out[k] = (sum(v) / len(v), len(v))
    return out


def read_trace(path):
    """Dispatches of a --kernel-trace CSV grouped by (kernel, grid, LDS, alone):
    the tuner's trial launches fall in groups of their own, and `alone` tells
    the dispatches that ran by themselves from those whose run overlapped
    another dispatch (the kernels of a captured batch that run side by side;
    their start-to-end times include the sharing)."""
    rows = []
    with open(path) as fh:
        for r in csv.DictReader(fh):
            grid = (int(r["Grid_Size_X"]), int(r["Grid_Size_Y"]), int(r["Grid_Size_Z"]))
            rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"], grid,
                         int(r["LDS_Block_Size"])))
    rows.sort()
    overlapped = [False] * len(rows)
    end_max, who = -1, -1
    for i, (t0, t1, _, _, _) in enumerate(rows):
        if t0 < end_max:
            overlapped[i] = True
            overlapped[who] = True
        if t1 > end_max:
            end_max, who = t1, i
    groups = collections.OrderedDict()
    for (t0, t1, name, grid, lds), ov in zip(rows, overlapped):
        groups.setdefault((name, grid, lds, not ov), []).append(t1 - t0)
    return groups


def cmd_stats(args):
    rows = read_stats(args.csv)
    print("# rocprofv3 kernel statistics\n")
    print("Command:\n\n    %s\n" % args.command)
    print("| kernel | calls | total ms | average us | % |")
    print("|---|---:|---:|---:|---:|")
    for name, calls, tot, avg, pct in rows[:args.top]:
        print("| `%s` | %d | %.3f | %.2f | %.2f |" % (short_name(name), calls, tot / 1e6,
                                                     avg / 1e3, pct))
    if args.trace:
        groups = read_trace(args.trace)
        print("\nBy launch shape (the same kernels, dispatches grouped by grid and LDS "
              "size, and by whether the dispatch ran alone or overlapped another one, as "
              "the kernels of consecutive frames do in a captured batch; the tuner's trial "
              "tiles form groups of their own):\n")
        print("| kernel | grid (threads) | LDS bytes | ran | calls | average us |")
        print("|---|---|---:|---|---:|---:|")
        keep = [k for k in groups if base_name(k[0]) in STAGE_OF]
        keep.sort(key=lambda k: -sum(groups[k]))
        for k in keep[:args.top]:
            d = groups[k]
            print("| `%s` | %s | %d | %s | %d | %.2f |" % (short_name(k[0]), "x".join(map(str, k[1])),
                                                           k[2], "alone" if k[3] else "overlapped",
                                                           len(d), sum(d) / len(d) / 1e3))


def cmd_traffic(args):
    fetch = read_counter(args.fetch, "FETCH_SIZE")
    write = read_counter(args.write, "WRITE_SIZE")
    copies_f = [v for k, (v, _) in fetch.items() if is_copy(k)]
    copies_w = [v for k, (v, _) in write.items() if is_copy(k)]
    if not copies_f or not copies_w:
        sys.exit("no calibration copy kernel in the counter files")
    kib = args.copy_bytes / 1024.0
    sf = kib / copies_f[0]
    sw = kib / copies_w[0]
    out = {"_calibration": {"copy_bytes": args.copy_bytes, "fetch_kib": copies_f[0],
                            "write_kib": copies_w[0], "fetch_scale": sf, "write_scale": sw}}
    print("# HBM traffic per dispatch (rocprofv3 FETCH_SIZE / WRITE_SIZE)\n")
    print("Commands:\n\n    %s\n" % args.command.replace(" && ", "\n    "))
    print("Calibration: elementwise copy of %d bytes in and out read FETCH_SIZE %.0f KiB, "
          "WRITE_SIZE %.0f KiB; scale factors fetch %.4f, write %.4f.\n"
          % (args.copy_bytes, copies_f[0], copies_w[0], sf, sw))
    print("| kernel | stage | dispatches | fetch MB | write MB | total MB (scaled) |")
    print("|---|---|---:|---:|---:|---:|")
    # a stage's bytes: the sum over its kernels, each taken in the compiled
    # instance that ran most often (the tuned plan's; the tuner's trials of
    # other tiles run other instances)
    chosen = {}
    for k, (_, n) in fetch.items():
        b = base_name(k)
        if b not in STAGE_OF or k not in write:
            continue
        if b not in chosen or n > fetch[chosen[b]][1]:
            chosen[b] = k
    per_stage = collections.defaultdict(float)
    for k in sorted(fetch, key=lambda k: -fetch[k][0]):
        b = base_name(k)
        if b not in STAGE_OF or k not in write:
            continue
        fb = fetch[k][0] * 1024.0 * sf
        wb = write[k][0] * 1024.0 * sw
        st = STAGE_OF[b]
        if chosen[b] == k:
            per_stage[st] += fb + wb
        print("| `%s` | %s | %d | %.1f | %.1f | %.1f |" % (short_name(k), st, fetch[k][1], fb / 1e6,
                                                          wb / 1e6, (fb + wb) / 1e6))
    for st, v in per_stage.items():
        out[st] = int(round(v))
    if args.json:
        with open(args.json, "w") as fh:
            json.dump(out, fh, indent=1)
            fh.write("\n")


def read_counters(paths):
    """{kernel: {counter: (mean per dispatch, dispatches)}} over every counter
    in the files, each kernel in its most frequent launch shape."""
    acc = collections.defaultdict(lambda: collections.defaultdict(
        lambda: collections.defaultdict(list)))
    for path in paths:
        with open(path) as fh:
            for r in csv.DictReader(fh):
                shape = (r.get("Grid_Size"), r.get("LDS_Block_Size"))
                acc[r["Kernel_Name"]][shape][r["Counter_Name"]].append(float(r["Counter_Value"]))
    out = {}
    for k, shapes in acc.items():
        best = max(shapes.values(), key=lambda c: max(len(v) for v in c.values()))
        out[k] = {c: (sum(v) / len(v), len(v)) for c, v in best.items()}
    return out


def cmd_counters(args):
    data = read_counters(args.csv)
    names = sorted({c for v in data.values() for c in v})
    print("# rocprofv3 counters per dispatch\n")
    print("Commands:\n\n    %s\n" % args.command.replace(" && ", "\n    "))
    print("Each kernel in its most frequent launch shape (the tuned plan's; the tuner's "
          "trial tiles are left out).  Derived metrics are rocprofv3's own: VALUBusy and "
          "LdsUtil in % of the CU cycles, LDSBankConflict in % of the CU cycles spent in "
          "bank conflicts, OccupancyPercent in % of the wave slots.\n")
    print("| kernel | dispatches | " + " | ".join(names) +
          " | VALU insts / wave | LDS insts / wave |")
    print("|---|---:|" + "---:|" * (len(names) + 2))
    keep = [k for k in data if base_name(k) in STAGE_OF]
    keep.sort(key=lambda k: (STAGE_OF[base_name(k)], short_name(k)))
    for k in keep:
        v = data[k]
        n = max(c[1] for c in v.values())
        get = lambda c: v[c][0] if c in v else float("nan")
        div = lambda a, b: a / b if b else float("nan")
        waves = get("SQ_WAVES")
        extra = [div(get("SQ_INSTS_VALU"), waves), div(get("SQ_INSTS_LDS"), waves)]
        cells = ["%.4g" % get(c) for c in names] + ["%.1f" % x for x in extra]
        print("| `%s` | %d | %s |" % (short_name(k), n, " | ".join(cells)))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("stats")
    a.add_argument("csv")
    a.add_argument("--command", default="")
    a.add_argument("--top", type=int, default=20)
    a.add_argument("--trace", help="the *_kernel_trace.csv of the same run")
    b = sub.add_parser("traffic")
    b.add_argument("--fetch", required=True)
    b.add_argument("--write", required=True)
    b.add_argument("--copy-bytes", type=int, required=True)
    b.add_argument("--json")
    b.add_argument("--command", default="")
    c = sub.add_parser("counters")
    c.add_argument("csv", nargs="+")
    c.add_argument("--command", default="")
    args = ap.parse_args(argv)
    {"stats": cmd_stats, "traffic": cmd_traffic, "counters": cmd_counters}[args.cmd](args)


if __name__ == "__main__":
    main()
