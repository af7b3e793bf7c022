"""Summaries of rocprofv3 output for the solver kernels.

    python -m cie_amd.profiling stats <kernel_stats.csv> <out.md> "<command>"
    python -m cie_amd.profiling traffic <fetch dir> <write dir> <out.json> <out.md> "<command>"
    python -m cie_amd.profiling merge <out.json> <out.md> "<command>" <traffic.json> ...

``stats`` turns a --kernel-trace --stats CSV into a short table.
``traffic`` reads two --pmc runs (FETCH_SIZE, then WRITE_SIZE; both in KiB per
dispatch) of a program that ends with a device-to-device copy of a known
size (bench.py --calibration-copy): the counters of that copy give the scale
factor of each counter, and the scaled mean per dispatch of every kernel is
its HBM traffic in bytes.
"""
import csv
import glob
import json
import re
import sys
from collections import defaultdict
from pathlib import Path

CALIBRATION_BYTES = 1 << 30  # bench.py --calibration-copy moves 1 GiB device to device
CALIBRATION_KERNEL = "__amd_rocclr_copyBuffer"


def _short(name):
    """Kernel name without namespaces, return type and parameter list."""
    name = re.sub(r"\(anonymous namespace\)::", "", name)
    name = name.replace("ciehip::", "").replace("void ", "")
    return name.split("(", 1)[0][:100]


LONG_FRACTION = 0.25  # a launch "long" if it ran at least this fraction of the kernel's longest launch


def _long_launches(directory):
    """Per kernel name, from the --kernel-trace CSV beside the stats: (count, mean ns) of the launches that ran at
    least LONG_FRACTION of that kernel's longest one, and the VGPR count. The solver loops enqueue a few trips past
    convergence whose kernels return at once; this mean leaves those out."""
    files = glob.glob(str(Path(directory) / "*kernel_trace.csv"))
    if not files:
        return {}
    durations, vgprs = defaultdict(list), {}
    for row in csv.DictReader(open(files[0])):
        durations[row["Kernel_Name"]].append(int(row["End_Timestamp"]) - int(row["Start_Timestamp"]))
        vgprs[row["Kernel_Name"]] = row.get("VGPR_Count", "")
    result = {}
    for name, values in durations.items():
        longest = max(values)
        kept = [v for v in values if v >= LONG_FRACTION * longest]
        result[name] = (len(kept), sum(kept) / len(kept), vgprs[name])
    return result


def kernel_stats(path, out, command):
    rows = list(csv.DictReader(open(path)))
    long = _long_launches(Path(path).parent)
    lines = [f"# Kernel time: `{command}`", "",
             f"long calls / long mean: the launches that ran at least {LONG_FRACTION:g} of the kernel's longest "
             "(the solver loops enqueue a few trips past convergence whose kernels return at once).", "",
             "| kernel | calls | mean (us) | long calls | long mean (us) | VGPRs | total share |",
             "|---|---:|---:|---:|---:|---:|---:|"]
    for r in rows[:48]:  # the workloads' set-up (autotuning, format builds) lists many short-lived kernels
        count, mean, vgpr = long.get(r["Name"], ("", None, ""))
        lines.append(f"| `{_short(r['Name'])}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | {count} | "
                     f"{'' if mean is None else f'{mean / 1e3:.1f}'} | {vgpr} | {float(r['Percentage']):.1f}% |")
    Path(out).write_text("\n".join(lines) + "\n")


def _counter_means(directory, counter):
    """Mean counter value per kernel name over its working dispatches (those at least
    LONG_FRACTION of the kernel's largest: the solver loops enqueue a few trips past
    convergence whose kernels return at once and move next to nothing, as in the kernel
    time summaries), and the value of the last dispatch of the copy kernel: the calibration
    copy ends the program, and the runtime launches other copies under that name before it."""
    files = glob.glob(str(Path(directory) / "*counter_collection.csv"))
    if not files:
        raise FileNotFoundError(f"no counter_collection.csv under {directory}")
    values = defaultdict(list)
    calibration = None
    for row in csv.DictReader(open(files[0])):
        if row["Counter_Name"] != counter:
            continue
        if row["Kernel_Name"] == CALIBRATION_KERNEL:
            dispatch = int(row["Dispatch_Id"])
            if calibration is None or dispatch > calibration[0]:
                calibration = (dispatch, float(row["Counter_Value"]))
        else:
            values[row["Kernel_Name"]].append(float(row["Counter_Value"]))
    if calibration is None:
        raise ValueError(f"{directory}: no {CALIBRATION_KERNEL} dispatch to calibrate {counter} with")
    means = {}
    for name, v in values.items():
        working = [x for x in v if x >= LONG_FRACTION * max(v)] or v
        means[name] = sum(working) / len(working)
    return means, calibration[1]


def traffic(fetch_dir, write_dir, out_json, out_md, command):
    fetch, copy_fetch = _counter_means(fetch_dir, "FETCH_SIZE")
    write, copy_write = _counter_means(write_dir, "WRITE_SIZE")
    copy_fetch *= 1024
    copy_write *= 1024
    fetch_scale = CALIBRATION_BYTES / copy_fetch
    write_scale = CALIBRATION_BYTES / copy_write
    kernels = {}
    for name in fetch:
        if name not in write:
            continue
        kernels[_short(name)] = {
            "fetch_bytes": fetch[name] * 1024 * fetch_scale,
            "write_bytes": write[name] * 1024 * write_scale,
            "traffic_bytes": fetch[name] * 1024 * fetch_scale + write[name] * 1024 * write_scale,
        }
    result = {"command": command, "calibration": {"bytes": CALIBRATION_BYTES, "fetch_scale": fetch_scale,
                                                  "write_scale": write_scale}, "kernels": kernels}
    Path(out_json).write_text(json.dumps(result, indent=1) + "\n")
    lines = [f"# HBM traffic per working dispatch: `{command}`", "",
             f"Working dispatches: those that moved at least {LONG_FRACTION:g} of the kernel's largest (the solver "
             "loops enqueue a few trips past convergence whose kernels return at once).", "",
             f"Calibration: a {CALIBRATION_BYTES} B device copy read FETCH_SIZE x{fetch_scale:.3f}, "
             f"WRITE_SIZE x{write_scale:.3f} of its true bytes; the figures below are scaled by those factors.", "",
             "| kernel | fetch (MB) | write (MB) | traffic (MB) |", "|---|---:|---:|---:|"]
    # solver kernels first, then the standalone/autotune SpMV candidates
    for name, k in sorted(kernels.items(), key=lambda kv: (kv[0].startswith("csrProductKernel"), kv[0])):
        lines.append(f"| `{name}` | {k['fetch_bytes'] / 1e6:.1f} | {k['write_bytes'] / 1e6:.1f} | "
                     f"{k['traffic_bytes'] / 1e6:.1f} |")
    Path(out_md).write_text("\n".join(lines) + "\n")


def merge(out_json, out_md, command, *inputs):
    """One traffic summary from several (profiled runs that each pinned another launch of the same workload):
    every kernel's entry from the first input that has it."""
    kernels, calibration = {}, None
    for path in inputs:
        data = json.loads(Path(path).read_text())
        calibration = calibration or data["calibration"]
        for name, entry in data["kernels"].items():
            kernels.setdefault(name, entry)
    result = {"command": command, "calibration": calibration, "kernels": kernels}
    Path(out_json).write_text(json.dumps(result, indent=1) + "\n")
    lines = [f"# HBM traffic per working dispatch: `{command}`", "",
             f"Working dispatches: those that moved at least {LONG_FRACTION:g} of the kernel's largest (the solver "
             "loops enqueue a few trips past convergence whose kernels return at once).", "",
             f"Calibration: a {CALIBRATION_BYTES} B device copy read FETCH_SIZE x{calibration['fetch_scale']:.3f}, "
             f"WRITE_SIZE x{calibration['write_scale']:.3f} of its true bytes in the first run (each run is scaled "
             "by its own copy); the figures below are scaled by those factors.", "",
             "| kernel | fetch (MB) | write (MB) | traffic (MB) |", "|---|---:|---:|---:|"]
    for name, k in sorted(kernels.items(), key=lambda kv: (kv[0].startswith("csrProductKernel"), kv[0])):
        lines.append(f"| `{name}` | {k['fetch_bytes'] / 1e6:.1f} | {k['write_bytes'] / 1e6:.1f} | "
                     f"{k['traffic_bytes'] / 1e6:.1f} |")
    Path(out_md).write_text("\n".join(lines) + "\n")


def main(argv):
    if len(argv) >= 4 and argv[0] == "stats":
        kernel_stats(argv[1], argv[2], argv[3])
    elif len(argv) >= 6 and argv[0] == "traffic":
        traffic(argv[1], argv[2], argv[3], argv[4], argv[5])
    elif len(argv) >= 5 and argv[0] == "merge":
        merge(argv[1], argv[2], argv[3], *argv[4:])
    else:
        print(__doc__)
        return 2
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
