"""Condenses rocprofv3 CSV output into the short summaries kept under profiles/.

    python rocprof_summary.py stats  <rocprof-dir> <out.txt> "<command that made it>"
    python rocprof_summary.py traffic <fetch-dir> <write-dir> <cal-fetch-dir> <cal-write-dir> \
        <copy-bytes> <out.json> "<commands>"
    python rocprof_summary.py counts <pmc-dir> <out.json> "<command that made it>" [<kernel-trace-dir>]

``stats`` reads run_kernel_stats.csv (rocprofv3 --kernel-trace --stats --output-format csv) and
writes one line per kernel: launches, average/min/max duration and share of GPU time.

``traffic`` reads the FETCH_SIZE and WRITE_SIZE passes (rocprofv3 --pmc <counter> --kernel-trace
--output-format csv, KiB per dispatch) of the workload and of a plain device copy of
``copy-bytes`` each way, scales each counter by (copy bytes / what the counter read for the copy)
and writes HBM bytes per launch for every kernel, the scale factors beside them.

``counts`` reads one --pmc pass with any counters and writes, per kernel and counter, the mean
value per dispatch; given the kernel-trace directory of a run of the same command it adds each
kernel's mean duration and every counter as a rate per second of kernel time.
"""
import collections
import csv
import json
import os
import sys

_SHORT = {
    "p2g3_wave_kernel": "p2g3_seg", "p2g3_atomic_kernel": "p2g3_atomic", "g2p3_tiled_kernel<4, true>": "g2p3_fast",
    "g2p3_tiled_kernel": "g2p3_tiled", "g2p3_kernel": "g2p3", "grid3_kernel": "grid3", "clear4_kernel": "mpm3_clear",
    "keys3_kernel": "keys3", "hist_kernel": "radix_hist", "scan_rows_kernel": "radix_scan",
    "scatter_kernel": "radix_scatter", "permute_kernel": "permute", "hll_step_kernel": "hll_step",
    "clear_rows_kernel": "mpm3_clear", "hll_step_kernel<false>": "hll_step", "hll_step_kernel<true>": "hll_step_fast", "flags_kernel": "grid3_flags", "batch2_p2g": "batch2_p2g",
    "batch2_grid": "batch2_grid", "batch2_g2p": "batch2_g2p", "batch2_clear": "batch2_clear",
    "batch2_keys": "batch2_keys", "batch2_starts": "batch2_starts", "conv3x3_lds_kernel": "conv3x3",
    "conv3x3_wgrad_kernel": "conv3x3_wgrad", "wgrad_reduce_kernel": "wgrad_reduce", "hll64_kernel": "hll64",
    "adam_kernel": "adam", "loss_kernel": "loss", "unpack_kernel": "unpack", "pack_kernel": "pack",
    "flip_kernel": "flip", "dam_bumps_kernel": "dam_bumps", "keys2_kernel": "keys2", "g2p2_kernel": "g2p2",
    "starts_kernel": "starts2", "grid_tension_rows_kernel": "grid3_tension", "blur_march_kernel": "tension3_blur",
    "stress_rows_kernel": "tension3_stress", "region_kernel": "tension3_region", "lod_target_kernel": "lod3_target",
    "lod_kernel": "lod3", "after_sort_kernel": "mpm3_after_sort", "slots_kernel": "mpm3_slots",
    "fill_zero_kernel": "hll_fill_zero", "f64_to_f32_kernel": "f64_to_f32", "batch2_flags_kernel": "batch2_flags",
    "p2g3_layer_kernel": "p2g3_layer", "group_count_kernel": "layers3_count", "group_scan_kernel": "layers3_scan",
    "group_list_kernel": "layers3_list", "group_order_kernel": "layers3_order", "p2g3_fixed_kernel": "p2g3_fixed",
    "fixed_rows_kernel": "fixed3_rows", "blur2_kernel": "tension2_blur", "stress2_kernel": "tension2_stress",
    "grid2_tension_kernel": "grid2_tension", "batch2_decode_kernel": "batch2_decode",
}


def short_name(kernel: str) -> str:
    """The library's timer name for a demangled kernel name (the names bench.py reports): the
    function's own name first (with, then without, its template arguments), then any key the
    name contains."""
    head = kernel.replace("(anonymous namespace)", "").split("(")[0]
    base = head.split("::")[-1]
    for candidate in (base, base.split("<")[0]):
        if candidate in _SHORT:
            return _SHORT[candidate]
    for key, name in _SHORT.items():
        if key in kernel:
            return name
    return head[-40:]


def _labels(kernels) -> dict:
    """A label for each demangled kernel name: its short name, with its template arguments
    added where two instantiations share one."""
    by_short = collections.defaultdict(list)
    for k in kernels:
        by_short[short_name(k)].append(k)
    labels = {}
    for short, names in by_short.items():
        for k in names:
            head = k.replace("(anonymous namespace)", "").split("(")[0]
            args = head[head.index("<"):] if len(names) > 1 and "<" in head else ""
            labels[k] = short + args.replace(" ", "")
    return labels


def _counters(directory):
    per = collections.defaultdict(list)
    with open(os.path.join(directory, "run_counter_collection.csv")) as f:
        for r in csv.DictReader(f):
            per[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in per.items()}, {k: len(v) for k, v in per.items()}


def stats(directory, out, command):
    with open(os.path.join(directory, "run_kernel_stats.csv")) as f:
        rows = list(csv.DictReader(f))
    lines = [f"# rocprofv3 kernel statistics", f"# command: {command}",
             "# name | launches | avg us | min us | max us | share of GPU time %"]
    for r in rows:
        lines.append(f"{short_name(r['Name']):14s} | {r['Name'][:90]} | {r['Calls']} | "
                     f"{float(r['AverageNs']) / 1e3:.1f} | {float(r['MinNs']) / 1e3:.1f} | "
                     f"{float(r['MaxNs']) / 1e3:.1f} | {float(r['Percentage']):.2f}")
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")


def traffic(fetch_dir, write_dir, cal_fetch_dir, cal_write_dir, copy_bytes, out, command):
    copy_bytes = float(copy_bytes)
    cf, _ = _counters(cal_fetch_dir)
    cw, _ = _counters(cal_write_dir)
    copy = [k for k in cf if "copyBuffer" in k or "copy" in k.lower()]
    if not copy:
        raise SystemExit("no copy kernel in the calibration passes")
    fetch_scale = copy_bytes / (cf[copy[0]] * 1024.0)
    write_scale = copy_bytes / (cw[copy[0]] * 1024.0)
    f, nf = _counters(fetch_dir)
    w, _ = _counters(write_dir)
    kernels = {}
    labels = _labels(f)
    for k in f:
        if k not in w or "rocclr" in k:
            continue
        fb = f[k] * 1024.0 * fetch_scale
        wb = w[k] * 1024.0 * write_scale
        kernels[labels[k]] = {"kernel": k, "dispatches": nf[k], "fetch_kib_raw": round(f[k], 1),
                                  "write_kib_raw": round(w[k], 1), "fetch_bytes": round(fb),
                                  "write_bytes": round(wb), "hbm_bytes_per_launch": round(fb + wb)}
    doc = {"command": command, "copy_bytes_each_way": copy_bytes, "fetch_scale": round(fetch_scale, 4),
           "write_scale": round(write_scale, 4), "kernels": kernels}
    with open(out, "w") as fh:
        json.dump(doc, fh, indent=1)
        fh.write("\n")


def counts(directory, out, command, trace_dir=None):
    per = collections.defaultdict(list)
    with open(os.path.join(directory, "run_counter_collection.csv")) as f:
        for r in csv.DictReader(f):
            per[(r["Kernel_Name"], r["Counter_Name"])].append(float(r["Counter_Value"]))
    durations = {}
    if trace_dir:
        with open(os.path.join(trace_dir, "run_kernel_stats.csv")) as f:
            for r in csv.DictReader(f):
                durations[r["Name"]] = float(r["AverageNs"]) * 1e-9
    kernels = {}
    labels = _labels({k for k, _ in per})
    for (k, counter), values in sorted(per.items()):
        if "rocclr" in k:
            continue
        entry = kernels.setdefault(labels[k], {"kernel": k, "dispatches": len(values)})
        mean = sum(values) / len(values)
        entry[counter] = round(mean, 1)
        if k in durations:
            entry["avg_seconds"] = durations[k]
            entry[counter + "_per_s"] = round(mean / durations[k], 1)
    with open(out, "w") as fh:
        json.dump({"command": command, "kernels": kernels}, fh, indent=1)
        fh.write("\n")


if __name__ == "__main__":
    if sys.argv[1] == "stats":
        stats(*sys.argv[2:5])
    elif sys.argv[1] == "traffic":
        traffic(*sys.argv[2:9])
    elif sys.argv[1] == "counts":
        counts(*sys.argv[2:6])
    else:
        raise SystemExit(__doc__)
