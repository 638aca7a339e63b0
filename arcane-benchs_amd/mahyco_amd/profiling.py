"""Summaries of rocprofv3 output for the committed profiles.

`stats`: per-kernel durations from a `--kernel-trace --stats` run.
`counters`: per-kernel means of the counters of one `--pmc` run (millions per
dispatch, summed over the device).
`traffic`: HBM bytes per dispatch from two `--pmc` runs (FETCH_SIZE, then
WRITE_SIZE, KiB per dispatch).  Each counter is scaled by what it reads for a
device-to-device copy of known size recorded in the same run (the largest
`__amd_rocclr_copyBuffer` dispatch, `bench.py --copy-calibration MIB`).

    python -m mahyco_amd.profiling stats <dir>/tp_kernel_stats.csv
    python -m mahyco_amd.profiling traffic <fetch csv> <write csv> MIB [label=prefix ...] [step=label,...]
    python -m mahyco_amd.profiling counters <counter csv>

The optional label=prefix arguments name the launch variant that the step
used for a kernel label (the autotuner's choice); the per-label table averages
the dispatches of that variant.  step=... lists the labels of the kernels the
step runs (default: every step label found), leaving out the autotuner's
candidate launches of kernels the step does not use.
"""
import collections
import csv
import sys

# step kernel label (as timed by the library) -> kernel name prefixes
STEP_KERNELS = {
    # with the previous step's dt kernel folded in (fused steps after the first)
    "node_lagrange": ["k_node_lagrange_2d", "k_node_lagrange_fold_2d"],
    "cell_lagrange": ["k_cell_lagrange_2d"],
    # first cell sweep with the Lagrange cell phase, last with the final one
    "cell_sweep_x_lag": ["k_cell_sweep_x_lag_2d"],
    "cell_sweep_y_lag": ["k_cell_sweep_y_lag_2d"],
    "cell_sweep_x_final": ["k_cell_sweep_x_final_2d"],
    "cell_sweep_y_final": ["k_cell_sweep_y_final_2d"],
    "cell_sweep_x": ["k_cell_sweep_lane_2d<0"],
    "cell_sweep_y": ["k_cell_sweep_ytile_2d", "k_cell_sweep_y_2d"],
    "node_sweep_x": ["k_node_sweep_x_2d", "k_node_sweep_2d<0"],
    "node_sweep_y": ["k_node_sweep_ytile_2d", "k_node_sweep_y_2d", "k_node_sweep_2d<1"],
    "cell_final": ["k_cell_final_2d"],
    "cfl": ["k_cfl_2d"],
    "dt": ["k_dt_finalize"],
}


# labels of the kernels outside the 2D single-mesh step (statistics only)
OTHER_KERNELS = {
    "node_lagrange_3d": ["k_node_lagrange_3d"],
    "cell_lagrange_3d": ["k_cell_lagrange_3d"],
    "cell_sweep_3d": ["k_cell_sweep_x_3d", "k_cell_sweep_tile_3d"],
    "node_sweep_3d": ["k_node_sweep_x_3d", "k_node_sweep_tile_3d"],
    "cell_final_3d": ["k_cell_final_3d"],
    "cell_geometry_3d": ["k_cell_geometry_3d"],
    "cfl_3d": ["k_cfl_3d"],
    "cell_stage": ["k_cell_stage_2d"],
    "eos_env": ["k_eos_env"],
    "pression_moyenne": ["k_pression_moyenne"],
    "batch_node_lagrange": ["kb_node_lagrange_2d"],
    "batch_cell_lagrange": ["kb_cell_lagrange_2d"],
    "batch_cell_sweep_x": ["kb_cell_sweep_x_2d"],
    "batch_cell_sweep_y": ["kb_cell_sweep_y_2d"],
    "batch_node_sweep_x": ["kb_node_sweep_x_2d"],
    "batch_node_sweep_y": ["kb_node_sweep_y_2d"],
    "batch_cell_final": ["kb_cell_final_2d"],
    "batch_dt": ["kb_dt_finalize"],
    "learned_alpha": ["k_learned_alpha_2d"],
    "mlp": ["k_mlp_"],
}


def short(name):
    return name.split("(")[0].replace("void ", "").replace("mhg::", "")


def untyped(kernel):
    """Kernel name without the leading storage-type template argument
    (k_x<double, 0, ...> -> k_x<0, ...>)."""
    for t in ("double, ", "float, "):
        kernel = kernel.replace("<" + t, "<")
    return kernel.replace("<double>", "").replace("<float>", "")


def label_of(kernel):
    k = untyped(kernel)
    for table in (STEP_KERNELS, OTHER_KERNELS):
        for lab, prefixes in table.items():
            if any(k.startswith(p) for p in prefixes):
                return lab
    return ""


def kernel_stats(path):
    rows = []
    for r in csv.DictReader(open(path)):
        rows.append((short(r["Name"]), int(r["Calls"]), float(r["AverageNs"]) / 1e3,
                     float(r["MinNs"]) / 1e3, float(r["MaxNs"]) / 1e3, float(r["Percentage"])))
    return rows


def _counter(path, name):
    per = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] == name:
            per[short(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return per


def traffic(fetch_csv, write_csv, calib_mib):
    """Per kernel: mean FETCH_SIZE and WRITE_SIZE (KiB per dispatch), the copy
    calibration factors and the scaled bytes per dispatch."""
    f = _counter(fetch_csv, "FETCH_SIZE")
    w = _counter(write_csv, "WRITE_SIZE")
    calib = calib_mib * 1024.0  # KiB moved each way by the calibration copy
    ff = calib / max(f["__amd_rocclr_copyBuffer"])
    wf = calib / max(w["__amd_rocclr_copyBuffer"])
    out = {}
    for k in sorted(set(f) & set(w)):
        if k.startswith("__amd") or k.startswith("at::"):
            continue
        fk = sum(f[k]) / len(f[k])
        wk = sum(w[k]) / len(w[k])
        out[k] = dict(n=len(f[k]), fetch_kib=fk, write_kib=wk,
                      bytes=(fk * ff + wk * wf) * 1024.0)
    return out, ff, wf


def main(argv):
    if argv[0] == "stats":
        print("| kernel | label | calls | avg us | min us | max us | % |")
        print("|---|---|---|---|---|---|---|")
        for k, n, a, lo, hi, pct in kernel_stats(argv[1]):
            print(f"| `{k}` | {label_of(k)} | {n} | {a:.2f} | {lo:.2f} | {hi:.2f} | {pct:.2f} |")
        return 0
    if argv[0] == "traffic":
        out, ff, wf = traffic(argv[1], argv[2], float(argv[3]))
        print(f"Copy calibration: FETCH_SIZE x {ff:.4f}, WRITE_SIZE x {wf:.4f}\n")
        print("| kernel | label | dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM bytes |")
        print("|---|---|---|---|---|---|")
        for k, v in out.items():
            print(f"| `{k}` | {label_of(k)} | {v['n']} | {v['fetch_kib']:.0f} | "
                  f"{v['write_kib']:.0f} | {v['bytes']:.0f} |")
        tuned = dict(a.split("=", 1) for a in argv[4:])
        step = tuned.pop("step").split(",") if "step" in tuned else list(STEP_KERNELS)
        print("\nStep kernels (launch variants used by the step; mean over their dispatches):\n")
        print("| label | HBM bytes per launch |")
        print("|---|---|")
        for lab, prefixes in STEP_KERNELS.items():
            if lab not in step:
                continue
            pre = [tuned[lab]] if lab in tuned else prefixes
            ks = [k for k in out if any(untyped(k).startswith(p) for p in pre) and "float" not in k]
            if not ks:
                continue
            b = sum(out[k]["bytes"] for k in ks) / len(ks)
            print(f"| {lab} | {b:.0f} |")
        return 0
    if argv[0] == "counters":
        per = collections.defaultdict(lambda: collections.defaultdict(list))
        for r in csv.DictReader(open(argv[1])):
            per[short(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
        names = sorted({c for d in per.values() for c in d})
        print("| kernel | label | dispatches | " + " | ".join(names) + " |")
        print("|---|---|---|" + "---|" * len(names))
        for k, d in sorted(per.items()):
            if k.startswith("__amd") or k.startswith("at::"):
                continue
            n = len(d[names[0]])
            vals = " | ".join(f"{sum(d[c]) / len(d[c]) / 1e6:.3f}" for c in names)
            print(f"| `{k}` | {label_of(k)} | {n} | {vals} |")
        return 0
    print(__doc__)
    return 2


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
