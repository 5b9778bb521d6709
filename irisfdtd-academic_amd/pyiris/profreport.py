#!/usr/bin/env python3
"""Summaries of rocprofv3 output for the engine's kernels.

  profreport.py stats KERNEL_STATS_CSV [--command CMD] [--steps-fp32 N] [--steps-fp64 N]
      Markdown table of a --kernel-trace --stats run (per kernel: calls,
      average and total time) and, for the precisions whose step count is
      given, the time per step of each kernel class.

  profreport.py pmc FETCH_DIR WRITE_DIR --steps N --out-dir DIR [--command-fetch CMD --command-write CMD]
      HBM bytes per step of each kernel class from two counter runs of
      pmcprobe.py (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch).  Each
      counter is scaled by what it reads for the probe's device-to-device copy
      of a known size, so the bytes are the copy's bytes by construction.
      Writes <class>_fp32_pmc.json / <class>_fp64_pmc.json and a markdown
      summary into DIR.

Kernel classes: "yee2" = the two-step pass over the plain z-chunks (one
launch per two steps); "yee2_pml" = its UPML variant; "yee" = both
instantiations of the single-step fused E+H pass (the lean one and the
general one); "yee2_batch" / "yee_batch" = their batched forms; "dft_fold",
"dft_sample", "epass", "hpass", "tfsf", "flux", "adjoint" (the gradient's
own kernels), "record" (its H records), "copy" (the calibration).
"""
import argparse
import csv
import glob
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from pmcprobe import COPY_BYTES  # noqa: E402

CLASSES = [
    ("yee2_batch", r"k_yee2_batch"),
    ("yee2_pml", r"k_yee2<\w+, \w+, true\b"),
    ("yee2", r"k_yee2<"),
    ("yee_batch", r"k_yee_fused_batch"),
    ("yee", r"k_yee_fused<"),
    ("epass", r"k_epass"),
    ("hpass", r"k_hpass"),
    ("tfsf", r"k_tfsf"),
    ("dft_fold", r"k_dft_fold|k_dft_accum"),
    ("dft_phase", r"k_dft_phase"),
    ("dft_sample", r"k_dft_sample"),
    ("flux", r"k_flux"),
    ("adjoint", r"k_adj_"),
    ("record", r"k_record"),
    ("copy", r"copyBuffer"),
]


def kernel_class(name):
    for cls, pat in CLASSES:
        if re.search(pat, name):
            return cls
    return None


def precision_of(name):
    """Field precision from the kernel's template arguments (None for the
    kernels that have none, such as the phase table)."""
    tmpl = name.split("(")[0]
    if "<float" in tmpl:
        return 4
    if "<double" in tmpl:
        return 8
    return None


def short(name):
    name = re.sub(r"\(.*$", "", name)
    return name.replace("void ", "").replace("iris::", "")


def stats(args):
    rows = list(csv.DictReader(open(args.csv)))
    out = []
    if args.command:
        out.append(f"Command: `{args.command}`\n")
    out.append("| kernel | calls | avg (us) | total (ms) |")
    out.append("|---|---:|---:|---:|")
    per_class = {}
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"])):
        calls = int(r["Calls"])
        avg = float(r["AverageNs"]) / 1e3
        tot = float(r["TotalDurationNs"]) / 1e6
        out.append(f"| `{short(r['Name'])}` | {calls} | {avg:.1f} | {tot:.2f} |")
        cls, prec = kernel_class(r["Name"]), precision_of(r["Name"])
        if cls and prec:
            key = (cls, prec)
            per_class[key] = per_class.get(key, 0.0) + tot
    nsteps = {4: args.steps_fp32, 8: args.steps_fp64}
    nsteps = {k: v for k, v in nsteps.items() if v}
    if nsteps:
        out.append("")
        out.append("Per step (all timed and warm-up steps of the run):")
        out.append("")
        out.append("| class | fields | steps | ms per step |")
        out.append("|---|---|---:|---:|")
        for (cls, prec), tot in sorted(per_class.items(), key=lambda kv: (kv[0][1], -kv[1])):
            if prec in nsteps:
                out.append(f"| {cls} | fp{32 if prec == 4 else 64} | {nsteps[prec]} | {tot / nsteps[prec]:.3f} |")
    print("\n".join(out))


def read_counter(d, counter):
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no counter_collection.csv under {d}")
    rows = []
    for f in files:
        rows += [r for r in csv.DictReader(open(f)) if r["Counter_Name"] == counter]
    return rows


def pmc(args):
    res = {}
    nlaunch = {}
    for counter, d in (("FETCH_SIZE", args.fetch), ("WRITE_SIZE", args.write)):
        rows = read_counter(d, counter)
        copies = [float(r["Counter_Value"]) for r in rows if kernel_class(r["Kernel_Name"]) == "copy"]
        if not copies:
            raise SystemExit(f"{counter}: no calibration copy found")
        ref_kib = max(copies)  # the 2 GiB copies; smaller ones are uploads
        scale = COPY_BYTES / (ref_kib * 1024.0)
        sums = {}
        for r in rows:
            cls, prec = kernel_class(r["Kernel_Name"]), precision_of(r["Kernel_Name"])
            if not cls or cls == "copy" or not prec:
                continue
            sums[(cls, prec)] = sums.get((cls, prec), 0.0) + float(r["Counter_Value"]) * 1024.0 * scale
            if counter == "FETCH_SIZE":
                nlaunch[(cls, prec)] = nlaunch.get((cls, prec), 0) + 1
        res[counter] = {"copy_kib": ref_kib, "scale": scale, "bytes": sums}
    os.makedirs(args.out_dir, exist_ok=True)
    md = ["Counter runs of `irisfdtd-academic_amd/pyiris/pmcprobe.py` (512x512x1024 slit-array workload,",
          f"{args.steps} steps per precision).", ""]
    if args.command_fetch:
        md += [f"FETCH_SIZE: `{args.command_fetch}`", ""]
    if args.command_write:
        md += [f"WRITE_SIZE: `{args.command_write}`", ""]
    f, w = res["FETCH_SIZE"], res["WRITE_SIZE"]
    md += [f"Calibration ({COPY_BYTES} B device-to-device copy): FETCH_SIZE reads {f['copy_kib']:.0f} KiB "
           f"(scale x{f['scale']:.3f}), WRITE_SIZE reads {w['copy_kib']:.0f} KiB (scale x{w['scale']:.3f}).", "",
           "| class | fields | launches | fetch GB/step | write GB/step | HBM GB/step |",
           "|---|---|---:|---:|---:|---:|"]
    keys = sorted(set(f["bytes"]) | set(w["bytes"]), key=lambda k: (k[1], k[0]))
    for cls, prec in keys:
        fb = f["bytes"].get((cls, prec), 0.0) / args.steps
        wb = w["bytes"].get((cls, prec), 0.0) / args.steps
        md.append(f"| {cls} | fp{32 if prec == 4 else 64} | {nlaunch.get((cls, prec), 0)} | {fb / 1e9:.3f} | "
                  f"{wb / 1e9:.3f} | {(fb + wb) / 1e9:.3f} |")
        # the single-step class is a per-step figure only when no two-step
        # pass ran (otherwise it also holds the other chunks' steps)
        if cls == "yee2" or (cls == "yee" and ("yee2", prec) not in nlaunch):
            name = f"{cls}_fp{32 if prec == 4 else 64}_pmc.json"
            launches = args.steps if cls == "yee" else nlaunch.get((cls, prec), 1)
            fb = f["bytes"].get((cls, prec), 0.0) / launches
            wb = w["bytes"].get((cls, prec), 0.0) / launches
            what = ("yee2 (two-step pass over the z-chunks between the UPML slabs and away from the TF/SF planes; per launch = two steps)" if cls == "yee2"
                    else "yee (fused E+H pass, all launches of one step)")
            with open(os.path.join(args.out_dir, name), "w") as fo:
                json.dump({"kernel": what,
                           "workload": "512x512x1024 slit-array film, nof=100",
                           "fetch_bytes_per_launch": round(fb), "write_bytes_per_launch": round(wb),
                           "hbm_bytes_per_launch": round(fb + wb),
                           "fetch_scale": f["scale"], "write_scale": w["scale"],
                           "commands": [args.command_fetch, args.command_write]}, fo, indent=1)
    with open(os.path.join(args.out_dir, "hbm_traffic_512x512x1024.md"), "w") as fo:
        fo.write("\n".join(md) + "\n")
    print("\n".join(md))


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("stats")
    a.add_argument("csv")
    a.add_argument("--command", default="")
    a.add_argument("--steps-fp32", type=int, default=0)
    a.add_argument("--steps-fp64", type=int, default=0)
    b = sub.add_parser("pmc")
    b.add_argument("fetch")
    b.add_argument("write")
    b.add_argument("--steps", type=int, required=True)
    b.add_argument("--out-dir", required=True)
    b.add_argument("--command-fetch", default="")
    b.add_argument("--command-write", default="")
    args = ap.parse_args()
    stats(args) if args.cmd == "stats" else pmc(args)


if __name__ == "__main__":
    main()
