#!/usr/bin/env python3
"""Throughput of the MI355X FDTD engine on the roofline-scale workload.

Workload: the slit-array film on a 512 x 512 x 1024 grid (dx = 5 nm),
periodic x/y, 20-cell z-UPML at each end, eps_r = 4.0 film 200 nm (40 cells)
thick cut by a 100 nm (20-cell) slit along y, normal-incidence TF/SF
Gaussian pulse covering 400-1000 nm, on-the-fly DFT at 100 wavelengths on a
full-lateral box (z = 25..1000), output cadence every nt/10 steps (nt =
20000).  A step is one pass of the time loop (E, H, TF/SF, DFT sample, DFT
fold every 64 samples with FP32 fields / 32 with FP64, face fluxes at the
cadence) over the whole grid; the engine advances the plain z-chunks two
steps per pass.

Prints one JSON line.  value = FP32-field Gcell-updates/s over the timed
steps (inputs resident in HBM); the FP64-field rate, per-kernel HIP-event
timings, the HBM roofline of the dominant kernel and a CPU baseline (the
reference binary on a bounded sample) ride along.

A second object, "adjoint", times the gradient workload (see ADJ_WORKLOAD):
the plain forward run against the gradient of its summed transmission.  A
third, "batch", times the batched Bloch sweep (run_batch) against the same
runs one after another.

Before the forward runs the engine times candidate layouts of its step
passes (iris_autotune; all give the same results bit for bit) and keeps the
fastest; --no-autotune keeps the default layout.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-fp64] [--no-cpu] [--no-adjoint]
                  [--no-batch] [--no-autotune]
  python bench.py --adjoint-only [--adjoint-steps N]
  python bench.py --batch-only [--batch-steps N]
"""
import argparse
import json
import os
import platform
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "irisfdtd-academic_amd")
sys.path.insert(0, PKG)

METRIC = "Gcell-updates/s & % HBM roofline, FP32 vs FP64 plain; adjoint/forward time"
HBM_PEAK_GBS = 8000.0
NX = NY = 512
NZ = 1024
NT_CONFIG = 20000
WORKLOAD = dict(description="slit-array film 512x512x1024 (roofline workload)",
                ixmax=NX, iymax=NY, izmax=NZ, pmlzi=20, pmlzf=1005, hi=492, dh=40, dx=20,
                eps1=1.0, eps_bbox=4.0, eps3=1.0, tfsf_o=30, wlpump=700.0, pump_dt=4.0, def_pulse=4,
                box=(1, NX, 1, NY, 25, 1000), nof=100, dlambda=600.0 / 99.0, lno=400.0,
                nt=NT_CONFIG, no_downloads=10, to_fourier=0, sample_fourier=1)


# gradient workload: a random design layer (eps_r uniform in [1, 4] per cell,
# seed 4) of 128 x 128 x 64 cells inside a 128 x 128 x 768 grid, objective
# the transmission T = J6/(J5+J6) summed over 10 wavelengths in 500-900 nm,
# FP32 fields, N = 12000 steps, forward snapshots every ceil(sqrt(N)) steps
ADJ_NT = 12000
ADJ_WORKLOAD = dict(description="random design layer 128x128x64 in 128x128x768 (gradient workload)",
                    ixmax=128, iymax=128, izmax=768, pmlzi=20, pmlzf=749, hi=353, dh=64, dx=0,
                    eps1=1.0, eps_bbox=2.5, eps3=1.0, tfsf_o=30, wlpump=700.0, pump_dt=4.0, def_pulse=4,
                    box=(1, 128, 1, 128, 25, 744), nof=10, dlambda=400.0 / 9.0, lno=500.0,
                    nt=ADJ_NT, no_downloads=ADJ_NT, to_fourier=0, sample_fourier=1)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_setup(precision, base=None, **over):
    from pyiris import inputs, load_input
    kw = dict(base or WORKLOAD)
    kw.update(over)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "inputFDTD.dat")
        with open(path, "w") as f:
            f.write(inputs.render(**kw))
        return load_input(path, precision=precision)


def algorithmic_bytes(p, precision, two_step_planes=0, upml_two_step_planes=0, upml_in_two_step=0):
    """Compulsory HBM bytes per launch (c = bytes per complex value, r = per
    real):
      yee (fused E+H pass, one step): read E and H, write E and H (12
        complex per cell); per UPML cell read+write D and B (12 complex); per
        structure cell eps_inv;
      yee2 (two-step pass over the plain z-chunks): read E and H of step n,
        write E and H of step n+2 (12 complex per cell of those planes);
      yee2_pml (its variant over the chunks reaching into the UPML): as yee2,
        plus read D and B of step n and write them for step n+2 (12 complex
        per UPML cell);
      yee_rest (one step of the other z-chunks inside a two-step pass): as
        yee over the remaining planes;
      in-place E or H pass: read E and H, write one of them (9 complex per
        cell); per UPML cell read+write D or B (6 complex);
      per step by the two-pass definition (144 B/cell FP32 + 96 B per UPML
        cell + 4 B per structure cell, doubled for FP64)."""
    c = 8 if precision == 4 else 16
    r = c // 2
    ncell = p.nx * p.ny * p.nz
    npml = p.pmlzi + (p.nz - p.pmlzf + 1)
    nxy = p.nx * p.ny
    nstr = nxy * (p.hf - p.hi + 1)
    rest = p.nz - two_step_planes - upml_two_step_planes
    return {"yee": 12 * c * ncell + 12 * c * nxy * npml + r * nstr,
            "yee2": 12 * c * nxy * two_step_planes,
            "yee2_pml": 12 * c * nxy * (upml_two_step_planes + upml_in_two_step),
            "yee_rest": 12 * c * nxy * rest + 12 * c * nxy * (npml - upml_in_two_step),
            "epass": 9 * c * ncell + 6 * c * nxy * npml + r * nstr,
            "hpass": 9 * c * ncell + 6 * c * nxy * npml,
            "step_two_pass": 18 * c * ncell + 12 * c * nxy * npml + r * nstr}


AUTOTUNE_STEPS = 8  # steps timed per candidate layout


def run_gpu(precision, steps, warmup, autotune=True):
    from pyiris import Engine
    s = make_setup(precision)
    p = s.params
    log(f"precision {precision}: creating engine for {p.nx}x{p.ny}x{p.nz}, nof={p.nof}")
    eng = Engine(s, precision=precision)
    layout = None
    if autotune:
        # layout of the step passes, chosen before the first step (untimed, like the set-up)
        layout = eng.autotune(AUTOTUNE_STEPS)
        log(f"precision {precision}: layout {layout}")
    nos = 1

    def advance(it_from, it_to, nos):
        it = it_from
        while it <= it_to:
            cand = p.to_fourier + p.no_downloads * nos
            stop = min(cand, it_to) if cand >= it else it_to
            eng.step(it, stop)
            if stop >= p.to_fourier and (stop == cand or stop == p.nt):
                eng.flux(1)
                nos += 1
            it = stop + 1
        return nos

    nos = advance(1, warmup, nos)
    eng.sync()
    eng.profile(True)
    eng.profile_read(reset=True)
    log(f"precision {precision}: timing {steps} steps")
    t0 = time.perf_counter()
    nos = advance(warmup + 1, warmup + steps, nos)
    eng.sync()
    t1 = time.perf_counter()
    prof = eng.profile_read(reset=True)
    eng.profile(False)
    info = eng.pass_info()
    eng.close()
    secs = t1 - t0
    gcells = p.nx * p.ny * p.nz * steps / secs / 1e9
    ab = algorithmic_bytes(p, precision, info["two_step_planes"], info["upml_two_step_planes"],
                           info["upml_planes_in_two_step"])
    kern = {}
    for k, (ms, n) in prof.items():
        if n:
            kern[k] = {"avg_us": 1e3 * ms / n, "launches": n, "total_ms": ms,
                       "algorithmic_bytes": ab.get(k)}
            if ab.get(k):
                kern[k]["gbs"] = ab[k] / (1e-3 * ms / n) / 1e9
    dom = max(kern, key=lambda k: kern[k]["total_ms"])
    ms_step = 1e3 * secs / steps
    yee_ms = sum(kern[k]["total_ms"] for k in ("yee", "yee2", "yee2_pml", "yee_rest") if k in kern) / steps
    return dict(gcells=gcells, ms_per_step=ms_step, secs=secs, kernels=kern, dominant=dom, passes=info, layout=layout,
                yee_ms_per_step=yee_ms,
                step_two_pass_frac=ab["step_two_pass"] / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS)


def run_adjoint(precision=4, nt=None):
    """Wall time of the plain forward run and of the gradient (forward run
    recording H of its last steps in the free device memory and keeping
    snapshots every ceil(sqrt(N)) steps before them, then the adjoint sweep
    with segment-wise recomputation) on the gradient workload; both with
    inputs resident on the device."""
    import math

    import numpy as np
    from pyiris import Engine
    from pyiris.adjoint import structure_of, with_structure
    over = {} if nt is None else dict(nt=nt, no_downloads=nt)
    s0 = make_setup(precision, ADJ_WORKLOAD, **over)
    rng = np.random.default_rng(4)
    s = with_structure(s0, 1.0 / rng.uniform(1.0, 4.0, size=structure_of(s0).shape))
    p = s.params
    seg = int(math.ceil(math.sqrt(p.nt)))
    log(f"adjoint: {p.nx}x{p.ny}x{p.nz}, design {p.nx}x{p.ny}x{p.hf - p.hi + 1}, nt={p.nt}, segment={seg}")
    with Engine(s, precision=precision) as eng:   # warm-up: module load, first launches
        eng.run_to(50)
        eng.sync()
    with Engine(s, precision=precision) as eng:
        eng.sync()
        t0 = time.perf_counter()
        eng.run_to(p.nt)
        j = eng.flux(1)
        t_fwd = time.perf_counter() - t0
    with Engine(s, precision=precision) as eng:
        eng.checkpoint(seg, -1)
        ck = eng.checkpoint_info()
        eng.sync()
        t0 = time.perf_counter()
        eng.run_to(p.nt)
        j2 = eng.flux(1)
        t_mid = time.perf_counter()
        eng.profile(True)
        # weights of sum_w T(w), T = J6/(J5+J6)
        den = j2[:, 4] + j2[:, 5]
        w = np.zeros_like(j2)
        w[:, 4] = -j2[:, 5] / den ** 2
        w[:, 5] = j2[:, 4] / den ** 2
        g = eng.adjoint(1, w)
        t_grad = time.perf_counter() - t0
        prof = eng.profile_read(reset=True)
    T = float((j[:, 5] / (j[:, 4] + j[:, 5])).sum())
    adj_ms = prof["adjoint"][0]
    out = {"workload": "random design layer 128x128x64 (eps_r U[1,4], seed 4) in 128x128x768, z-UPML 20, "
                       "objective sum T over 10 wavelengths 500-900 nm",
           "grid": [p.nx, p.ny, p.nz], "design": [p.nx, p.ny, p.hf - p.hi + 1], "nt": p.nt, "segment": seg,
           "snapshots": ck["snapshots"], "recorded_steps": ck["recorded_steps"],
           "fields": "fp32" if precision == 4 else "fp64",
           "forward_s": round(t_fwd, 4), "gradient_s": round(t_grad, 4),
           "gradient_forward_part_s": round(t_mid - t0, 4),
           "ratio": round(t_grad / t_fwd, 3),
           "adjoint_ms_per_step": round(adj_ms / p.nt, 4),
           "objective_sum_T": T, "forward_flux_equal": bool(np.array_equal(j, j2)),
           "grad_max_abs": float(np.abs(g).max())}
    log(f"adjoint: forward {t_fwd:.2f} s, gradient {t_grad:.2f} s, ratio {t_grad / t_fwd:.2f} "
        f"({ck['recorded_steps']} steps recorded, {ck['snapshots']} snapshots)")
    return out


BATCH_N = 64
BATCH_STEPS = 1000


def run_batch(precision=4, n=BATCH_N, steps=BATCH_STEPS):
    """The batched Bloch sweep: n random 2D-periodic gratings (nx, ny drawn
    from 32..128, nz = 768, a 30-cell design layer with eps_r per column
    U[1,4] smoothed by a 3-cell box filter, incidence angle U[0, 60] degrees
    with the matching Bloch phase, seed 3, DFT at 50 wavelengths on a full
    box), `steps` steps each.  Timed once as n runs one after another and
    once as one batch; the fluxes of the two must be bitwise equal."""
    import numpy as np
    from pyiris import Batch, Engine
    from pyiris.sweep import grating_sweep
    setups, drawn = grating_sweep(n=n, seed=3, nz=768, nt=steps, nof=50, precision=precision)
    cells = sum(s.params.nx * s.params.ny * s.params.nz for s in setups)
    log(f"batch: {n} gratings, {cells / 1e6:.1f} M cells in all, {steps} steps")
    with Engine(setups[0], precision=precision) as eng:   # warm-up: module load, first launches
        eng.run_to(min(20, steps))
        eng.sync()
    t_seq, f_seq = 0.0, []
    for s in setups:
        with Engine(s, precision=precision) as eng:
            eng.sync()
            t0 = time.perf_counter()
            eng.run_to(steps)
            eng.sync()
            t_seq += time.perf_counter() - t0
            f_seq.append(eng.flux(1))
    engines = [Engine(s, precision=precision) for s in setups]
    try:
        b = Batch(engines)
        info = b.info()
        for e in engines:
            e.sync()
        t0 = time.perf_counter()
        b.step(1, steps)
        b.sync()
        t_bat = time.perf_counter() - t0
        f_bat = [e.flux(1) for e in engines]
        b.close()
    finally:
        for e in engines:
            e.close()
    same = all(np.array_equal(a, c) for a, c in zip(f_seq, f_bat))
    log(f"batch: one after another {t_seq:.2f} s, batched {t_bat:.2f} s, speed-up {t_seq / t_bat:.2f}, "
        f"bitwise equal {same}")
    return {"workload": f"{n} random 2D-periodic gratings (nx, ny in 32..128, nz 768, 30-cell layer eps_r U[1,4] "
                        f"box-filtered, angle U[0,60] deg with Bloch phase, seed 3), DFT at 50 wavelengths",
            "instances": n, "steps": steps, "cells": cells, "fields": "fp32" if precision == 4 else "fp64",
            "sequential_s": round(t_seq, 4), "batched_s": round(t_bat, 4), "speedup": round(t_seq / t_bat, 3),
            "batched_gcells": round(cells * steps / t_bat / 1e9, 3),
            "sequential_gcells": round(cells * steps / t_seq / 1e9, 3),
            "batch_zchunk": info["zchunk"], "two_step_handles": info["two_step"],
            "flux_bitwise_equal": bool(same)}


def measured_traffic(kernel, precision):
    """HBM bytes per launch from the committed rocprofv3 PMC summary, if any."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_fp{32 if precision == 4 else 64}_pmc.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get("hbm_bytes_per_launch")


def cpu_baseline():
    """The reference binary (oracle/_ref/irisfdtd_ref, serial Fortran) on a
    bounded sample of the workload: same z-structure, source and slit, 128 x
    128 lateral cells, stepping only (nof=1 on a 1x1 box, as the reference's
    stepping-only rate is defined), two runs of S1 and S2 steps differenced to
    remove set-up.  Pinned to one core."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from pyiris import inputs
    if not pyoracle.ref_available():
        return None
    lat = 128
    s1, s2 = 3, 15
    kw = dict(WORKLOAD)
    kw.update(ixmax=lat, iymax=lat, box=(1, 1, 1, 1, 25, 1000), nof=1, no_downloads=1)
    times = {}
    base = tempfile.mkdtemp(prefix="iris_cpu_")
    try:
        for n in (s1, s2):
            kw["nt"] = n
            text = inputs.render(**kw)
            t0 = time.perf_counter()
            pyoracle.run_reference(text, os.path.join(base, f"r{n}"), timeout=600, pin_core=0)
            times[n] = time.perf_counter() - t0
    finally:
        shutil.rmtree(base, ignore_errors=True)
    dt = times[s2] - times[s1]
    rate = lat * lat * NZ * (s2 - s1) / dt / 1e9
    cpu = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": rate, "unit": "Gcell-updates/s", "cores": 1, "kind": "reference",
            "sample": f"reference Fortran (amdflang -O2) on {lat}x{lat}x{NZ} of the same z-structure, stepping "
                      f"only (nof=1, 1x1 box), ({s2}-{s1}) steps by differencing two runs "
                      f"({times[s2]:.1f}s - {times[s1]:.1f}s), taskset core 0; host {cpu}, "
                      f"{os.cpu_count()} cpus visible"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--fp64-steps", type=int, default=None)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default layout of the step passes")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="run the gradient workload only and print its JSON object (profiling)")
    ap.add_argument("--adjoint-steps", type=int, default=None,
                    help="steps of the gradient workload (default 12000; fewer for profiling)")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-only", action="store_true",
                    help="run the batched sweep only and print its JSON object")
    ap.add_argument("--batch-steps", type=int, default=BATCH_STEPS)
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this engine runs on one GPU; a run on several GPUs (one z-slab per GPU with "
              "plane exchanges) is not implemented", file=sys.stderr)
        sys.exit(2)

    if a.adjoint_only:
        print(json.dumps(run_adjoint(4, a.adjoint_steps)), flush=True)
        return
    if a.batch_only:
        print(json.dumps(run_batch(4, BATCH_N, a.batch_steps)), flush=True)
        return
    r32 = run_gpu(4, a.steps, a.warmup, not a.no_autotune)
    d32 = r32["kernels"][r32["dominant"]]
    log(f"fp32: {r32['gcells']:.2f} Gcell/s, {r32['ms_per_step']:.3f} ms/step, {r32['dominant']} "
        f"{d32.get('gbs', 0):.0f} GB/s")
    r64 = None
    if not a.no_fp64:
        r64 = run_gpu(8, a.fp64_steps or max(a.steps // 2, 1), a.warmup, not a.no_autotune)
        d64 = r64["kernels"][r64["dominant"]]
        log(f"fp64: {r64['gcells']:.2f} Gcell/s, {r64['ms_per_step']:.3f} ms/step, {r64['dominant']} "
            f"{d64.get('gbs', 0):.0f} GB/s")
    adj = None if a.no_adjoint else run_adjoint(4, a.adjoint_steps)
    bat = None if a.no_batch else run_batch(4, BATCH_N, a.batch_steps)
    cpu = None if a.no_cpu else cpu_baseline()
    if cpu:
        log(f"cpu baseline: {cpu['value'] * 1e3:.2f} Mcell/s")

    k = r32["kernels"]
    dom = r32["dominant"]
    achieved = k[dom]["gbs"]
    traffic = measured_traffic(dom, 4)
    line = {
        "metric": METRIC,
        "value": round(r32["gcells"], 4),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(r32["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "slit-array film 512x512x1024, 20-cell z-UPML, TF/SF Gaussian 400-1000 nm, "
                               "DFT at 100 wavelengths on a full-lateral box, cadence nt/10 (nt=20000)",
                   "grid": [NX, NY, NZ], "nof": 100, "fields": "fp32", "dft_accumulators": "fp64"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes_per_launch": k[dom]["algorithmic_bytes"],
                     "avg_launch_us": round(k[dom]["avg_us"], 2)},
        "cpu_baseline": cpu,
        "kernels_fp32_us": {n: round(v["avg_us"], 2) for n, v in k.items()},
        "kernels_fp32_launches": {n: v["launches"] for n, v in k.items()},
        "yee_ms_per_step_fp32": round(r32["yee_ms_per_step"], 4),
        "passes": r32["passes"],
        "layout": r32["layout"],
        # per-step bytes of the two-pass algorithm (144 B/cell FP32 + UPML +
        # eps) over the measured step time, against 8 TB/s
        "two_pass_bytes_roofline_frac_fp32": round(r32["step_two_pass_frac"], 4),
    }
    if r64:
        line["fp64"] = {"value": round(r64["gcells"], 4), "ms_per_step": round(r64["ms_per_step"], 4),
                        "dominant": r64["dominant"],
                        "dominant_gbs": round(r64["kernels"][r64["dominant"]].get("gbs", 0), 1),
                        "yee_ms_per_step": round(r64["yee_ms_per_step"], 4), "layout": r64["layout"],
                        "two_pass_bytes_roofline_frac": round(r64["step_two_pass_frac"], 4),
                        "steps": a.fp64_steps or max(a.steps // 2, 1),
                        "kernels_us": {n: round(v["avg_us"], 2) for n, v in r64["kernels"].items()}}
    if adj:
        line["adjoint"] = adj
    if bat:
        line["batch"] = bat
    if cpu:
        line["speedup_vs_cpu"] = round(r32["gcells"] / cpu["value"], 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
