#!/usr/bin/env python3
"""Throughput of the MI355X FDTD engine on the roofline-scale workload.

Workload: the slit-array film on a 512 x 512 x 1024 grid (dx = 5 nm),
periodic x/y, 20-cell z-UPML at each end, eps_r = 4.0 film 200 nm (40 cells)
thick cut by one 100 nm (20-cell) air slit along y per x-period, normal-
incidence TF/SF Gaussian pulse covering 400-1000 nm, on-the-fly DFT at 100
wavelengths on a full-lateral box (z = 25..1000), output cadence every nt/10
steps (nt = 20000).  The x-period is the grid's, 512 cells = 2560 nm: the
reference's slit() (code/mesh.f90:57-69) cuts one slit per grid period, and
a 600 nm pitch does not tile 512 cells.  One 600 nm period (120 x 1 x 1024)
is a parity case of the tests instead (reference goldens).

A bench step is one period of the DFT sample ring: P time steps (64 with
FP32 fields, 32 with FP64), in which the ring fills once and folds once, so
every step holds the same work -- E, H, TF/SF and the DFT samples of each
time step, one fold of the buffered samples into the FP64 accumulators, and
the face fluxes when the output cadence falls inside.  The timed window
starts at a ring boundary (W warm-up periods after step 0) and covers
exactly K periods.  value = FP32-field Gcell-updates/s over that window
(cells x time steps / wall seconds, inputs resident in HBM), timed without
per-launch events so that the steps run as the captured HIP graph they
normally take.  A second window of K periods right after it runs with HIP
events around every launch on the engine's stream: the per-kernel times and
the roofline of the dominant kernel come from it.

Also reported: the FP64-field rate; the stepping-only rate (nof = 1 on a
1x1 column box); "adjoint", the gradient workload (128 x 128 x 768 with a
random 128 x 128 x 64 design, plain forward against the gradient of its
summed transmission); "adjoint_angles", the angle-averaged gradient of a
256 x 256 x 1024 voxel metasurface over 16 incidence angles (truncated
steps, one angle at a time); "batch", the 64-grating Bloch sweep (one
batch against the same runs one after another); and "cpu_baseline", the
reference Fortran pinned to one core on bounded samples.

Before the forward runs the engine times candidate layouts of its step
passes (iris_autotune; all give the same results bit for bit) and keeps the
fastest; --no-autotune keeps the default layout.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-fp64] [--no-cpu] [--no-adjoint]
                  [--no-angles] [--no-batch] [--no-autotune]
  python bench.py --adjoint-only [--adjoint-steps N]
  python bench.py --angles-only [--angle-steps N] [--angles A]
  python bench.py --batch-only [--batch-steps N]
  python bench.py --pmc-probe [--fp64] [--pmc-steps N]     workload of the counter runs
  python bench.py --bw-probe                               streaming bandwidth of plain kernels
  python bench.py --report-stats CSV [--command CMD] [--steps-fp32 N] [--steps-fp64 N]
  python bench.py --report-phases TRACE_CSV [--split-at REGEX]   kernel totals before/after a launch
  python bench.py --report-sq PMC_DIR [--kernels REGEX]          mean SQ counters per dispatch
  python bench.py --report-pmc FETCH_DIR WRITE_DIR --pmc-steps N --out-dir DIR
                  [--command-fetch CMD --command-write CMD]

--pmc-probe copies a known number of bytes device to device (the
calibration of the counters) and runs --pmc-steps steps of the workload per
precision.  --report-stats turns the kernel_stats.csv of a rocprofv3
--kernel-trace --stats run into a markdown table; --report-pmc turns two
counter runs of --pmc-probe (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch)
into HBM bytes per launch of each kernel class, each counter scaled by what
it reads for the calibration copy.
"""
import argparse
import csv
import glob
import json
import math
import os
import re
import shutil
import statistics
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
WORKLOAD_TEXT = ("slit-array film 512x512x1024 (dx 5 nm), eps_r 4.0 film 200 nm thick with one 100 nm air slit "
                 "along y per 2560 nm x-period (the grid's), periodic x/y, 20-cell z-UPML, normal-incidence TF/SF "
                 "Gaussian 400-1000 nm, DFT at 100 wavelengths on a full-lateral box, cadence nt/10 (nt=20000)")
# stepping only: the same grid with one wavelength on a 1x1 column box
STEPPING = dict(box=(1, 1, 1, 1, 25, 1000), nof=1, dlambda=1.0, lno=700.0)

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

# angle-averaged gradient workload: a 256 x 256 x 1024 grid whose 160 nm
# (32-cell) thick design layer has eps_r uniform in [1, 6] per 32 x 32 x 32
# voxel (seed 5, 8 x 8 voxels over the 1280 nm period), 16 incidence angles
# uniform in 0..45 degrees (xz plane, matching Bloch phase), objective the
# transmission summed over 20 wavelengths in 450-950 nm, FP32 fields.  The
# run length N = 16000; the default bench times ANGLES_TIMED of the 16 angles
# (every fourth) for ANGLE_STEPS steps (light that crossed the design
# reaches both flux faces from about step 3900), with the H records of each
# angle's last steps limited to the share of its steps that the device
# memory holds at N = 16000, so that the truncated run keeps the full run's
# balance of recorded and recomputed steps; --angles 16 --angle-steps 16000
# runs the workload as specified.
ANGLE_NT = 16000
ANGLE_STEPS = 4500
ANGLES = 16
ANGLES_TIMED = 4
ANGLE_WORKLOAD = dict(description="voxel metasurface 256x256x1024 (angle-averaged gradient)",
                      ixmax=256, iymax=256, izmax=1024, pmlzi=20, pmlzf=1005, hi=496, dh=32, dx=0,
                      eps1=1.0, eps_bbox=2.5, eps3=1.0, tfsf_o=30, wlpump=700.0, pump_dt=4.0, def_pulse=4,
                      box=(1, 256, 1, 256, 25, 1000), nof=20, dlambda=500.0 / 19.0, lno=450.0,
                      nt=ANGLE_NT, no_downloads=ANGLE_NT, to_fourier=0, sample_fourier=1)


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
        write E and H of step n+2 (12 complex per cell of those planes), plus
        eps_inv once per structure cell among them;
      yee2_pml (its variant over the chunks reaching into the UPML): as yee2,
        plus read D and B of step n and write them for step n+2 (12 complex
        per UPML cell);
      yee_rest (one step of the other z-chunks inside a two-step pass): as
        yee over the remaining planes;
      in-place E or H pass: read E and H, write one of them (9 complex per
        cell); per UPML cell read+write D or B (6 complex);
      per time step of the two-step algorithm (step_blocked): half of yee2
        and yee2_pml, plus yee_rest."""
    c = 8 if precision == 4 else 16
    r = c // 2
    ncell = p.nx * p.ny * p.nz
    npml = p.pmlzi + (p.nz - p.pmlzf + 1)
    nxy = p.nx * p.ny
    nstr = nxy * (p.hf - p.hi + 1)
    rest = p.nz - two_step_planes - upml_two_step_planes
    out = {"yee": 12 * c * ncell + 12 * c * nxy * npml + r * nstr,
           "yee2": 12 * c * nxy * two_step_planes + (r * nstr if two_step_planes else 0),
           "yee2_pml": 12 * c * nxy * (upml_two_step_planes + upml_in_two_step),
           "yee_rest": 12 * c * nxy * rest + 12 * c * nxy * (npml - upml_in_two_step),
           "epass": 9 * c * ncell + 6 * c * nxy * npml + r * nstr,
           "hpass": 9 * c * ncell + 6 * c * nxy * npml}
    if two_step_planes:
        out["step_blocked"] = out["yee2"] / 2 + out["yee2_pml"] / 2 + out["yee_rest"]
    else:
        out["step_blocked"] = out["yee"]
    return out


AUTOTUNE_STEPS = 8  # steps timed per candidate layout


def run_gpu(precision, periods, warmup, autotune=True, over=None, profiled=True, label=None):
    """K = `periods` periods of the DFT sample ring after `warmup` of them;
    returns the rates of an unprofiled window and (profiled=True) the
    per-kernel times of a second, profiled window of the same length."""
    from pyiris import Engine
    s = make_setup(precision, **(over or {}))
    p = s.params
    label = label or f"fp{32 if precision == 4 else 64}"
    log(f"{label}: creating engine for {p.nx}x{p.ny}x{p.nz}, nof={p.nof}")
    eng = Engine(s, precision=precision)
    layout = None
    if autotune:
        # layout of the step passes, chosen before the first step (untimed, like the set-up)
        layout = eng.autotune(AUTOTUNE_STEPS)
        log(f"{label}: layout {layout}")
    P = eng.graph_info()["period"]
    nos = [1]
    fluxes = [0]

    def advance(it_from, it_to):
        it = it_from
        while it <= it_to:
            cand = p.to_fourier + p.no_downloads * nos[0]
            stop = min(cand, it_to) if cand >= it else it_to
            eng.step(it, stop)
            if stop >= p.to_fourier and (stop == cand or stop == p.nt):
                eng.flux(1)
                nos[0] += 1
                fluxes[0] += 1
            it = stop + 1

    w_steps, k_steps = warmup * P, periods * P
    if 2 * k_steps + w_steps > p.nt:
        raise SystemExit(f"{label}: {warmup} + 2 x {periods} periods of {P} steps exceed nt = {p.nt}")
    advance(1, w_steps)
    eng.sync()
    log(f"{label}: timing {periods} periods of {P} steps (steps {w_steps + 1}..{w_steps + k_steps})")
    f0 = fluxes[0]
    t0 = time.perf_counter()
    advance(w_steps + 1, w_steps + k_steps)
    eng.sync()
    secs = time.perf_counter() - t0
    g = eng.graph_info()
    out = dict(period=P, periods=periods, time_steps=k_steps, secs=secs,
               gcells=p.nx * p.ny * p.nz * k_steps / secs / 1e9, ms_per_period=1e3 * secs / periods,
               ms_per_time_step=1e3 * secs / k_steps, fluxes_in_window=fluxes[0] - f0,
               graph_replays=g["replays"], layout=layout)
    info = eng.pass_info()
    out["passes"] = info
    if profiled:
        eng.profile(True)
        eng.profile_read(reset=True)
        t0 = time.perf_counter()
        advance(w_steps + k_steps + 1, w_steps + 2 * k_steps)
        eng.sync()
        psecs = time.perf_counter() - t0
        prof = eng.profile_read(reset=True)
        eng.profile(False)
        ab = algorithmic_bytes(p, precision, info["two_step_planes"], info["upml_two_step_planes"],
                               info["upml_planes_in_two_step"])
        kern = {}
        for k, (ms, n) in prof.items():
            if n:
                kern[k] = {"avg_us": 1e3 * ms / n, "launches": n, "total_ms": ms, "algorithmic_bytes": ab.get(k)}
                if ab.get(k):
                    kern[k]["gbs"] = ab[k] / (1e-3 * ms / n) / 1e9
        dom = max(kern, key=lambda k: kern[k]["total_ms"])
        out.update(kernels=kern, dominant=dom, profiled_gcells=p.nx * p.ny * p.nz * k_steps / psecs / 1e9,
                   kernel_ms_per_time_step=sum(v["total_ms"] for v in kern.values()) / k_steps,
                   step_blocked_bytes=ab["step_blocked"])
    eng.close()
    return out


def run_adjoint(precision=4, nt=None):
    """Wall time of the plain forward run and of the gradient (forward run
    recording H of its last steps in the free device memory and keeping
    snapshots every ceil(sqrt(N)) steps before them, then the adjoint sweep
    with segment-wise recomputation) on the gradient workload; both with
    inputs resident on the device.  The gradient is reported with respect to
    eps_r (dT/d eps = -dT/d eps_inv / eps^2)."""
    import numpy as np
    from pyiris import Engine
    from pyiris.adjoint import structure_of, with_structure
    over = {} if nt is None else dict(nt=nt, no_downloads=nt)
    s0 = make_setup(precision, ADJ_WORKLOAD, **over)
    rng = np.random.default_rng(4)
    eps = rng.uniform(1.0, 4.0, size=structure_of(s0).shape)
    s = with_structure(s0, 1.0 / eps)
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
    g_eps = -g.real / eps ** 2
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
           "dT_deps_max_abs": float(np.abs(g_eps).max())}
    log(f"adjoint: forward {t_fwd:.2f} s, gradient {t_grad:.2f} s, ratio {t_grad / t_fwd:.2f} "
        f"({ck['recorded_steps']} steps recorded, {ck['snapshots']} snapshots)")
    return out


def angle_thetas(n=ANGLES):
    """n incidence angles uniform in [0, 45] degrees (seed 5)."""
    import numpy as np
    return sorted(float(t) for t in np.random.default_rng(5).uniform(0.0, 45.0, size=n))


def run_angles(precision=4, steps=ANGLE_STEPS, n=ANGLES_TIMED):
    """The angle-averaged gradient workload (ANGLE_WORKLOAD) for `steps` of
    its N steps over n of its 16 angles (every 16/n-th): per angle a plain
    forward run, then the gradient (forward with snapshots and records,
    adjoint sweep), one angle at a time, each engine freed before the next.
    Returns the summed times, their ratio and the angle-averaged dT/d eps_r
    over those angles."""
    import numpy as np
    import torch
    from pyiris import Engine
    from pyiris.adjoint import flux, structure_of, transmission
    from pyiris.sweep import metasurface, oblique
    s0 = make_setup(precision, ANGLE_WORKLOAD, nt=steps, no_downloads=steps)
    s = metasurface(s0, block=32, eps_range=(1.0, 6.0), seed=5)
    p = s.params
    eps = 1.0 / structure_of(s).real
    thetas = angle_thetas(ANGLES)[::max(1, ANGLES // n)][:n]
    n = len(thetas)
    # the share of the steps of a full-length run whose H records fit beside
    # its snapshots: records take what the engine's automatic budget gives
    with Engine(oblique(make_setup(precision, ANGLE_WORKLOAD), thetas[0]), precision=precision) as eng:
        eng.checkpoint(-1, -1)
        full = eng.checkpoint_info()
    share = full["recorded_steps"] / ANGLE_NT
    rec_bytes = int(share * steps) * full["record_step_bytes"]
    log(f"angles: {n} angles, {p.nx}x{p.ny}x{p.nz}, {steps} of {ANGLE_NT} steps; at full length "
        f"{full['recorded_steps']} steps recorded ({share:.2f}), segment {full['segment']}")
    with Engine(oblique(s, thetas[0]), precision=precision) as eng:   # warm-up
        eng.run_to(20)
        eng.sync()
    # both sides include each engine's creation (the gradient's is inside flux())
    t_fwd = 0.0
    j_first = None
    for th in thetas:
        t0 = time.perf_counter()
        with Engine(oblique(s, th), precision=precision) as eng:
            eng.run_to(p.nt)
            j = eng.flux(1)
        t_fwd += time.perf_counter() - t0
        if j_first is None:
            j_first = j
    # the first angle's forward T/R against the same run with FP64 fields
    # (whose T/R match the reference's FP64 program to 1e-10 on the goldens),
    # on the wavelengths whose flux has arrived (J5+J6 above 1e-3 of its largest)
    tr_err = None
    if precision == 4:
        with Engine(oblique(s, thetas[0]), precision=8) as eng:
            eng.run_to(p.nt)
            j64 = eng.flux(1)
        den64 = j64[:, 4] + j64[:, 5]
        live = np.abs(den64) > 1e-3 * np.abs(den64).max()
        den32 = j_first[:, 4] + j_first[:, 5]
        dR = np.abs(j_first[:, 4] / den32 - j64[:, 4] / den64)[live]
        dT = np.abs(j_first[:, 5] / den32 - j64[:, 5] / den64)[live]
        tr_err = {"angle_deg": round(float(thetas[0]), 3), "wavelengths": int(live.sum()),
                  "max_abs_dR": float(dR.max()) if live.any() else None,
                  "max_abs_dT": float(dT.max()) if live.any() else None}
        if live.any():
            log(f"angles: FP32 against FP64 T/R at {thetas[0]:.1f} deg: max |dR| {tr_err['max_abs_dR']:.2e}, "
                f"max |dT| {tr_err['max_abs_dT']:.2e} over {tr_err['wavelengths']} wavelengths")
    leaf = torch.tensor(eps, requires_grad=True)
    t_grad, total, ck = 0.0, 0.0, None
    for th in thetas:
        t0 = time.perf_counter()
        obj = transmission(flux(1.0 / leaf, oblique(s, th), precision, 1, None, rec_bytes)).sum() / n
        obj.backward()
        t_grad += time.perf_counter() - t0
        total += float(obj.detach())
        del obj
    g = leaf.grad.numpy()
    out = {"workload": "voxel metasurface 256x256x1024, 32-cell (160 nm) design layer, eps_r U[1,6] per 32^3 voxel "
                       "(seed 5), 16 angles U[0,45] deg (seed 5) with Bloch phase, objective mean over angles of "
                       "sum T over 20 wavelengths 450-950 nm",
           "grid": [p.nx, p.ny, p.nz], "angles": n, "angles_drawn": ANGLES,
           "thetas_deg": [round(t, 3) for t in thetas], "steps": steps, "full_length_steps": ANGLE_NT,
           "recorded_share": round(share, 4), "record_bytes_per_angle": rec_bytes,
           "fields": "fp32" if precision == 4 else "fp64",
           "forward_s": round(t_fwd, 4), "gradient_s": round(t_grad, 4), "ratio": round(t_grad / t_fwd, 3),
           "objective": total, "dT_deps_max_abs": float(np.abs(g).max()),
           "fp32_vs_fp64_tr": tr_err}
    log(f"angles: forwards {t_fwd:.2f} s, gradients {t_grad:.2f} s, ratio {t_grad / t_fwd:.2f}")
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
    # the batch takes the UPML slabs into the two-step pass, which with a
    # Bloch phase rounds differently from the single steps a small grid
    # alone gives them: the largest difference of any face flux, relative to
    # the largest flux of its run
    rel = max(float(np.abs(a - c).max() / np.abs(a).max()) for a, c in zip(f_seq, f_bat))
    log(f"batch: one after another {t_seq:.2f} s, batched {t_bat:.2f} s, speed-up {t_seq / t_bat:.2f}, "
        f"bitwise equal {same}, largest flux difference {rel:.2e} of the run's largest flux")
    return {"workload": f"{n} random 2D-periodic gratings (nx, ny in 32..128, nz 768, 30-cell layer eps_r U[1,4] "
                        f"box-filtered, angle U[0,60] deg with Bloch phase, seed 3), DFT at 50 wavelengths",
            "instances": n, "steps": steps, "cells": cells, "fields": "fp32" if precision == 4 else "fp64",
            "sequential_s": round(t_seq, 4), "batched_s": round(t_bat, 4), "speedup": round(t_seq / t_bat, 3),
            "batched_gcells": round(cells * steps / t_bat / 1e9, 3),
            "sequential_gcells": round(cells * steps / t_seq / 1e9, 3),
            "batch_zchunk": info["zchunk"], "two_step_handles": info["two_step"],
            "flux_bitwise_equal": bool(same), "flux_max_rel_diff": rel,
            "layout": "the batch's own: 64-plane two-step z-chunks, UPML slabs in the two-step pass (one after "
                      "another: each grid's default layout)"}


def measured_traffic(kernel, precision):
    """HBM bytes per launch from the committed rocprofv3 PMC summary, if any."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_fp{32 if precision == 4 else 64}_pmc.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get("hbm_bytes_per_launch")


# ---- CPU baseline: the reference binary ------------------------------------

def face_points(box, nx, ny, nz):
    """Point-components of a full 3D Fourier box (all three face pairs, six
    components each; reference code/driver.f90:208-219)."""
    sxi, sxf, syi, syf, szi, szf = box
    ax, ay, az = sxf - sxi + 1, syf - syi + 1, szf - szi + 1
    return 2 * 6 * (ay * az + ax * az + ax * ay)


def _ref_runs(kw, counts, pin_core=0, jobs=1):
    """Wall seconds of reference runs of each step count in `counts` (jobs >
    1: that many identical runs at once on cores 0..jobs-1; the slowest
    counts)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from pyiris import inputs
    base = tempfile.mkdtemp(prefix="iris_cpu_")
    times = {}
    try:
        for n in counts:
            k = dict(kw)
            k["nt"] = n
            text = inputs.render(**k)
            if jobs == 1:
                t0 = time.perf_counter()
                pyoracle.run_reference(text, os.path.join(base, f"r{n}"), timeout=600, pin_core=pin_core)
                times[n] = time.perf_counter() - t0
                continue
            procs = []
            t0 = time.perf_counter()
            for j in range(jobs):
                wd = os.path.join(base, f"r{n}_{j}")
                os.makedirs(wd)
                with open(os.path.join(wd, "inputFDTD.dat"), "w") as f:
                    f.write(text)
                cmd = [pyoracle.REF_BIN]
                if shutil.which("taskset"):
                    cmd = ["taskset", "-c", str(j)] + cmd
                procs.append(subprocess.Popen(cmd, cwd=wd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL))
            for pr in procs:
                if pr.wait(timeout=600) != 0:
                    raise RuntimeError("reference run failed")
            times[n] = time.perf_counter() - t0
    finally:
        shutil.rmtree(base, ignore_errors=True)
    return times


def cpu_baseline(reps=3):
    """The reference binary (oracle/_ref/irisfdtd_ref, serial Fortran built
    with amdflang -O2) pinned to one core, on bounded samples of the
    workload (same z-structure, source and slit), each rate the median of
    `reps` differences of a short and a long run (set-up removed):
      stepping only: 128 x 128 x 1024, nof = 1 on a 1x1 box, 2 and 10 steps;
      full loop: 16 x 16 x 1024 with the configured DFT (100 wavelengths on
        the full-lateral box), 1 and 2 steps; its DFT time per step (full
        minus stepping) over its point-components x wavelengths gives the
        reference's DFT rate;
      full loop at the workload: 512 x 512 x 1024 time per step composed as
        cells / stepping rate + point-components x 100 / DFT rate (the
        reference's DFT is a loop over points and wavelengths, linear in
        both; the full grid needs ~80 GB and ~30 s per step);
      aggregate: P concurrent stepping-only runs on cores 0..P-1 (2 and 14
        steps), P = min(cores, 16, free memory / 4 GB), labelled as such.
    One untimed run first pages in the binary and its libraries."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    if not pyoracle.ref_available():
        return None
    kw = dict(WORKLOAD)
    lat = 128
    step_kw = dict(kw, ixmax=lat, iymax=lat, **STEPPING, no_downloads=1)
    s1, s2 = 2, 10
    _ref_runs(step_kw, (1,))
    step_rates, step_secs = [], []
    for _ in range(reps):
        t = _ref_runs(step_kw, (s1, s2))
        step_secs.append((t[s1], t[s2]))
        step_rates.append(lat * lat * NZ * (s2 - s1) / (t[s2] - t[s1]))
    rate_step = statistics.median(step_rates)
    flat = 16
    full_kw = dict(kw, ixmax=flat, iymax=flat, box=(1, flat, 1, flat, 25, 1000), no_downloads=1)
    f1, f2 = 1, 2
    full_rates, full_secs = [], []
    for _ in range(reps):
        t = _ref_runs(full_kw, (f1, f2))
        full_secs.append((t[f1], t[f2]))
        full_rates.append(flat * flat * NZ * (f2 - f1) / (t[f2] - t[f1]))
    rate_full_sample = statistics.median(full_rates)
    cells_s = flat * flat * NZ
    t_dft = cells_s / rate_full_sample - cells_s / rate_step
    pcs = face_points(full_kw["box"], flat, flat, NZ)
    dft_rate = pcs * kw["nof"] / t_dft
    cells = NX * NY * NZ
    t_w = cells / rate_step + face_points(WORKLOAD["box"], NX, NY, NZ) * kw["nof"] / dft_rate
    rate_full = cells / t_w
    ncpu = os.cpu_count() or 1
    avail = None
    try:
        with open("/proc/meminfo") as f:
            for line in f:
                if line.startswith("MemAvailable"):
                    avail = int(line.split()[1]) * 1024
    except OSError:
        pass
    jobs = min(ncpu, 16, max(1, int((avail or 8 << 30) // (4 << 30))))
    agg = None
    if jobs > 1:
        a1, a2 = 2, 14
        t = _ref_runs(step_kw, (a1, a2), jobs=jobs)
        agg = {"value": jobs * lat * lat * NZ * (a2 - a1) / (t[a2] - t[a1]) / 1e9, "unit": "Gcell-updates/s",
               "processes": jobs, "kind": "aggregate of independent concurrent stepping-only runs, not one run",
               "sample": f"{jobs} copies of the stepping-only sample at once, cores 0..{jobs - 1}, "
                         f"({a2}-{a1}) steps differenced ({t[a2]:.1f}s - {t[a1]:.1f}s)"}
    cpu = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": rate_full / 1e9, "unit": "Gcell-updates/s", "cores": 1, "kind": "reference",
            "sample": f"full loop of the 512x512x1024 workload composed from two measured rates of the reference "
                      f"Fortran (amdflang -O2, taskset core 0, medians of {reps}): stepping "
                      f"{rate_step / 1e9:.4f} Gcell/s on {lat}x{lat}x{NZ} with nof=1 on a 1x1 box "
                      f"(({s2}-{s1}) steps differenced), DFT {dft_rate / 1e6:.1f} M point-wavelength updates/s "
                      f"from {flat}x{flat}x{NZ} with the configured DFT (({f2}-{f1}) steps differenced); host "
                      f"{cpu}, {ncpu} cpus visible, 1 used",
            "stepping_only": {"value": rate_step / 1e9, "runs": [round(r / 1e9, 5) for r in step_rates],
                              "grid": [lat, lat, NZ], "steps": [s1, s2],
                              "secs": [[round(a, 2), round(b, 2)] for a, b in step_secs]},
            "full_loop_sample": {"value": rate_full_sample / 1e9,
                                 "runs": [round(r / 1e9, 6) for r in full_rates], "grid": [flat, flat, NZ],
                                 "steps": [f1, f2], "point_components": pcs, "nof": kw["nof"],
                                 "secs": [[round(a, 2), round(b, 2)] for a, b in full_secs]},
            "dft_point_wavelength_per_s": dft_rate,
            "full_loop_workload_s_per_step": round(t_w, 3),
            "aggregate": agg}


# ---- counter runs and profile summaries --------------------------------------

COPY_BYTES = 1 << 31

CLASSES = [
    ("yee2_batch", r"k_yee2_batch"),
    ("yee2_pml", r"k_yee2<\w+, \w+, true\b"),
    ("yee2", r"k_yee2<|k_yee2p<"),
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


def bw_probe(gib=3, reps=5):
    """Device-memory bandwidth a plain streaming kernel reaches on this part:
    torch's elementwise kernels on `gib`-GiB buffers (no engine in this
    process), best of `reps` of each: read only (a sum), copy (one read, one
    write stream), and six read and six write streams at once (as the Yee
    passes have).  The ceiling the Yee kernels are held against in practice;
    the roofline fraction is taken against the 8 TB/s peak regardless."""
    import torch
    n = (gib << 30) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    six_a = [t for t in a.view(6, -1)]
    six_b = [t for t in b.view(6, -1)]
    def timed(fn, moved):
        best = None
        for _ in range(reps + 1):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1)
            best = ms if best is None or ms < best else best
        return round(moved / (best * 1e-3) / 1e9, 1)
    out = {"buffer_bytes": n * 4,
           "read_gbs": timed(lambda: a.sum(), n * 4),
           "copy_gbs": timed(lambda: b.copy_(a), 2 * n * 4),
           "six_streams_gbs": timed(lambda: [y.copy_(x) for x, y in zip(six_a, six_b)], 2 * n * 4),
           "note": "best of %d; six streams: six copies of a sixth each, one after another" % reps}
    del a, b
    return out


def pmc_probe(fp64, steps):
    """Workload of the counter runs: the calibration copy (COPY_BYTES device
    to device with hipMemcpy, which runs the runtime's copy kernel, three
    times), then `steps` steps of the workload with FP32 fields (and with
    FP64 fields when fp64)."""
    import ctypes as ct
    from pyiris import Engine
    hip = ct.CDLL("libamdhip64.so")
    a, b = ct.c_void_p(), ct.c_void_p()
    assert hip.hipMalloc(ct.byref(a), ct.c_size_t(COPY_BYTES)) == 0
    assert hip.hipMalloc(ct.byref(b), ct.c_size_t(COPY_BYTES)) == 0
    assert hip.hipMemset(a, 1, ct.c_size_t(COPY_BYTES)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, ct.c_size_t(COPY_BYTES), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"pmc probe: copied {COPY_BYTES} bytes x3")
    for prec in ([4, 8] if fp64 else [4]):
        s = make_setup(prec)
        with Engine(s, precision=prec) as eng:
            eng.run_to(steps)
            eng.sync()
        log(f"pmc probe: {steps} steps precision {prec}")


def kernel_class(name):
    for cls, pat in CLASSES:
        if re.search(pat, name):
            return cls
    return None


def precision_of(name):
    """Field precision from the kernel's template arguments (None for the
    kernels that have none, such as the phase table; k_yee2p holds FP32
    fields only, and k_dft_fold_mfma and k_dft_fold_mfma_p fold FP32 samples)."""
    tmpl = name.split("(")[0]
    if "<float" in tmpl or "k_yee2p<" in tmpl or "k_dft_fold_mfma<" in tmpl or "k_dft_fold_mfma_p<" in tmpl:
        return 4
    if "<double" in tmpl:
        return 8
    return None


def short(name):
    name = re.sub(r"\(.*$", "", name)
    return name.replace("void ", "").replace("iris::", "")


def report_stats(path, command, steps32, steps64):
    rows = list(csv.DictReader(open(path)))
    out = []
    if command:
        out.append(f"Command: `{command}`\n")
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
            per_class[(cls, prec)] = per_class.get((cls, prec), 0.0) + tot
    nsteps = {k: v for k, v in {4: steps32, 8: steps64}.items() if v}
    if nsteps:
        out += ["", "Per step (all timed and warm-up steps of the run):", "",
                "| class | fields | steps | ms per step |", "|---|---|---:|---:|"]
        for (cls, prec), tot in sorted(per_class.items(), key=lambda kv: (kv[0][1], -kv[1])):
            if prec in nsteps:
                out.append(f"| {cls} | fp{32 if prec == 4 else 64} | {nsteps[prec]} | {tot / nsteps[prec]:.3f} |")
    print("\n".join(out))


def _read_counter(d, counter):
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no counter_collection.csv under {d}")
    rows = []
    for f in files:
        rows += [r for r in csv.DictReader(open(f)) if r["Counter_Name"] == counter]
    return rows


def report_phases(path, split_at):
    """Per-kernel totals of a rocprofv3 kernel trace (kernel_trace.csv), in
    two phases split at the first launch whose kernel name matches the
    regular expression split_at (kernel end minus start, summed)."""
    rows = sorted(csv.DictReader(open(path)), key=lambda r: int(r["Start_Timestamp"]))
    starts = [int(r["Start_Timestamp"]) for r in rows if re.search(split_at, r["Kernel_Name"])]
    if not starts:
        raise SystemExit(f"no kernel matches {split_at!r}")
    t0 = starts[0]
    phases = [{}, {}]
    for r in rows:
        k = 1 if int(r["Start_Timestamp"]) >= t0 else 0
        e = phases[k].setdefault(short(r["Kernel_Name"]), [0, 0.0])
        e[0] += 1
        e[1] += (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6
    out = []
    for k, title in ((0, f"Before the first `{split_at}` launch"), (1, f"From the first `{split_at}` launch")):
        tot = sum(v[1] for v in phases[k].values())
        out += [f"## {title} (kernel time {tot:.0f} ms)", "", "| kernel | calls | total (ms) |", "|---|---:|---:|"]
        for n, v in sorted(phases[k].items(), key=lambda kv: -kv[1][1]):
            out.append(f"| `{n}` | {v[0]} | {v[1]:.1f} |")
        out.append("")
    print("\n".join(out))


SQ_COUNTERS = ("SQ_WAVES", "SQ_WAVE_CYCLES", "SQ_WAIT_ANY", "SQ_WAIT_INST_ANY", "SQ_ACTIVE_INST_VALU",
               "SQ_INSTS_VALU", "SQ_ACTIVE_INST_LDS")


def report_sq(d, pattern):
    """Mean SQ counters per dispatch of the kernels whose name matches the
    regular expression `pattern`, from a rocprofv3 --pmc run, with the
    wait and VALU fractions of the waves' cycles."""
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no counter_collection.csv under {d}")
    acc = {}
    for f in files:
        for r in csv.DictReader(open(f)):
            if re.search(pattern, r["Kernel_Name"]) and r["Counter_Name"] in SQ_COUNTERS:
                acc.setdefault(short(r["Kernel_Name"]), {}).setdefault(r["Counter_Name"], []).append(
                    float(r["Counter_Value"]))
    out = ["| kernel | " + " | ".join(SQ_COUNTERS) + " |", "|---|" + "---:|" * len(SQ_COUNTERS)]
    ratios = []
    for n, c in sorted(acc.items()):
        m = {k: sum(v) / len(v) for k, v in c.items()}
        out.append(f"| `{n}` | " + " | ".join(f"{m.get(k, 0):.0f}" for k in SQ_COUNTERS) + " |")
        if m.get("SQ_WAVE_CYCLES") and m.get("SQ_WAVES"):
            wc = m["SQ_WAVE_CYCLES"]
            ratios.append(f"`{n}`: waves wait (any) {100 * m.get('SQ_WAIT_ANY', 0) / wc:.1f}% of their cycles, "
                          f"wait to issue {100 * m.get('SQ_WAIT_INST_ANY', 0) / wc:.1f}%, VALU active "
                          f"{100 * m.get('SQ_ACTIVE_INST_VALU', 0) / wc:.1f}%, "
                          f"{m.get('SQ_INSTS_VALU', 0) / m['SQ_WAVES']:.0f} VALU instructions a wave")
    print("\n".join(out + [""] + ratios))


def report_pmc(fetch_dir, write_dir, steps, out_dir, cmd_fetch, cmd_write):
    # launches of a class: the dispatches of its most frequent kernel (a
    # class whose launch is several kernels counts the most frequent one)
    res, per_kernel = {}, {}
    for counter, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        rows = _read_counter(d, counter)
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
                k = per_kernel.setdefault((cls, prec), {})
                k[r["Kernel_Name"]] = k.get(r["Kernel_Name"], 0) + 1
        res[counter] = {"copy_kib": ref_kib, "scale": scale, "bytes": sums}
    nlaunch = {k: max(v.values()) for k, v in per_kernel.items()}
    os.makedirs(out_dir, exist_ok=True)
    md = ["Counter runs of `python3 bench.py --pmc-probe` (512x512x1024 slit-array workload,",
          f"{steps} steps per precision).", ""]
    if cmd_fetch:
        md += [f"FETCH_SIZE: `{cmd_fetch}`", ""]
    if cmd_write:
        md += [f"WRITE_SIZE: `{cmd_write}`", ""]
    f, w = res["FETCH_SIZE"], res["WRITE_SIZE"]
    md += [f"Calibration ({COPY_BYTES} B device-to-device copy): FETCH_SIZE reads {f['copy_kib']:.0f} KiB "
           f"(scale x{f['scale']:.3f}), WRITE_SIZE reads {w['copy_kib']:.0f} KiB (scale x{w['scale']:.3f}).", "",
           "| class | fields | launches | fetch GB/launch | write GB/launch | HBM GB/launch | HBM GB/step |",
           "|---|---|---:|---:|---:|---:|---:|"]
    for cls, prec in sorted(set(f["bytes"]) | set(w["bytes"]), key=lambda k: (k[1], k[0])):
        n = max(nlaunch.get((cls, prec), 1), 1)
        fb = f["bytes"].get((cls, prec), 0.0)
        wb = w["bytes"].get((cls, prec), 0.0)
        md.append(f"| {cls} | fp{32 if prec == 4 else 64} | {n} | {fb / n / 1e9:.3f} | {wb / n / 1e9:.3f} | "
                  f"{(fb + wb) / n / 1e9:.3f} | {(fb + wb) / steps / 1e9:.3f} |")
        if cls == "yee2":
            name = f"{cls}_fp{32 if prec == 4 else 64}_pmc.json"
            with open(os.path.join(out_dir, name), "w") as fo:
                json.dump({"kernel": "yee2 (two-step pass over the z-chunks between the UPML slabs and away from "
                                     "the TF/SF planes; per launch = two steps)",
                           "workload": "512x512x1024 slit-array film, nof=100",
                           "fetch_bytes_per_launch": round(fb / n), "write_bytes_per_launch": round(wb / n),
                           "hbm_bytes_per_launch": round((fb + wb) / n),
                           "fetch_scale": f["scale"], "write_scale": w["scale"],
                           "commands": [cmd_fetch, cmd_write]}, fo, indent=1)
    with open(os.path.join(out_dir, "hbm_traffic_512x512x1024.md"), "w") as fo:
        fo.write("\n".join(md) + "\n")
    print("\n".join(md))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6, help="timed bench steps (periods of the DFT sample ring)")
    ap.add_argument("--warmup", type=int, default=2, help="warm-up bench steps (periods)")
    ap.add_argument("--fp64-steps", type=int, default=None)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-stepping", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default layout of the step passes")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="run the gradient workload only and print its JSON object (profiling)")
    ap.add_argument("--adjoint-steps", type=int, default=None,
                    help="steps of the gradient workload (default 12000; fewer for profiling)")
    ap.add_argument("--no-angles", action="store_true")
    ap.add_argument("--angles-only", action="store_true")
    ap.add_argument("--angle-steps", type=int, default=ANGLE_STEPS)
    ap.add_argument("--angles", type=int, default=ANGLES_TIMED, help="angles timed, of the 16 (every 16/A-th)")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-only", action="store_true",
                    help="run the batched sweep only and print its JSON object")
    ap.add_argument("--batch-steps", type=int, default=BATCH_STEPS)
    ap.add_argument("--cpu-only", action="store_true")
    ap.add_argument("--pmc-probe", action="store_true")
    ap.add_argument("--bw-probe", action="store_true", help="device-memory bandwidth of torch's streaming kernels")
    ap.add_argument("--fp64", action="store_true")
    ap.add_argument("--pmc-steps", type=int, default=40)
    ap.add_argument("--report-stats", metavar="CSV")
    ap.add_argument("--command", default="")
    ap.add_argument("--steps-fp32", type=int, default=0)
    ap.add_argument("--steps-fp64", type=int, default=0)
    ap.add_argument("--report-pmc", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"))
    ap.add_argument("--out-dir", default=None)
    ap.add_argument("--command-fetch", default="")
    ap.add_argument("--command-write", default="")
    ap.add_argument("--report-phases", metavar="TRACE_CSV")
    ap.add_argument("--split-at", default=r"_batch<.*|k_yee2p_batch")
    ap.add_argument("--report-sq", metavar="PMC_DIR")
    ap.add_argument("--kernels", default=r"k_yee2p<|k_yee2<float, false, true")
    a = ap.parse_args()
    if a.report_phases:
        report_phases(a.report_phases, a.split_at)
        return
    if a.report_sq:
        report_sq(a.report_sq, a.kernels)
        return
    if a.report_stats:
        report_stats(a.report_stats, a.command, a.steps_fp32, a.steps_fp64)
        return
    if a.report_pmc:
        report_pmc(a.report_pmc[0], a.report_pmc[1], a.pmc_steps, a.out_dir or ".", a.command_fetch,
                   a.command_write)
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this engine runs on one GPU; a run on several GPUs (one z-slab per GPU with "
              "plane exchanges) is not implemented", file=sys.stderr)
        sys.exit(2)
    if a.pmc_probe:
        pmc_probe(a.fp64, a.pmc_steps)
        return
    if a.bw_probe:
        print(json.dumps(bw_probe()), flush=True)
        return
    if a.cpu_only:
        print(json.dumps(cpu_baseline()), flush=True)
        return
    if a.adjoint_only:
        print(json.dumps(run_adjoint(4, a.adjoint_steps)), flush=True)
        return
    if a.angles_only:
        print(json.dumps(run_angles(4, a.angle_steps, a.angles)), flush=True)
        return
    if a.batch_only:
        print(json.dumps(run_batch(4, BATCH_N, a.batch_steps)), flush=True)
        return
    r32 = run_gpu(4, a.steps, a.warmup, not a.no_autotune)
    d32 = r32["kernels"][r32["dominant"]]
    log(f"fp32: {r32['gcells']:.2f} Gcell/s ({r32['profiled_gcells']:.2f} with per-launch events), "
        f"{r32['ms_per_time_step']:.3f} ms per time step, {r32['dominant']} {d32.get('gbs', 0):.0f} GB/s")
    r64 = None
    if not a.no_fp64:
        r64 = run_gpu(8, a.fp64_steps or max(a.steps // 2, 1), a.warmup, not a.no_autotune)
        log(f"fp64: {r64['gcells']:.2f} Gcell/s, {r64['ms_per_time_step']:.3f} ms per time step")
    rst = None
    if not a.no_stepping:
        rst = run_gpu(4, a.steps, a.warmup, not a.no_autotune, over=STEPPING, profiled=False, label="fp32 stepping")
        log(f"fp32 stepping only: {rst['gcells']:.2f} Gcell/s")
    adj = None if a.no_adjoint else run_adjoint(4, a.adjoint_steps)
    ang = None if a.no_angles else run_angles(4, a.angle_steps, a.angles)
    bat = None if a.no_batch else run_batch(4, BATCH_N, a.batch_steps)
    cpu = None if a.no_cpu else cpu_baseline()
    if cpu:
        log(f"cpu baseline: full loop {cpu['value'] * 1e3:.3f} Mcell/s, stepping only "
            f"{cpu['stepping_only']['value'] * 1e3:.2f} Mcell/s")

    k = r32["kernels"]
    dom = r32["dominant"]
    achieved = k[dom]["gbs"]
    traffic = measured_traffic(dom, 4)
    P = r32["period"]
    line = {
        "metric": METRIC,
        "value": round(r32["gcells"], 4),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(r32["ms_per_period"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "step": f"one period of the DFT sample ring = {P} time steps (FP32 fields): E, H, TF/SF and DFT samples of "
                f"each, one fold into the FP64 accumulators, fluxes at the output cadence; the window starts at a "
                f"ring boundary",
        "ms_per_time_step": round(r32["ms_per_time_step"], 5),
        "time_steps_timed": r32["time_steps"],
        "fluxes_in_window": r32["fluxes_in_window"],
        "graph_replays": r32["graph_replays"],
        "config": {"workload": WORKLOAD_TEXT, "grid": [NX, NY, NZ], "nof": 100, "fields": "fp32",
                   "dft_accumulators": "fp64"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes_per_launch": k[dom]["algorithmic_bytes"],
                     "avg_launch_us": round(k[dom]["avg_us"], 2),
                     "window": "the profiled window (HIP events around every launch on the engine's stream)"},
        "profiled_window_value": round(r32["profiled_gcells"], 4),
        # compulsory bytes of one time step of the two-step algorithm (half a
        # two-step launch's plus the single-step chunks') over the measured
        # time per step, against 8 TB/s
        "step_roofline_frac_fp32": round(r32["step_blocked_bytes"] / (r32["ms_per_time_step"] * 1e-3) / 1e9
                                         / HBM_PEAK_GBS, 4),
        "cpu_baseline": cpu,
        "kernels_fp32_us": {n: round(v["avg_us"], 2) for n, v in k.items()},
        "kernels_fp32_launches": {n: v["launches"] for n, v in k.items()},
        "kernel_ms_per_time_step_fp32": round(r32["kernel_ms_per_time_step"], 4),
        "passes": r32["passes"],
        "layout": r32["layout"],
    }
    if r64:
        k64 = r64["kernels"]
        line["fp64"] = {"value": round(r64["gcells"], 4), "ms_per_step": round(r64["ms_per_period"], 4),
                        "period": r64["period"], "steps": r64["periods"],
                        "ms_per_time_step": round(r64["ms_per_time_step"], 5),
                        "profiled_window_value": round(r64["profiled_gcells"], 4),
                        "dominant": r64["dominant"],
                        "dominant_gbs": round(k64[r64["dominant"]].get("gbs", 0), 1),
                        "dominant_frac": round(k64[r64["dominant"]].get("gbs", 0) / HBM_PEAK_GBS, 4),
                        "step_roofline_frac": round(r64["step_blocked_bytes"] / (r64["ms_per_time_step"] * 1e-3)
                                                    / 1e9 / HBM_PEAK_GBS, 4),
                        "layout": r64["layout"],
                        "kernels_us": {n: round(v["avg_us"], 2) for n, v in k64.items()}}
    if rst:
        line["stepping_only"] = {"value": round(rst["gcells"], 4), "ms_per_time_step": round(rst["ms_per_time_step"], 5),
                                 "what": "the same grid and steps with nof=1 on a 1x1 column box"}
    if adj:
        line["adjoint"] = adj
    if ang:
        line["adjoint_angles"] = ang
    if bat:
        line["batch"] = bat
    if cpu:
        line["speedup_vs_cpu"] = round(r32["gcells"] / cpu["value"], 1)
        if rst:
            line["speedup_vs_cpu_stepping_only"] = round(rst["gcells"] / cpu["stepping_only"]["value"], 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
