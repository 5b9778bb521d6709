"""Sweeps over incidence angles and structures: setups for batched runs.

oblique(setup, theta_deg) tilts the TF/SF plane wave in the xz plane the way
the reference's source does for RATIO_ANGLE /= 1 (code/tfsf.f90:280-284:
DELTA = RATIO_ANGLE*Pi/2, u = (cos DELTA, 0, sin DELTA)), with
RATIO_ANGLE = 1 - 2*theta/pi, and sets the Bloch phase across the x period
that matches it, akx = ko * sqrt(eps(1)) * sin(theta) * nx * Q at the pulse's
centre wavelength (ko of code/tfsf.f90:269).  The reference hard-wires
RATIO_ANGLE = 1 and akx = 0 (code/files.f90:268, code/parameters.f90:85);
the engine takes both from iris_params.

grating_sweep() builds the random 2D-periodic gratings of a batched Bloch
sweep: per instance a grid of nx, ny cells drawn from {32..128}, a design
layer whose eps_r per (x, y) column is drawn uniform in [1, 4] and smoothed by
a 3-cell periodic box filter in x and y, and an incidence angle drawn uniform
in [0, 60] degrees.

metasurface() and angle_averaged_gradient() are the pieces of an
angle-averaged gradient: a structure whose eps_r is drawn per voxel block,
and the mean over incidence angles of the summed transmission with its
gradient with respect to the structure's eps_inv or eps_r (one forward run
and one adjoint per angle, one angle at a time).
"""
import math
import os
import tempfile

import numpy as np

from . import inputs
from .host import Setup, load_input


def oblique(setup, theta_deg):
    """Copy of `setup` with the plane wave at theta degrees from the z axis
    (in the xz plane) and the matching Bloch phase in x."""
    p = setup.params.copy()
    theta = math.radians(theta_deg)
    ratio = 1.0 - 2.0 * theta / math.pi
    delta = ratio * p.pi / 2.0
    p.ratio_angle = ratio
    p.ux = math.cos(delta)
    p.uz = math.sin(delta)
    ko = (2.0 * p.pi) / (p.wlpump * p.f_abohr)
    p.ak[0][0] = ko * math.sqrt(p.eps[0]) * math.sin(theta) * p.nx * p.q
    p.ak[0][1] = 0.0
    return Setup(p, setup.eps_inv, setup.fpzmn, setup.fmzmn, setup.fpzmx, setup.fmzmx, setup.fw, setup.spectra,
                 dict(setup.extra))


def box_filter3(a):
    """3-cell periodic box average along both axes of a 2D array."""
    a = (np.roll(a, 1, 0) + a + np.roll(a, -1, 0)) / 3.0
    return (np.roll(a, 1, 1) + a + np.roll(a, -1, 1)) / 3.0


def grating_sweep(n=64, seed=3, nz=768, nt=10000, nof=50, layer=30, precision=4, sizes=(32, 128),
                  angles=(0.0, 60.0)):
    """n setups of the batched Bloch sweep (see the module docstring); also
    returns the per-instance (nx, ny, theta) drawn."""
    rng = np.random.default_rng(seed)
    pmlzi, pmlzf = 20, nz - 19
    hi = nz // 2 - layer // 2
    out, drawn = [], []
    with tempfile.TemporaryDirectory() as d:
        for i in range(n):
            nx = int(rng.integers(sizes[0], sizes[1] + 1))
            ny = int(rng.integers(sizes[0], sizes[1] + 1))
            theta = float(rng.uniform(*angles))
            eps = box_filter3(rng.uniform(1.0, 4.0, size=(ny, nx)))
            text = inputs.render(description=f"grating {i}", ixmax=nx, iymax=ny, izmax=nz, pmlzi=pmlzi, pmlzf=pmlzf,
                                 hi=hi, dh=layer, dx=0, eps1=1.0, eps_bbox=2.5, eps3=1.0, tfsf_o=30, wlpump=700.0,
                                 pump_dt=4.0, def_pulse=4, box=(1, nx, 1, ny, 25, nz - 24), nof=nof,
                                 dlambda=600.0 / max(nof - 1, 1), lno=400.0, nt=nt, no_downloads=nt, to_fourier=0,
                                 sample_fourier=1)
            path = os.path.join(d, f"in{i}.dat")
            with open(path, "w") as f:
                f.write(text)
            s = load_input(path, precision=precision)
            p = s.params
            e = np.array(s.eps_inv, dtype=np.complex128, copy=True)
            e.reshape(p.hf - p.hi + 3, p.ny + 2, p.nx + 2)[1:-1, 1:-1, 1:-1] = (1.0 / eps)[None, :, :]
            s = Setup(p.copy(), e, s.fpzmn, s.fmzmn, s.fpzmx, s.fmzmx, s.fw, s.spectra, dict(s.extra))
            out.append(oblique(s, theta))
            drawn.append((nx, ny, theta))
    return out, drawn


def metasurface(setup, block=32, eps_range=(1.0, 6.0), seed=5):
    """Copy of `setup` whose structure eps_r (x 1..nx, y 1..ny, z hi..hf) is
    drawn uniform in eps_range per block x block x block voxel (seed); the
    voxels tile the period from its corner."""
    from .adjoint import structure_of, with_structure
    nz_s, ny, nx = structure_of(setup).shape
    rng = np.random.default_rng(seed)
    nb = [-(-n // block) for n in (nz_s, ny, nx)]
    eps = rng.uniform(eps_range[0], eps_range[1], size=nb)
    eps = np.repeat(np.repeat(np.repeat(eps, block, 0), block, 1), block, 2)[:nz_s, :ny, :nx]
    return with_structure(setup, 1.0 / eps)


def angle_averaged_gradient(setup, thetas, precision=4, nobox=1, segment=None, record_bytes=-1, wrt="eps_inv"):
    """Mean over the incidence angles thetas (degrees, xz plane, matching
    Bloch phase) of sum_w T(w), T = J6/(J5+J6) of box nobox, and its gradient
    with respect to the structure cells, (hf-hi+1, ny, nx): of their real
    eps_inv (wrt="eps_inv") or of their eps_r (wrt="eps", dT/d eps =
    -dT/d eps_inv / eps^2 by the chain rule through eps_inv = 1/eps).

    The angles run one after another: each forward run and its adjoint
    (pyiris.adjoint.flux) finish, and the engine with its snapshots and
    records is freed, before the next angle's engine is made, so the device
    memory held does not grow with the number of angles.  The gradients add
    up in the leaf tensor."""
    import torch
    from .adjoint import flux, structure_of, transmission
    if wrt not in ("eps_inv", "eps"):
        raise ValueError(f"wrt must be 'eps_inv' or 'eps', got {wrt!r}")
    w0 = structure_of(setup).real
    leaf = torch.tensor(w0 if wrt == "eps_inv" else 1.0 / w0, requires_grad=True)
    total = 0.0
    for th in thetas:
        w = leaf if wrt == "eps_inv" else 1.0 / leaf
        obj = transmission(flux(w, oblique(setup, th), precision, nobox, segment, record_bytes)).sum() / len(thetas)
        obj.backward()
        total += float(obj.detach())
        del obj, w
    return total, leaf.grad.numpy().copy()
