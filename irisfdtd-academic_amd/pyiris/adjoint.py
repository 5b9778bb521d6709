"""Differentiable face fluxes: torch.autograd over the engine's adjoint.

The forward run is the plain engine run of a setup whose structure eps_inv
(reference code/mesh.f90:98-106, applied by the E update of
code/fdtd.f90:115-117) is replaced by a tensor; its output is the Jabs table
of reference code/dataprocess.f90:241-277 (iris_flux).  The backward runs
iris_adjoint with the incoming gradient as the face weights, which gives
dObjective/d eps_inv for every structure cell.  Any objective built in torch
from Jabs -- e.g. the transmission T = J6/(J5+J6) of the reference's
post-processing -- is differentiated this way.

The engine holds the forward snapshots (iris_checkpoint; by default every
ceil(sqrt(nt)) steps or more, where the snapshots and a segment's records
take the least memory) and the H records of the run's last steps (by default
as many as fit in device memory) until the backward has run.
"""
import math

import numpy as np
import torch

from .capi import Engine
from .host import Setup


def with_structure(setup, w):
    """Copy of `setup` whose eps_inv on the structure cells x in 1..nx,
    y in 1..ny, z in hi..hf is w (array (hf-hi+1, ny, nx), real or complex)."""
    p = setup.params
    w = np.asarray(w)
    shape = (p.hf - p.hi + 1, p.ny, p.nx)
    if w.shape != shape:
        raise ValueError(f"structure eps_inv must have shape {shape}, got {w.shape}")
    eps = np.array(setup.eps_inv, dtype=np.complex128, copy=True)
    eps.reshape(p.hf - p.hi + 3, p.ny + 2, p.nx + 2)[1:-1, 1:-1, 1:-1] = w
    return Setup(p.copy(), eps, setup.fpzmn, setup.fmzmn, setup.fpzmx, setup.fmzmx, setup.fw, setup.spectra,
                 dict(setup.extra))


def structure_of(setup):
    """eps_inv on the structure cells as complex (hf-hi+1, ny, nx)."""
    p = setup.params
    eps = np.asarray(setup.eps_inv, dtype=np.complex128)
    return eps.reshape(p.hf - p.hi + 3, p.ny + 2, p.nx + 2)[1:-1, 1:-1, 1:-1].copy()


def default_segment(nt):
    return max(1, int(math.ceil(math.sqrt(nt))))


class FluxFunction(torch.autograd.Function):
    """Jabs (nof, 6) of box `nobox` as a function of the structure's eps_inv."""

    @staticmethod
    def forward(ctx, w, setup, precision, nobox, segment, record_bytes):
        s = with_structure(setup, w.detach().cpu().numpy())
        eng = Engine(s, precision=precision)
        try:
            eng.checkpoint(segment or -1, record_bytes)
            eng.run_to(s.params.nt)
            jabs = eng.flux(nobox)
        except Exception:
            eng.close()
            raise
        ctx.eng = eng
        ctx.nobox = nobox
        ctx.complex_in = w.is_complex()
        ctx.in_dtype = w.dtype
        ctx.in_device = w.device
        return torch.from_numpy(jabs).to(device=w.device)

    @staticmethod
    def backward(ctx, g):
        eng = ctx.eng
        try:
            grad = eng.adjoint(ctx.nobox, g.detach().cpu().numpy().astype(np.float64))
        finally:
            eng.close()
            ctx.eng = None
        out = torch.from_numpy(grad if ctx.complex_in else grad.real.copy())
        return out.to(device=ctx.in_device, dtype=ctx.in_dtype), None, None, None, None, None


def flux(w, setup, precision=4, nobox=1, segment=None, record_bytes=-1):
    """Differentiable Jabs (nof, 6): w is the structure's eps_inv as a tensor
    (hf-hi+1, ny, nx), real or complex; precision 4 (FP32 fields, FP64 DFT)
    or 8; segment = steps between forward snapshots (default: chosen by the
    engine, at least ceil(sqrt(nt)), larger where snapshots would crowd the
    device memory);
    record_bytes = device memory for H records of the run's last steps
    (-1: as much as fits, 0: none)."""
    return FluxFunction.apply(w, setup, precision, nobox, segment, record_bytes)


def transmission(jabs):
    """T = J6/(J5+J6) per wavelength (the reference's post-processing,
    irisfdtd_scatt.py:37-42)."""
    return jabs[:, 5] / (jabs[:, 4] + jabs[:, 5])
