"""Python host of the MI355X FDTD engine for IrisFDTD-Academic's time loop.

  capi    ctypes binding of include/iris.h (lib/libiris.so); Engine class
  host    parameter set-up through the Fortran host library (lib/libiris_host.so)
  inputs  inputFDTD.dat writer for workloads
  sdl     reader of the field-snapshot files (the reference's SDL records)
  adjoint differentiable face fluxes (torch.autograd over iris_adjoint)
  sweep   oblique incidence, grating sweeps, angle-averaged gradients
"""
from .capi import Batch, Engine, IrisError, IrisParams, load_library  # noqa: F401
from .host import Setup, load_input  # noqa: F401
from . import inputs  # noqa: F401
