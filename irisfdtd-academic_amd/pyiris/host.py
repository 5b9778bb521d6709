"""Host-side set-up through the Fortran host library ``lib/libiris_host.so``.

The Fortran module ``iris_setup`` (fortran/iris_setup.f90) is the one place
that turns an ``inputFDTD.dat`` into engine parameters, with the reference's
arithmetic (reference code/files.f90:39-344, code/mesh.f90:75-109,
code/boundary_II.f90:224-276).  Python calls it here so that tests, the
benchmark and the Fortran program all use the same derived numbers.
"""
import ctypes as ct
import os
import threading
from dataclasses import dataclass, field

import numpy as np

from .capi import LIB_DIR, IrisError, IrisParams

_host = None
_lock = threading.Lock()


def load_host_library(path=None):
    global _host
    if _host is not None:
        return _host
    path = path or os.path.join(LIB_DIR, "libiris_host.so")
    if not os.path.exists(path):
        raise IrisError(f"Fortran host library not built: {path} (run irisfdtd-academic_amd/build.py)")
    lib = ct.CDLL(path)
    lib.iris_host_load.argtypes = [ct.c_char_p, ct.c_int32, ct.c_int32, ct.POINTER(IrisParams)]
    lib.iris_host_load.restype = ct.c_int
    lib.iris_host_arrays.argtypes = [ct.c_void_p] * 7
    lib.iris_host_arrays.restype = ct.c_int
    lib.iris_host_extra.argtypes = [ct.c_void_p, ct.c_void_p]
    lib.iris_host_extra.restype = ct.c_int
    lib.iris_host_error.argtypes = [ct.c_char_p, ct.c_int32]
    lib.iris_host_error.restype = ct.c_int
    lib.iris_host_sdl_config.argtypes = [ct.c_void_p]
    lib.iris_host_sdl_config.restype = ct.c_int
    lib.iris_host_sdl_geometry.argtypes = [ct.c_int32, ct.c_int32, ct.c_void_p]
    lib.iris_host_sdl_geometry.restype = ct.c_int
    lib.iris_host_sdl_write.argtypes = [ct.c_int32, ct.c_int32, ct.c_int32, ct.c_void_p, ct.c_void_p]
    lib.iris_host_sdl_write.restype = ct.c_int
    _host = lib
    return lib


@dataclass
class Setup:
    """Parameters and arrays of one run (what iris_create consumes)."""
    params: IrisParams
    eps_inv: np.ndarray      # complex128, Fortran order flattened ((nx+2)(ny+2)(hf-hi+3))
    fpzmn: np.ndarray
    fmzmn: np.ndarray
    fpzmx: np.ndarray
    fmzmx: np.ndarray
    fw: np.ndarray
    spectra: np.ndarray      # (nof, 3): lambda, R, T of the analytic film
    extra: dict = field(default_factory=dict)

    @property
    def ncells(self):
        p = self.params
        return p.nx * p.ny * p.nz

    def with_precision(self, precision):
        p = self.params.copy()
        p.precision = int(precision)
        return Setup(p, self.eps_inv, self.fpzmn, self.fmzmn, self.fpzmx, self.fmzmx, self.fw, self.spectra,
                     dict(self.extra))

    def with_steps(self, nt):
        p = self.params.copy()
        p.nt = int(nt)
        return Setup(p, self.eps_inv, self.fpzmn, self.fmzmn, self.fpzmx, self.fmzmx, self.fw, self.spectra,
                     dict(self.extra))


def load_input(path, precision=8):
    """Parse an inputFDTD.dat with the Fortran host set-up."""
    lib = load_host_library()
    p = IrisParams()
    b = os.fsencode(path)
    with _lock:
        rc = lib.iris_host_load(b, len(b), int(precision), ct.byref(p))
        if rc != 0:
            buf = ct.create_string_buffer(512)
            lib.iris_host_error(buf, 512)
            raise IrisError(f"{path}: {buf.value.decode(errors='replace')} (code {rc})")
        neps = (p.nx + 2) * (p.ny + 2) * (p.hf - p.hi + 3)
        eps = np.zeros(neps, dtype=np.complex128)
        fpzmn = np.zeros(3 * p.pmlzi)
        fmzmn = np.zeros(3 * p.pmlzi)
        ntop = 3 * (p.nz - p.pmlzf + 1)
        fpzmx = np.zeros(ntop)
        fmzmx = np.zeros(ntop)
        fw = np.zeros(p.nof)
        spectra = np.zeros(3 * p.nof)
        ptr = lambda a: a.ctypes.data_as(ct.c_void_p)  # noqa: E731
        rc = lib.iris_host_arrays(ptr(eps), ptr(fpzmn), ptr(fmzmn), ptr(fpzmx), ptr(fmzmx), ptr(fw), ptr(spectra))
        if rc != 0:
            raise IrisError("iris_host_arrays failed")
        out = np.zeros(8)
        iout = np.zeros(8, dtype=np.int32)
        lib.iris_host_extra(ptr(out), ptr(iout))
        sdl = _sdl_config(lib)
    extra = dict(zip(["q_nm", "texit_fs", "dlamda", "lno", "lnf", "pump_dt_fs", "def_pulse", "wzpml"],
                     out.tolist()))
    extra.update(zip(["no_downloads_in", "output_time_step", "dh", "dx", "aix", "afx", "sdl_on", "tfsf_zi"],
                     iout.tolist()))
    extra["sdl"] = sdl
    return Setup(p, eps, fpzmn, fmzmn, fpzmx, fmzmx, fw, spectra.reshape(p.nof, 3), extra)


def _sdl_config(lib):
    """Snapshot options of the run the host library holds: the parsed flags
    and, per enabled plane and layer, (fixed, a0, a1, b0, b1)."""
    iout = np.zeros(16, dtype=np.int32)
    lib.iris_host_sdl_config(iout.ctypes.data_as(ct.c_void_p))
    v = iout.tolist()
    cfg = dict(plane_on=v[0:3], copy_flag=v[3], copy_no=v[4], module_e=v[5], fields=v[6:12],
               nlayers=v[12:15], enabled=bool(v[15]), planes={})
    if cfg["enabled"]:
        geo = np.zeros(5, dtype=np.int32)
        for layer in (1, 2, 3):
            if cfg["plane_on"][layer - 1] != 1:
                continue
            for il in range(1, cfg["nlayers"][layer - 1] + 1):
                if lib.iris_host_sdl_geometry(layer, il, geo.ctypes.data_as(ct.c_void_p)) != 0:
                    raise IrisError("iris_host_sdl_geometry failed")
                cfg["planes"][(layer, il)] = tuple(geo.tolist())
    return cfg


def sdl_write(path, layer, ilayer, ino, e, hp):
    """Write the snapshot files of one plane with the program's writer
    (fortran/iris_sdl.f90) into the current directory, for the input at
    `path`: e and hp are E after the step and H of the step before on that
    plane, complex arrays (3, nb, na) (component, second, first free
    coordinate)."""
    lib = load_host_library()
    p = IrisParams()
    b = os.fsencode(path)
    e = np.ascontiguousarray(e, dtype=np.complex128)
    hp = np.ascontiguousarray(hp, dtype=np.complex128)
    with _lock:
        rc = lib.iris_host_load(b, len(b), 8, ct.byref(p))
        if rc != 0:
            raise IrisError(f"{path}: cannot load (code {rc})")
        geo = np.zeros(5, dtype=np.int32)
        if lib.iris_host_sdl_geometry(int(layer), int(ilayer), geo.ctypes.data_as(ct.c_void_p)) != 0:
            raise IrisError(f"no snapshot plane {layer} layer {ilayer}")
        _, a0, a1, b0, b1 = geo.tolist()
        shape = (3, b1 - b0 + 1, a1 - a0 + 1)
        if e.shape != shape or hp.shape != shape:
            raise IrisError(f"plane arrays must have shape {shape}")
        rc = lib.iris_host_sdl_write(int(layer), int(ilayer), int(ino), e.ctypes.data_as(ct.c_void_p),
                                     hp.ctypes.data_as(ct.c_void_p))
        if rc != 0:
            buf = ct.create_string_buffer(512)
            lib.iris_host_error(buf, 512)
            raise IrisError(f"snapshot write failed: {buf.value.decode(errors='replace')}")
