"""ctypes binding of the engine's C ABI (include/iris.h).

The engine library ``lib/libiris.so`` must have been built (``build.py``);
loading fails loudly otherwise -- there is no CPU fallback.
"""
import ctypes as ct
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(PKG_DIR, "lib")

IRIS_ABI_VERSION = 1


class IrisParams(ct.Structure):
    """Mirror of ``iris_params`` (include/iris.h)."""
    _fields_ = [
        ("abi_version", ct.c_int32), ("precision", ct.c_int32),
        ("nx", ct.c_int32), ("ny", ct.c_int32), ("nz", ct.c_int32),
        ("pmlzi", ct.c_int32), ("pmlzf", ct.c_int32),
        ("hi", ct.c_int32), ("hf", ct.c_int32),
        ("i_pol", ct.c_int32), ("em_source", ct.c_int32), ("plane_wave", ct.c_int32),
        ("cosine", ct.c_int32), ("iz_min", ct.c_int32), ("iz_max", ct.c_int32),
        ("ixo", ct.c_int32), ("nt", ct.c_int32), ("to_fourier", ct.c_int32),
        ("sample_fourier", ct.c_int32), ("no_downloads", ct.c_int32), ("nof", ct.c_int32),
        ("nbox", ct.c_int32), ("box", (ct.c_int32 * 6) * 2),
        ("q", ct.c_double), ("dt", ct.c_double), ("c0", ct.c_double), ("eps0", ct.c_double),
        ("mu0", ct.c_double), ("pi", ct.c_double), ("f_abohr", ct.c_double),
        ("eps", ct.c_double * 3), ("eps_in", ct.c_double * 3),
        ("epsa_inv", ((ct.c_double * 3) * 3) * 2),
        ("eamp", ct.c_double), ("wlpump", ct.c_double), ("gpump", ct.c_double), ("zopump", ct.c_double),
        ("ratio_angle", ct.c_double), ("ux", ct.c_double), ("uz", ct.c_double), ("win", ct.c_double),
        ("ak", (ct.c_double * 2) * 3),
    ]

    def copy(self):
        c = IrisParams()
        ct.pointer(c)[0] = self
        return c

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            if hasattr(v, "__len__"):
                v = np.array(v).tolist()
            out[name] = v
        return out


class IrisError(RuntimeError):
    pass


_lib = None


def _p(a):
    return a.ctypes.data_as(ct.c_void_p)


def load_library(path=None):
    """Load libiris.so once; raises if it is missing (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or os.path.join(LIB_DIR, "libiris.so")
    if not os.path.exists(path):
        raise IrisError(f"HIP engine library not built: {path} (run irisfdtd-academic_amd/build.py)")
    lib = ct.CDLL(path)
    P = ct.POINTER(IrisParams)
    H = ct.c_void_p
    lib.iris_create.argtypes = [P, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                ct.POINTER(H)]
    lib.iris_create.restype = ct.c_int
    lib.iris_step.argtypes = [H, ct.c_int32, ct.c_int32]
    lib.iris_step.restype = ct.c_int
    lib.iris_sync.argtypes = [H]
    lib.iris_sync.restype = ct.c_int
    lib.iris_flux.argtypes = [H, ct.c_int32, ct.c_void_p]
    lib.iris_flux.restype = ct.c_int
    lib.iris_dft_points.argtypes = [H, ct.c_int32, ct.c_int32, ct.POINTER(ct.c_int64)]
    lib.iris_dft_points.restype = ct.c_int
    lib.iris_get_dft.argtypes = [H, ct.c_int32, ct.c_int32, ct.c_void_p]
    lib.iris_get_dft.restype = ct.c_int
    lib.iris_get_field.argtypes = [H, ct.c_int32, ct.c_void_p]
    lib.iris_get_field.restype = ct.c_int
    lib.iris_set_field.argtypes = [H, ct.c_int32, ct.c_void_p]
    lib.iris_set_field.restype = ct.c_int
    lib.iris_get_plane.argtypes = [H, ct.c_int32, ct.c_int32, ct.c_int32, ct.c_void_p, ct.c_void_p]
    lib.iris_get_plane.restype = ct.c_int
    lib.iris_get_plane_noalloc.argtypes = [H, ct.c_int32, ct.c_int32, ct.c_int32, ct.c_void_p, ct.c_void_p]
    lib.iris_get_plane_noalloc.restype = ct.c_int
    lib.iris_diagnostics.argtypes = [H, ct.c_int32, ct.c_int32, ct.c_void_p]
    lib.iris_diagnostics.restype = ct.c_int
    lib.iris_pass_info.argtypes = [H, ct.c_void_p]
    lib.iris_pass_info.restype = ct.c_int
    lib.iris_autotune.argtypes = [H, ct.c_int32, ct.c_void_p]
    lib.iris_autotune.restype = ct.c_int
    lib.iris_graph_info.argtypes = [H, ct.c_void_p]
    lib.iris_graph_info.restype = ct.c_int
    lib.iris_checkpoint.argtypes = [H, ct.c_int32, ct.c_int64]
    lib.iris_checkpoint.restype = ct.c_int
    lib.iris_checkpoint_info.argtypes = [H, ct.c_void_p]
    lib.iris_checkpoint_info.restype = ct.c_int
    lib.iris_adjoint.argtypes = [H, ct.c_int32, ct.c_void_p, ct.c_void_p]
    lib.iris_adjoint.restype = ct.c_int
    lib.iris_batch_create.argtypes = [ct.POINTER(H), ct.c_int32, ct.POINTER(H)]
    lib.iris_batch_create.restype = ct.c_int
    lib.iris_batch_step.argtypes = [H, ct.c_int32, ct.c_int32]
    lib.iris_batch_step.restype = ct.c_int
    lib.iris_batch_info.argtypes = [H, ct.c_void_p]
    lib.iris_batch_info.restype = ct.c_int
    lib.iris_batch_destroy.argtypes = [H]
    lib.iris_batch_destroy.restype = ct.c_int
    lib.iris_batch_last_error.argtypes = [H]
    lib.iris_batch_last_error.restype = ct.c_char_p
    lib.iris_current_step.argtypes = [H, ct.POINTER(ct.c_int32)]
    lib.iris_current_step.restype = ct.c_int
    lib.iris_destroy.argtypes = [H]
    lib.iris_destroy.restype = ct.c_int
    lib.iris_last_error.argtypes = [H]
    lib.iris_last_error.restype = ct.c_char_p
    lib.iris_profile.argtypes = [H, ct.c_int32]
    lib.iris_profile.restype = ct.c_int
    lib.iris_profile_read.argtypes = [H, ct.c_void_p, ct.c_void_p, ct.c_int32]
    lib.iris_profile_read.restype = ct.c_int
    lib.iris_version.argtypes = []
    lib.iris_version.restype = ct.c_char_p
    lib.iris_incident_table.argtypes = [P, ct.c_void_p]
    lib.iris_pair_layout.argtypes = [ct.c_int32, ct.c_int32, ct.c_int32, ct.c_void_p]
    lib.iris_pair_layout.restype = ct.c_int
    lib.iris_incident_table.restype = ct.c_int
    _lib = lib
    return lib


EXPORTED = ["iris_create", "iris_step", "iris_sync", "iris_flux", "iris_dft_points", "iris_get_dft",
            "iris_get_field", "iris_set_field", "iris_get_plane", "iris_diagnostics", "iris_current_step",
            "iris_destroy", "iris_last_error", "iris_version", "iris_profile", "iris_profile_read",
            "iris_pass_info", "iris_autotune", "iris_graph_info", "iris_checkpoint", "iris_checkpoint_info",
            "iris_adjoint",
            "iris_batch_create", "iris_batch_step", "iris_batch_info", "iris_batch_destroy", "iris_batch_last_error",
            "iris_incident_table", "iris_get_plane_noalloc", "iris_pair_layout"]

KERNELS = ["epass", "hpass", "tfsf", "dft_sample", "dft_accum", "flux", "yee", "yee2", "yee_rest", "adjoint",
           "checkpoint", "yee2_pml"]


def pair_layout(nx, ny, only_ox=0):
    """Tiles of the two-cells-a-lane two-step pass over an nx x ny plane
    (iris_pair_layout; host arithmetic, no device): dict of the main and
    narrow tile shapes, or None where none fits."""
    lib = load_library()
    out = (ct.c_int32 * 11)()
    if lib.iris_pair_layout(nx, ny, only_ox, out) != 0:
        return None
    keys = ["ox", "sl", "ns", "tx", "ty", "oxn", "sln", "nsn", "txn", "tyn", "x0n"]
    return dict(zip(keys, list(out)))


def incident_table(params):
    """The engine's incident field for an oblique source (ratio_angle != 1;
    iris_incident_table): complex (nt+1, 2, 3) -- step, plane (iz_min-1,
    iz_min), components 1,3,5 (i_pol=2) or 2,4,6 (i_pol=1) at ix = 0.  Host
    computation only."""
    lib = load_library()
    out = np.zeros(2 * (params.nt + 1) * 6, dtype=np.float64)
    rc = lib.iris_incident_table(ct.byref(params), _p(out))
    if rc:
        raise IrisError(f"iris_incident_table failed ({rc}): {lib.iris_last_error(None).decode()}")
    return out.view(np.complex128).reshape(params.nt + 1, 2, 3)


class Engine:
    """One simulation on the GPU, driven through the C ABI.

    ``setup`` is a :class:`pyiris.host.Setup` (parameters and arrays derived
    from an inputFDTD.dat by the Fortran host library).
    """

    def __init__(self, setup, precision=None):
        self.lib = load_library()
        p = setup.params.copy()
        if precision is not None:
            p.precision = int(precision)
        self.params = p
        self.setup = setup
        self._h = ct.c_void_p()
        rc = self.lib.iris_create(ct.byref(p), _p(setup.eps_inv), _p(setup.fpzmn), _p(setup.fmzmn),
                                  _p(setup.fpzmx), _p(setup.fmzmx), _p(setup.fw), ct.byref(self._h))
        if rc != 0:
            raise IrisError(f"iris_create failed ({rc}): {self.lib.iris_last_error(None).decode()}")

    def _check(self, rc, what):
        if rc != 0:
            raise IrisError(f"{what} failed ({rc}): {self.lib.iris_last_error(self._h).decode()}")

    @property
    def it(self):
        v = ct.c_int32()
        self._check(self.lib.iris_current_step(self._h, ct.byref(v)), "iris_current_step")
        return v.value

    def step(self, it_first, it_last):
        self._check(self.lib.iris_step(self._h, int(it_first), int(it_last)), "iris_step")

    def run_to(self, it_last):
        self.step(self.it + 1, it_last)

    def sync(self):
        self._check(self.lib.iris_sync(self._h), "iris_sync")

    def flux(self, nobox=1):
        """Jabs as (nof, 6): column j-1 is J_j (reference Jabs(j,iw))."""
        out = np.zeros(6 * self.params.nof, dtype=np.float64)
        self._check(self.lib.iris_flux(self._h, int(nobox), _p(out)), "iris_flux")
        return out.reshape(self.params.nof, 6)

    def dft_points(self, nobox, pair):
        n = ct.c_int64()
        self._check(self.lib.iris_dft_points(self._h, int(nobox), int(pair), ct.byref(n)), "iris_dft_points")
        return n.value

    def dft(self, nobox, pair):
        """emg(pair)%fdata as complex (nof, npts)."""
        n = self.dft_points(nobox, pair)
        out = np.zeros(2 * n * self.params.nof, dtype=np.float64)
        self._check(self.lib.iris_get_dft(self._h, int(nobox), int(pair), _p(out)), "iris_get_dft")
        return out.view(np.complex128).reshape(self.params.nof, n)

    def field(self, which):
        """E (0) or H (1) as complex (nz+2, ny+2, nx+2, 3) (C order of the Fortran (3,0:nx+1,0:ny+1,0:nz+1))."""
        p = self.params
        out = np.zeros(2 * 3 * (p.nx + 2) * (p.ny + 2) * (p.nz + 2), dtype=np.float64)
        self._check(self.lib.iris_get_field(self._h, int(which), _p(out)), "iris_get_field")
        return out.view(np.complex128).reshape(p.nz + 2, p.ny + 2, p.nx + 2, 3)

    # iris_get_plane selectors: base + component (0: x, 1: y, 2: z)
    PLANE_E, PLANE_H_PREV, PLANE_H = 0, 3, 6

    def plane(self, what, axis, index, a, b):
        """One field component on one plane as complex (nb, na): axis 1 is
        x = index over (y, z), 2 is y = index over (x, z), 3 is z = index over
        (x, y); a = (a0, a1) and b = (b0, b1) are the inclusive ranges of the
        two free coordinates.  what = PLANE_E/PLANE_H_PREV/PLANE_H + component."""
        na, nb = a[1] - a[0] + 1, b[1] - b[0] + 1
        rng = (ct.c_int32 * 4)(a[0], a[1], b[0], b[1])
        out = np.zeros(2 * max(na, 0) * max(nb, 0), dtype=np.float64)
        self._check(self.lib.iris_get_plane(self._h, int(what), int(axis), int(index), rng, _p(out)),
                    "iris_get_plane")
        return out.view(np.complex128).reshape(nb, na)

    def plane_noalloc(self, what, axis, index, a, b, out=None):
        """plane() through iris_get_plane_noalloc: no allocation in the
        library (planes of at most 2^20 values); `out` (float64, 2*na*nb) may
        be given to avoid one here too."""
        na, nb = a[1] - a[0] + 1, b[1] - b[0] + 1
        rng = (ct.c_int32 * 4)(a[0], a[1], b[0], b[1])
        if out is None:
            out = np.zeros(2 * max(na, 0) * max(nb, 0), dtype=np.float64)
        self._check(self.lib.iris_get_plane_noalloc(self._h, int(what), int(axis), int(index), rng, _p(out)),
                    "iris_get_plane_noalloc")
        return out.view(np.complex128).reshape(nb, na)

    def diagnostics(self, zi=None, zf=None):
        """[W_E, W_H, max |div D|, max |div B|] over z in zi..zf (default: the
        whole region between the UPML slabs)."""
        p = self.params
        zi = p.pmlzi + 1 if zi is None else zi
        zf = p.pmlzf - 1 if zf is None else zf
        out = np.zeros(4)
        self._check(self.lib.iris_diagnostics(self._h, int(zi), int(zf), _p(out)), "iris_diagnostics")
        return out

    def pass_info(self):
        """{zchunk, two_step_planes, two_step_chunks, two_step, upml_two_step_planes,
        upml_planes_in_two_step}: the layout of the step passes."""
        v = np.zeros(6, dtype=np.int64)
        self._check(self.lib.iris_pass_info(self._h, _p(v)), "iris_pass_info")
        return dict(zchunk=int(v[0]), two_step_planes=int(v[1]), two_step_chunks=int(v[2]), two_step=bool(v[3]),
                    upml_two_step_planes=int(v[4]), upml_planes_in_two_step=int(v[5]))

    def graph_info(self):
        """{captured, replays, period}: steps run as replays of a captured
        HIP graph of one period of `period` steps."""
        v = np.zeros(3, dtype=np.int64)
        self._check(self.lib.iris_graph_info(self._h, _p(v)), "iris_graph_info")
        return dict(captured=bool(v[0]), replays=int(v[1]), period=int(v[2]))

    def autotune(self, steps=8):
        """Time candidate layouts of the step passes for `steps` steps each and
        keep the fastest (before the first step; results unchanged bit for
        bit).  Returns {two_step, upml_planes, zchunk, zchunk2, xcd_rows,
        xcd_rows2, layouts, ns_per_step, tile_width}."""
        v = np.zeros(9, dtype=np.int64)
        self._check(self.lib.iris_autotune(self._h, int(steps), _p(v)), "iris_autotune")
        return dict(two_step=bool(v[0]), upml_planes=int(v[1]), zchunk=int(v[2]), zchunk2=int(v[3]),
                    xcd_rows=int(v[4]), xcd_rows2=int(v[5]), layouts=int(v[6]), ns_per_step=int(v[7]),
                    tile_width=int(v[8]))

    def checkpoint(self, segment, record_bytes=0):
        """Prepare the coming forward run for adjoint(): H of its last steps
        recorded in up to record_bytes of device memory (-1: as much as fits),
        a snapshot of the state every `segment` steps before them (< 0: the
        engine chooses, see iris_checkpoint).  Call before the first step;
        segment = 0 turns it off."""
        self._check(self.lib.iris_checkpoint(self._h, int(segment), int(record_bytes)), "iris_checkpoint")

    def checkpoint_info(self):
        """{segment, snapshots, recorded_steps, first_recorded,
        record_step_bytes, snapshot_bytes}."""
        v = np.zeros(6, dtype=np.int64)
        self._check(self.lib.iris_checkpoint_info(self._h, _p(v)), "iris_checkpoint_info")
        return dict(segment=int(v[0]), snapshots=int(v[1]), recorded_steps=int(v[2]), first_recorded=int(v[3]),
                    record_step_bytes=int(v[4]), snapshot_bytes=int(v[5]))

    def adjoint(self, nobox, weights):
        """dJ/d eps_inv on the structure cells for J = sum(weights * Jabs)
        (weights shaped like flux(): (nof, 6)), as complex (hf-hi+1, ny, nx):
        real part dJ/d Re eps_inv, imaginary part dJ/d Im eps_inv.  Needs the
        forward run at nt with checkpoint() called before it."""
        p = self.params
        w = np.ascontiguousarray(weights, dtype=np.float64).reshape(p.nof, 6)
        out = np.zeros(2 * p.nx * p.ny * (p.hf - p.hi + 1), dtype=np.float64)
        self._check(self.lib.iris_adjoint(self._h, int(nobox), _p(w), _p(out)), "iris_adjoint")
        return out.view(np.complex128).reshape(p.hf - p.hi + 1, p.ny, p.nx)

    def set_field(self, which, arr):
        a = np.ascontiguousarray(arr, dtype=np.complex128)
        self._check(self.lib.iris_set_field(self._h, int(which), _p(a.view(np.float64))), "iris_set_field")

    def profile(self, enable=True):
        self._check(self.lib.iris_profile(self._h, 1 if enable else 0), "iris_profile")

    def profile_read(self, reset=False):
        """{kernel: (total_ms, launches)} since the last reset."""
        ms = np.zeros(len(KERNELS))
        n = np.zeros(len(KERNELS), dtype=np.int64)
        self._check(self.lib.iris_profile_read(self._h, _p(ms), _p(n), 1 if reset else 0), "iris_profile_read")
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(KERNELS)}

    def close(self):
        if self._h:
            self.lib.iris_destroy(self._h)
            self._h = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Batch:
    """Engines stepped together (iris_batch_*): one launch of the fused pass
    per step over the tiles of every engine.  The engines stay usable on
    their own and must outlive the batch."""

    def __init__(self, engines):
        self.lib = load_library()
        self.engines = list(engines)
        arr = (ct.c_void_p * len(self.engines))(*[e._h.value for e in self.engines])
        self._b = ct.c_void_p()
        rc = self.lib.iris_batch_create(ct.cast(arr, ct.POINTER(ct.c_void_p)), len(self.engines), ct.byref(self._b))
        if rc != 0:
            raise IrisError(f"iris_batch_create failed ({rc}): {self.lib.iris_batch_last_error(None).decode()}")

    def step(self, it_first, it_last):
        rc = self.lib.iris_batch_step(self._b, int(it_first), int(it_last))
        if rc != 0:
            raise IrisError(f"iris_batch_step failed ({rc}): {self.lib.iris_batch_last_error(self._b).decode()}")

    def run_to(self, it_last):
        self.step(self.engines[0].it + 1, it_last)

    def info(self):
        v = np.zeros(5, dtype=np.int64)
        self.lib.iris_batch_info(self._b, _p(v))
        return dict(instances=int(v[0]), zchunk=int(v[1]), lean_workgroups=int(v[2]), general_workgroups=int(v[3]),
                    two_step=int(v[4]))

    def sync(self):
        for e in self.engines:
            e.sync()

    def close(self):
        if self._b:
            self.lib.iris_batch_destroy(self._b)
            self._b = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
