"""ctypes binding of the oracle (liboracle.so, the C restatement) and a runner
for the reference binary (oracle/_ref/irisfdtd_ref).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  It mirrors pyiris.capi.Engine so that a test
drives the engine and the oracle with the same calls.
"""
import ctypes as ct
import os
import shutil
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_BIN = os.path.join(HERE, "_ref", "irisfdtd_ref")

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    path = os.path.join(HERE, "liboracle.so")
    if not os.path.exists(path):
        subprocess.run(["make", "-C", HERE, "liboracle.so"], check=True, stdout=subprocess.DEVNULL)
    lib = ct.CDLL(path)
    V = ct.c_void_p
    lib.oracle_create.argtypes = [V] * 7
    lib.oracle_create.restype = V
    lib.oracle_step.argtypes = [V, ct.c_int, ct.c_int]
    lib.oracle_step.restype = None
    lib.oracle_flux.argtypes = [V, ct.c_int, V]
    lib.oracle_flux.restype = None
    lib.oracle_dft_points.argtypes = [V, ct.c_int, ct.c_int]
    lib.oracle_dft_points.restype = ct.c_int64
    lib.oracle_get_dft.argtypes = [V, ct.c_int, ct.c_int, V]
    lib.oracle_get_dft.restype = None
    lib.oracle_get_field.argtypes = [V, ct.c_int, V]
    lib.oracle_get_field.restype = None
    lib.oracle_set_field.argtypes = [V, ct.c_int, V]
    lib.oracle_set_field.restype = None
    lib.oracle_diagnostics.argtypes = [V, ct.c_int, ct.c_int, V]
    lib.oracle_diagnostics.restype = None
    lib.oracle_current_step.argtypes = [V]
    lib.oracle_current_step.restype = ct.c_int
    lib.oracle_set_incident.argtypes = [V, V]
    lib.oracle_set_incident.restype = None
    lib.oracle_destroy.argtypes = [V]
    lib.oracle_destroy.restype = None
    _lib = lib
    return lib


def _p(a):
    return a.ctypes.data_as(ct.c_void_p)


class Oracle:
    """CPU FP64 restatement with the Engine interface."""

    def __init__(self, setup):
        self.lib = load()
        self.params = setup.params.copy()
        self.setup = setup
        self._h = self.lib.oracle_create(ct.addressof(self.params), _p(setup.eps_inv), _p(setup.fpzmn),
                                         _p(setup.fmzmn), _p(setup.fpzmx), _p(setup.fmzmx), _p(setup.fw))

    @property
    def it(self):
        return self.lib.oracle_current_step(self._h)

    def step(self, it_first, it_last):
        self.lib.oracle_step(self._h, int(it_first), int(it_last))

    def run_to(self, it_last):
        self.step(self.it + 1, it_last)

    def sync(self):
        pass

    def flux(self, nobox=1):
        out = np.zeros(6 * self.params.nof)
        self.lib.oracle_flux(self._h, int(nobox), _p(out))
        return out.reshape(self.params.nof, 6)

    def dft_points(self, nobox, pair):
        return self.lib.oracle_dft_points(self._h, int(nobox), int(pair))

    def dft(self, nobox, pair):
        n = self.dft_points(nobox, pair)
        out = np.zeros(2 * n * self.params.nof)
        self.lib.oracle_get_dft(self._h, int(nobox), int(pair), _p(out))
        return out.view(np.complex128).reshape(self.params.nof, n)

    def field(self, which):
        p = self.params
        out = np.zeros(2 * 3 * (p.nx + 2) * (p.ny + 2) * (p.nz + 2))
        self.lib.oracle_get_field(self._h, int(which), _p(out))
        return out.view(np.complex128).reshape(p.nz + 2, p.ny + 2, p.nx + 2, 3)

    def set_field(self, which, arr):
        a = np.ascontiguousarray(arr, dtype=np.complex128)
        self.lib.oracle_set_field(self._h, int(which), _p(a.view(np.float64)))

    def diagnostics(self, zi, zf):
        """[W_E, W_H, max |div D|, max |div B|] over z in zi..zf (the
        oracle_diagnostics definitions)."""
        out = np.zeros(4)
        self.lib.oracle_diagnostics(self._h, int(zi), int(zf), _p(out))
        return out

    def set_incident(self, table):
        """Tabulated incident field ((nt+1, 2, 3) complex, the layout of
        iris_incident_table) in place of the analytic pulse."""
        t = np.ascontiguousarray(table, dtype=np.complex128)
        assert t.shape == (self.params.nt + 1, 2, 3), t.shape
        self._tab = t
        self.lib.oracle_set_incident(self._h, _p(t.view(np.float64)))

    def close(self):
        if self._h:
            self.lib.oracle_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ref_available():
    return os.path.exists(REF_BIN)


def run_reference(input_text, workdir, dump_emg=False, timeout=3600, pin_core=None, check=True):
    """Run the reference binary in `workdir` on `input_text`; returns the
    workdir (check=False: returns (workdir, exit status, stderr text)).
    Raises if the binary is absent."""
    if not ref_available():
        raise FileNotFoundError(REF_BIN)
    if os.path.exists(workdir):
        shutil.rmtree(workdir)
    os.makedirs(workdir)
    with open(os.path.join(workdir, "inputFDTD.dat"), "w") as f:
        f.write(input_text)
    env = dict(os.environ)
    if dump_emg:
        env["IRIS_REF_DUMP"] = "1"
    cmd = [REF_BIN]
    if pin_core is not None and shutil.which("taskset"):
        cmd = ["taskset", "-c", str(pin_core)] + cmd
    r = subprocess.run(cmd, cwd=workdir, env=env, check=check, stdout=subprocess.DEVNULL,
                       stderr=subprocess.PIPE, timeout=timeout)
    if not check:
        return workdir, r.returncode, r.stderr.decode(errors="replace")
    return workdir


def read_scatt(path):
    """Blocks of a <box>_scatt<j>.dat file: list of (it, values)."""
    blocks = []
    cur = None
    with open(path) as f:
        for line in f:
            t = line.split()
            if not t:
                continue
            if t[0] == "#":
                cur = (int(t[1]), [])
                blocks.append(cur)
            else:
                cur[1].append(float(t[0]))
    return [(it, np.array(v)) for it, v in blocks]


def read_emg_dump(path):
    """emg dump of ref_dump.f90: complex (n2, n1) = (nof, npts)."""
    with open(path, "rb") as f:
        n1, n2 = np.frombuffer(f.read(8), dtype=np.int32)
        data = np.frombuffer(f.read(), dtype=np.complex128)
    return data.reshape(n2, n1)
