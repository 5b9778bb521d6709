"""Host logic on the CPU: input parsing and derived parameters (Fortran host
library), the C ABI library's exports and argument checking, the input
writer.  No GPU call is made here."""
import ctypes as ct
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import PKG, ROOT, golden_path, write_input
from pyiris import IrisError, inputs, load_input
from pyiris import capi


def bits(x):
    return np.float64(x).view(np.int64)


def test_constants_match_reference_fp64_values():
    """The reference's FP64 constants carry float32 rounding (reference
    code/physconst.f90:31-36); values as the reference binary prints them."""
    s = load_input(golden_path("cfg1_short", "dat"))
    p = s.params
    assert bits(p.pi) == bits(3.14159274101257324e+00)
    assert bits(p.c0) == bits(1.37035995483398438e+02)
    assert bits(p.eps0) == bits(7.95774693315028370e-02)
    assert bits(p.mu0) == bits(6.69176310406942288e-04)
    assert bits(p.f_abohr) == bits(1.88972606658935547e+01)
    assert bits(p.q) == bits(9.44863033294677734e+01)
    assert bits(p.dt) == bits(1.99041472532702712e-01)
    dtcq2 = ((p.dt * p.c0) / p.q) ** 2
    assert bits(dtcq2) == bits(8.33333333333333287e-02)
    assert bits(s.fw[0]) == bits(2.26724926838472902e-02)
    # float-pi normal incidence: ux = cos(pi_f32/2) is not zero
    assert p.ux == pytest.approx(-4.371139000186241e-08, rel=1e-12)


def test_example_1_parameters():
    """log.dat of the reference for example_1: dt = 0.00770333 fs,
    texit = 25962, no_downloads = 519 steps, slit absent (aix > afx)."""
    s = load_input(golden_path("example_1", "dat"))
    p = s.params
    assert abs(p.dt * 2.41888426244258881e-02 - 0.00770333) < 5e-9
    assert p.nt == 25962
    assert p.no_downloads == 519
    assert (p.nx, p.ny, p.nz, p.pmlzi, p.pmlzf, p.hi, p.hf) == (1, 1, 400, 75, 325, 150, 299)
    assert p.nbox == 1 and list(p.box[0]) == [1, 1, 1, 1, 90, 320]
    assert s.extra["aix"] > s.extra["afx"]
    assert p.iz_min == 105 and p.ixo == 25


def test_upml_tables_profile():
    """set_sz: only the z column is graded, quadratically from the outer edge
    (reference code/boundary_II.f90:224-276)."""
    s = load_input(golden_path("cfg1_short", "dat"))
    p = s.params
    fp = s.fpzmn.reshape(p.pmlzi, 3)
    fm = s.fmzmn.reshape(p.pmlzi, 3)
    base = 2 * p.eps0 * p.eps[0]
    assert np.all(fp[:, 0] == base) and np.all(fp[:, 1] == base) and np.all(fm[:, 0] == base)
    sig = (fp[:, 2] - fm[:, 2]) / (2 * p.dt)
    assert sig[-1] == 0.0 and np.all(np.diff(sig) < 0)
    ratio = sig[:-1] / sig[0]
    i = np.arange(1, p.pmlzi)
    np.testing.assert_allclose(ratio, ((p.pmlzi - i) / (p.pmlzi - 1)) ** 2, rtol=1e-12)


def test_eps_inv_and_slit():
    s = load_input(golden_path("small3d_ipol2", "dat"))
    p = s.params
    e = s.eps_inv.reshape(p.hf - p.hi + 3, p.ny + 2, p.nx + 2)
    aix, afx = s.extra["aix"], s.extra["afx"]
    assert (aix, afx) == (3, 5)  # struct_limit: dx=3 on 8 cells (reference code/files.f90:403-405)
    assert np.all(e[1:-1, :, aix:afx + 1] == 1.0)
    assert np.all(e[1:-1, :, :aix].real == 1 / 2.25) or np.allclose(e[1:-1, :, :aix].real, 1 / 2.25, rtol=1e-15)
    assert np.all(e[0] == 1.0 / p.eps[0]) and np.all(e[-1] == 1.0 / p.eps[2])


def test_complex_eps_literal_is_accepted(tmp_path):
    s = load_input(write_input(tmp_path, inputs.render(eps_bbox=complex(2.25, 0.1))))
    assert s.eps_inv.imag.min() < 0 and s.params.eps[1] == 2.25


@pytest.mark.parametrize("mutate,msg", [
    (lambda t: t.replace('"Mesh options"', '"Mesh opts"'), "Mesh options"),
    (lambda t: re.sub(r"^0\.5\tCFL", "-1.0\tCFL", t, flags=re.M), "CFL"),
    (lambda t: re.sub(r"^1005\tpmlzf", "2000\tpmlzf", t, flags=re.M), "pmlzi"),
    (lambda t: re.sub(r"^1 1\tsxi mxi", "0 0\tsxi mxi", t, flags=re.M), "Fourier box"),
    (lambda t: re.sub(r"^1\tplane_wave", "0\tplane_wave", t, flags=re.M), "plane_wave"),
])
def test_bad_input_is_an_error_not_a_stop(tmp_path, mutate, msg):
    text = mutate(inputs.render())
    with pytest.raises(IrisError, match=msg):
        load_input(write_input(tmp_path, text))


def test_missing_file_is_an_error(tmp_path):
    with pytest.raises(IrisError, match="cannot open"):
        load_input(str(tmp_path / "nope.dat"))


def test_render_round_trip(tmp_path):
    s = load_input(write_input(tmp_path, inputs.render(nt=1234, nof=7, ixmax=9, iymax=3)))
    assert s.params.nt == 1234 and s.params.nof == 7 and (s.params.nx, s.params.ny) == (9, 3)


def _declared_symbols():
    with open(os.path.join(ROOT, "include", "iris.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(iris_\w+)\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = capi.load_library()
    declared = _declared_symbols()
    assert len(declared) >= 12
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(capi.EXPORTED)
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(PKG, "lib", "libiris.so")],
                         capture_output=True, text=True, check=True).stdout
    for name in declared:
        assert re.search(rf"\bT {name}$", out, flags=re.M), name


def test_library_rejects_bad_params_before_touching_a_device():
    lib = capi.load_library()
    s = load_input(golden_path("cfg1_short", "dat"))
    h = ct.c_void_p()

    def create(p):
        ptr = lambda a: a.ctypes.data_as(ct.c_void_p)  # noqa: E731
        return lib.iris_create(ct.byref(p), ptr(s.eps_inv), ptr(s.fpzmn), ptr(s.fmzmn), ptr(s.fpzmx),
                               ptr(s.fmzmx), ptr(s.fw), ct.byref(h))

    p = s.params.copy()
    p.abi_version = 99
    assert create(p) == -1 and b"abi_version" in lib.iris_last_error(None)
    p = s.params.copy()
    p.precision = 2
    assert create(p) == -1
    p = s.params.copy()
    p.pmlzf = p.nz
    assert create(p) == -1 and b"pmlzf" in lib.iris_last_error(None)
    p = s.params.copy()
    p.box[0][1] = p.nx + 1
    assert create(p) == -1 and b"box" in lib.iris_last_error(None)
    p = s.params.copy()
    p.i_pol = 3
    assert create(p) == -5
    assert lib.iris_step(None, 1, 1) == -1
    assert lib.iris_destroy(None) == 0
    assert lib.iris_version().startswith(b"iris-mi355x")


def test_host_library_exports():
    from pyiris import host
    lib = host.load_host_library()
    for name in ("iris_host_load", "iris_host_arrays", "iris_host_extra", "iris_host_error"):
        assert hasattr(lib, name)
