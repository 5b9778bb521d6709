import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "irisfdtd-academic_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN_CASES = ["example_1", "example_2_short", "cfg1_short", "cfg1_full", "cfg2_period2d", "small3d_ipol2",
                "small3d_ipol1_cos_2box", "slit2d_x", "slit2d_y", "substrate3d", "sdl3d", "sdl3d_hmod"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: longer CPU test")


def golden_path(name, ext):
    return os.path.join(GOLDEN, f"{name}.{ext}")


def load_golden(name):
    return dict(np.load(golden_path(name, "npz")))


def write_input(tmp_path, text, name="inputFDTD.dat"):
    p = os.path.join(str(tmp_path), name)
    with open(p, "w", encoding="latin-1") as f:
        f.write(text)
    return p


def run_with_cadence(sim, params, nobox_count):
    """Step `sim` (Engine or Oracle) through the reference's output cadence
    (code/driver.f90:455-466) and return ({box: [(it, Jabs(nof,6))]})."""
    out = {b: [] for b in range(1, nobox_count + 1)}
    nos, it = 1, 1
    nt = params.nt
    while it <= nt:
        cand = params.to_fourier + params.no_downloads * nos
        stop = min(cand, nt) if cand >= it else nt
        sim.step(it, stop)
        if stop >= params.to_fourier and (stop == cand or stop == nt):
            for b in out:
                out[b].append((stop, sim.flux(b)))
            nos += 1
        it = stop + 1
    return out


def face_rel_err(j_test, j_ref):
    """Normwise relative error per face (max over faces): |dJ|_inf over the
    wavelengths of a face, divided by that face's |J|_inf, floored at 1e-3 of
    the largest face so that faces whose flux cancels to ~0 are judged on
    the scale of the incident flux."""
    j_test = np.asarray(j_test)
    j_ref = np.asarray(j_ref)
    glob = np.abs(j_ref).max()
    errs = []
    for f in range(j_ref.shape[-1]):
        scale = max(np.abs(j_ref[..., f]).max(), 1e-3 * glob, 1e-300)
        errs.append(np.abs(j_test[..., f] - j_ref[..., f]).max() / scale)
    return max(errs)


def dft_rel_err(a, b):
    """Normwise relative error of DFT fields per wavelength (max over wavelengths)."""
    num = np.linalg.norm(a - b, axis=-1)
    den = np.maximum(np.linalg.norm(b, axis=-1), 1e-300)
    return float((num / den).max())


def oracle_for(setup):
    """The oracle of `setup`, driven off normal incidence (ratio_angle != 1)
    by the engine's matched incident field (iris_incident_table) as the
    engine is."""
    import pyoracle
    o = pyoracle.Oracle(setup)
    if setup.params.ratio_angle != 1.0:
        from pyiris.capi import incident_table
        o.set_incident(incident_table(setup.params))
    return o


@pytest.fixture(scope="session")
def ref_bin_available():
    import pyoracle
    return pyoracle.ref_available()
