"""Pin the oracle (oracle/iris_oracle.c) before trusting it.

1. Golden vectors: the reference binary's own outputs on the fixtures of
   tests/golden/ (make_golden.py) -- the oracle must match them bit for bit.
2. The reference binary itself, when it is built here, on a seeded random
   small configuration.
3. Closed forms and invariants that a plausible mistake would break: the Airy
   spectrum of a lossless film, exact linearity in the source amplitude,
   finite propagation speed, J1 = -J2 on coincident faces.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN_CASES, golden_path, load_golden, run_with_cadence, write_input
import pyoracle
from pyiris import inputs, load_input


# the full-length goldens (thousands of steps with a 100-200 wavelength DFT):
# the oracle is run to the end of their first output block only, to keep
# the CPU suite short; the GPU tests compare the engine with every block
FIRST_BLOCK_ONLY = {"cfg1_full", "cfg2_period2d"}


def _oracle_run(name, nt=None):
    s = load_input(golden_path(name, "dat"))
    if nt is not None:
        s.params.nt = nt
    o = pyoracle.Oracle(s)
    res = run_with_cadence(o, s.params, s.params.nbox)
    return s, o, res


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_oracle_matches_reference_golden_bitwise(name):
    g = load_golden(name)
    short = name in FIRST_BLOCK_ONLY
    nblk = 1 if short else len(g["it_b1"])
    s, o, res = _oracle_run(name, int(g["it_b1"][0]) if short else None)
    for b in range(1, s.params.nbox + 1):
        its = [it for it, _ in res[b]]
        assert its == g[f"it_b{b}"].tolist()[:nblk]
        jab = np.stack([j for _, j in res[b]])
        ref = g[f"jabs_b{b}"][:nblk]
        assert np.array_equal(jab, ref), f"box {b}: max diff {np.abs(jab - ref).max()}"
        for key in g:
            if key.startswith(f"emg_b{b}_p") and not short:
                pair = int(key[-1])
                assert np.array_equal(o.dft(b, pair), g[key]), key
    o.close()


def test_oracle_matches_reference_binary_random(tmp_path, ref_bin_available):
    """A seeded random small 3D case through the reference binary itself."""
    if not ref_bin_available:
        pytest.skip("reference binary not built (oracle/_ref/irisfdtd_ref)")
    rng = np.random.default_rng(20240)
    nx, ny = int(rng.integers(3, 9)), int(rng.integers(3, 9))
    nz = int(rng.integers(60, 90))
    pmlzi = int(rng.integers(6, 11))
    pmlzf = nz - int(rng.integers(6, 11))
    hi = int(rng.integers(pmlzi + 12, pmlzi + 20))
    kw = dict(i_pol=int(rng.integers(1, 3)), ixmax=nx, iymax=ny, izmax=nz, pmlzi=pmlzi, pmlzf=pmlzf, hi=hi,
              dh=int(rng.integers(3, 8)), dx=int(rng.integers(0, nx)), eps_bbox=float(rng.uniform(1.5, 6)),
              eps1=float(rng.uniform(1, 2)), eps3=float(rng.uniform(1, 3)), tfsf_o=pmlzi + 5,
              box=(1, nx, 1, ny, pmlzi + 2, pmlzf - 2), nof=3, nt=150, no_downloads=3, cfl=float(rng.uniform(0.3, 0.9)),
              wlpump=float(rng.uniform(500, 900)))
    text = inputs.render(**kw)
    wd = pyoracle.run_reference(text, str(tmp_path / "ref"), dump_emg=True)
    s = load_input(os.path.join(wd, "inputFDTD.dat"))
    o = pyoracle.Oracle(s)
    o.run_to(s.params.nt)
    J = o.flux(1)
    for j in range(1, 7):
        r = pyoracle.read_scatt(os.path.join(wd, f"1_scatt{j}.dat"))[-1][1]
        assert np.array_equal(r, J[:, j - 1]), j
    for pair in (1, 2, 3):
        assert np.array_equal(pyoracle.read_emg_dump(os.path.join(wd, f"emg1_{pair}.bin")), o.dft(1, pair))


def test_airy_spectrum_of_lossless_film(tmp_path):
    """Laterally uniform film: T and R against the closed-form Airy formula
    (reference code/common.f90:104-131).  With a 60-cell UPML the deviation
    is the discretisation error, below 0.01 at 5 nm cells and eps=2.25 (the
    20-cell UPML of the reference's own setups adds ~0.05 of spurious
    reflection); a wrong sign, index or dropped term gives O(0.1)."""
    text = inputs.render(ixmax=1, iymax=1, nt=16000, nof=13, dlambda=50.0, lno=400.0, no_downloads=1,
                         pmlzi=60, pmlzf=965, tfsf_o=70, box=(1, 1, 1, 1, 65, 960))
    s = load_input(write_input(tmp_path, text))
    o = pyoracle.Oracle(s)
    o.run_to(s.params.nt)
    J = o.flux(1)
    r, t = J[:, 4], J[:, 5]
    R, T = r / (r + t), t / (r + t)
    spec = s.spectra
    assert np.abs(R - spec[:, 1]).max() < 0.012
    assert np.abs(T - spec[:, 2]).max() < 0.012
    assert np.abs(spec[:, 1] + spec[:, 2] - 1).max() < 1e-12


def test_film_spectrum_matches_reference_file():
    g = load_golden("example_1")
    s = load_input(golden_path("example_1", "dat"))
    np.testing.assert_allclose(s.spectra, g["film_spectrum"], rtol=1e-14, atol=1e-15)


def test_linearity_in_source_amplitude_is_exact(tmp_path):
    """Every operation of the loop is linear in the incident field, and
    doubling Epump doubles every intermediate exactly."""
    base = dict(ixmax=5, iymax=4, izmax=64, pmlzi=8, pmlzf=56, hi=30, dh=5, dx=2, tfsf_o=14,
                box=(1, 5, 1, 4, 10, 50), nof=3, nt=120, no_downloads=6)
    s1 = load_input(write_input(tmp_path, inputs.render(epump=1.0, **base), "a.dat"))
    s2 = load_input(write_input(tmp_path, inputs.render(epump=2.0, **base), "b.dat"))
    o1, o2 = pyoracle.Oracle(s1), pyoracle.Oracle(s2)
    o1.run_to(120)
    o2.run_to(120)
    for w in (0, 1):
        assert np.array_equal(o2.field(w), 2.0 * o1.field(w))
    for pair in (1, 2, 3):
        assert np.array_equal(o2.dft(1, pair), 2.0 * o1.dft(1, pair))
    assert np.array_equal(o2.flux(1), 4.0 * o1.flux(1))


def test_finite_propagation_speed(tmp_path):
    """The source enters at iz_min; after n steps nothing beyond ~n cells from
    it has been touched (one E and one H half step per step)."""
    text = inputs.render(ixmax=4, iymax=4, izmax=200, pmlzi=10, pmlzf=190, hi=150, dh=10, tfsf_o=40,
                         box=(1, 4, 1, 4, 20, 180), nof=2, nt=30, no_downloads=30)
    s = load_input(write_input(tmp_path, text))
    o = pyoracle.Oracle(s)
    o.run_to(30)
    e = o.field(0)
    h = o.field(1)
    z = s.params.iz_min
    assert np.abs(e[z + 32:, 1:-1, 1:-1]).max() == 0.0
    assert np.abs(h[z + 32:, 1:-1, 1:-1]).max() == 0.0
    assert np.abs(e[z - 1:z + 2]).max() > 0.0


def test_coincident_faces_give_opposite_flux():
    """sxi == sxf: both x faces are the same plane, so J1 = -J2 exactly
    (seen in the reference's example_2)."""
    g = load_golden("slit2d_x")
    j = g["jabs_b1"]
    assert np.array_equal(j[..., 0], -j[..., 1])
    assert np.abs(j[..., 0]).max() > 0
