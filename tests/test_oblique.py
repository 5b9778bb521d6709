"""Oblique incidence: the engine's matched incident field against closed forms.

The reference only runs RATIO_ANGLE = 1 (code/files.f90:267).  Off normal
incidence the engine drives its TF/SF planes from iris_incident_table: the
grid's own update for the Bloch mode exp(i ak_x ix / nx) in the source medium,
so that every frequency keeps the grid's kx = ak_x / (nx Q) and travels at its
own angle, sin(theta(lambda)) = kx lambda / (2 pi n1).  These tests step the
oracle (whose stepping is pinned bit for bit to the reference binary) with
that table on a laterally uniform film and compare R(lambda) =
J5/(J5+J6) with the thin-film (Airy/Fresnel) reflectance of s (i_pol = 1,
E along y) and p (i_pol = 2, E along x, in the plane of incidence xz)
polarisation at theta(lambda).

Tolerance 0.006 absolute: the deviation is the discretisation and the
spurious UPML reflection (a 100-cell z-UPML; at normal incidence, with the
reference's own pulse, the same geometry deviates by up to 0.006 from its
Airy spectrum).  A wrong
polarisation, the angle of the carrier at every wavelength, or the
reference's analytic pulse (whose frequencies each have their own kx) miss by
0.02-0.08.  Wavelengths where the table's drive is filtered (lambda >
lambda_c / 1.4, lambda_c = 2 pi / kx) are not compared.
"""
import math
import os

import numpy as np
import pytest

from conftest import write_input
import pyoracle
from pyiris import IrisError, inputs, load_input
from pyiris.capi import incident_table
from pyiris.sweep import oblique

LAM = 400.0 + 50.0 * np.arange(13)


def film_reflectance(lam_nm, n1, n2, n3, d_nm, sin1, pol):
    """Reflectance of a film n2 (thickness d) between n1 and n3 at incidence
    sin1 in n1, from its characteristic matrix (s: admittance n cos, p:
    n / cos)."""
    c = [np.sqrt(1 - (n1 * sin1 / n) ** 2 + 0j) for n in (n1, n2, n3)]
    y = [n * ci if pol == "s" else n / ci for n, ci in zip((n1, n2, n3), c)]
    delta = 2 * np.pi * n2 * d_nm * c[1] / lam_nm
    m11, m12, m21, m22 = np.cos(delta), 1j * np.sin(delta) / y[1], 1j * y[1] * np.sin(delta), np.cos(delta)
    B, C = m11 + m12 * y[2], m21 + m22 * y[2]
    r = (y[0] * B - C) / (y[0] * B + C)
    return float(abs(r) ** 2)


def _film_setup(tmp_path, i_pol, theta, nt=6000):
    # 4 x 2 x 600 cells, 100-cell z-UPML, eps 2.25 film 100 nm (20 cells)
    # thick at z = 290..309, TF/SF plane at z = 110, 1 x 1 column box
    text = inputs.render(i_pol=i_pol, ixmax=4, iymax=2, izmax=600, nt=nt, nof=13, dlambda=50.0, lno=400.0,
                         no_downloads=nt, pmlzi=100, pmlzf=501, tfsf_o=110, hi=290, dh=20, eps_bbox=2.25,
                         box=(1, 1, 1, 1, 105, 495))
    s = load_input(write_input(tmp_path, text))
    return oblique(s, theta) if theta else s


def _reflectance(s, table=True):
    o = pyoracle.Oracle(s)
    if table:
        o.set_incident(incident_table(s.params))
    o.run_to(s.params.nt)
    J = o.flux(1)
    o.close()
    return J[:, 4] / (J[:, 4] + J[:, 5])


def _angles(s):
    p = s.params
    kx = p.ak[0][0] / (p.nx * p.q)
    return kx * LAM * p.f_abohr / (2 * np.pi), 2 * np.pi / kx / p.f_abohr


@pytest.mark.parametrize("i_pol,pol", [(1, "s"), (2, "p")])
def test_matched_source_gives_the_film_reflectance_at_each_wavelengths_angle(tmp_path, i_pol, pol):
    s = _film_setup(tmp_path, i_pol, 30.0)
    R = _reflectance(s)
    sin1, lam_c = _angles(s)
    assert lam_c / 1.4 > LAM.max() - 0.01     # the whole band has drive power at 30 degrees
    Ra = np.array([film_reflectance(l, 1.0, 1.5, 1.0, 100.0, sn, pol) for l, sn in zip(LAM, sin1)])
    err = np.abs(R - Ra).max()
    assert err < 0.006, (err, R - Ra)
    # negative controls: the carrier's angle at every wavelength, the other polarisation
    Rf = np.array([film_reflectance(l, 1.0, 1.5, 1.0, 100.0, 0.5, pol) for l in LAM])
    Ro = np.array([film_reflectance(l, 1.0, 1.5, 1.0, 100.0, sn, "p" if pol == "s" else "s")
                   for l, sn in zip(LAM, sin1)])
    assert np.abs(R - Rf).max() > 0.03
    assert np.abs(R - Ro).max() > 0.05


def test_matched_source_at_60_degrees_below_the_cutoff(tmp_path):
    """At 60 degrees lambda_c = 808 nm lies inside the band: the wavelengths
    with the full drive (lambda < lambda_c / 1.4) match the film reflectance."""
    s = _film_setup(tmp_path, 2, 60.0, nt=8000)
    R = _reflectance(s)
    sin1, lam_c = _angles(s)
    ok = LAM < lam_c / 1.4
    assert ok.sum() >= 4
    Ra = np.array([film_reflectance(l, 1.0, 1.5, 1.0, 100.0, sn, "p") for l, sn in zip(LAM, sin1)])
    assert np.abs(R - Ra)[ok].max() < 0.006, (R - Ra)[ok]


def test_reference_pulse_off_normal_incidence_misses(tmp_path):
    """Why the table: the reference's analytic pulse evaluated off normal
    incidence (every frequency with its own kx) on the same film and grid."""
    s = _film_setup(tmp_path, 2, 30.0)
    R = _reflectance(s, table=False)
    sin1, _ = _angles(s)
    Ra = np.array([film_reflectance(l, 1.0, 1.5, 1.0, 100.0, sn, "p") for l, sn in zip(LAM, sin1)])
    assert np.abs(R - Ra).max() > 0.02


def test_normal_incidence_geometry_matches_airy(tmp_path):
    """The same grid at normal incidence (the reference's own pulse): the
    bound the oblique cases are held to."""
    s = _film_setup(tmp_path, 1, 0.0)
    R = _reflectance(s, table=False)
    Ra = np.array([film_reflectance(l, 1.0, 1.5, 1.0, 100.0, 0.0, "s") for l in LAM])
    assert np.abs(R - Ra).max() < 0.008   # measured 0.0060
    np.testing.assert_allclose(Ra, s.spectra[:, 1], atol=2e-3)   # the host's own film spectrum


def test_empty_grid_stays_empty_on_the_scattered_side(tmp_path):
    """No film: the reflected flux is the UPML's spurious reflection only
    (the TF/SF plane passes the tabulated wave), both polarisations."""
    for i_pol in (1, 2):
        text = inputs.render(i_pol=i_pol, ixmax=4, iymax=2, izmax=600, nt=5000, nof=13, dlambda=50.0, lno=400.0,
                             no_downloads=5000, pmlzi=100, pmlzf=501, tfsf_o=110, hi=290, dh=20, eps_bbox=1.0,
                             box=(1, 1, 1, 1, 105, 495))
        s = oblique(load_input(write_input(tmp_path, text, f"e{i_pol}.dat")), 30.0)
        o = pyoracle.Oracle(s)
        o.set_incident(incident_table(s.params))
        o.run_to(s.params.nt)
        J = o.flux(1)
        o.close()
        assert np.abs(J[:, 4] / J[:, 5]).max() < 1e-4, i_pol


def test_incident_table_shape_and_refusals(tmp_path):
    s = _film_setup(tmp_path, 2, 30.0, nt=300)
    t = incident_table(s.params)
    assert t.shape == (301, 2, 3) and np.all(t[0] == 0) and np.abs(t).max() > 0
    # peak tangential E at iz_min scaled to the reference pulse's
    p = s.params
    assert abs(np.abs(t[:, 1, 0]).max() - p.eamp * p.q * p.uz) < 1e-12 * p.eamp * p.q
    n = load_input(write_input(tmp_path, inputs.render(nt=10), "n.dat"))
    with pytest.raises(IrisError, match="normal incidence"):
        incident_table(n.params)
    q = s.params.copy()
    q.ak[1][0] = 0.1
    with pytest.raises(IrisError, match="x only"):
        incident_table(q)
