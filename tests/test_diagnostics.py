"""Field monitors (energy and divergence), restating the reference's
uncalled diagnostics code/energy.f90:57-201 (see oracle_diagnostics).

Pinned by what the update itself fixes: in a lossless, source-free region
with periodic (Bloch) x/y boundaries the leapfrog conserves
  W = sum Re(eps_i conj(E_i^n) E_i^(n-1)) + sum |H_i^(n-1)|^2 / dtcq2
exactly (eps*A = B^T/dtcq2 for E += A H, H -= B E), and the backward
divergence of eps E and the forward divergence of H cell by cell
(div o curl = 0 for commuting one-sided differences).  A wrong weight
(eps_inv in place of its inverse) or the plain |E^n|^2 in place of the
staggered product is not conserved, which the negative controls show.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import PKG, write_input
import pyoracle
from pyiris import Engine, IrisError, inputs, load_input
from pyiris.host import Setup

AK = (0.3, -0.2, 0.0)


def source_free_setup(tmp_path, **kw):
    """Structure, three permittivities, Bloch phases, no source (Epump = 0)."""
    base = dict(ixmax=6, iymax=5, izmax=120, pmlzi=10, pmlzf=110, hi=50, dh=12, dx=2, eps_bbox=3.0, eps1=1.5,
                eps3=2.0, tfsf_o=20, epump=0.0, box=(1, 6, 1, 5, 20, 100), nof=2, nt=200, no_downloads=100)
    base.update(kw)
    s = load_input(write_input(tmp_path, inputs.render(**base)))
    p = s.params.copy()
    for i, v in enumerate(AK):
        p.ak[i][0] = v
        p.ak[i][1] = 0.0
    return Setup(p, s.eps_inv, s.fpzmn, s.fmzmn, s.fpzmx, s.fmzmx, s.fw, s.spectra, s.extra)


def bloch_fields(p, zlo, zhi, seed):
    """Random E and H on z in zlo..zhi with the x/y ghost planes set to the
    Bloch images the wraps would write (code/boundary.f90:163-348)."""
    rng = np.random.default_rng(seed)
    out = []
    fx, fy = np.exp(1j * AK[0]), np.exp(1j * AK[1])
    for _ in range(2):
        f = np.zeros((p.nz + 2, p.ny + 2, p.nx + 2, 3), dtype=np.complex128)
        shape = (zhi - zlo + 1, p.ny, p.nx, 3)
        f[zlo:zhi + 1, 1:p.ny + 1, 1:p.nx + 1] = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        f[:, 1:p.ny + 1, 0] = f[:, 1:p.ny + 1, p.nx] / fx
        f[:, 1:p.ny + 1, p.nx + 1] = f[:, 1:p.ny + 1, 1] * fx
        f[:, 0, :] = f[:, p.ny, :] / fy
        f[:, p.ny + 1, :] = f[:, 1, :] * fy
        out.append(f)
    return out


def eps_weights(s):
    """eps_i per cell (nz+2, ny+2, nx+2, 3) as the E update defines it."""
    p = s.params
    w = np.zeros((p.nz + 2, p.ny + 2, p.nx + 2, 3))
    for i in range(3):
        w[:p.hi, ..., i] = 1.0 / p.epsa_inv[0][i][i]
        w[p.hf + 1:, ..., i] = 1.0 / p.epsa_inv[1][i][i]
    einv = s.eps_inv.reshape(p.hf - p.hi + 3, p.ny + 2, p.nx + 2)[1:-1]
    w[p.hi:p.hf + 1] = (1.0 / einv).real[..., None]
    return w


def run_source_free(s, steps, seed=5):
    p = s.params
    e0, h0 = bloch_fields(p, 40, 75, seed)
    o = pyoracle.Oracle(s)
    o.set_field(0, e0)
    o.set_field(1, h0)
    diags, fields = [], []
    prev_e = e0
    prev_h = h0
    for it in range(1, steps + 1):
        o.step(it, it)
        diags.append(o.diagnostics(p.pmlzi + 1, p.pmlzf - 1))
        e, h = o.field(0), o.field(1)
        fields.append((e, prev_e, prev_h))
        prev_e, prev_h = e, h
    o.close()
    return np.array(diags), fields


def test_energy_and_divergence_are_invariants(tmp_path):
    s = source_free_setup(tmp_path)
    d, _ = run_source_free(s, 25)
    w = d[:, 0] + d[:, 1]
    assert d[0, 0] > 0 and d[0, 1] > 0
    assert np.abs(w / w[0] - 1).max() <= 1e-12, np.abs(w / w[0] - 1).max()
    # the electric and magnetic parts do exchange energy
    assert np.abs(d[:, 0] / d[0, 0] - 1).max() > 1e-3
    for k in (2, 3):
        assert d[0, k] > 1.0
        assert np.abs(d[:, k] / d[0, k] - 1).max() <= 1e-12, (k, np.abs(d[:, k] / d[0, k] - 1).max())


def test_wrong_energy_forms_are_not_conserved(tmp_path):
    s = source_free_setup(tmp_path)
    d, fields = run_source_free(s, 25)
    p = s.params
    dtcq2 = ((p.dt * p.c0) / p.q) ** 2
    wgt = eps_weights(s)
    zs = slice(p.pmlzi + 1, p.pmlzf)
    inner = (zs, slice(1, p.ny + 1), slice(1, p.nx + 1))

    def energy(e, ep, hp, weight, staggered=True):
        prod = (np.conj(e) * ep).real if staggered else np.abs(e) ** 2
        return (weight[inner] * prod[inner]).sum() + (np.abs(hp[inner]) ** 2).sum() / dtcq2

    right = np.array([energy(*f, wgt) for f in fields])
    np.testing.assert_allclose(right, d[:, 0] + d[:, 1], rtol=1e-13)
    for wrong in (np.array([energy(*f, 1.0 / wgt) for f in fields]),
                  np.array([energy(*f, wgt, staggered=False) for f in fields])):
        assert np.abs(wrong / wrong[0] - 1).max() > 1e-6


def test_divergence_of_a_wrong_stencil_is_not_conserved(tmp_path):
    """The forward divergence of E (the stencil of H) is not an invariant of
    the E update."""
    s = source_free_setup(tmp_path)
    _, fields = run_source_free(s, 10)
    p = s.params
    wgt = eps_weights(s)
    vals = []
    for e, _, _ in fields:
        d = e * wgt
        div = (d[1:-1, 1:-1, 2:, 0] - d[1:-1, 1:-1, 1:-1, 0]) + (d[1:-1, 2:, 1:-1, 1] - d[1:-1, 1:-1, 1:-1, 1]) \
            + (d[2:, 1:-1, 1:-1, 2] - d[1:-1, 1:-1, 1:-1, 2])
        vals.append(np.abs(div[p.pmlzi:p.pmlzf - 1]).max())
    assert np.abs(np.array(vals) / vals[0] - 1).max() > 1e-3


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [8, 4])
def test_engine_diagnostics_match_oracle(tmp_path, precision):
    print(f"[gpu-test] diagnostics precision {precision}", flush=True)
    s = source_free_setup(tmp_path)
    d_ref, _ = run_source_free(s, 20)
    p = s.params
    e0, h0 = bloch_fields(p, 40, 75, 5)
    tol = 1e-12 if precision == 8 else 2e-5
    with Engine(s, precision=precision) as eng:
        eng.set_field(0, e0)
        eng.set_field(1, h0)
        got = []
        for it in range(1, 21):
            eng.step(it, it)
            got.append(eng.diagnostics())
        got = np.array(got)
        np.testing.assert_allclose(got, d_ref, rtol=tol)
        w = got[:, 0] + got[:, 1]
        assert np.abs(w / w[0] - 1).max() <= (1e-12 if precision == 8 else 1e-5)
        with pytest.raises(IrisError):
            eng.diagnostics(p.pmlzi, p.pmlzf - 1)   # reaches into the UPML


@pytest.mark.gpu
def test_engine_diagnostics_with_source(tmp_path):
    """A driven run (TF/SF source, no Bloch phase): the monitors agree with
    the oracle's at several steps -- the energy sums to 1e-11; max |div D|,
    which sits at the TF/SF plane where Ez is a difference of nearly equal
    numbers (float32 pi in the source direction, ux = cos(pi_f32/2)), to the
    ~1e-9 that one ulp of the device's cosine moves Ez by."""
    print("[gpu-test] diagnostics with source", flush=True)
    s = load_input(write_input(tmp_path, inputs.render(
        ixmax=7, iymax=6, izmax=90, pmlzi=10, pmlzf=80, hi=40, dh=8, dx=3, eps_bbox=2.5, tfsf_o=20,
        box=(1, 7, 1, 6, 15, 70), nof=2, nt=150, no_downloads=50)))
    o = pyoracle.Oracle(s)
    with Engine(s, precision=8) as eng:
        for n in (40, 90, 150):
            o.run_to(n)
            eng.run_to(n)
            ref = o.diagnostics(s.params.pmlzi + 1, s.params.pmlzf - 1)
            got = eng.diagnostics()
            np.testing.assert_allclose(got[:2], ref[:2], rtol=1e-11)
            np.testing.assert_allclose(got[2:], ref[2:], rtol=1e-7, atol=1e-300)
    o.close()


@pytest.mark.gpu
def test_program_writes_monitor(tmp_path):
    """IRIS_MONITOR=1: bin/irisfdtd appends the monitors of every output step
    to monitor.dat, equal to the Python binding's at the same steps."""
    print("[gpu-test] program monitor.dat", flush=True)
    text = inputs.render(ixmax=7, iymax=6, izmax=90, pmlzi=10, pmlzf=80, hi=40, dh=8, dx=3, eps_bbox=2.5,
                         tfsf_o=20, box=(1, 7, 1, 6, 15, 70), nof=2, nt=150, no_downloads=3)
    path = write_input(tmp_path, text)
    env = dict(os.environ, IRIS_MONITOR="1")
    subprocess.run([os.path.join(PKG, "bin", "irisfdtd")], cwd=tmp_path, check=True, timeout=300,
                   stdout=subprocess.DEVNULL, env=env)
    rows = np.loadtxt(tmp_path / "monitor.dat", ndmin=2)
    assert rows[:, 0].tolist() == [50, 100, 150]
    s = load_input(path)
    with Engine(s, precision=8) as eng:
        for row in rows:
            eng.run_to(int(row[0]))
            d = eng.diagnostics()
            np.testing.assert_allclose(row[1:], [d[0], d[1], d[0] + d[1], d[2], d[3]], rtol=1e-14)
