"""Gradient of face-flux objectives (iris_checkpoint / iris_adjoint and the
torch.autograd binding pyiris.adjoint).

The engine's gradient is the discrete adjoint of its own time step, so it is
pinned by two independent routes:
  * PyTorch autograd through tests/torch_fdtd.py, a differentiable FP64
    restatement of the step whose flux is first checked against the oracle
    (CPU tests below), on the same grids;
  * central finite differences of the plain engine run (h = 1e-4 on eps_inv).
A transposed curl, a dropped UPML term, a wrap taken the wrong way or a
missing conjugate each move the gradient by far more than the tolerances
used (1e-9 relative against autograd; 1e-6 against the differences, whose
truncation error is O(h^2)).  Size-independent properties cover the rest:
linearity in the weights, bitwise independence of the snapshot interval and
of the two-step pass, and the state at nt restored after the adjoint.
"""
import os

import numpy as np
import pytest
import torch

from conftest import write_input
import pyoracle
from pyiris import Engine, IrisError, inputs, load_input
from pyiris.adjoint import flux as flux_fn, structure_of, transmission, with_structure
from pyiris.host import Setup
from torch_fdtd import TorchFDTD

AK = (0.3, -0.2, 0.0)


def make_setup(tmp_path, case, **over):
    """Small 3D grids with a structure layer (eps_inv drawn per cell); `over`
    replaces input values."""
    if case == "slab_ipol2":
        kw = dict(ixmax=8, iymax=6, izmax=64, pmlzi=8, pmlzf=57, hi=28, dh=8, dx=3, eps_bbox=2.5, tfsf_o=14,
                  box=(1, 8, 1, 6, 12, 50), nof=3, nt=160, no_downloads=160, lno=450.0, dlambda=100.0)
    elif case == "bloch_lossy_ipol1":
        kw = dict(i_pol=1, ixmax=7, iymax=6, izmax=64, pmlzi=8, pmlzf=57, hi=26, dh=6, dx=2, eps_bbox=3.0 + 0.4j,
                  eps3=1.8, tfsf_o=14, box=(2, 6, 1, 5, 12, 48), nof=2, nt=150, no_downloads=150, lno=500.0,
                  dlambda=150.0)
    elif case == "grid16":
        kw = dict(ixmax=16, iymax=16, izmax=64, pmlzi=8, pmlzf=57, hi=26, dh=10, dx=0, eps_bbox=2.0, tfsf_o=14,
                  box=(1, 16, 1, 16, 12, 50), nof=4, nt=160, no_downloads=160, lno=450.0, dlambda=100.0)
    elif case == "two_step":
        kw = dict(ixmax=70, iymax=21, izmax=96, pmlzi=10, pmlzf=86, hi=44, dh=10, dx=8, eps_bbox=2.25, tfsf_o=20,
                  box=(1, 70, 1, 21, 15, 80), nof=4, nt=200, no_downloads=200, lno=450.0, dlambda=100.0)
    else:
        raise KeyError(case)
    kw.update(over)
    s = load_input(write_input(tmp_path, inputs.render(**kw)))
    if case == "bloch_lossy_ipol1":
        p = s.params.copy()
        for i, v in enumerate(AK):
            p.ak[i][0] = v
            p.ak[i][1] = 0.0
        s = Setup(p, s.eps_inv, s.fpzmn, s.fmzmn, s.fpzmx, s.fmzmx, s.fw, s.spectra, s.extra)
    # a design layer: eps_r drawn uniform in [1, 4] per cell (lossy case: plus
    # the structure's own imaginary part)
    rng = np.random.default_rng(4)
    w0 = structure_of(s)
    w = 1.0 / rng.uniform(1.0, 4.0, size=w0.shape)
    if case == "bloch_lossy_ipol1":
        w = w + 1j * w0.imag
    return with_structure(s, w)


def weights_for(p, seed=1):
    return np.random.default_rng(seed).normal(size=(p.nof, 6))


def torch_gradient(s, w):
    """(dJ/d Re eps_inv + i dJ/d Im eps_inv, Jabs) by autograd."""
    tf = TorchFDTD(s)
    w0 = structure_of(s)
    wr = torch.tensor(w0.real.copy(), requires_grad=True)
    wi = torch.tensor(w0.imag.copy(), requires_grad=True)
    J = tf.run(torch.complex(wr, wi))
    (J * torch.tensor(w)).sum().backward()
    return wr.grad.numpy() + 1j * wi.grad.numpy(), J.detach().numpy()


def engine_gradient(s, w, precision=8, segment=7):
    with Engine(s, precision=precision) as eng:
        eng.checkpoint(segment, 0)
        eng.run_to(s.params.nt)
        j = eng.flux(1)
        g = eng.adjoint(1, w)
    return g, j


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


# ---- CPU: the PyTorch restatement reproduces the oracle ----------------------

@pytest.mark.parametrize("case", ["slab_ipol2", "bloch_lossy_ipol1"])
def test_torch_restatement_matches_oracle(tmp_path, case):
    s = make_setup(tmp_path, case)
    o = pyoracle.Oracle(s)
    o.run_to(s.params.nt)
    jo = o.flux(1)
    o.close()
    jt = TorchFDTD(s).run(torch.tensor(structure_of(s))).numpy()
    assert np.abs(jo).max() > 0
    assert rel(jt, jo) < 1e-12, rel(jt, jo)


def test_torch_restatement_sees_the_structure(tmp_path):
    """Negative control: the restatement's flux moves when one design cell
    does (so the gradient it gives is not trivially zero)."""
    s = make_setup(tmp_path, "slab_ipol2")
    w = structure_of(s)
    j0 = TorchFDTD(s).run(torch.tensor(w)).numpy()
    w[3, 2, 4] *= 1.1
    j1 = TorchFDTD(s).run(torch.tensor(w)).numpy()
    assert rel(j1, j0) > 1e-6


# ---- GPU --------------------------------------------------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("case", ["slab_ipol2", "bloch_lossy_ipol1", "grid16"])
def test_adjoint_matches_autograd(tmp_path, case):
    print(f"[gpu-test] adjoint vs autograd {case}", flush=True)
    s = make_setup(tmp_path, case)
    w = weights_for(s.params)
    g_ref, j_ref = torch_gradient(s, w)
    g, j = engine_gradient(s, w)
    assert rel(j, j_ref) < 1e-10
    assert np.abs(g_ref).max() > 0
    assert rel(g, g_ref) < 1e-9, rel(g, g_ref)


@pytest.mark.gpu
def test_adjoint_finite_differences(tmp_path):
    """d(sum T)/d eps_inv through the torch binding against central
    differences of the plain FP64 engine run on 16 random cells."""
    print("[gpu-test] adjoint vs finite differences", flush=True)
    s = make_setup(tmp_path, "slab_ipol2")
    w0 = structure_of(s).real
    wt = torch.tensor(w0, requires_grad=True)
    T = transmission(flux_fn(wt, s, precision=8, segment=13)).sum()
    T.backward()
    g = wt.grad.numpy()

    def objective(w):
        with Engine(with_structure(s, w), precision=8) as eng:
            eng.run_to(s.params.nt)
            j = eng.flux(1)
        return (j[:, 5] / (j[:, 4] + j[:, 5])).sum()

    assert abs(objective(w0) - T.item()) < 1e-13
    rng = np.random.default_rng(16)
    cells = [tuple(rng.integers(0, n) for n in w0.shape) for _ in range(16)]
    h = 1e-4
    fd = []
    for c in cells:
        wp, wm = w0.copy(), w0.copy()
        wp[c] += h
        wm[c] -= h
        fd.append((objective(wp) - objective(wm)) / (2 * h))
    fd = np.array(fd)
    ga = np.array([g[c] for c in cells])
    scale = np.abs(ga).max()
    assert scale > 0
    assert np.abs(fd - ga).max() < 1e-6 * scale, (fd, ga)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [8, 4])
def test_adjoint_segment_and_linearity(tmp_path, precision):
    """Bitwise the same gradient for any snapshot interval; linear in the
    weights; the state at nt comes back after the adjoint."""
    print(f"[gpu-test] adjoint segments precision={precision}", flush=True)
    s = make_setup(tmp_path, "slab_ipol2")
    p = s.params
    w1, w2 = weights_for(p, 1), weights_for(p, 2)
    grads = []
    for seg in (1, 7, p.nt):
        g, _ = engine_gradient(s, w1, precision, seg)
        grads.append(g)
    assert np.array_equal(grads[0], grads[1]) and np.array_equal(grads[0], grads[2])
    with Engine(s, precision=precision) as eng:
        eng.checkpoint(11)
        eng.run_to(p.nt)
        e0, h0, j0 = eng.field(0), eng.field(1), eng.flux(1)
        ga = eng.adjoint(1, w1)
        gb = eng.adjoint(1, w2)
        gc = eng.adjoint(1, w1 + 2.0 * w2)
        assert np.array_equal(ga, grads[0])
        assert np.array_equal(eng.field(0), e0) and np.array_equal(eng.field(1), h0)
        assert np.array_equal(eng.flux(1), j0)
        assert eng.it == p.nt
    # exact up to the rounding of the field precision
    assert rel(gc, ga + 2.0 * gb) < (1e-10 if precision == 8 else 1e-5)


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["slab_ipol2", "bloch_lossy_ipol1", "two_step"])
def test_fused_adjoint_matches_part_kernels(tmp_path, monkeypatch, case):
    """The fused adjoint sweep (default) against the one-kernel-per-part
    adjoint (IRIS_ADJ_FUSED=0), which forms the wraps' transposes as ghost
    deposits instead of ghost images: equal up to rounding order."""
    print(f"[gpu-test] fused adjoint vs part kernels {case}", flush=True)
    s = make_setup(tmp_path, case)
    w = weights_for(s.params)
    for prec, tol in ((8, 1e-12), (4, 1e-5)):
        g_f, _ = engine_gradient(s, w, prec, 9)
        monkeypatch.setenv("IRIS_ADJ_FUSED", "0")
        g_p, _ = engine_gradient(s, w, prec, 9)
        monkeypatch.delenv("IRIS_ADJ_FUSED")
        assert np.abs(g_p).max() > 0
        assert rel(g_f, g_p) < tol, (prec, rel(g_f, g_p))


@pytest.mark.gpu
def test_adjoint_fp32_close_to_fp64(tmp_path):
    print("[gpu-test] adjoint fp32 vs fp64", flush=True)
    s = make_setup(tmp_path, "grid16")
    w = weights_for(s.params)
    g8, _ = engine_gradient(s, w, 8)
    g4, _ = engine_gradient(s, w, 4)
    assert rel(g4, g8) < 1e-4, rel(g4, g8)


@pytest.mark.gpu
def test_adjoint_with_two_step_forward(tmp_path, monkeypatch):
    """On a grid where the forward run takes the two-step pass, the gradient
    equals the one of a single-step forward run bitwise."""
    print("[gpu-test] adjoint two-step forward", flush=True)
    monkeypatch.setenv("IRIS_TWO_STEP_UPML", "1")   # the UPML slabs in the two-step pass too
    s = make_setup(tmp_path, "two_step")
    w = weights_for(s.params)
    for prec in (4, 8):
        monkeypatch.setenv("IRIS_TWO_STEP", "0")
        with Engine(s, precision=prec) as eng:
            assert not eng.pass_info()["two_step"]
        g1, _ = engine_gradient(s, w, prec, 17)
        monkeypatch.delenv("IRIS_TWO_STEP")
        for width in ("64", "32"):  # tiles of 64 or 32 cells a row (plain and UPML variant)
            monkeypatch.setenv("IRIS_TILE_W2", width)
            monkeypatch.setenv("IRIS_TILE_W2P", width)
            with Engine(s, precision=prec) as eng:
                assert eng.pass_info()["two_step"]
            g2, _ = engine_gradient(s, w, prec, 17)
            assert np.array_equal(g1, g2), width
        monkeypatch.delenv("IRIS_TILE_W2")
        monkeypatch.delenv("IRIS_TILE_W2P")
        assert np.abs(g1).max() > 0


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [8, 4])
def test_two_step_adjoint_sweep(tmp_path, monkeypatch, precision):
    """With only the z-faces weighted (a transmission objective) the
    gradient sweeps advance the plain planes two steps per pass, forward and
    adjoint.  The result is bitwise that of single steps (snapshot interval 1,
    or no two-step pass at all) and matches the one-kernel-per-part adjoint."""
    print(f"[gpu-test] two-step adjoint sweep precision={precision}", flush=True)
    monkeypatch.setenv("IRIS_TWO_STEP_UPML", "1")   # the UPML slabs in the two-step pass too
    s = make_setup(tmp_path, "two_step")
    p = s.params
    w = np.zeros((p.nof, 6))
    w[:, 4:] = weights_for(p)[:, 4:]
    g1, _ = engine_gradient(s, w, precision, 1)
    with Engine(s, precision=precision) as eng:
        eng.checkpoint(17)
        eng.run_to(p.nt)
        eng.profile(True)
        g17 = eng.adjoint(1, w)
        prof = eng.profile_read(reset=True)
    assert prof["yee2"][1] > 0, "the recomputation did not take the two-step pass"
    gnt, _ = engine_gradient(s, w, precision, p.nt)
    assert np.array_equal(g1, g17) and np.array_equal(g1, gnt)
    monkeypatch.setenv("IRIS_TWO_STEP", "0")
    g_single, _ = engine_gradient(s, w, precision, 17)
    monkeypatch.setenv("IRIS_ADJ_FUSED", "0")
    g_part, _ = engine_gradient(s, w, precision, 17)
    monkeypatch.delenv("IRIS_TWO_STEP")
    monkeypatch.delenv("IRIS_ADJ_FUSED")
    assert np.array_equal(g1, g_single)
    assert np.abs(g1).max() > 0
    assert rel(g1, g_part) < (1e-12 if precision == 8 else 1e-5)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [8, 4])
def test_two_step_adjoint_with_bloch_phase(tmp_path, monkeypatch, precision):
    """Oblique incidence with the matching Bloch phase on a grid the two-step
    pass covers: the gradient through two-step sweeps (forward, recomputation
    and adjoint) equals that of single steps up to rounding, and the latter
    matches the one-kernel-per-part adjoint."""
    from pyiris.sweep import oblique
    print(f"[gpu-test] two-step adjoint bloch precision={precision}", flush=True)
    monkeypatch.setenv("IRIS_TWO_STEP_UPML", "1")   # the UPML slabs in the two-step pass too
    # 700 steps: the pulse has crossed the structure and reached both flux
    # faces (the objective's weights sit on J5 and J6)
    s = oblique(make_setup(tmp_path, "two_step", nt=700, no_downloads=700), 30.0)
    p = s.params
    w = np.zeros((p.nof, 6))
    w[:, 4:] = weights_for(p)[:, 4:]
    with Engine(s, precision=precision) as eng:
        assert eng.pass_info()["two_step"]
        eng.checkpoint(17, 0)
        eng.run_to(p.nt)
        eng.profile(True)
        g2 = eng.adjoint(1, w)
        assert eng.profile_read()["yee2"][1] > 0
    monkeypatch.setenv("IRIS_TWO_STEP", "0")
    g1, _ = engine_gradient(s, w, precision, 17)
    monkeypatch.setenv("IRIS_ADJ_FUSED", "0")
    gp, _ = engine_gradient(s, w, precision, 17)
    tol = 1e-11 if precision == 8 else 1e-4
    assert np.abs(g1).max() > 0
    assert rel(g2, g1) < tol, rel(g2, g1)
    assert rel(g1, gp) < tol


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [8, 4])
def test_recorded_tail(tmp_path, precision):
    """H recorded during the forward run for its last steps (some, or all)
    gives bitwise the gradient of pure segment-wise recomputation, with the
    forward flux unchanged."""
    print(f"[gpu-test] recorded tail precision={precision}", flush=True)
    s = make_setup(tmp_path, "two_step")
    p = s.params
    w = np.zeros((p.nof, 6))
    w[:, 4:] = weights_for(p)[:, 4:]
    g0, j0 = engine_gradient(s, w, precision, 17)
    for steps in (1, 58, 120, p.nt, -1):
        with Engine(s, precision=precision) as eng:
            if steps > 0:
                eng.checkpoint(17, 0)
                rstep = eng.checkpoint_info()["record_step_bytes"]
                eng.checkpoint(17, steps * rstep)
            else:
                eng.checkpoint(17, -1)        # as many as fit: the whole run here
            info = eng.checkpoint_info()
            want = p.nt if steps < 0 else steps
            assert info["recorded_steps"] == want
            assert info["first_recorded"] == p.nt - want
            assert info["snapshots"] == -(-(p.nt - want) // 17)
            eng.run_to(p.nt)
            assert np.array_equal(eng.flux(1), j0)
            g = eng.adjoint(1, w)
        assert np.array_equal(g, g0), steps


def test_metasurface_voxels(tmp_path):
    """metasurface() draws one eps_r per block^3 voxel, uniform in the range."""
    from pyiris.sweep import metasurface
    kw = dict(ixmax=10, iymax=7, izmax=64, pmlzi=8, pmlzf=57, hi=26, dh=6, dx=0, eps_bbox=2.0, tfsf_o=14,
              box=(1, 10, 1, 7, 12, 48), nof=2, nt=20, no_downloads=20)
    d = tmp_path / "m"
    d.mkdir()
    s = metasurface(load_input(write_input(d, inputs.render(**kw))), block=4, eps_range=(1.0, 6.0), seed=5)
    eps = 1.0 / structure_of(s).real
    assert eps.shape == (6, 7, 10)
    assert eps.min() >= 1.0 and eps.max() <= 6.0
    for z0 in (0, 4):
        for y0 in (0, 4):
            for x0 in (0, 4, 8):
                blk = eps[z0:z0 + 4, y0:y0 + 4, x0:x0 + 4]
                assert np.all(blk == blk.flat[0])
    assert len(np.unique(eps)) == 2 * 2 * 3


@pytest.mark.gpu
def test_angle_averaged_gradient_finite_differences(tmp_path):
    """Angle-averaged gradient (three incidence angles, each with its Bloch
    phase) of the summed transmission of a voxel metasurface, against central
    differences of the averaged objective at two cells (FP64 fields)."""
    from pyiris.adjoint import structure_of, with_structure
    from pyiris.sweep import angle_averaged_gradient, metasurface, oblique
    print("[gpu-test] angle-averaged gradient", flush=True)
    kw = dict(ixmax=16, iymax=16, izmax=96, pmlzi=10, pmlzf=86, hi=40, dh=8, dx=0, eps_bbox=2.0, tfsf_o=14,
              box=(1, 16, 1, 16, 14, 80), nof=3, nt=260, no_downloads=260, lno=500.0, dlambda=150.0)
    d = tmp_path / "meta"
    d.mkdir()
    s = metasurface(load_input(write_input(d, inputs.render(**kw))), block=4, seed=5)
    eps = 1.0 / structure_of(s).real
    assert eps.min() >= 1.0 and eps.max() <= 6.0 and np.array_equal(eps[0, :4, :4], np.full((4, 4), eps[0, 0, 0]))
    thetas = (0.0, 20.0, 40.0)
    obj, g = angle_averaged_gradient(s, thetas, precision=8, record_bytes=0)

    def objective(w):
        tot = 0.0
        for th in thetas:
            with Engine(oblique(with_structure(s, w), th), precision=8) as eng:
                eng.run_to(s.params.nt)
                j = eng.flux(1)
            tot += (j[:, 5] / (j[:, 4] + j[:, 5])).sum()
        return tot / len(thetas)

    w0 = structure_of(s).real
    assert abs(objective(w0) - obj) < 1e-12 * abs(obj)
    h = 1e-4
    for cell in [(2, 5, 7), (6, 11, 3)]:
        wp, wm = w0.copy(), w0.copy()
        wp[cell] += h
        wm[cell] -= h
        fd = (objective(wp) - objective(wm)) / (2 * h)
        assert abs(fd - g[cell]) < 1e-5 * np.abs(g).max() + 1e-4 * abs(fd), (cell, fd, g[cell])
    assert np.abs(g).max() > 0


@pytest.mark.gpu
def test_automatic_snapshot_interval(tmp_path):
    """checkpoint(-1) picks an interval of at least ceil(sqrt(nt)) and gives
    the gradient of an explicit interval bit for bit."""
    import math
    print("[gpu-test] automatic snapshot interval", flush=True)
    s = make_setup(tmp_path, "two_step")
    p = s.params
    w = weights_for(p)
    g17, _ = engine_gradient(s, w, 4, 17)
    with Engine(s, precision=4) as eng:
        eng.checkpoint(-1, 0)
        info = eng.checkpoint_info()
        seg = info["segment"]
        assert math.ceil(math.sqrt(p.nt)) <= seg <= p.nt
        want = math.sqrt(p.nt * info["snapshot_bytes"] / info["record_step_bytes"])
        assert seg == max(math.ceil(math.sqrt(p.nt)), min(p.nt, math.floor(want)))
        assert info["snapshots"] == -(-p.nt // seg)
        eng.run_to(p.nt)
        g = eng.adjoint(1, w)
    assert np.array_equal(g, g17)


@pytest.mark.gpu
def test_adjoint_call_order(tmp_path):
    print("[gpu-test] adjoint call order", flush=True)
    s = make_setup(tmp_path, "slab_ipol2")
    w = weights_for(s.params)
    with Engine(s, precision=8) as eng:
        eng.run_to(10)
        with pytest.raises(IrisError):
            eng.checkpoint(5)          # after the first step
        eng.run_to(s.params.nt)
        with pytest.raises(IrisError):
            eng.adjoint(1, w)          # no snapshots
    with Engine(s, precision=8) as eng:
        eng.checkpoint(5)
        eng.run_to(50)
        with pytest.raises(IrisError):
            eng.adjoint(1, w)          # not at nt
