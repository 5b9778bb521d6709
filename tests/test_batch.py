"""Batched runs (iris_batch_*): many independent grids advanced by one
launch of the fused pass per step.

The batched pass performs, per workgroup, exactly the work of the single-grid
pass, so every instance's fields, DFT accumulators and fluxes must equal those
of stepping that handle alone, bit for bit, whatever mix of grid shapes,
polarisations, Bloch phases, structures and boxes the batch holds (and
whether the handle alone would take the two-step pass).  A batch lays its
handles out with long z-chunks (which change no result) and the UPML slabs in
the two-step pass; with a Bloch phase the latter rounds differently from
single steps, so the tests either run the handles alone with the UPML
variant too (IRIS_TWO_STEP_UPML=1) or keep each handle's own choice in the
batch (IRIS_BATCH_UPML=0).
"""
import numpy as np
import pytest

from conftest import write_input
from pyiris import Batch, Engine, IrisError, inputs, load_input
from pyiris.sweep import box_filter3, grating_sweep, oblique


def make(tmp_path, name, **kw):
    base = dict(nt=150, no_downloads=150, tfsf_o=14, lno=450.0, dlambda=100.0, nof=3)
    base.update(kw)
    d = tmp_path / name
    d.mkdir()
    return load_input(write_input(d, inputs.render(**base)))


def mixed_setups(tmp_path):
    return [
        oblique(make(tmp_path, "a", ixmax=8, iymax=6, izmax=64, pmlzi=8, pmlzf=57, hi=28, dh=8, dx=3, eps_bbox=2.5,
                     box=(1, 8, 1, 6, 12, 50)), 25.0),
        make(tmp_path, "b", i_pol=1, ixmax=7, iymax=6, izmax=64, pmlzi=8, pmlzf=57, hi=26, dh=6, dx=2,
             eps_bbox=3.0 + 0.4j, eps3=1.8, box=(2, 6, 1, 5, 12, 48), mbox=(1, 7, 1, 6, 20, 40), nof=2),
        make(tmp_path, "c", ixmax=70, iymax=21, izmax=96, pmlzi=10, pmlzf=86, hi=44, dh=10, dx=8, eps_bbox=2.25,
             tfsf_o=20, box=(1, 70, 1, 21, 15, 80), nof=4),
        make(tmp_path, "d", ixmax=20, iymax=1, izmax=80, pmlzi=10, pmlzf=70, hi=36, dh=8, dx=5, eps_bbox=4.0,
             box=(1, 20, 1, 1, 14, 60)),
        oblique(make(tmp_path, "e", ixmax=33, iymax=9, izmax=70, pmlzi=9, pmlzf=62, hi=30, dh=7, dx=0, eps_bbox=1.7,
                     box=(3, 30, 2, 8, 13, 55)), 50.0),
        # two-step grids with a Bloch phase (and one more shape without)
        oblique(make(tmp_path, "f", ixmax=70, iymax=21, izmax=96, pmlzi=10, pmlzf=86, hi=44, dh=10, dx=8,
                     eps_bbox=2.25, tfsf_o=20, box=(1, 70, 1, 21, 15, 80), nof=4), 20.0),
        oblique(make(tmp_path, "g", ixmax=130, iymax=40, izmax=80, pmlzi=9, pmlzf=71, hi=36, dh=8, dx=20,
                     eps_bbox=3.0, tfsf_o=18, box=(1, 130, 1, 40, 13, 66), nof=3), 40.0),
        make(tmp_path, "h", ixmax=190, iymax=33, izmax=88, pmlzi=12, pmlzf=76, hi=40, dh=6, dx=10, eps_bbox=2.0,
             tfsf_o=20, box=(1, 190, 1, 33, 15, 72), nof=3),
    ]


def alone(s, precision, steps):
    with Engine(s, precision=precision) as eng:
        for a, b in steps:
            eng.step(a, b)
        return eng.flux(1), (eng.flux(2) if s.params.nbox == 2 else None), eng.field(0), eng.field(1)


def test_box_filter_is_periodic_mean():
    a = np.arange(20.0).reshape(4, 5)
    f = box_filter3(a)
    assert np.isclose(f.mean(), a.mean())
    assert np.isclose(f[0, 0], np.mean([a[i % 4, j % 5] for i in (-1, 0, 1) for j in (-1, 0, 1)]))


def test_grating_sweep_draws(tmp_path):
    setups, drawn = grating_sweep(n=3, seed=3, nz=120, nt=20, nof=2)
    for s, (nx, ny, th) in zip(setups, drawn):
        p = s.params
        assert (p.nx, p.ny, p.nz) == (nx, ny, 120)
        assert 32 <= nx <= 128 and 32 <= ny <= 128 and 0.0 <= th <= 60.0
        assert np.isclose(p.ux, np.sin(np.radians(th)), atol=1e-7)
        eps = 1.0 / s.eps_inv.reshape(p.hf - p.hi + 3, ny + 2, nx + 2)[1:-1, 1:-1, 1:-1].real
        assert eps.min() >= 1.0 and eps.max() <= 4.0
        assert np.array_equal(eps[0], eps[-1])      # a column is uniform through the layer


@pytest.mark.gpu
@pytest.mark.parametrize("width", ["auto", "32", "32/64", "128"])
@pytest.mark.parametrize("upml", ["auto", "1"])
@pytest.mark.parametrize("precision", [8, 4])
def test_batch_is_bitwise_each_run_alone(tmp_path, monkeypatch, precision, upml, width):
    """Single-step grids, and two-step grids with and without a Bloch phase
    (with the UPML slabs in the two-step pass where the engine would put them
    there, or wherever it applies; two-step tiles of 64 or 32 cells a
    row, for the plain pass / its UPML variant)."""
    print(f"[gpu-test] batch vs alone precision={precision} upml={upml} width={width}", flush=True)
    if upml != "auto":
        monkeypatch.setenv("IRIS_TWO_STEP_UPML", upml)
    else:
        # the batch keeps each handle's own use of the UPML variant (by
        # default it takes the slabs wherever the variant applies)
        monkeypatch.setenv("IRIS_BATCH_UPML", "0")
    if width != "auto":
        w, wp = (width.split("/") + [width])[:2]
        monkeypatch.setenv("IRIS_TILE_W2", w)
        monkeypatch.setenv("IRIS_TILE_W2P", wp)
    setups = mixed_setups(tmp_path)
    nt = setups[0].params.nt
    steps = [(1, 37), (38, 38), (39, nt)]
    ref = [alone(s, precision, steps) for s in setups]
    engines = [Engine(s, precision=precision) for s in setups]
    two = [e.pass_info() for e in engines]
    assert sum(t["two_step"] for t in two) >= 3
    if upml == "1":
        assert sum(t["upml_two_step_planes"] > 0 for t in two) >= 3
    try:
        with Batch(engines) as b:
            assert b.info()["two_step"] == sum(t["two_step"] for t in two)
            for a, c in steps:
                b.step(a, c)
        for s, e, r in zip(setups, engines, ref):
            assert e.it == nt
            assert np.array_equal(e.flux(1), r[0])
            if r[1] is not None:
                assert np.array_equal(e.flux(2), r[1])
            assert np.array_equal(e.field(0), r[2]) and np.array_equal(e.field(1), r[3])
            assert np.abs(r[0]).max() > 0
    finally:
        for e in engines:
            e.close()


@pytest.mark.gpu
def test_batch_then_alone(tmp_path, monkeypatch):
    """Handles leave a batch in a state they can go on from alone: the same
    calls alone give the same results bit for bit.  (Where a Bloch phase
    meets the two-step pass, as on grid "a", where the calls cut the pairs
    of steps changes the rounding, so the reference cuts its calls alike.)"""
    print("[gpu-test] batch then alone", flush=True)
    monkeypatch.setenv("IRIS_BATCH_UPML", "0")
    setups = mixed_setups(tmp_path)[:3]
    nt = setups[0].params.nt
    ref = [alone(s, 4, [(1, 61), (62, nt)]) for s in setups]
    engines = [Engine(s, precision=4) for s in setups]
    try:
        b = Batch(engines)
        b.step(1, 61)
        b.close()
        for e, r in zip(engines, ref):
            e.run_to(nt)
            assert np.array_equal(e.flux(1), r[0])
    finally:
        for e in engines:
            e.close()


@pytest.mark.gpu
def test_batch_rejects(tmp_path):
    print("[gpu-test] batch rejects", flush=True)
    setups = mixed_setups(tmp_path)[:2]
    e8 = Engine(setups[0], precision=8)
    e4 = Engine(setups[1], precision=4)
    try:
        with pytest.raises(IrisError):
            Batch([e8, e4])                       # mixed precision
        with pytest.raises(IrisError):
            Batch([e8, e8])                       # the same handle twice
        e4b = Engine(setups[0], precision=4)
        e4.run_to(3)
        b = Batch([e4, e4b])
        with pytest.raises(IrisError):
            b.step(4, 10)                          # e4b is at step 0
        b.close()
        e4b.close()
    finally:
        e8.close()
        e4.close()


@pytest.mark.gpu
def test_tuned_handles_in_a_batch(tmp_path, monkeypatch):
    """Handles whose layouts iris_autotune chose (each its own) batch
    together and still give the fluxes of the default runs alone (the batch
    keeping each handle's own use of the UPML variant)."""
    print("[gpu-test] tuned handles in a batch", flush=True)
    monkeypatch.setenv("IRIS_BATCH_UPML", "0")
    setups = mixed_setups(tmp_path)
    nt = setups[0].params.nt
    ref = [alone(s, 4, [(1, nt)]) for s in setups]
    engines = [Engine(s, precision=4) for s in setups]
    try:
        for e in engines:
            e.autotune(4)
        with Batch(engines) as b:
            b.step(1, nt)
        for e, r in zip(engines, ref):
            assert np.array_equal(e.flux(1), r[0])
            assert np.array_equal(e.field(0), r[2]) and np.array_equal(e.field(1), r[3])
    finally:
        for e in engines:
            e.close()
