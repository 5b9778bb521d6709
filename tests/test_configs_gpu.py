"""GPU checks of the sweep and gradient workloads at their own sizes, and of
the oblique source against the oracle.

  - oblique incidence (the matched incident field of iris_incident_table):
    the FP64 engine against the oracle driven by the same table, on a
    single-step and a two-step grid; fields <= 1e-12, fluxes <= 1e-10;
  - the batched Bloch sweep of 64 random gratings (nx, ny in 32..128,
    nz = 768): one batch bitwise each run alone over 200 steps, and the two
    smallest drawn gratings (FP64, 5 wavelengths, 600 steps) against the
    oracle;
  - the gradient workload (a 128 x 128 x 64 random design in a 128 x 128 x
    768 grid, FP64, 3600 of its 12000 steps): dT/d eps against central
    differences (h = 1e-4 on eps) on 16 random design cells;
  - the angle-averaged gradient at 256 x 256 x 1024 (two angles, FP32): each
    angle's engine is gone before the next is made (device memory at every
    engine creation the same), and the average equals the mean of the
    angles' gradients taken one by one;
  - a batch refuses to step after one of its handles was retuned.
"""

import numpy as np
import pytest

from conftest import oracle_for, write_input
from pyiris import Batch, Engine, IrisError, inputs, load_input
from pyiris.sweep import grating_sweep, oblique

pytestmark = pytest.mark.gpu


def _announce(msg):
    print(f"[gpu-test] {msg}", flush=True)


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


OBLIQUE_CASES = {
    "single_step": dict(ixmax=9, iymax=7, izmax=80, pmlzi=10, pmlzf=70, hi=35, dh=8, dx=3, eps_bbox=3.0, tfsf_o=20,
                        eps3=1.5, box=(2, 8, 1, 7, 14, 64), nof=5, nt=240, no_downloads=80, lno=450.0, dlambda=100.0),
    "two_step": dict(ixmax=70, iymax=21, izmax=96, pmlzi=10, pmlzf=86, hi=44, dh=10, dx=8, eps_bbox=2.25, tfsf_o=20,
                     box=(1, 70, 1, 21, 15, 80), nof=4, nt=200, no_downloads=200, lno=450.0, dlambda=100.0),
}


@pytest.mark.parametrize("i_pol", [1, 2])
@pytest.mark.parametrize("case", list(OBLIQUE_CASES))
def test_oblique_engine_matches_oracle(tmp_path, case, i_pol):
    _announce(f"oblique engine vs oracle {case} i_pol={i_pol}")
    s = oblique(load_input(write_input(tmp_path, inputs.render(i_pol=i_pol, **OBLIQUE_CASES[case]))), 35.0)
    p = s.params
    o = oracle_for(s)
    o.run_to(p.nt)
    with Engine(s, precision=8) as eng:
        eng.run_to(p.nt)
        if case == "two_step":
            assert eng.pass_info()["two_step"]
        inner = (slice(1, -1),) * 3
        for w in (0, 1):
            a, b = eng.field(w), o.field(w)
            assert np.abs(b[inner]).max() > 0
            assert _rel(a[inner], b[inner]) <= 1e-12, (w, _rel(a[inner], b[inner]))
        jo = o.flux(1)
        assert np.abs(eng.flux(1) - jo).max() <= 1e-10 * np.abs(jo).max()
        for pair in (1, 2, 3):
            assert _rel(eng.dft(1, pair), o.dft(1, pair)) <= 1e-10
    with Engine(s, precision=4) as eng:   # FP32 fields on the same table
        eng.run_to(p.nt)
        assert _rel(eng.field(0)[inner], o.field(0)[inner]) <= 1e-4
    o.close()


def test_grating_sweep_of_64_batched_is_each_alone(monkeypatch):
    """Config-3 sweep: 64 gratings with their Bloch phases and matched oblique
    sources, FP32, 200 steps: the batch's fluxes and fields equal each run's
    alone bit for bit, the run alone taking its UPML slabs into the two-step
    pass as the batch does (with a Bloch phase that rounds differently from
    single steps; the batch's long z-chunks change no result)."""
    _announce("64-grating sweep: batch vs alone")
    setups, drawn = grating_sweep(n=64, seed=3, nz=768, nt=200, nof=50, precision=4)
    assert len({(nx, ny) for nx, ny, _ in drawn}) > 40
    ref = []
    monkeypatch.setenv("IRIS_TWO_STEP_UPML", "1")
    for s in setups:
        with Engine(s, precision=4) as eng:
            eng.run_to(100)
            eng.flux(1)
            eng.run_to(200)
            ref.append((eng.flux(1), eng.field(0)))
    monkeypatch.delenv("IRIS_TWO_STEP_UPML")
    engines = [Engine(s, precision=4) for s in setups]
    try:
        with Batch(engines) as b:
            assert b.info()["two_step"] >= 60
            b.step(1, 100)
            for e in engines:
                e.flux(1)
            b.step(101, 200)
        for e, (j, f) in zip(engines, ref):
            assert np.array_equal(e.flux(1), j)
            assert np.array_equal(e.field(0), f)
        assert all(np.abs(f).max() > 0 for _, f in ref)
    finally:
        for e in engines:
            e.close()


def test_grating_sweep_instances_match_oracle():
    """Two drawn gratings of the sweep (the two smallest), FP64 fields, 5
    wavelengths, 600 steps, against the oracle with the same source table."""
    _announce("64-grating sweep: two instances vs oracle")
    setups, drawn = grating_sweep(n=64, seed=3, nz=768, nt=600, nof=5, precision=8)
    order = sorted(range(64), key=lambda i: drawn[i][0] * drawn[i][1])
    for i in order[:2]:
        s = setups[i]
        o = oracle_for(s)
        o.run_to(s.params.nt)
        with Engine(s, precision=8) as eng:
            eng.run_to(s.params.nt)
            inner = (slice(1, -1),) * 3
            for w in (0, 1):
                assert _rel(eng.field(w)[inner], o.field(w)[inner]) <= 1e-12, (i, w)
            # judged on the scale of the DFT fields: at 600 steps the z faces
            # hold only the TF/SF plane's leakage (~1e-12 of the side faces)
            ref = {pair: o.dft(1, pair) for pair in (1, 2, 3)}
            scale = max(np.abs(v).max() for v in ref.values())
            for pair in (1, 2, 3):
                assert np.abs(eng.dft(1, pair) - ref[pair]).max() <= 1e-11 * scale, (i, pair)
        o.close()


def test_gradient_workload_finite_differences():
    """Config-4 design (128 x 128 x 64 cells, eps_r U[1,4], seed 4, in a
    128 x 128 x 768 grid), FP64, the first 3600 of its steps (light that
    crossed the design reaches both flux faces from about step 2900):
    d(sum T)/d eps through the torch binding against central differences
    (h = 1e-4 on eps) of the plain engine run on 16 random design cells."""
    import torch
    import bench
    from pyiris.adjoint import flux, structure_of, transmission, with_structure
    _announce("gradient workload: dT/deps vs finite differences")
    nt = 3600
    s0 = bench.make_setup(8, bench.ADJ_WORKLOAD, nt=nt, no_downloads=nt)
    eps = np.random.default_rng(4).uniform(1.0, 4.0, size=structure_of(s0).shape)
    assert eps.shape == (64, 128, 128)
    leaf = torch.tensor(eps, requires_grad=True)
    T = transmission(flux(1.0 / leaf, s0, 8)).sum()
    T.backward()
    g = leaf.grad.numpy()

    def objective(e):
        with Engine(with_structure(s0, 1.0 / e), precision=8) as eng:
            eng.run_to(nt)
            j = eng.flux(1)
        return float((j[:, 5] / (j[:, 4] + j[:, 5])).sum())

    assert abs(objective(eps) - float(T)) < 1e-12 * abs(float(T))
    rng = np.random.default_rng(16)
    cells = [tuple(int(rng.integers(0, n)) for n in eps.shape) for _ in range(16)]
    h = 1e-4
    fd = []
    for c in cells:
        ep, em = eps.copy(), eps.copy()
        ep[c] += h
        em[c] -= h
        fd.append((objective(ep) - objective(em)) / (2 * h))
    fd = np.array(fd)
    ga = np.array([g[c] for c in cells])
    _announce(f"gradient workload: max |g| {np.abs(ga).max():.3e}, max |fd - g| {np.abs(fd - ga).max():.3e}")
    assert np.abs(ga).max() > 0
    assert np.abs(fd - ga).max() < 1e-3 * np.abs(ga).max(), (fd, ga)


def _free_device_bytes():
    """Free device memory, read through the HIP runtime the engine's library
    is linked against (libamdhip64.so.7, already loaded with it).  torch
    ships a HIP runtime of its own, and a second runtime started in a
    process where the engine's has started sees no device."""
    import ctypes as ct
    hip = ct.CDLL("libamdhip64.so.7")
    free, total = ct.c_size_t(), ct.c_size_t()
    assert hip.hipMemGetInfo(ct.byref(free), ct.byref(total)) == 0
    return free.value


def test_angle_averaged_gradient_at_full_size(monkeypatch):
    """Config-5 grid (256 x 256 x 1024, voxel metasurface), two angles, FP32,
    4500 steps (light that crossed the design reaches both flux faces from
    about step 3900), 48 GiB of H records per angle: the free device memory
    is the same when each angle's engine is made, the averaged gradient is
    the mean of the two angles' own gradients, and a gradient taken twice is
    the same bit for bit."""
    import bench
    import pyiris.adjoint as adj
    from pyiris.sweep import angle_averaged_gradient, metasurface
    _announce("angle-averaged gradient at 256x256x1024")
    nt = 4500
    s = metasurface(bench.make_setup(4, bench.ANGLE_WORKLOAD, nt=nt, no_downloads=nt), block=32,
                    eps_range=(1.0, 6.0), seed=5)
    seen = []

    class Watched(adj.Engine):
        def __init__(self, *a, **k):
            seen.append(_free_device_bytes())
            super().__init__(*a, **k)

    monkeypatch.setattr(adj, "Engine", Watched)
    thetas = (10.0, 35.0)
    # a fixed record budget: the automatic one (90% of the free memory) moves
    # with the free memory by a record or so, and with a Bloch phase the
    # recorded and the recomputed FP32 steps agree to rounding, not bitwise
    rec = 48 << 30
    obj, g = angle_averaged_gradient(s, thetas, precision=4, record_bytes=rec, wrt="eps")
    assert len(seen) == 2
    assert abs(seen[1] - seen[0]) < 256 << 20, seen      # nothing of the first angle is held
    parts = [angle_averaged_gradient(s, (th,), precision=4, record_bytes=rec, wrt="eps") for th in thetas]
    assert abs(obj - sum(o for o, _ in parts) / 2) < 1e-12 * abs(obj)
    mean = (parts[0][1] + parts[1][1]) / 2
    assert np.abs(mean).max() > 0
    # the average hands each angle half the weights: the FP32 adjoint is
    # linear in them to rounding (2 g(w/2) = g(w) within ~1.5e-7, measured),
    # while the same call twice is bitwise the same (the engine is
    # deterministic)
    assert _rel(g, mean) < 1e-6, _rel(g, mean)
    again = angle_averaged_gradient(s, thetas[:1], precision=4, record_bytes=rec, wrt="eps")
    assert again[0] == parts[0][0] and np.array_equal(again[1], parts[0][1])


def test_batch_refuses_after_autotune(tmp_path):
    """A handle retuned after it joined a batch has new pass lists; the
    batch's tables hold the old ones, so stepping the batch is an error."""
    _announce("batch refuses after autotune")
    kw = OBLIQUE_CASES["two_step"]
    s1 = load_input(write_input(tmp_path, inputs.render(**kw), "a.dat"))
    s2 = load_input(write_input(tmp_path, inputs.render(**dict(kw, eps_bbox=3.0)), "b.dat"))
    with Engine(s1, precision=4) as e1, Engine(s2, precision=4) as e2:
        with Batch([e1, e2]) as b:
            e1.autotune(4)
            with pytest.raises(IrisError, match="changed after iris_batch_create"):
                b.step(1, 10)
        assert e1.it == 0 and e2.it == 0
        with Batch([e1, e2]) as b:   # a batch made after the tuning steps
            b.step(1, 10)
        assert e1.it == 10
