"""Steps as a captured HIP graph (iris_graph_info): whole periods of steps
(the DFT sample ring fills and folds once, the field buffers return to their
assignment) run as replays of one captured graph, with the TF/SF step shift
set per replay.  A run with graphs must equal the run with the launches one
by one (IRIS_GRAPH=0) bit for bit, whatever the calls' step ranges, with
eager work and synchronisations between the replays."""
import numpy as np
import pytest

from conftest import write_input
from pyiris import Engine, inputs, load_input
from pyiris.sweep import oblique

pytestmark = pytest.mark.gpu


def make(tmp_path, name, **kw):
    base = dict(nt=700, no_downloads=700, tfsf_o=20, lno=450.0, dlambda=100.0, nof=3)
    base.update(kw)
    d = tmp_path / name
    d.mkdir()
    return load_input(write_input(d, inputs.render(**base)))


def setups(tmp_path):
    two = make(tmp_path, "two", ixmax=70, iymax=21, izmax=96, pmlzi=10, pmlzf=86, hi=44, dh=10, dx=8,
               eps_bbox=2.25, box=(1, 70, 1, 21, 15, 80))
    return {
        "two_step": two,
        "two_step_bloch": oblique(two, 30.0),
        "two_step_cadence": make(tmp_path, "cad", ixmax=70, iymax=21, izmax=96, pmlzi=10, pmlzf=86, hi=44, dh=10,
                                 dx=8, eps_bbox=2.25, box=(1, 70, 1, 21, 15, 80), to_fourier=37, sample_fourier=3,
                                 nt=1400, no_downloads=1400),
        "single": make(tmp_path, "one", ixmax=20, iymax=12, izmax=80, pmlzi=9, pmlzf=71, hi=36, dh=8, dx=6,
                       eps_bbox=3.0, box=(1, 20, 1, 12, 14, 66), mbox=(2, 19, 2, 11, 20, 60)),
    }


def run(s, precision, cuts):
    out = []
    with Engine(s, precision=precision) as eng:
        a = 1
        for c in cuts + [s.params.nt]:
            eng.step(a, c)
            out.append(eng.flux(1))          # synchronises between the calls
            a = c + 1
        out += [eng.field(0), eng.field(1)] + [eng.dft(1, j) for j in (1, 2, 3)]
        if s.params.nbox == 2:
            out.append(eng.flux(2))
        info = eng.graph_info()
    return out, info


@pytest.mark.parametrize("case", ["two_step", "two_step_bloch", "two_step_cadence", "single"])
@pytest.mark.parametrize("precision", [8, 4])
def test_graph_replays_are_bitwise_the_launches(tmp_path, monkeypatch, case, precision):
    print(f"[gpu-test] graph {case} precision={precision}", flush=True)
    s = setups(tmp_path)[case]
    cuts = [250, 251, 333]
    monkeypatch.setenv("IRIS_GRAPH", "0")
    ref, info0 = run(s, precision, cuts)
    assert not info0["captured"] and info0["replays"] == 0
    monkeypatch.setenv("IRIS_GRAPH", "1")
    got, info = run(s, precision, cuts)
    assert info["period"] == (64 if precision == 4 else 32) * s.params.sample_fourier
    assert info["captured"] and info["replays"] >= 2, info
    for k, (a, b) in enumerate(zip(got, ref)):
        assert np.array_equal(a, b), k
    assert np.abs(ref[0]).max() > 0


def test_no_graph_while_profiling(tmp_path):
    print("[gpu-test] graph off while profiling", flush=True)
    s = setups(tmp_path)["two_step"]
    with Engine(s, precision=4) as eng:
        eng.profile(True)
        eng.run_to(s.params.nt)
        assert eng.graph_info()["replays"] == 0
    with Engine(s, precision=4) as eng:
        eng.run_to(s.params.nt)
        assert eng.graph_info()["replays"] >= 1


def _interplay(s, precision, tune, batch_with, poke):
    """Steps 1..nt with a tuner call first, an excursion through a batch, and
    a field write between calls, as chosen; returns the end state."""
    from pyiris import Batch
    with Engine(s, precision=precision) as eng:
        if tune:
            eng.autotune(4)
        eng.step(1, 300)
        if poke:
            e = eng.field(0)
            e[s.params.nz // 2, 3, 4, 0] += 1e-3           # any change the later steps carry
            eng.set_field(0, e)
        if batch_with is not None:
            with Engine(batch_with, precision=precision) as other:
                other.run_to(300)
                with Batch([eng, other]) as b:
                    b.step(301, 340)
        eng.step(eng.it + 1, s.params.nt)
        return [eng.flux(1), eng.field(0), eng.field(1)], eng.graph_info()


@pytest.mark.parametrize("what", ["autotune", "batch", "set_field"])
def test_graph_with_tuner_batch_and_field_writes(tmp_path, monkeypatch, what):
    """A graph captured before iris_autotune's new layout, before a batch
    ran the handle, or before a field write is replaced or replayed correctly:
    bitwise the launches one by one."""
    print(f"[gpu-test] graph interplay {what}", flush=True)
    ss = setups(tmp_path)
    s, other = ss["two_step"], ss["two_step_bloch"]
    kw = dict(tune=what == "autotune", batch_with=other if what == "batch" else None, poke=what == "set_field")
    monkeypatch.setenv("IRIS_GRAPH", "0")
    ref, _ = _interplay(s, 4, **kw)
    monkeypatch.setenv("IRIS_GRAPH", "1")
    got, info = _interplay(s, 4, **kw)
    assert info["replays"] >= 1, info
    for k, (a, b) in enumerate(zip(got, ref)):
        assert np.array_equal(a, b), k


def _batch_run(group, precision, cuts):
    from pyiris import Batch
    engines = [Engine(s, precision=precision) for s in group]
    try:
        out = []
        with Batch(engines) as b:
            a = 1
            for c in cuts + [group[0].params.nt]:
                b.step(a, c)
                out.append([e.flux(1) for e in engines])     # synchronises between the calls
                a = c + 1
        for e in engines:
            out.append([e.field(0), e.field(1)])
        return out, [e.graph_info() for e in engines], [e.pass_info()["two_step"] for e in engines]
    finally:
        for e in engines:
            e.close()


@pytest.mark.parametrize("precision", [8, 4])
def test_batch_graph_replays_are_bitwise_the_launches(tmp_path, monkeypatch, precision):
    """A batch's pairs of steps as replays of one captured graph per batch
    equal the launches one by one, across calls cut at arbitrary steps."""
    print(f"[gpu-test] batch graph precision={precision}", flush=True)
    ss = setups(tmp_path)
    group = [ss["two_step"], ss["two_step_bloch"], ss["single"]]
    cuts = [250, 251, 333]
    monkeypatch.setenv("IRIS_GRAPH", "0")
    ref, info0, _ = _batch_run(group, precision, cuts)
    assert all(i["replays"] == 0 for i in info0)
    monkeypatch.setenv("IRIS_GRAPH", "1")
    got, info, two = _batch_run(group, precision, cuts)
    assert two[0] and two[1]
    assert all(i["replays"] >= 1 for i, t in zip(info, two) if t), info
    for k, (ra, ga) in enumerate(zip(ref, got)):
        for a, b in zip(ra, ga):
            assert np.array_equal(a, b), k


@pytest.mark.parametrize("precision", [8, 4])
def test_graph_with_the_in_place_passes(tmp_path, monkeypatch, precision):
    """The in-place two-pass engine (IRIS_FUSED=0: E pass, H pass, TF/SF
    pass per step, no buffer rotation) replays its periods bitwise too."""
    print(f"[gpu-test] graph in-place passes precision={precision}", flush=True)
    s = setups(tmp_path)["single"]
    monkeypatch.setenv("IRIS_FUSED", "0")
    monkeypatch.setenv("IRIS_GRAPH", "0")
    ref, _ = run(s, precision, [250, 251, 333])
    monkeypatch.setenv("IRIS_GRAPH", "1")
    got, info = run(s, precision, [250, 251, 333])
    assert info["replays"] >= 2, info
    for k, (a, b) in enumerate(zip(got, ref)):
        assert np.array_equal(a, b), k


@pytest.mark.parametrize("seed", range(6))
def test_graph_random_cadences_and_cuts(tmp_path, monkeypatch, seed):
    """Seeded draws of the DFT start step, sampling cadence, precision, grid
    and call cuts: graph replays equal the launches one by one."""
    rng = np.random.default_rng(100 + seed)
    to_fourier = int(rng.integers(0, 90))
    sample = int(rng.integers(1, 4))
    precision = int(rng.choice([4, 8]))
    two = bool(rng.integers(0, 2))
    nt = 400 + 64 * sample * int(rng.integers(2, 5))
    cuts = sorted(int(c) for c in rng.choice(np.arange(5, nt - 5), size=3, replace=False))
    print(f"[gpu-test] graph random seed={seed} to_fourier={to_fourier} sample={sample} prec={precision} "
          f"two={two} nt={nt} cuts={cuts}", flush=True)
    if two:
        s = make(tmp_path, "r", ixmax=70, iymax=21, izmax=96, pmlzi=10, pmlzf=86, hi=44, dh=10, dx=8, eps_bbox=2.25,
                 box=(1, 70, 1, 21, 15, 80), to_fourier=to_fourier, sample_fourier=sample, nt=nt, no_downloads=nt)
        if seed % 2:
            s = oblique(s, 15.0 + 5 * seed)
    else:
        s = make(tmp_path, "r", ixmax=20, iymax=12, izmax=80, pmlzi=9, pmlzf=71, hi=36, dh=8, dx=6, eps_bbox=3.0,
                 box=(1, 20, 1, 12, 14, 66), to_fourier=to_fourier, sample_fourier=sample, nt=nt, no_downloads=nt)
    monkeypatch.setenv("IRIS_GRAPH", "0")
    ref, _ = run(s, precision, cuts)
    monkeypatch.setenv("IRIS_GRAPH", "1")
    got, info = run(s, precision, cuts)
    for k, (a, b) in enumerate(zip(got, ref)):
        assert np.array_equal(a, b), (k, info)


def test_graph_bitwise_at_full_size(monkeypatch):
    """512 x 512 x 1024, FP32: 200 steps with a captured period replayed
    (and the UPML variant, the single-step length groups and the fold inside
    the graph) equal the launches one by one."""
    print("[gpu-test] graph full size", flush=True)
    import bench
    s = bench.make_setup(4, nof=4, nt=200, no_downloads=200)
    out = {}
    for g in ("0", "1"):
        monkeypatch.setenv("IRIS_GRAPH", g)
        with Engine(s, precision=4) as eng:
            eng.run_to(200)
            p = s.params
            out[g] = (eng.flux(1), eng.graph_info(),
                      eng.plane(Engine.PLANE_E + 0, 3, 500, (1, p.nx), (1, p.ny)),
                      eng.plane(Engine.PLANE_H + 1, 2, 256, (1, p.nx), (1, p.nz)))
    assert out["1"][1]["replays"] >= 1, out["1"][1]
    assert np.abs(out["0"][0]).max() > 0
    for k in (0, 2, 3):
        assert np.array_equal(out["0"][k], out["1"][k]), k
