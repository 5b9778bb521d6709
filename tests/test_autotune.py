"""Layout tuner (iris_autotune): it times candidate layouts of the step
passes before the first step and keeps the fastest.  Every layout it
compares gives the same results bit for bit, so a tuned run must equal the
default run exactly, and the handle must come back at step 0 with its
initial state."""
import numpy as np
import pytest

from conftest import write_input
from pyiris import Engine, IrisError, inputs, load_input
from pyiris.sweep import oblique


def make(tmp_path, name, **kw):
    base = dict(nt=90, no_downloads=30, tfsf_o=20, lno=450.0, dlambda=100.0, nof=3)
    base.update(kw)
    d = tmp_path / name
    d.mkdir()
    return load_input(write_input(d, inputs.render(**base)))


def setups(tmp_path):
    two = make(tmp_path, "two", ixmax=70, iymax=21, izmax=96, pmlzi=10, pmlzf=86, hi=44, dh=10, dx=8,
               eps_bbox=2.25, box=(1, 70, 1, 21, 15, 80))
    return {
        "two_step": two,
        "two_step_bloch": oblique(two, 25.0),
        "single": make(tmp_path, "one", ixmax=20, iymax=12, izmax=80, pmlzi=9, pmlzf=71, hi=36, dh=8, dx=6,
                       eps_bbox=3.0, box=(1, 20, 1, 12, 14, 66)),
    }


def run(s, precision, tune):
    with Engine(s, precision=precision) as eng:
        info = eng.autotune(6) if tune else None
        assert eng.it == 0
        out = dict(info=info, e0=eng.field(0), pass_info=eng.pass_info())
        eng.run_to(s.params.nt)
        out.update(flux=eng.flux(1), e=eng.field(0), h=eng.field(1))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["two_step", "two_step_bloch", "single"])
@pytest.mark.parametrize("precision", [8, 4])
def test_tuned_run_is_bitwise_the_default_run(tmp_path, case, precision):
    print(f"[gpu-test] autotune {case} precision={precision}", flush=True)
    s = setups(tmp_path)[case]
    ref = run(s, precision, False)
    tuned = run(s, precision, True)
    info = tuned["info"]
    assert info["layouts"] >= 3 and info["ns_per_step"] > 0
    assert info["two_step"] == tuned["pass_info"]["two_step"]
    assert info["zchunk"] == tuned["pass_info"]["zchunk"]
    if case == "two_step_bloch":  # the tuner keeps the two-step choice with a Bloch phase
        assert tuned["pass_info"]["two_step"] == ref["pass_info"]["two_step"]
    assert np.array_equal(tuned["e0"], ref["e0"])  # the initial state is back
    assert np.array_equal(tuned["flux"], ref["flux"])
    assert np.array_equal(tuned["e"], ref["e"]) and np.array_equal(tuned["h"], ref["h"])
    assert np.abs(ref["flux"]).max() > 0


@pytest.mark.gpu
def test_autotune_refuses(tmp_path):
    print("[gpu-test] autotune refuses", flush=True)
    s = setups(tmp_path)["two_step"]
    with Engine(s, precision=4) as eng:
        with pytest.raises(IrisError):
            eng.autotune(1)
        eng.checkpoint(10)
        with pytest.raises(IrisError):
            eng.autotune(4)  # after iris_checkpoint
    with Engine(s, precision=4) as eng:
        eng.run_to(1)
        with pytest.raises(IrisError):
            eng.autotune(4)  # after the first step
