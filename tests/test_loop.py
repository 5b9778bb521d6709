"""<<LOOP>> inputs (reference code/program.f90:34-57).

A first line <<LOOP>> followed by io_main, if_main and three unused deltas
repeats the whole run for i = io_main..if_main with an empty description
(code/main.f90:108-113); every run appends its '#',it blocks to the same
<box>_scatt<j>.dat files (code/dataprocess.f90:279-299, position='append').

The Flang build of the reference completes the first run and then aborts
when the second run reopens log.dat's unit (code/main.f90:115: "ACTION= may
not be changed on an open unit"), so the reference pins the first run only.
The later runs are pinned by the loop's own definition: the index changes
nothing but the description, so each run repeats the first bit for bit.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import PKG, face_rel_err, golden_path, load_golden, run_with_cadence
import pyoracle
from pyiris import load_input


def test_reference_runs_first_iteration_then_aborts():
    g = load_golden("loop_2runs")
    assert bytes(g["ref_abort"]).decode() == "ACTION= may not be changed on an open unit"
    # one run's worth of output steps (nt=80, cadence 40)
    assert g["it_b1"].tolist() == [40, 80]


def test_oracle_reproduces_first_iteration():
    g = load_golden("loop_2runs")
    s = load_input(golden_path("loop_2runs", "dat"))
    o = pyoracle.Oracle(s)
    res = run_with_cadence(o, s.params, 1)
    o.close()
    assert [it for it, _ in res[1]] == g["it_b1"].tolist()
    assert np.array_equal(np.stack([j for _, j in res[1]]), g["jabs_b1"])


@pytest.mark.gpu
def test_program_repeats_the_run(tmp_path):
    exe = os.path.join(PKG, "bin", "irisfdtd")
    shutil.copy(golden_path("loop_2runs", "dat"), tmp_path / "inputFDTD.dat")
    print("[gpu-test] fortran program <<LOOP>> 1..2", flush=True)
    subprocess.run([exe], cwd=tmp_path, check=True, timeout=300, stdout=subprocess.DEVNULL)
    g = load_golden("loop_2runs")
    n = len(g["it_b1"])
    blocks = [pyoracle.read_scatt(str(tmp_path / f"1_scatt{j}.dat")) for j in range(1, 7)]
    its = [it for it, _ in blocks[0]]
    assert its == g["it_b1"].tolist() * 2
    jab = np.stack([np.stack([v for _, v in blk]) for blk in blocks], axis=-1)
    assert face_rel_err(jab[:n], g["jabs_b1"]) <= 1e-10
    assert np.array_equal(jab[:n], jab[n:]), "the second run differs from the first"
    tot = np.stack([v for _, v in pyoracle.read_scatt(str(tmp_path / "1_scatt.dat"))])
    assert np.array_equal(tot[:n], tot[n:])
    np.testing.assert_allclose(tot[:n], g["jtot_b1"], rtol=1e-9, atol=1e-20)
    log = open(tmp_path / "log.dat").read()
    assert "Description     :" in log and "<<LOOP>>" not in log


@pytest.mark.gpu
@pytest.mark.parametrize("io, if_, runs", [(3, 3, 1), (4, 2, 0)])
def test_program_loop_bounds(tmp_path, io, if_, runs):
    """io_main == if_main runs once; io_main > if_main runs nothing (an empty
    Fortran DO range)."""
    text = open(golden_path("loop_2runs", "dat")).read().split("\n", 6)
    text[1], text[2] = str(io), str(if_)
    (tmp_path / "inputFDTD.dat").write_text("\n".join(text))
    exe = os.path.join(PKG, "bin", "irisfdtd")
    subprocess.run([exe], cwd=tmp_path, check=True, timeout=300, stdout=subprocess.DEVNULL)
    path = tmp_path / "1_scatt5.dat"
    nblocks = len(pyoracle.read_scatt(str(path))) if path.exists() else 0
    assert nblocks == 2 * runs
