"""Field snapshots (SDL planes): option parsing and the snapshot writer.

The writer (fortran/iris_sdl.f90, the same code the program runs) is fed the
oracle's E after each output step and H of the step before -- the fields
the reference passes (e_cur, h_prev, reference code/driver.f90:468-495) --
and every file it writes must equal, byte for byte, the file the reference
binary wrote for the same input (tests/golden/<case>.npz, keys
"sdl__<file name>").  The oracle's fields are bitwise the reference's, so
this pins the values (modulus/Re/Im/|Re|/|Im|, float32 normalisation), the
copy layout, the record framing and the file names.
"""
import os

import numpy as np
import pytest

from conftest import golden_path, load_golden, write_input
import pyoracle
from pyiris import IrisError, inputs, load_input
from pyiris.host import sdl_write
from pyiris.sdl import read_plane, read_record

SDL_CASES = ["sdl3d", "sdl3d_hmod", "example_1"]


def test_example_1_options_parsed():
    s = load_input(golden_path("example_1", "dat"))
    c = s.extra["sdl"]
    assert c["enabled"] and c["plane_on"] == [0, 1, 0]
    assert (c["copy_flag"], c["copy_no"], c["module_e"]) == (1, 300, 1)
    assert c["fields"] == [0] * 6 and c["nlayers"] == [1, 1, 1]
    # Y plane y=1 over x 1..1, z 75..325 (example_1/inputFDTD.dat:66-72)
    assert c["planes"] == {(2, 1): (1, 1, 1, 75, 325)}


def test_extra_layers_parsed():
    s = load_input(golden_path("sdl3d", "dat"))
    c = s.extra["sdl"]
    assert c["nlayers"] == [2, 1, 2]
    assert c["planes"][(1, 1)] == (3, 1, 6, 5, 40)
    assert c["planes"][(1, 2)] == (5, 1, 6, 5, 40)
    assert c["planes"][(2, 1)] == (4, 1, 8, 10, 30)
    assert c["planes"][(3, 1)] == (30, 1, 8, 1, 6)
    assert c["planes"][(3, 2)] == (20, 1, 8, 1, 6)


def test_no_snapshots_when_planes_off():
    s = load_input(golden_path("small3d_ipol2", "dat"))
    assert not s.extra["sdl"]["enabled"] and s.extra["sdl"]["planes"] == {}


@pytest.mark.parametrize("opts, match", [
    (dict(xi=(9, 1, 1), xf=(9, 1, 1)), "outside the grid"),        # x=9 on an 8-wide grid
    (dict(zi=(5, 1, 1), zf=(70, 1, 1)), "outside the grid"),       # z past nz
    (dict(yi=(4, 1, 1), yf=(2, 1, 1)), "outside the grid"),        # reversed range
    (dict(copy=(1, 0)), "SDL_copy_no"),
    (dict(layers=((1, 1, 1), [7], [0], [0])), "outside the grid"),  # x=3+7
])
def test_bad_snapshot_options_rejected(tmp_path, opts, match):
    base = dict(xi=(3, 1, 1), xf=(3, 1, 1), yi=(1, 1, 1), yf=(6, 1, 1), zi=(5, 1, 1), zf=(40, 1, 1))
    base.update(opts)
    text = inputs.render(ixmax=8, iymax=6, izmax=64, pmlzi=8, pmlzf=56, hi=28, dh=6, dx=3, tfsf_o=14,
                         box=(1, 8, 1, 6, 10, 50), nof=2, nt=10, no_downloads=2, sdl=(1, 0, 0), sdl_opts=base)
    with pytest.raises(IrisError, match=match):
        load_input(write_input(tmp_path, text))


def test_missing_snapshot_section_rejected(tmp_path):
    # the reference would search for the section forever (code/sdl.f90:872-874)
    text = inputs.render(ixmax=4, iymax=4, izmax=48, pmlzi=8, pmlzf=40, hi=20, dh=4, tfsf_o=12,
                         box=(1, 4, 1, 4, 10, 30), nof=2, nt=10, no_downloads=2, sdl=(0, 0, 1))
    with pytest.raises(IrisError, match="Time evolution options"):
        load_input(write_input(tmp_path, text))


def plane_fields(field, layer, geo):
    """(3, nb, na) slice of a field array (nz+2, ny+2, nx+2, 3) on a plane."""
    fixed, a0, a1, b0, b1 = geo
    if layer == 1:
        v = field[b0:b1 + 1, a0:a1 + 1, fixed, :]
    elif layer == 2:
        v = field[b0:b1 + 1, fixed, a0:a1 + 1, :]
    else:
        v = field[fixed, b0:b1 + 1, a0:a1 + 1, :]
    return np.moveaxis(v, -1, 0)


def oracle_snapshots(name, outdir):
    """Step the oracle through the reference's output cadence and write the
    snapshot files of every output step into outdir."""
    path = golden_path(name, "dat")
    s = load_input(path)
    cfg = s.extra["sdl"]
    p = s.params
    o = pyoracle.Oracle(s)
    cwd = os.getcwd()
    os.chdir(outdir)
    try:
        nos, it = 1, 1
        while it <= p.nt:
            cand = p.to_fourier + p.no_downloads * nos
            stop = min(cand, p.nt) if cand >= it else p.nt
            if stop >= p.to_fourier and (stop == cand or stop == p.nt):
                if stop - 1 >= it:
                    o.step(it, stop - 1)
                h_prev = o.field(1)
                o.step(stop, stop)
                e = o.field(0)
                for (layer, il), geo in sorted(cfg["planes"].items()):
                    sdl_write(path, layer, il, nos + 1, plane_fields(e, layer, geo), plane_fields(h_prev, layer, geo))
                nos += 1
            else:
                o.step(it, stop)
            it = stop + 1
    finally:
        os.chdir(cwd)
        o.close()


@pytest.mark.parametrize("name", SDL_CASES)
def test_writer_reproduces_reference_files(tmp_path, name):
    g = load_golden(name)
    want = {k[len("sdl__"):]: bytes(v) for k, v in g.items() if k.startswith("sdl__")}
    assert want, "golden has no snapshot files"
    oracle_snapshots(name, tmp_path)
    got = {fn: open(tmp_path / fn, "rb").read() for fn in os.listdir(tmp_path)}
    assert sorted(got) == sorted(want)
    for fn in sorted(want):
        assert got[fn] == want[fn], f"{name}: {fn} differs"


def test_reader_frames_records():
    g = load_golden("sdl3d")
    raw = bytes(g["sdl__Z_SDL30layer_E2.e"])
    import tempfile
    with tempfile.NamedTemporaryFile(suffix=".e", delete=False) as f:
        f.write(raw)
    try:
        v, _ = read_record(f.name)
        # Z plane 8 x 6 with two copies along both axes: 16 x 12
        assert v.size == 16 * 12
        img = read_plane(f.name, 16, 12)
        np.testing.assert_array_equal(img[:6, :8], img[6:, 8:])
        np.testing.assert_array_equal(img[:6, :8], img[:6, 8:])
        assert np.all(img >= 0) and img.max() > 0
    finally:
        os.unlink(f.name)
