"""iris_get_plane_noalloc: the allocation-free form of iris_get_plane.

Through the C ABI, against iris_get_plane on the same handle: bitwise the
same values for E, H of the previous step and H on planes of all three
axes, two calls in a row with the second smaller, on both precisions and on
a second handle of the same device; planes above its fixed staging array
are refused with an error before any device work.
"""
import numpy as np
import pytest

from conftest import write_input
from pyiris import Engine, IrisError, inputs, load_input

pytestmark = pytest.mark.gpu


def _setup(tmp_path, name="in.dat", **kw):
    base = dict(ixmax=40, iymax=30, izmax=80, pmlzi=10, pmlzf=70, hi=35, dh=8, dx=3, eps_bbox=3.0, tfsf_o=20,
                box=(2, 38, 1, 30, 14, 64), nof=3, nt=80, no_downloads=80, lno=450.0, dlambda=100.0)
    base.update(kw)
    return load_input(write_input(tmp_path, inputs.render(**base), name))


@pytest.mark.parametrize("precision", [8, 4])
def test_plane_noalloc_is_bitwise_the_plane(tmp_path, precision):
    print(f"[gpu-test] plane noalloc precision={precision}", flush=True)
    s = _setup(tmp_path)
    cases = [  # (what, axis, index, a, b): a large plane first, then smaller ones
        (Engine.PLANE_E + 0, 3, 40, (1, 40), (1, 30)),
        (Engine.PLANE_H_PREV + 1, 1, 5, (3, 10), (20, 50)),
        (Engine.PLANE_H + 1, 2, 17, (4, 4), (1, 80)),
    ]
    with Engine(s, precision=precision) as eng:
        eng.run_to(60)
        for c in cases:
            a = eng.plane_noalloc(*c)
            b = eng.plane(*c)
            assert a.shape == b.shape
            assert np.array_equal(a, b), c
            assert np.abs(b).max() > 0, c
        eng.run_to(70)   # the array keeps the last call's values: a new step, a new read
        a, b = eng.plane_noalloc(*cases[1]), eng.plane(*cases[1])
        assert np.array_equal(a, b)
    with Engine(_setup(tmp_path, "b.dat", eps_bbox=2.0), precision=precision) as e2:   # same device, same array
        e2.run_to(30)
        assert np.array_equal(e2.plane_noalloc(*cases[0]), e2.plane(*cases[0]))


def test_plane_noalloc_refuses_planes_above_its_array(tmp_path):
    print("[gpu-test] plane noalloc limit", flush=True)
    s = _setup(tmp_path, ixmax=4, iymax=1030, izmax=1030, pmlzi=10, pmlzf=1020, hi=500, dh=8, dx=0,
               box=(1, 4, 1, 1030, 14, 1010), nt=4, no_downloads=4)
    with Engine(s, precision=4) as eng:
        eng.run_to(2)
        big = (Engine.PLANE_E + 0, 1, 2, (1, 1030), (1, 1030))     # 1030 * 1030 > 2^20 values
        with pytest.raises(IrisError, match="exceed"):
            eng.plane_noalloc(*big)
        ok = (Engine.PLANE_E + 0, 1, 2, (1, 1024), (1, 1024))      # exactly 2^20
        assert np.array_equal(eng.plane_noalloc(*ok), eng.plane(*ok))
