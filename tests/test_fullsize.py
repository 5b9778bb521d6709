"""The roofline workload at its own grid and length: FP32 fields against FP64.

512 x 512 x 1024 cells, the slit-array film (eps_r 4.0, 200 nm, one 100 nm
slit per 2560 nm x-period), normal-incidence TF/SF pulse over 400-1000 nm,
DFT at 100 wavelengths on the full-lateral box, all 20000 steps, fluxes at
the reference's output cadence (every 2000 steps).  The FP64-field run is
pinned to the reference at reduced sizes (goldens, oracle); here the
FP32-field run with FP64 DFT accumulators must give T(lambda) =
J6/(J5+J6) and R(lambda) = J5/(J5+J6) within 1e-5 (absolute) of it at every
output block whose flux is live (J5+J6 above 1e-3 of its largest value).

The FP64 run takes about 2.5 minutes, the FP32 run about 1.5; each test
prints its progress at every output block (past pytest's capture, so that a
long run is visibly alive).
"""
import numpy as np
import pytest

from conftest import run_with_cadence

pytestmark = pytest.mark.gpu

_RESULTS = {}


def _run(precision, request):
    if precision in _RESULTS:
        return _RESULTS[precision]
    import bench
    from pyiris import Engine
    capman = request.config.pluginmanager.getplugin("capturemanager")

    def say(msg):
        with capman.global_and_fixture_disabled():
            print(f"[gpu-test] full size fp{32 if precision == 4 else 64}: {msg}", flush=True)

    s = bench.make_setup(precision)
    p = s.params
    say(f"{p.nx}x{p.ny}x{p.nz}, {p.nt} steps, nof={p.nof}")

    class Loud:
        def __init__(self, eng):
            self.eng = eng

        def step(self, a, b):
            self.eng.step(a, b)

        def flux(self, b):
            j = self.eng.flux(b)
            say(f"step {self.eng.it}")
            return j

    with Engine(s, precision=precision) as eng:
        res = run_with_cadence(Loud(eng), p, 1)[1]
    jab = np.stack([j for _, j in res])
    assert np.all(np.isfinite(jab))
    _RESULTS[precision] = jab
    return jab


def _tr(jab):
    den = jab[..., 4] + jab[..., 5]
    return jab[..., 4] / den, jab[..., 5] / den, den


@pytest.mark.timeout(900)
def test_fullsize_fp64_run(request):
    jab = _run(8, request)
    R, T, den = _tr(jab)
    assert jab.shape == (10, 100, 6)
    live = den > 1e-3 * den.max()
    assert live[-1].all()                          # the last block is live at every wavelength
    assert np.all((T[-1] > 0) & (T[-1] < 1.05))    # lossless film: T, R in [0, 1] up to TF/SF leakage


@pytest.mark.timeout(900)
def test_fullsize_fp32_tr_within_1e5_of_fp64(request):
    j64 = _run(8, request)
    j32 = _run(4, request)
    R8, T8, den = _tr(j64)
    R4, T4, _ = _tr(j32)
    live = den > 1e-3 * den.max()
    dR = float(np.abs(R4 - R8)[live].max())
    dT = float(np.abs(T4 - T8)[live].max())
    print(f"[gpu-test] full size: max |dR| {dR:.3e}, max |dT| {dT:.3e} over {int(live.sum())} live block-wavelengths")
    assert dR <= 1e-5 and dT <= 1e-5, (dR, dT)
