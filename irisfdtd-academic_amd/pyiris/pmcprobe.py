#!/usr/bin/env python3
"""Workload for hardware-counter runs (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE).

1. Calibration: a plain device-to-device copy of a known number of bytes
   (hipMemcpy, which runs the runtime's copy kernel), three times.
2. 40 steps of the 512x512x1024 slit-array workload with FP32 fields and, if
   asked, with FP64 fields (one FP64 DFT fold happens at step 32).

Comparing each kernel's counter with the copy's turns the counters into
bytes: HBM bytes of a kernel = counter(kernel) * copy_bytes / counter(copy).

  rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT -- python3 pmcprobe.py [--fp64]
"""
import ctypes as ct
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

COPY_BYTES = 1 << 31


def copy_calibration():
    hip = ct.CDLL("libamdhip64.so")
    a, b = ct.c_void_p(), ct.c_void_p()
    assert hip.hipMalloc(ct.byref(a), ct.c_size_t(COPY_BYTES)) == 0
    assert hip.hipMalloc(ct.byref(b), ct.c_size_t(COPY_BYTES)) == 0
    assert hip.hipMemset(a, 1, ct.c_size_t(COPY_BYTES)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, ct.c_size_t(COPY_BYTES), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(f"[pmcprobe] copied {COPY_BYTES} bytes x3", flush=True)


def steps(precision, n=40):
    import bench
    from pyiris import Engine
    s = bench.make_setup(precision)
    with Engine(s, precision=precision) as eng:
        eng.run_to(n)
        eng.sync()
    print(f"[pmcprobe] {n} steps precision {precision}", flush=True)


if __name__ == "__main__":
    copy_calibration()
    steps(4)
    if "--fp64" in sys.argv:
        steps(8)
