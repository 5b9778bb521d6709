"""Reader of the field-snapshot (SDL) files.

Each file holds one Fortran unformatted sequential record of float32 values
(reference code/sdl.f90:474-481): a 4-byte length, the values with the first
free coordinate fastest, the length again.  The reference's own reader skips
the leading marker and reads nx*ny float32 values
(IrisFDTD_time_evolution.py:52-69).
"""
import numpy as np


def read_record(path):
    """(values float32 1-D, raw bytes) of the single record in `path`;
    raises ValueError when the record markers do not frame the file."""
    raw = open(path, "rb").read()
    if len(raw) < 8:
        raise ValueError(f"{path}: too short for a record")
    n = int(np.frombuffer(raw[:4], dtype="<i4")[0])
    m = int(np.frombuffer(raw[-4:], dtype="<i4")[0])
    if n != m or n != len(raw) - 8 or n % 4:
        raise ValueError(f"{path}: record markers {n}, {m} do not match {len(raw) - 8} data bytes")
    return np.frombuffer(raw[4:-4], dtype="<f4").copy(), raw


def read_plane(path, na, nb):
    """Snapshot values as (nb, na): na values of the first free coordinate
    per row."""
    v, _ = read_record(path)
    if v.size != na * nb:
        raise ValueError(f"{path}: {v.size} values, expected {na}x{nb}")
    return v.reshape(nb, na)
