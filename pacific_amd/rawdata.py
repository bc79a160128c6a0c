"""Readers of the time loop's output files (host side).

RawData (RawDataPostProcessingWriter.cpp:152-600): one file per field,
``<root>_position_{x,y,z}.dat``, ``_translational_velocity_{x,y,z}.dat``,
``_angular_velocity_{x,y,z}.dat``, ``_coordinationNumber.dat`` and
``_particleType.dat``.  Text mode: one line per frame, the time then one
value per particle in id order (the coordination-number lines carry no
time; the particle-type file holds the last frame only).  Binary mode: per
frame the time (f64; twice in the z angular-velocity file, not at all in
the coordination file) then one f64 (int32 for coordination numbers and
types) per particle.

``<root>_VelocityMaxMean.dat`` (Grains.cpp:136-146): tab-separated
``time  max|v|  mean|v|`` per output time.

Numbers are written as GrainsExec::doubleToString(ios::scientific, 6, x)
(GrainsExec.cpp:212-224): ``%.6e``, except zero, which keeps the default
stream format (``0``).
"""
import os

import numpy as np

VECTOR_FIELDS = ["position_x", "position_y", "position_z", "translational_velocity_x",
                 "translational_velocity_y", "translational_velocity_z", "angular_velocity_x",
                 "angular_velocity_y", "angular_velocity_z"]


def format_number(x):
    """GrainsExec::doubleToString(ios::scientific, 6, x)."""
    x = float(x)
    if x != 0.0:
        return f"{x:.6e}"
    return "-0" if np.signbit(x) else "0"


def velocity_line(t, vmax, vmean):
    """One line of <root>_VelocityMaxMean.dat (without the newline)."""
    return "\t".join(format_number(v) for v in (t, vmax, vmean))


def read_velocity_max_mean(path):
    """(n, 3) array of (time, max |v|, mean |v|)."""
    return np.loadtxt(path, ndmin=2)


def read_field(path, n=None, timed=True):
    """Text RawData field file: (times, values (frames, n)).  timed=False for
    the coordination-number file (no time column); then times is None."""
    rows = [line.split() for line in open(path) if line.strip()]
    if timed:
        t = np.array([float(r[0]) for r in rows])
        v = np.array([[float(x) for x in r[1:]] for r in rows])
    else:
        t = None
        v = np.array([[int(x) for x in r] for r in rows])
    if n is not None and v.shape[1] != n:
        raise ValueError(f"{path}: {v.shape[1]} values per frame, expected {n}")
    return t, v


def read_rawdata(root, n, binary=False):
    """All fields of a RawData output: dict field -> (frames, n) array, plus
    'time' (frames,), 'coordination' (frames, n) int and 'type' (n,) int."""
    out = {}
    if binary:
        for k, f in enumerate(VECTOR_FIELDS):
            raw = np.fromfile(f"{root}_{f}.dat", dtype=np.float64)
            lead = 2 if f == "angular_velocity_z" else 1
            frame = lead + n
            if raw.size % frame:
                raise ValueError(f"{root}_{f}.dat: size not a whole number of frames")
            raw = raw.reshape(-1, frame)
            if k == 0:
                out["time"] = raw[:, 0].copy()
            out[f] = raw[:, lead:].copy()
        out["coordination"] = np.fromfile(f"{root}_coordinationNumber.dat", dtype=np.int32).reshape(-1, n)
        out["type"] = np.fromfile(f"{root}_particleType.dat", dtype=np.int32)
        return out
    for k, f in enumerate(VECTOR_FIELDS):
        t, v = read_field(f"{root}_{f}.dat", n)
        if k == 0:
            out["time"] = t
        out[f] = v
    out["coordination"] = read_field(f"{root}_coordinationNumber.dat", n, timed=False)[1]
    path = f"{root}_particleType.dat"
    out["type"] = np.array([int(x) for x in open(path).read().split()]) if os.path.exists(path) else None
    return out
