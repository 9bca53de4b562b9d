"""Restart files in the reference's layout (wrt-rst.f:30-45, read by
read-rst.f:37-49), so that a state advanced on the GPU can be handed to the
reference's own driver and back.

Unformatted sequential records, little endian, each framed by its 4-byte
length as gfortran and flang write them:

    mxa                                     (integer, the writer's capacity)
    ekin, epot, ekin, epot, ekin            (5 doubles)
    ewall, ewallsum, fwall, fwallsum        (4 doubles)
    x0, y0, z0                              (3 x mxa doubles: positions)
    x1, y1, z1                              (velocities)
    x2, y2, z2                              (accelerations)
    x3 .. z5                                (higher derivatives, zero here)

Arrays are mxa long whatever natom is (atom i in element i-1).  The file
holds neither PSI nor s (Nose-Hoover state): read-rst.f leaves them to the
driver (main.f:96-104 recomputes PSI).
"""
from __future__ import annotations

import struct

import numpy as np

KEYS = ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az")


def _rec(f, payload: bytes):
    f.write(struct.pack("<i", len(payload)))
    f.write(payload)
    f.write(struct.pack("<i", len(payload)))


def write_restart(path, state: dict, mxa: int, ekin: float = 0.0, epot: float = 0.0,
                  ewall=(0.0, 0.0, 0.0, 0.0)):
    """state: x, y, z, vx, vy, vz, ax, ay, az of natom <= mxa atoms."""
    n = len(state["x"])
    if n > mxa:
        raise ValueError(f"natom {n} exceeds mxa {mxa}")

    def arr(v):
        a = np.zeros(mxa)
        a[:n] = v
        return a.astype("<f8").tobytes()

    with open(path, "wb") as f:
        _rec(f, struct.pack("<i", mxa))
        _rec(f, struct.pack("<5d", ekin, epot, ekin, epot, ekin))
        _rec(f, struct.pack("<4d", *ewall))
        for k in range(3):
            _rec(f, arr(state[KEYS[3 * k]]) + arr(state[KEYS[3 * k + 1]]) + arr(state[KEYS[3 * k + 2]]))
        z = arr(np.zeros(n))
        for _ in range(3):
            _rec(f, z + z + z)


def read_restart(path, natom: int) -> dict:
    """The first natom atoms of a restart file, with mxa, ekin and epot."""
    buf = open(path, "rb").read()
    off = 0
    recs = []
    while off < len(buf):
        (ln,) = struct.unpack_from("<i", buf, off)
        recs.append(buf[off + 4:off + 4 + ln])
        (ln2,) = struct.unpack_from("<i", buf, off + 4 + ln)
        if ln2 != ln:
            raise ValueError(f"{path}: record framing broken at byte {off}")
        off += ln + 8
    if len(recs) != 9:
        raise ValueError(f"{path}: {len(recs)} records, expected 9")
    (mxa,) = struct.unpack("<i", recs[0])
    e = struct.unpack("<5d", recs[1])
    out = {"mxa": mxa, "ekin": e[0], "epot": e[1]}
    for k in range(3):
        a = np.frombuffer(recs[3 + k], dtype="<f8").reshape(3, mxa)
        for c in range(3):
            out[KEYS[3 * k + c]] = a[c, :natom].copy()
    return out
