"""Input decks (.3d, version 1.30) and restart files for the reference's own
driver (main.f), so the reference program and the drop-in program can run
the same seeded system.

Deck layout: cluster/read-3d.f:57-257 (values at column 42, `ntim` read with
40X); example cluster/sample.3d.  Restart layout: cluster/wrt-rst.f:30-45 /
read-rst.f:37-49 (unformatted sequential; the record arrays are mxa long).
"""
from __future__ import annotations

import os

import numpy as np

from fcmd import restart

CIKB = 1.38066e-6
CIMU = 1.66056


def _val(label, v):
    label = (label + " " * 40)[:40] + ":"
    return f"{label}{v}\n"


def _f(v):
    s = f"{v:14.8f}"
    if len(s) > 14:
        s = f"{v:14.6f}"
    assert len(s) == 14, s
    return s


def _i(v):
    return f"{v:10d}"


def write_deck(path, case, ntim, ftcsi, ftave, sctim, ethst="-", q_star=100.0, fpath="./",
               rst="start.rst", nbtim=10, ftxyz=0, ftvel=0, ftacc=0, ftrst=0, restart=True,
               rndctrl="-", nucleation=None):
    """case in internal units (tests/cases.py layout); written back in deck
    units (nm, u, K, fs) so that read-3d.f reproduces the internal values.
    restart: start from the restart file rst (read-rst.f), else from the
    linked initconf; nucleation: (csover, clwait) switches the
    chk-nucleation.f abort on."""
    sig1 = case["sig"][0]
    rule = "---------1---------2---------3---------4---------5---------6---------7---------8\n"
    L = []
    L.append(rule)
    L.append(_val("Version number Program/Inputfile", " 1.30"))
    L.append(rule)
    L.append(_val("path for following files (relative)", fpath))
    for name, ft in (("test.xyz", ftxyz), ("test.vel", ftvel), ("test.acc", ftacc), ("test.rst", ftrst),
                     ("test.csi", ftcsi), ("test.ave", ftave), ("test.rat", 0)):
        L.append(_val("filename", name))
        L.append(_val("output interval", _i(ft)))
    L.append(rule)
    L.append(_val("timestep length                    [fs]", _f(case["dt"])))
    L.append((("total number of timesteps" + " " * 40)[:39] + ":") + _i(ntim) + "\n")
    L.append(_val("start timestep number", _i(0)))
    L.append(_val("screen output interval", _i(sctim)))
    L.append(rule)
    for k in range(3):
        L.append(_val("box size [nm]", _f(case["box"][k] / 10.0)))
    for _ in range(3):
        L.append(_val("periodic", "X"))
    L.append(_val("potential exponent for wall", _f(9.0)))
    L.append(_val("potential parameter for wall", _f(120.0)))
    L.append(_val("show box corners", "-"))
    L.append(_val("resize Box", "-"))
    for k in range(3):
        L.append(_val("new size of box", _f(case["box"][k] / 10.0)))
    for k in range(3):
        L.append(_val("maximum resize-step", _f(0.04)))
    L.append(_val("time steps between resizes", _i(50)))
    L.append(rule)
    L.append(_val("use restartfile? (X = yes, - = no)", "X" if restart else "-"))
    L.append(_val("file name of Input-Restartfile", rst))
    L.append(_val("thermostat? (X,N = yes, - = no)", ethst))
    L.append(_val("Nose-Hoover coupling constant Q", _f(q_star)))
    L.append(_val("pure random velocities after init", rndctrl))
    L.append(_val("Cluster Thermostat", "-"))
    L.append(_val("Cluster Thermostat Timer", _i(200)))
    L.append(rule)
    L.append(_val("system temperature                  [K]", _f(case["att"])))
    L.append(_val("cluster distance criterion   (in Sigma)", _f(case["atskin"] / sig1)))
    L.append(_val("number of neighbors for Frenkel", _i(int(case["atnenum"]))))
    L.append(_val("potential cutoff-Radius      (in Sigma)", _f(case["atrcut"] / sig1)))
    L.append(_val("skin-Radius of neighborlist  (in Sigma)", _f(case["atrskin"] / sig1)))
    L.append(_val("update-Time for Neighbor", _i(nbtim)))
    L.append(_val("update-Time for Natom-change", _i(500)))
    L.append(rule)
    csover, clwait = nucleation if nucleation else (15, 10000)
    L.append(_val("check for nucleation?", "X" if nucleation else "-"))
    L.append(_val("minimum cluster size to abort", _i(csover)))
    L.append(_val("time steps the cluster must exist", _i(clwait)))
    L.append(rule)
    L.append(_val("number of atomsorts", _i(int(case["nsorts"]))))
    L.append(rule)
    for s in range(int(case["nsorts"])):
        L.append(_val("atom symbol", "Ar" if s == 0 else "He"))
        L.append(_val("number of atoms", _i(int(case["num"][s]))))
        L.append(_val("atom mass                           [u]", _f(case["mas"][s] / CIMU)))
        L.append(_val("Lennard-Jones Sigma                [nm]", _f(case["sig"][s] / 10.0)))
        L.append(_val("Lennard-Jones Epsilon               [K]", _f(case["eps"][s] / CIKB)))
        L.append(_val("potential exponent", _f(12.0)))
        L.append(_val("potentialexponent attractive", _f(6.0)))
        L.append(_val("thermostat?", "X"))
        L.append(_val("Change number", "-"))
        L.append(_val("New number of atoms", _i(int(case["num"][s]))))
        L.append(rule)
    L.append(rule)
    with open(path, "w") as f:
        f.writelines(L)


def deck_units(case):
    """The internal values read-3d.f produces from the written deck (the deck
    has limited digits): apply the same conversions to the written text."""
    c = dict(case)
    sig = float(_f(case["sig"][0] / 10.0)) * 10.0
    c["sig"] = np.array([float(_f(v / 10.0)) * 10.0 for v in case["sig"]])
    c["mas"] = np.array([float(_f(v / CIMU)) * CIMU for v in case["mas"]])
    c["eps"] = np.array([float(_f(v / CIKB)) * CIKB for v in case["eps"]])
    c["box"] = np.array([float(_f(v / 10.0)) * 10.0 for v in case["box"]])
    c["dt"] = float(_f(case["dt"]))
    c["att"] = float(_f(case["att"]))
    c["atskin"] = float(_f(case["atskin"] / case["sig"][0])) * sig
    c["atrcut"] = float(_f(case["atrcut"] / case["sig"][0])) * sig
    c["atrskin"] = float(_f(case["atrskin"] / case["sig"][0])) * sig
    return c


def write_restart(path, case, mxa=6000, ekin=0.0, epot=0.0):
    """The package's restart writer (wrt-rst.f:30-45 layout) on a case."""
    restart.write_restart(path, {k: case[k] for k in restart.KEYS}, mxa, ekin=ekin, epot=epot)


def read_csi(path):
    """{step: {size: (count, ke_per_atom)}} from a wrt-csi.f output file."""
    out = {}
    step = None
    for line in open(path):
        parts = line.split()
        if len(parts) == 1:
            step = int(parts[0])
            out[step] = {}
        elif len(parts) >= 3 and step is not None:
            out[step][int(parts[0])] = (int(parts[1]), float(parts[2]))
    return out


def read_ave(path):
    rows = []
    for line in open(path):
        p = line.split()
        if p:
            rows.append((int(p[0]), float(p[1]), float(p[2])))
    return rows


def run_program(exe, deck, cwd, timeout=600, env=None):
    import resource
    import subprocess

    def lim():
        resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))

    e = dict(os.environ)
    if env:
        e.update(env)
    return subprocess.run([exe, deck], cwd=cwd, capture_output=True, text=True, timeout=timeout,
                          preexec_fn=lim, env=e)
