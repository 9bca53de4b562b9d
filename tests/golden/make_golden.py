"""Writes the golden vectors under tests/golden/ from the compiled reference.

Each fixture is one small seeded system (inputs) plus what the reference's own
subroutines produce from it, run through the harness in oracle/ (built from
/root/reference/cluster by `make -C oracle`).  Run from the repository root:

    python tests/golden/make_golden.py

The reference publishes no test vectors of its own; these are the vectors.
Only data is stored (inputs and outputs), never reference source.
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import cases  # noqa: E402
import pyoracle as po  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))

# (name, harness variant, case kwargs, actions); actions as pyoracle.run_harness
FIXTURES = [
    ("mav_n200", "mav", dict(n=200, rho=0.03, seed=11, dense_frac=0.4),
     [(1, 0, 0), (2, 0, 0), (4, 0, 0), (3, 1, 40), (4, 40, 0)]),
    ("vv_n128", "vv", dict(n=128, rho=0.02, seed=12, dense_frac=0.3),
     [(1, 0, 0), (2, 0, 0), (3, 1, 20), (4, 20, 0)]),
    ("nose_n200", "nose", dict(n=200, rho=0.03, seed=13, dense_frac=0.4, ethst_nh=1),
     [(1, 0, 0), (2, 0, 0), (3, 1, 30), (4, 30, 0)]),
    ("nuc_n240", "nuc", dict(n=240, rho=0.03, seed=14, dense_frac=0.4, nsorts=2, frac_a=0.5),
     [(1, 0, 0), (2, 0, 0), (3, 1, 30), (4, 30, 0)]),
    # stoddard-pbc-n-m-fre.f is only well defined on its first call in a
    # process (its cvtl() is never zeroed), so each fre fixture ends with one
    ("fre_n240_s0", "fre", dict(n=240, rho=0.03, seed=15, dense_frac=0.5, nsorts=2, frac_a=0.6),
     [(1, 0, 0), (2, 0, 0), (4, 0, 0)]),
    ("fre_n240_s25", "fre", dict(n=240, rho=0.03, seed=15, dense_frac=0.5, nsorts=2, frac_a=0.6),
     [(1, 0, 0), (2, 0, 0), (3, 1, 25), (4, 25, 0)]),
]

CASE_KEYS = ("natom", "nsorts", "api", "box", "mas", "sig", "eps", "num", "atrcut",
             "atrskin", "atskin", "dt", "Q", "att", "nbtim", "btim", "atnenum",
             "ethst_nh", "x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "psi", "s",
             "dmmh", "dmms", "dmmsu", "dmmax", "dmmav", "dmmst", "dmmi")
REC_KEYS = ("code", "step", "nlen", "x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az",
            "epot", "ekin", "sumv2", "psi", "s", "atrskin", "dmmh", "dmms", "dmmsu",
            "dmmax", "dmmav", "dmmst", "dmmi", "atnlist", "atnidx", "cnum", "cnumf",
            "cbiggest", "cbiggestf", "clookfor", "clookforf", "cl", "cs", "cc", "csi",
            "cek", "ceks", "clf", "csf", "ccf", "csif", "cekf", "ceksf", "cgs", "cgsf",
            "cgsfa", "cfpt", "cfptf")


def main():
    for name, variant, kw, actions in FIXTURES:
        case = cases.make_case(**kw)
        recs = po.run_harness(variant, case, actions)
        out = {"variant": np.array(variant), "actions": np.array(actions, dtype=np.int32)}
        for k in CASE_KEYS:
            out["in_" + k] = np.asarray(case[k])
        for i, r in enumerate(recs):
            for k in REC_KEYS:
                out[f"r{i}_{k}"] = np.asarray(r[k])
        path = os.path.join(OUT, name + ".npz")
        np.savez_compressed(path, **out)
        print(f"{path}: {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
