"""Python access to the checker: the C restatement (liboracle.so) and the
compiled reference harness (oracle/_ref/bin/*/harness_*).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by __graft_entry__.smoke() and
by the CPU-baseline leg of bench.py; never by the product.

A *case* is a dict describing one system in the reference's internal units
(Angstrom, fs, 1e-27 kg, 1e-17 J, K), with 0-based numpy arrays:

    natom, nsorts, api (int32, species 1..), box (3,), mas/sig/eps/num (nsorts,)
    atrcut, atrskin, atskin, dt, Q, att, nbtim, btim, atnenum, ethst_nh
    x, y, z, vx, vy, vz, ax, ay, az (float64, natom)
    psi, s, dmmh (256,), dmms, dmmsu, dmmax, dmmav, dmmst, dmmi
"""
from __future__ import annotations

import ctypes as C
import os
import resource
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")
CIKB = 1.38066e-6  # const.inc:17
CIMU = 1.66056     # const.inc:18

NB_PBC, NB_MAV = 0, 1
ACC_1, ACC_M, ACC_TABLE = 0, 1, 2
INT_VV, INT_MAV, INT_NH = 0, 1, 2
ST_N, ST_NM, ST_FRE = 0, 1, 2

# plugin sets of the harness executables (oracle/Makefile SET_*), expressed
# as the oracle's variant codes: (integrator, neighbor, AccAtom, stoddard)
VARIANTS = {
    "vv": (INT_VV, NB_PBC, ACC_1, ST_N),
    "mav": (INT_MAV, NB_MAV, ACC_1, ST_N),
    "nuc": (INT_MAV, NB_MAV, ACC_M, ST_NM),
    "nose": (INT_NH, NB_MAV, ACC_M, ST_NM),
    "fre": (INT_MAV, NB_MAV, ACC_M, ST_FRE),
}

ERR_TEXT = {
    1: "*** FATAL ERROR! ***  Skin radius exceeded!!",
    2: "*** FATAL ERROR *** Neighborlist is nearly full!",
    3: "Numerical integration did not converge!",
}


class OrcState(C.Structure):
    _fields_ = [
        ("natom", C.c_int), ("nsorts", C.c_int),
        ("box", C.c_double * 3),
        ("mas", C.c_double * 10), ("sig", C.c_double * 10), ("eps", C.c_double * 10),
        ("num", C.c_int * 10),
        ("atrcut", C.c_double), ("atrskin", C.c_double), ("atskin", C.c_double),
        ("dt", C.c_double), ("Q", C.c_double), ("att", C.c_double),
        ("nbtim", C.c_int), ("btim", C.c_int), ("atnenum", C.c_int), ("ethst_nh", C.c_int),
        ("dmmh", C.c_double * 256), ("dmms", C.c_double), ("dmmsu", C.c_double),
        ("dmmax", C.c_double), ("dmmav", C.c_double),
        ("dmmst", C.c_int), ("dmmi", C.c_int),
        ("psi", C.c_double), ("s", C.c_double), ("ekin", C.c_double),
        ("epot", C.c_double), ("sumv2", C.c_double),
        ("cnum", C.c_int), ("cnumf", C.c_int), ("cbiggest", C.c_int),
        ("cbiggestf", C.c_int), ("clookfor", C.c_int), ("clookforf", C.c_int),
        ("cstep", C.c_int), ("nh_iters", C.c_int), ("npairs", C.c_int),
        ("virial", C.c_double), ("pressure", C.c_double),
        ("eps_ab", C.c_double * 100), ("sig_ab", C.c_double * 100),
    ]


_PD = C.POINTER(C.c_double)
_PI = C.POINTER(C.c_int)


class OrcArrays(C.Structure):
    _fields_ = [
        ("x", _PD), ("y", _PD), ("z", _PD), ("vx", _PD), ("vy", _PD), ("vz", _PD),
        ("ax", _PD), ("ay", _PD), ("az", _PD),
        ("api", _PI), ("atnlist", _PI), ("list_cap", C.c_long), ("atnidx", _PI),
        ("cl", _PI), ("cs", _PI), ("cc", _PI), ("csi", _PI),
        ("cek", _PD), ("ceks", _PD),
        ("clf", _PI), ("csf", _PI), ("ccf", _PI), ("csif", _PI),
        ("cekf", _PD), ("ceksf", _PD),
        ("cgs", _PI), ("cgsf", _PI), ("cgsfa", _PI),
        ("cfpt", _PI), ("cfptf", _PI),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(REF_DIR, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run make -C oracle)")
        _lib = C.CDLL(path)
        for name in ("orc_neighbor", "orc_accatom", "orc_integrator", "orc_stoddard"):
            getattr(_lib, name).restype = C.c_int
    return _lib


class Oracle:
    """A system held by the C restatement; mirrors the reference COMMON state."""

    def __init__(self, case: dict, variant: str = "mav", list_cap: int | None = None,
                 eps_ab=None, sig_ab=None):
        """eps_ab, sig_ab (nsorts x nsorts): the explicit pair-table form of
        AccAtom (ORC_ACC_TABLE) instead of the variant's."""
        self.variant = variant
        self.ivar, self.nbvar, self.accvar, self.stvar = VARIANTS[variant]
        n = int(case["natom"])
        self.n = n
        st = OrcState()
        if eps_ab is not None:
            self.accvar = ACC_TABLE
            ea = np.asarray(eps_ab, dtype=np.float64)
            sa = np.asarray(sig_ab, dtype=np.float64)
            for a in range(ea.shape[0]):
                for b in range(ea.shape[1]):
                    st.eps_ab[a * 10 + b] = ea[a, b]
                    st.sig_ab[a * 10 + b] = sa[a, b]
        st.natom = n
        st.nsorts = int(case["nsorts"])
        for k in range(3):
            st.box[k] = float(case["box"][k])
        for s in range(st.nsorts):
            st.mas[s] = float(case["mas"][s])
            st.sig[s] = float(case["sig"][s])
            st.eps[s] = float(case["eps"][s])
            st.num[s] = int(case["num"][s])
        for f in ("atrcut", "atrskin", "atskin", "dt", "Q", "att", "psi", "s",
                  "dmms", "dmmsu", "dmmax", "dmmav"):
            setattr(st, f, float(case[f]))
        for f in ("nbtim", "btim", "atnenum", "ethst_nh", "dmmst", "dmmi"):
            setattr(st, f, int(case[f]))
        for k in range(256):
            st.dmmh[k] = float(case["dmmh"][k])
        st.clookfor = 1
        st.clookforf = 1
        st.cbiggest = 0
        st.cbiggestf = 0
        st.cstep = -1
        self.st = st
        self.arr = {}
        for f in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az"):
            self.arr[f] = np.ascontiguousarray(np.array(case[f], dtype=np.float64))
        self.arr["api"] = np.ascontiguousarray(np.array(case["api"], dtype=np.int32))
        cap = list_cap if list_cap is not None else max(4 * n, n * (n + 1) // 2 + n + 8)
        self.arr["atnlist"] = np.zeros(cap, dtype=np.int32)
        self.arr["atnidx"] = np.zeros(n, dtype=np.int32)
        for f in ("cl", "cs", "cc", "clf", "ccf", "cgs", "cfpt", "cfptf"):
            self.arr[f] = np.zeros(n, dtype=np.int32)
        self.arr["cfpt"][:] = -1
        self.arr["cfptf"][:] = -1
        for f in ("csi", "csf", "csif", "cgsf", "cgsfa"):
            self.arr[f] = np.zeros(n + 1, dtype=np.int32)
        self.arr["cek"] = np.zeros(n, dtype=np.float64)
        for f in ("ceks", "cekf", "ceksf"):
            self.arr[f] = np.zeros(n + 1, dtype=np.float64)
        a = OrcArrays()
        for name, _t in OrcArrays._fields_:
            if name == "list_cap":
                a.list_cap = cap
            else:
                arr = self.arr[name]
                ctype = _PD if arr.dtype == np.float64 else _PI
                setattr(a, name, arr.ctypes.data_as(ctype))
        self.a = a

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(ERR_TEXT.get(rc, f"oracle error {rc}"))

    def neighbor(self, step: int) -> int:
        s = C.c_int(int(step))
        self._check(lib().orc_neighbor(self.nbvar, C.byref(self.st), C.byref(self.a), C.byref(s)))
        return s.value

    def accatom(self):
        """a = 0, epot = 0, AccAtom() -- what every integrator does before forces."""
        for f in ("ax", "ay", "az"):
            self.arr[f][:] = 0.0
        self.st.epot = 0.0
        self._check(lib().orc_accatom(self.accvar, C.byref(self.st), C.byref(self.a)))

    def integrator(self, step: int) -> int:
        s = C.c_int(int(step))
        self._check(lib().orc_integrator(self.ivar, self.nbvar, self.accvar,
                                         C.byref(self.st), C.byref(self.a), C.byref(s)))
        return s.value

    def stoddard(self, step: int):
        self._check(lib().orc_stoddard(self.stvar, C.byref(self.st), C.byref(self.a), int(step)))

    def halflist(self):
        """(atnlist[:len], atnidx) in the reference format."""
        n = self.n
        ln = int(self.arr["atnidx"][n - 1])
        return self.arr["atnlist"][:ln].copy(), self.arr["atnidx"].copy()

    def pairs(self):
        return halflist_pairs(*self.halflist())


def halflist_pairs(atnlist, atnidx):
    """Set of (i, j), i<j, 1-based, from a reference-format half list."""
    out = []
    i = None
    for v in atnlist:
        if v < 0:
            i = -int(v)
        else:
            out.append((i, int(v)))
    return out


# --------------------------------------------------------------------------
# compiled reference harness
def harness_path(variant: str, mxa: int = 6000) -> str:
    return os.path.join(REF_DIR, "bin", f"mxa{mxa}", f"harness_{variant}")


def have_harness(variant: str = "mav", mxa: int = 6000) -> bool:
    return os.path.exists(harness_path(variant, mxa))


def _write_input(path, case, actions):
    n = int(case["natom"])
    ns = int(case["nsorts"])
    with open(path, "wb") as f:
        f.write(struct.pack("<3i", n, ns, len(actions)))
        f.write(np.asarray(case["api"], dtype="<i4").tobytes())
        f.write(struct.pack("<3d", *[float(v) for v in case["box"]]))
        for s in range(ns):
            f.write(struct.pack("<3di", float(case["mas"][s]), float(case["sig"][s]),
                                float(case["eps"][s]), int(case["num"][s])))
        f.write(struct.pack("<6d", *[float(case[k]) for k in
                                     ("atrcut", "atrskin", "atskin", "dt", "Q", "att")]))
        f.write(struct.pack("<4i", int(case["nbtim"]), int(case["btim"]),
                            int(case["atnenum"]), int(case["ethst_nh"])))
        for k in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az"):
            f.write(np.asarray(case[k], dtype="<f8").tobytes())
        f.write(struct.pack("<2d", float(case["psi"]), float(case["s"])))
        f.write(np.asarray(case["dmmh"], dtype="<f8").tobytes())
        f.write(struct.pack("<4d", *[float(case[k]) for k in ("dmms", "dmmsu", "dmmax", "dmmav")]))
        f.write(struct.pack("<2i", int(case["dmmst"]), int(case["dmmi"])))
        for act in actions:
            code, a1, a2 = (list(act) + [0, 0])[:3]
            f.write(struct.pack("<3i", int(code), int(a1), int(a2)))


def _read_records(path, n):
    buf = open(path, "rb").read()
    off = 0

    def take(dtype, count):
        nonlocal off
        arr = np.frombuffer(buf, dtype=dtype, count=count, offset=off)
        off += arr.nbytes
        return arr.copy()

    recs = []
    while off < len(buf):
        code, step, natom, nlen = take("<i4", 4)
        r = {"code": int(code), "step": int(step), "nlen": int(nlen)}
        for k in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az"):
            r[k] = take("<f8", n)
        for k, v in zip(("epot", "ekin", "sumv2", "psi", "s", "atrskin"), take("<f8", 6)):
            r[k] = float(v)
        r["dmmh"] = take("<f8", 256)
        for k, v in zip(("dmms", "dmmsu", "dmmax", "dmmav"), take("<f8", 4)):
            r[k] = float(v)
        r["dmmst"], r["dmmi"] = (int(v) for v in take("<i4", 2))
        r["atnlist"] = take("<i4", nlen)
        r["atnidx"] = take("<i4", n)
        for k, v in zip(("cnum", "cnumf", "cbiggest", "cbiggestf", "clookfor", "clookforf"),
                        take("<i4", 6)):
            r[k] = int(v)
        r["cl"] = take("<i4", n)
        r["cs"] = take("<i4", n)
        r["cc"] = take("<i4", n)
        r["csi"] = take("<i4", n + 1)
        r["cek"] = take("<f8", n)
        r["ceks"] = take("<f8", n + 1)
        r["clf"] = take("<i4", n)
        r["csf"] = take("<i4", n + 1)
        r["ccf"] = take("<i4", n)
        r["csif"] = take("<i4", n + 1)
        r["cekf"] = take("<f8", n + 1)
        r["ceksf"] = take("<f8", n + 1)
        r["cgs"] = take("<i4", n)
        r["cgsf"] = take("<i4", n + 1)
        r["cgsfa"] = take("<i4", n + 1)
        r["cfpt"] = take("<i4", n)
        r["cfptf"] = take("<i4", n)
        recs.append(r)
    return recs


def _unlimited_stack():
    # the reference keeps stoddard's mxnlist-sized work arrays on the stack
    resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))


def run_harness(variant: str, case: dict, actions, mxa: int = 6000, timeout: float = 600):
    """Run the compiled reference on `case`; actions are (code, arg1, arg2):
    1 neighbor(arg1), 2 AccAtom, 3 integrator(arg1..arg2), 4 stoddard(arg1)."""
    exe = harness_path(variant, mxa)
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "in.bin")
        fout = os.path.join(td, "out.bin")
        _write_input(fin, case, actions)
        p = subprocess.run([exe, fin, fout], capture_output=True, text=True,
                           preexec_fn=_unlimited_stack, timeout=timeout)
        if p.returncode != 0 or not os.path.exists(fout):
            raise RuntimeError(f"harness failed ({p.returncode}): {p.stdout[-2000:]} {p.stderr[-2000:]}")
        recs = _read_records(fout, int(case["natom"]))
        for r in recs:
            r["stdout"] = p.stdout
        return recs
