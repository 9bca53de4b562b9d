"""Python host mirror of the reference's plugin interface, over the C ABI
(include/fcmd.h) of the HIP engine in ../lib/libfcmd.so.

`System` holds one or more trajectories on the GPU and offers the reference's
entry points with the reference's semantics:

    neighbor(step)     neighbor-pbc.f:11 / neighbor-pbc-mav.f:13
    accatom()          for-a-lj-n-pbc.f:14 / for-a-lj-n-pbc-m.f:18 (a += F/m)
    integrator(step)   int-verlet.f:16 / int-verlet-mav.f:18 / int-verlet-nh-mav.f:13
    stoddard(step)     stoddard-pbc-n.f:10 / -n-m.f:11 / -n-m-fre.f:11

Errors come back as FcmdError carrying the reference's STOP message.  There is
no CPU fallback: without the library (or a GPU) every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "lib", "libfcmd.so")

CIKB = 1.38066e-6  # const.inc:17, Boltzmann constant in internal units
CIMU = 1.66056     # const.inc:18

FP64, MIXED, MIXED_KAHAN = 0, 1, 2
INT_VV, INT_MAV, INT_NH = 0, 1, 2
NB_PBC, NB_MAV = 0, 1
ACC_1, ACC_M, ACC_TABLE = 0, 1, 2
ST_N, ST_NM, ST_FRE = 0, 1, 2
MAXS = 10

# reference build targets (cluster/Makefile:27-95) as plugin choices
VARIANTS = {
    "vv": (INT_VV, NB_PBC, ACC_1, ST_N),
    "mav": (INT_MAV, NB_MAV, ACC_1, ST_N),
    "nuc": (INT_MAV, NB_MAV, ACC_M, ST_NM),
    "nose": (INT_NH, NB_MAV, ACC_M, ST_NM),
    "fre": (INT_MAV, NB_MAV, ACC_M, ST_FRE),
}


class FcmdError(RuntimeError):
    def __init__(self, code, text):
        super().__init__(text)
        self.code = code


class Params(C.Structure):
    _fields_ = [
        ("natom", C.c_int), ("nsorts", C.c_int),
        ("box", C.c_double * 3),
        ("mas", C.c_double * MAXS), ("sig", C.c_double * MAXS), ("eps", C.c_double * MAXS),
        ("num", C.c_int * MAXS),
        ("atrcut", C.c_double), ("atrskin", C.c_double), ("atskin", C.c_double),
        ("dt", C.c_double), ("Q", C.c_double), ("att", C.c_double),
        ("nbtim", C.c_int), ("btim", C.c_int), ("atnenum", C.c_int), ("ethst_nh", C.c_int),
        ("integrator", C.c_int), ("neighbor", C.c_int), ("accatom", C.c_int), ("stoddard", C.c_int),
        ("eps_ab", C.c_double * (MAXS * MAXS)), ("sig_ab", C.c_double * (MAXS * MAXS)),
        ("nh_mask", C.c_int), ("pad", C.c_int),
    ]


class Scalars(C.Structure):
    _fields_ = [
        ("atrskin", C.c_double), ("psi", C.c_double), ("s", C.c_double),
        ("ekin", C.c_double), ("epot", C.c_double), ("sumv2", C.c_double),
        ("dmmh", C.c_double * 256),
        ("dmms", C.c_double), ("dmmsu", C.c_double), ("dmmax", C.c_double), ("dmmav", C.c_double),
        ("dmmsr", C.c_double),
        ("dmmst", C.c_int), ("dmmi", C.c_int), ("step", C.c_int), ("npairs", C.c_int),
        ("nrebuild", C.c_int), ("nh_iters", C.c_int), ("err", C.c_int), ("nlist", C.c_int),
        ("virial", C.c_double), ("pressure", C.c_double),
    ]


_PD = C.POINTER(C.c_double)
_PI = C.POINTER(C.c_int)
_lib = None

# fcmd_time_kernels entries, in the order of include/fcmd.h (FCMD_NTIMED)
TIMED_KERNELS = ("drift", "traj_pre", "rebuild", "force", "traj_post", "kick")

EXPORTS = ("fcmd_device_count", "fcmd_error_text", "fcmd_create", "fcmd_destroy",
           "fcmd_set_params", "fcmd_upload", "fcmd_download", "fcmd_set_scalars",
           "fcmd_get_scalars", "fcmd_neighbor", "fcmd_accatom", "fcmd_integrator",
           "fcmd_stoddard", "fcmd_halflist", "fcmd_run", "fcmd_init_random",
           "fcmd_last_run_ms", "fcmd_time_kernels", "fcmd_list_entries", "fcmd_set_tuning",
           "fcmd_get_tuning", "fcmd_autotune", "fcmd_size_counts", "fcmd_get_size_counts",
           "fcmd_last_run_analysis_ms", "fcmd_run_noalloc")


def lib():
    """Load libfcmd.so; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FcmdError(-1, f"HIP engine not built: {LIB_PATH} missing "
                                "(run __graft_entry__.build() or make -C fortran-cluster-md_amd)")
        L = C.CDLL(LIB_PATH)
        for name in EXPORTS:
            getattr(L, name).restype = C.c_int
        L.fcmd_error_text.restype = C.c_char_p
        L.fcmd_destroy.restype = None
        _lib = L
    return _lib


def _p(a, dt):
    return None if a is None else a.ctypes.data_as(dt)


def check(rc):
    if rc != 0:
        raise FcmdError(rc, lib().fcmd_error_text(rc).decode())


def device_count() -> int:
    return lib().fcmd_device_count()


def init_random(seed, natom, box, rmin, att, mas, api):
    """Seeded vapour: positions with min distance rmin, MB velocities at att."""
    n = int(natom)
    out = [np.zeros(n) for _ in range(6)]
    box = np.ascontiguousarray(box, dtype=np.float64)
    mas = np.ascontiguousarray(mas, dtype=np.float64)
    api = np.ascontiguousarray(api, dtype=np.int32)
    check(lib().fcmd_init_random(C.c_ulonglong(seed), n, _p(box, _PD), C.c_double(rmin),
                                 C.c_double(att), _p(mas, _PD), _p(api, _PI),
                                 *[_p(a, _PD) for a in out]))
    return out


def make_params(case: dict, variant: str = "mav", eps_ab=None, sig_ab=None, nh_mask=0) -> Params:
    """Params from a case dict in the reference's internal units (keys as the
    COMMON variables: natom, nsorts, box, mas, sig, eps, num, atrcut, atrskin,
    atskin, dt, Q, att, nbtim, btim, atnenum, ethst_nh)."""
    p = Params()
    p.natom = int(case["natom"])
    p.nsorts = int(case["nsorts"])
    for k in range(3):
        p.box[k] = float(case["box"][k])
    for s in range(p.nsorts):
        p.mas[s] = float(case["mas"][s])
        p.sig[s] = float(case["sig"][s])
        p.eps[s] = float(case["eps"][s])
        p.num[s] = int(case["num"][s])
    for f in ("atrcut", "atrskin", "atskin", "dt", "Q", "att"):
        setattr(p, f, float(case[f]))
    for f in ("nbtim", "btim", "atnenum", "ethst_nh"):
        setattr(p, f, int(case[f]))
    p.integrator, p.neighbor, p.accatom, p.stoddard = VARIANTS[variant]
    if eps_ab is not None:
        p.accatom = ACC_TABLE
        ea = np.asarray(eps_ab, dtype=np.float64)
        sa = np.asarray(sig_ab, dtype=np.float64)
        for a in range(p.nsorts):
            for b in range(p.nsorts):
                p.eps_ab[a * MAXS + b] = ea[a, b]
                p.sig_ab[a * MAXS + b] = sa[a, b]
    p.nh_mask = int(nh_mask)
    return p


class System:
    """One fcmd_system: an ensemble of trajectories resident on the GPU."""

    def __init__(self, params, precision=FP64, list_width=0):
        if isinstance(params, Params):
            params = [params]
        self.ntraj = len(params)
        arr = (Params * self.ntraj)(*params)
        self.params = list(params)
        self.natom = [p.natom for p in params]
        self.precision = precision
        h = C.c_void_p()
        check(lib().fcmd_create(C.byref(h), self.ntraj, arr, int(precision), int(list_width)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().fcmd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ----------------------------------------------------------- state
    def upload(self, case: dict, traj: int = 0):
        f = [np.ascontiguousarray(case[k], dtype=np.float64) if k in case else None
             for k in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az")]
        api = np.ascontiguousarray(case["api"], dtype=np.int32) if "api" in case else None
        check(lib().fcmd_upload(self.h, traj, *[_p(a, _PD) for a in f], _p(api, _PI)))

    def download(self, traj: int = 0) -> dict:
        n = self.natom[traj]
        out = {k: np.zeros(n) for k in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az")}
        check(lib().fcmd_download(self.h, traj, *[_p(out[k], _PD) for k in
                                                   ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az")]))
        return out

    def scalars(self, traj: int = 0) -> Scalars:
        s = Scalars()
        check(lib().fcmd_get_scalars(self.h, traj, C.byref(s)))
        return s

    def set_scalars(self, traj: int = 0, **kw):
        s = self.scalars(traj)
        for k, v in kw.items():
            if k == "dmmh":
                for i in range(256):
                    s.dmmh[i] = float(v[i])
            else:
                setattr(s, k, v)
        check(lib().fcmd_set_scalars(self.h, traj, C.byref(s)))

    # ----------------------------------------------------------- entry points
    def neighbor(self, step: int, traj: int = 0) -> int:
        s = C.c_int(int(step))
        check(lib().fcmd_neighbor(self.h, traj, C.byref(s)))
        return s.value

    def accatom(self):
        check(lib().fcmd_accatom(self.h))

    def integrator(self, step: int) -> int:
        s = C.c_int(int(step))
        check(lib().fcmd_integrator(self.h, C.byref(s)))
        return s.value

    def stoddard(self, step: int, traj: int = 0, state: dict | None = None) -> dict:
        """Cluster analysis; returns the CVTPP arrays.  `state` carries csi(0),
        ceks(0), cfpt/cfptf and the counters between calls as COMMON would."""
        n = self.natom[traj]
        st = state if state is not None else {}
        I = lambda m: np.zeros(m, dtype=np.int32)  # noqa: E731
        D = lambda m: np.zeros(m, dtype=np.float64)  # noqa: E731
        out = dict(cl=I(n), cs=I(n), cc=I(n), csi=I(n + 1), cek=D(n), ceks=D(n + 1),
                   clf=I(n), csf=I(n + 1), ccf=I(n), csif=I(n + 1), cekf=D(n + 1),
                   ceksf=D(n + 1), cgs=I(n), cgsf=I(n + 1), cgsfa=I(n + 1))
        out["csi"][0] = st.get("csi0", 0)
        out["ceks"][0] = st.get("ceks0", 0.0)
        cfpt = st.setdefault("cfpt", np.full(n, -1, dtype=np.int32))
        cfptf = st.setdefault("cfptf", np.full(n, -1, dtype=np.int32))
        cnt = st.setdefault("counters", np.array([0, 0, 0, 0, 1, 1], dtype=np.int32))
        check(lib().fcmd_stoddard(
            self.h, traj, int(step),
            *[_p(out[k], _PI) for k in ("cl", "cs", "cc", "csi")],
            _p(out["cek"], _PD), _p(out["ceks"], _PD),
            *[_p(out[k], _PI) for k in ("clf", "csf", "ccf", "csif")],
            _p(out["cekf"], _PD), _p(out["ceksf"], _PD),
            *[_p(out[k], _PI) for k in ("cgs", "cgsf", "cgsfa")],
            _p(cfpt, _PI), _p(cfptf, _PI), _p(cnt, _PI)))
        st["csi0"] = int(out["csi"][0])
        st["ceks0"] = float(out["ceks"][0])
        out["cfpt"] = cfpt.copy()
        out["cfptf"] = cfptf.copy()
        for k, v in zip(("cnum", "cnumf", "cbiggest", "cbiggestf", "clookfor", "clookforf"), cnt):
            out[k] = int(v)
        return out

    def write_restart(self, path, traj: int = 0, mxa: int | None = None):
        """Trajectory traj as a restart file the reference's readRST reads
        (wrt-rst.f:30-45 layout, mxa-long records; mxa defaults to natom)."""
        from . import restart
        st = self.download(traj)
        sc = self.scalars(traj)
        restart.write_restart(path, st, mxa or self.natom[traj], ekin=sc.ekin, epot=sc.epot)

    def halflist(self, traj: int = 0):
        n = self.natom[traj]
        sc = self.scalars(traj)
        cap = int(n + max(sc.npairs, 0) + 16)
        atnlist = np.zeros(cap, dtype=np.int32)
        atnidx = np.zeros(n, dtype=np.int32)
        ln = C.c_longlong()
        check(lib().fcmd_halflist(self.h, traj, _p(atnlist, _PI), C.c_longlong(cap),
                                  _p(atnidx, _PI), C.byref(ln)))
        return atnlist[:ln.value], atnidx

    def run(self, nsteps, cluster_every=0, fpt_size=50, use_graph=True):
        """Device-driven run; returns (largest-cluster series [nsnap, ntraj], fpt)."""
        nsnap = (nsteps + cluster_every - 1) // cluster_every if cluster_every else 0
        series = np.zeros((max(nsnap, 1), self.ntraj), dtype=np.int32)
        fpt = np.full(self.ntraj, -1, dtype=np.int32)
        check(lib().fcmd_run(self.h, int(nsteps), int(cluster_every), int(fpt_size),
                             _p(series, _PI), _p(fpt, _PI), int(bool(use_graph))))
        return series[:nsnap], fpt

    def size_counts(self, nmax: int, warmup: int = 0):
        """Start sizer-style cluster-size counting in later run() analyses
        (helpers/sizer.f:59-103); nmax <= 0 stops it."""
        check(lib().fcmd_size_counts(self.h, int(nmax), int(warmup)))

    def get_size_counts(self, traj: int = 0):
        """(counts[natom+1], stopped) of one trajectory."""
        counts = np.zeros(self.natom[traj] + 1, dtype=np.int32)
        stopped = np.zeros(1, dtype=np.int32)
        check(lib().fcmd_get_size_counts(self.h, traj, _p(counts, _PI), _p(stopped, _PI)))
        return counts, bool(stopped[0])

    def last_run_ms(self) -> float:
        v = C.c_double()
        check(lib().fcmd_last_run_ms(self.h, C.byref(v)))
        return v.value

    def last_run_analysis_ms(self):
        """(mean ms of one cluster analysis of the last run, analyses timed)."""
        v = C.c_double()
        n = C.c_int()
        check(lib().fcmd_last_run_analysis_ms(self.h, C.byref(v), C.byref(n)))
        return v.value, n.value

    def time_kernels(self, nsteps: int) -> dict:
        """Mean ms per step of each kernel class (HIP events, engine stream)."""
        out = np.zeros(len(TIMED_KERNELS))
        check(lib().fcmd_time_kernels(self.h, int(nsteps), _p(out, _PD)))
        return dict(zip(TIMED_KERNELS, out.tolist()))

    def tuning(self) -> dict:
        """Current launch tunables (force_u, nbr_cu, wl_grid)."""
        t = np.zeros(3, dtype=np.int32)
        check(lib().fcmd_get_tuning(self.h, _p(t, _PI)))
        return dict(zip(("force_u", "nbr_cu", "wl_grid"), t.tolist()))

    def set_tuning(self, force_u: int, nbr_cu: int, wl_grid: int):
        t = np.array([force_u, nbr_cu, wl_grid], dtype=np.int32)
        check(lib().fcmd_set_tuning(self.h, _p(t, _PI)))

    def autotune(self, rounds: int = 5) -> dict:
        """Time every tunable candidate over rounds x 6 real steps and keep
        the fastest (the steps advance the state like run())."""
        t = np.zeros(3, dtype=np.int32)
        check(lib().fcmd_autotune(self.h, int(rounds), _p(t, _PI)))
        return dict(zip(("force_u", "nbr_cu", "wl_grid"), t.tolist()))

    def list_entries(self) -> int:
        v = C.c_longlong()
        check(lib().fcmd_list_entries(self.h, C.byref(v)))
        return v.value
