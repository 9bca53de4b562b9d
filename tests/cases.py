"""Small seeded LJ systems in the reference's internal units, for the tests.

Units follow cluster/read-3d.f:152-261: lengths in Angstrom, time in fs, mass in
1e-27 kg, energy in 1e-17 J.  Species 1 is argon-like (sigma 0.3405 nm,
eps 120 K, 40 u) as in cluster/sample.3d:65-71.
"""
from __future__ import annotations

import numpy as np

CIKB = 1.38066e-6  # const.inc:17
CIMU = 1.66056     # const.inc:18

SIG_AR = 3.405            # Angstrom
EPS_AR = 120.0 * CIKB     # 1e-17 J
MAS_AR = 40.0 * CIMU      # 1e-27 kg
TAU_AR = SIG_AR * np.sqrt(MAS_AR / EPS_AR)  # fs


def _place(rng, n, box, rmin, centers=None, dense_frac=0.0, dense_r=0.0):
    """Uniform placement with rejection closer than rmin (minimum image); a
    fraction `dense_frac` of the atoms is drawn inside balls of radius dense_r
    around `centers` so that clusters and liquid-like atoms exist."""
    pos = np.zeros((n, 3))
    k = 0
    tries = 0
    while k < n:
        tries += 1
        if tries > 200000 * max(n, 1):
            raise RuntimeError("placement failed")
        if centers is not None and k < int(dense_frac * n):
            c = centers[k % len(centers)]
            d = rng.normal(size=3)
            d *= dense_r * rng.random() ** (1 / 3) / np.linalg.norm(d)
            p = (c + d) % box
        else:
            p = rng.random(3) * box
        if k:
            dd = pos[:k] - p
            dd -= box * np.round(dd / box)
            if np.min(np.einsum("ij,ij->i", dd, dd)) < rmin * rmin:
                continue
        pos[k] = p
        k += 1
    return pos


def make_case(n=200, rho=0.02, tstar=0.7, seed=1, nsorts=1, frac_a=1.0,
              dense_frac=0.0, ncenters=2, dense_rho=0.6, q_star=100.0,
              ethst_nh=0, rc_sig=2.5, rskin_sig=2.8, atskin_sig=1.5,
              nbtim=10, btim=0, atnenum=5, dt_tau=0.005, with_accel=False,
              eps2_ratio=0.1, rmin_sig=0.9):
    """One system; species 1 gets the lowest atom numbers (read-3d.f:249-253)."""
    rng = np.random.default_rng(seed)
    box = (n / rho) ** (1.0 / 3.0) * SIG_AR
    centers = rng.random((ncenters, 3)) * box if dense_frac > 0 else None
    per_ball = max(dense_frac * n / ncenters, 1.0)
    dense_r = (3.0 * per_ball / (4.0 * np.pi * dense_rho)) ** (1.0 / 3.0) * SIG_AR
    pos = _place(rng, n, box, rmin_sig * SIG_AR, centers, dense_frac, dense_r)
    if nsorts == 1:
        num = [n]
        mas, sig, eps = [MAS_AR], [SIG_AR], [EPS_AR]
    else:
        na = int(round(frac_a * n))
        num = [na, n - na]
        mas, sig, eps = [MAS_AR, MAS_AR], [SIG_AR, SIG_AR], [EPS_AR, eps2_ratio * EPS_AR]
    api = np.concatenate([np.full(c, s + 1, dtype=np.int32) for s, c in enumerate(num)])
    m = np.array(mas)[api - 1]
    T = tstar * 120.0
    v = rng.normal(size=(n, 3)) * np.sqrt(CIKB * T / m)[:, None]
    v -= (m[:, None] * v).sum(0) / m.sum() / 1.0
    case = dict(natom=n, nsorts=nsorts, api=api, box=np.array([box, box, box]),
                mas=np.array(mas), sig=np.array(sig), eps=np.array(eps), num=np.array(num),
                atrcut=rc_sig * SIG_AR, atrskin=rskin_sig * SIG_AR, atskin=atskin_sig * SIG_AR,
                dt=dt_tau * TAU_AR, Q=q_star * MAS_AR * SIG_AR * SIG_AR, att=T,
                nbtim=nbtim, btim=btim, atnenum=atnenum, ethst_nh=ethst_nh,
                x=pos[:, 0].copy(), y=pos[:, 1].copy(), z=pos[:, 2].copy(),
                vx=v[:, 0].copy(), vy=v[:, 1].copy(), vz=v[:, 2].copy(),
                ax=np.zeros(n), ay=np.zeros(n), az=np.zeros(n),
                psi=0.0, s=0.0)
    init_mav(case)
    if with_accel:
        case["ax"] = rng.normal(size=n) * 1e-6
        case["ay"] = rng.normal(size=n) * 1e-6
        case["az"] = rng.normal(size=n) * 1e-6
    return case


def init_mav(case):
    """DMMPP initial values as read-3d.f:263-268 leaves them."""
    w = float(np.float32(case["nbtim"]))
    skin = case["atrskin"] - case["atrcut"]
    case["dmmh"] = np.full(256, skin / w)
    case["dmmst"] = 0
    case["dmmi"] = 0
    case["dmmsu"] = 256.0 * skin / w
    case["dmms"] = 0.0
    case["dmmax"] = 0.0
    case["dmmav"] = 0.0
    return case


def nh_initial_psi(case, ekin):
    """main.f:96-104 initial PSI (ethst='N')."""
    sumv2 = 2.0 * ekin / case["mas"][0]
    return (case["mas"][0] * sumv2 - 3.0 * case["natom"] * CIKB * case["att"]) / case["Q"]
