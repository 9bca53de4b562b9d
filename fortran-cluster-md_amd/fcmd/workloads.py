"""The nucleation workloads, built in the reference's internal units.

Reduced units map through argon as in cluster/sample.3d:65-71 (sigma 0.3405
nm, eps 120 K, 40 u); the deck-to-internal conversions follow
cluster/read-3d.f:152-261 (lengths x10 to Angstrom, masses x cimu, energies
x cikb, radii x sigma_1, Q x m1 sigma_1^2).  Initial states come from the
seeded generator fcmd_init_random (uniform vapour, 0.9 sigma hard minimum,
Maxwell-Boltzmann at T, zero momentum, exact T).

    lj_vapour    N=1000, rho*=0.02, T*=0.7, NVE, Stillinger every 100 steps
    lj_nh_262k   N=262144, rho*=0.03, T*=0.65, Nose-Hoover Q*=100,
                 Stillinger + ten Wolde-Frenkel every 200 steps
    lj_nh_2m     N=2097152, rho*=0.05, T*=0.7, Nose-Hoover, FP64 vs mixed
    ensemble512  512 trajectories, N log-uniform 500..32768, rho* U(0.01,0.05),
                 T* U(0.6,0.8), Nose-Hoover, Stillinger every 100, FPT to 50
    mixture64    64 x 65536 atoms, 20 % A / 80 % B, eps_AA 1, eps_BB 0.1,
                 eps_AB 0.3, Nose-Hoover on B only, A-only Stillinger
"""
from __future__ import annotations

import numpy as np

from . import (ACC_1, ACC_TABLE, CIKB, CIMU, INT_MAV, INT_NH, NB_MAV, ST_FRE, ST_N, ST_NM,
               Params, init_random, MAXS)

SIG = 3.405               # Angstrom (0.3405 nm)
EPS = 120.0 * CIKB        # 1e-17 J
MAS = 40.0 * CIMU         # 1e-27 kg
TAU = SIG * np.sqrt(MAS / EPS)  # fs


def _params(n, rho, tstar, integrator, stoddard, accatom=ACC_1, q_star=100.0, nsorts=1,
            num=None, eps=None, nbtim=10, eps_ab=None, sig_ab=None, nh_mask=0):
    box = (n / rho) ** (1.0 / 3.0) * SIG
    p = Params()
    p.natom = n
    p.nsorts = nsorts
    for k in range(3):
        p.box[k] = box
    eps = eps if eps is not None else [1.0] * nsorts
    num = num if num is not None else [n]
    for s in range(nsorts):
        p.mas[s] = MAS
        p.sig[s] = SIG
        p.eps[s] = eps[s] * EPS
        p.num[s] = num[s]
    p.atrcut = 2.5 * SIG
    p.atrskin = 2.8 * SIG       # cutoff + 0.3 sigma skin
    p.atskin = 1.5 * SIG        # Stillinger distance
    p.dt = 0.005 * TAU
    p.Q = q_star * MAS * SIG * SIG
    p.att = tstar * 120.0
    p.nbtim = nbtim
    p.btim = 0
    p.atnenum = 5
    p.ethst_nh = 1 if integrator == INT_NH else 0
    p.integrator = integrator
    p.neighbor = NB_MAV
    p.accatom = accatom
    p.stoddard = stoddard
    if eps_ab is not None:
        for a in range(nsorts):
            for b in range(nsorts):
                p.eps_ab[a * MAXS + b] = eps_ab[a][b] * EPS
                p.sig_ab[a * MAXS + b] = sig_ab[a][b] * SIG
    p.nh_mask = nh_mask
    return p


def _state(p: Params, seed: int):
    n = p.natom
    api = np.concatenate([np.full(p.num[s], s + 1, dtype=np.int32) for s in range(p.nsorts)])
    mas = np.array([p.mas[s] for s in range(p.nsorts)])
    x, y, z, vx, vy, vz = init_random(seed, n, np.array(p.box[:]), 0.9 * SIG, p.att, mas, api)
    return dict(x=x, y=y, z=z, vx=vx, vy=vy, vz=vz, ax=np.zeros(n), ay=np.zeros(n),
                az=np.zeros(n), api=api)


def lj_vapour(seed=1):
    p = _params(1000, 0.02, 0.7, INT_MAV, ST_N)
    return [p], [_state(p, seed)], dict(steps=1000, cluster_every=100)


def lj_nh_262k(seed=2, n=262144):
    p = _params(n, 0.03, 0.65, INT_NH, ST_FRE)
    return [p], [_state(p, seed)], dict(steps=20000, cluster_every=200)


def lj_nh_2m(seed=3, n=2097152):
    p = _params(n, 0.05, 0.7, INT_NH, ST_N)
    return [p], [_state(p, seed)], dict(steps=10000, cluster_every=200)


def ensemble512(seed=4, ntraj=512, nmin=500, nmax=32768):
    rng = np.random.default_rng(seed)
    ns = np.rint(np.exp(rng.uniform(np.log(nmin), np.log(nmax), ntraj))).astype(int)
    rhos = rng.uniform(0.01, 0.05, ntraj)
    ts = rng.uniform(0.6, 0.8, ntraj)
    ps, st = [], []
    for k in range(ntraj):
        p = _params(int(ns[k]), float(rhos[k]), float(ts[k]), INT_NH, ST_N)
        ps.append(p)
        st.append(_state(p, seed * 100003 + k))
    return ps, st, dict(steps=50000, cluster_every=100, fpt_size=50)


def mixture64(seed=5, ntraj=64, n=65536):
    na = int(round(0.2 * n))
    ps, st = [], []
    for k in range(ntraj):
        p = _params(n, 0.04, 0.7, INT_NH, ST_NM, accatom=ACC_TABLE, nsorts=2, num=[na, n - na],
                    eps=[1.0, 0.1], eps_ab=[[1.0, 0.3], [0.3, 0.1]],
                    sig_ab=[[1.0, 1.0], [1.0, 1.0]], nh_mask=0b10)
        ps.append(p)
        st.append(_state(p, seed * 100003 + k))
    return ps, st, dict(steps=20000, cluster_every=100)


WORKLOADS = {
    "lj_vapour": lj_vapour,
    "lj_nh_262k": lj_nh_262k,
    "lj_nh_2m": lj_nh_2m,
    "ensemble512": ensemble512,
    "mixture64": mixture64,
}


def nh_initial_psi(p: Params, st: dict) -> float:
    """main.f:96-104: PSI = (m1 sum v^2 - 3 N k T)/Q with sum v^2 = 2 ekin/m1."""
    m = np.array([p.mas[s] for s in range(p.nsorts)])[st["api"] - 1]
    if p.nh_mask:
        sel = ((p.nh_mask >> (st["api"] - 1)) & 1).astype(bool)
    else:
        sel = np.ones(p.natom, dtype=bool)
        m = np.full(p.natom, p.mas[0])
    ek = 0.5 * (m * (st["vx"] ** 2 + st["vy"] ** 2 + st["vz"] ** 2))[sel].sum()
    return (2.0 * ek - 3.0 * sel.sum() * CIKB * p.att) / p.Q


def kinetic_energy(p: Params, st: dict) -> float:
    """ekin as the integrators define it: 0.5 m1 sum v^2 for the reference
    Nose-Hoover (int-verlet-nh-mav.f:180), per-species masses otherwise
    (int-verlet-mav.f:91)."""
    v2 = st["vx"] ** 2 + st["vy"] ** 2 + st["vz"] ** 2
    if p.integrator == INT_NH and not p.nh_mask:
        return 0.5 * p.mas[0] * v2.sum()
    m = np.array([p.mas[s] for s in range(p.nsorts)])[st["api"] - 1]
    return 0.5 * (m * v2).sum()


def setup(system, params, states):
    """Upload states, build every list (neighbor(0)), compute a(0)
    (AccAtom), and set ekin and the initial thermostat friction."""
    for t, (p, st) in enumerate(zip(params, states)):
        system.upload(st, traj=t)
        system.neighbor(0, traj=t)
    system.accatom()
    for t, (p, st) in enumerate(zip(params, states)):
        kw = dict(ekin=kinetic_energy(p, st))
        if p.integrator == INT_NH and p.ethst_nh:
            kw.update(psi=nh_initial_psi(p, st), s=0.0)
        system.set_scalars(t, **kw)
