"""Energy conservation over 1e4 steps, against the reference's own (needs a
GPU; the reference is the compiled Fortran of oracle/_ref, run on the host).

Both sides start from the same seeded state and run 10 000 steps in FP64.
They follow the same algorithm but sum in different orders, so the two
trajectories part after a few thousand steps (LJ dynamics is chaotic); what
must match is the quality of conservation of

* velocity Verlet with the adaptive skin (int-verlet-mav.f, NVE):
  ekin + epot;
* Nose-Hoover (int-verlet-nh-mav.f, ethst = 'N'): the extended Hamiltonian
  ekin + epot + psi^2 Q/2 + 3 N k T s.

Drift per tau is the slope of a least-squares line through the conserved
quantity (relative to its start) sampled every 500 steps.  Three seeds per
side; the bar is the GPU's mean |drift| no worse than the reference's mean
plus twice the reference's spread over its seeds (two sets of chaotic
realisations of the same statistics; the unshifted cutoff makes the energy
jump whenever a pair crosses it, on both sides alike).  The largest
excursion from the start is reported beside it."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import fcmd
import pyoracle as po
from fcmd import workloads as W

pytestmark = pytest.mark.gpu

NSTEPS, EVERY = 10000, 500


def _harness_case(p, st, psi):
    w = float(np.float32(p.nbtim))
    return dict(natom=p.natom, nsorts=p.nsorts, api=st["api"], box=np.array(p.box[:]),
                mas=np.array(p.mas[:p.nsorts]), sig=np.array(p.sig[:p.nsorts]),
                eps=np.array(p.eps[:p.nsorts]), num=np.array(p.num[:p.nsorts]),
                atrcut=p.atrcut, atrskin=p.atrskin, atskin=p.atskin, dt=p.dt, Q=p.Q, att=p.att,
                nbtim=p.nbtim, btim=p.btim, atnenum=p.atnenum, ethst_nh=p.ethst_nh, psi=psi,
                s=0.0, dmmh=np.full(256, (p.atrskin - p.atrcut) / w),
                dmmsu=256.0 * (p.atrskin - p.atrcut) / w, dmms=0.0, dmmax=0.0, dmmav=0.0,
                dmmst=0, dmmi=0,
                **{k: st[k] for k in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az")})


def _conserved(p, ekin, epot, psi, s):
    if p.integrator != W.INT_NH:
        return ekin + epot
    return ekin + epot + psi ** 2 * p.Q / 2 + 3 * p.natom * fcmd.CIKB * p.att * s


SEEDS = (12, 13, 14)


def _system(variant, seed):
    if variant == "mav":
        params, states, _ = W.lj_vapour(seed=seed)   # N=1000, rho* 0.02, T* 0.7, NVE
    else:
        p = W._params(1000, 0.03, 0.65, W.INT_NH, W.ST_N)
        params, states = [p], [W._state(p, seed)]
    return params, states


def _reference(variant, seed):
    params, states = _system(variant, seed)
    p, st = params[0], states[0]
    psi0 = W.nh_initial_psi(p, st) if p.integrator == W.INT_NH else 0.0
    ekin0 = W.kinetic_energy(p, st)
    actions = [(1, 0, 0), (2, 0, 0)] + [(3, k + 1, k + EVERY) for k in range(0, NSTEPS, EVERY)]
    recs = po.run_harness(variant, _harness_case(p, st, psi0), actions, timeout=900)
    e0 = _conserved(p, ekin0, recs[1]["epot"], psi0, 0.0)
    return e0, np.array([_conserved(p, r["ekin"], r["epot"], r["psi"], r["s"]) for r in recs[2:]])


def _gpu(variant, seed):
    params, states = _system(variant, seed)
    p = params[0]
    with fcmd.System(params, fcmd.FP64) as s:
        W.setup(s, params, states)
        sc = s.scalars()
        e0 = _conserved(p, sc.ekin, sc.epot, sc.psi, sc.s)
        e = []
        for _ in range(NSTEPS // EVERY):
            s.run(EVERY, use_graph=True)
            sc = s.scalars()
            e.append(_conserved(p, sc.ekin, sc.epot, sc.psi, sc.s))
    return e0, np.array(e)


def _drift_per_tau(e0, e, dt):
    """Slope of (H - H0)/|H0| against time in tau (least squares, through
    the start and the samples)."""
    t = np.arange(len(e) + 1) * EVERY * dt / W.TAU
    y = np.concatenate([[0.0], (e - e0) / abs(e0)])
    return float(np.polyfit(t, y, 1)[0])


@pytest.mark.skipif(not (po.have_harness("mav") and po.have_harness("nose")),
                    reason="reference harness not built")
@pytest.mark.parametrize("variant", ["mav", "nose"])
def test_energy_drift_1e4_steps_no_worse_than_reference(variant):
    dt = _system(variant, SEEDS[0])[0][0].dt
    with ThreadPoolExecutor(len(SEEDS)) as ex:  # the reference runs on host cores
        refs = list(ex.map(lambda sd: _reference(variant, sd), SEEDS))
    gpus = [_gpu(variant, sd) for sd in SEEDS]
    for (e0r, er), (e0g, eg) in zip(refs, gpus):
        # same start, same first energies (the trajectories part later)
        assert e0g == pytest.approx(e0r, rel=1e-12)
        assert eg[0] == pytest.approx(er[0], rel=1e-6)
    d_ref = np.array([abs(_drift_per_tau(e0, e, dt)) for e0, e in refs])
    d_gpu = np.array([abs(_drift_per_tau(e0, e, dt)) for e0, e in gpus])
    x_ref = max(np.abs(e - e0).max() / abs(e0) for e0, e in refs)
    x_gpu = max(np.abs(e - e0).max() / abs(e0) for e0, e in gpus)
    print(f"{variant}: |drift| per tau over {NSTEPS} steps, seeds {SEEDS}: reference {d_ref}, GPU {d_gpu}; "
          f"largest excursion: reference {x_ref:.3e}, GPU {x_gpu:.3e}")
    assert d_gpu.mean() <= d_ref.mean() + 2.0 * d_ref.std(ddof=1)
