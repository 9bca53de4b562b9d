"""Energy conservation over 1e4 steps, against the reference's own (needs a
GPU; the reference is the compiled Fortran of oracle/_ref, run on the host).

Both sides start from the same seeded state and run 10 000 steps in FP64.
They follow the same algorithm but sum in different orders, so the two
trajectories part after a few thousand steps (LJ dynamics is chaotic); what
must match is the quality of conservation.  The measure is the largest
excursion of the conserved quantity from its start, relative to it:

* velocity Verlet with the adaptive skin (int-verlet-mav.f, NVE):
  ekin + epot;
* Nose-Hoover (int-verlet-nh-mav.f, ethst = 'N'): the extended Hamiltonian
  ekin + epot + psi^2 Q/2 + 3 N k T s.

The bar is the GPU's excursion within 2x the reference's (two chaotic
realisations of the same statistics; the unshifted cutoff makes the energy
jump whenever a pair crosses it, on both sides alike)."""
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


@pytest.mark.skipif(not (po.have_harness("mav") and po.have_harness("nose")),
                    reason="reference harness not built")
@pytest.mark.parametrize("variant", ["mav", "nose"])
def test_energy_excursion_1e4_steps_within_reference(variant):
    if variant == "mav":
        params, states, _ = W.lj_vapour()   # N=1000, rho* 0.02, T* 0.7, NVE
    else:
        p = W._params(1000, 0.03, 0.65, W.INT_NH, W.ST_N)
        params, states = [p], [W._state(p, 12)]
    p, st = params[0], states[0]
    psi0 = W.nh_initial_psi(p, st) if p.integrator == W.INT_NH else 0.0
    ekin0 = W.kinetic_energy(p, st)

    actions = [(1, 0, 0), (2, 0, 0)] + [(3, k + 1, k + EVERY) for k in range(0, NSTEPS, EVERY)]
    recs = po.run_harness(variant, _harness_case(p, st, psi0), actions, timeout=900)
    e0_ref = _conserved(p, ekin0, recs[1]["epot"], psi0, 0.0)
    e_ref = np.array([_conserved(p, r["ekin"], r["epot"], r["psi"], r["s"]) for r in recs[2:]])

    with fcmd.System(params, fcmd.FP64) as s:
        W.setup(s, params, states)
        sc = s.scalars()
        e0 = _conserved(p, sc.ekin, sc.epot, sc.psi, sc.s)
        e_gpu = []
        for _ in range(NSTEPS // EVERY):
            s.run(EVERY, use_graph=True)
            sc = s.scalars()
            e_gpu.append(_conserved(p, sc.ekin, sc.epot, sc.psi, sc.s))
    e_gpu = np.array(e_gpu)
    # same start, same first energies (the start is identical and the two
    # trajectories only part later)
    assert e0 == pytest.approx(e0_ref, rel=1e-12)
    assert e_gpu[0] == pytest.approx(e_ref[0], rel=1e-6)
    d_ref = np.abs(e_ref - e0_ref).max() / abs(e0_ref)
    d_gpu = np.abs(e_gpu - e0).max() / abs(e0)
    print(f"{variant}: max |H - H0| / |H0| over {NSTEPS} steps: reference {d_ref:.3e}, GPU {d_gpu:.3e}")
    assert d_gpu <= 2.0 * d_ref
