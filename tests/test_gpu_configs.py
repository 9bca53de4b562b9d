"""The batched workloads at their real shapes, in the precision they are run
in (needs a GPU).

* mixture64: 64 trajectories of 65 536 atoms, 20 % condensable A and 80 %
  carrier B (eps_AA 1, eps_BB 0.1, eps_AB 0.3 from an explicit pair table),
  Nose-Hoover on B only, A-only Stillinger analysis, mixed precision.  Every
  trajectory's forces, energy and virial against the oracle's pair-table
  AccAtom on the GPU's own snapshot and list; with the table set to the
  Lorentz-Berthelot values, against the reference's own two-species form
  (for-a-lj-n-pbc-m.f:61-152); A-only cluster labels and histograms bit-exact
  against stoddard-pbc-n-m.f on the same snapshots; the extended Hamiltonian
  of the B-coupled thermostat conserved.
* The A-B dimer at eps_AB = 0.3 against its closed forms on both precision
  paths.
* ensemble512 (64 of its trajectories, the same size, density and
  temperature distributions) in mixed precision: every trajectory's forces,
  energy and virial against the oracle on its snapshot, RMS relative force
  error <= 1e-5, and labels bit-exact.

The oracle's pair-table form is pinned in tests/test_oracle.py (the A-B dimer
closed form at eps_AB = 0.3 and bit-equality with the reference form when the
table holds the Lorentz-Berthelot values).
"""
import numpy as np
import pytest

import cases
import fcmd
import pyoracle as po
from fcmd import workloads as W
from test_gpu_scale import _case_from, _oracle_on_snapshot

pytestmark = pytest.mark.gpu

SIG, EPS, MAS = W.SIG, W.EPS, W.MAS


def _table(p):
    ns = p.nsorts
    ea = np.array([[p.eps_ab[a * fcmd.MAXS + b] for b in range(ns)] for a in range(ns)])
    sa = np.array([[p.sig_ab[a * fcmd.MAXS + b] for b in range(ns)] for a in range(ns)])
    return ea, sa


def _force_err(got, ref):
    num = sum(((got[k] - ref[k]) ** 2).sum() for k in ("ax", "ay", "az"))
    den = sum((ref[k] ** 2).sum() for k in ("ax", "ay", "az"))
    return float(np.sqrt(num / den))


def _h_b(s, p, t):
    """Extended Hamiltonian with the thermostat on species 2 only:
    ekin + epot + psi^2 Q/2 + 3 N_B k T s."""
    sc = s.scalars(t)
    return sc.ekin + sc.epot + sc.psi ** 2 * p.Q / 2 + 3 * p.num[1] * fcmd.CIKB * p.att * sc.s


@pytest.mark.parametrize("precision", [fcmd.MIXED, fcmd.MIXED_KAHAN])
def test_mixture64_full_size_mixed(precision):
    params, states, meta = W.mixture64()
    assert len(params) == 64 and all(p.natom == 65536 for p in params)
    with fcmd.System(params, precision) as s:
        W.setup(s, params, states)
        h0 = np.array([_h_b(s, p, t) for t, p in enumerate(params)])
        series, fpt = s.run(200, cluster_every=100, use_graph=True)
        assert series.shape == (2, 64) and series.min() >= 1
        h1 = np.array([_h_b(s, p, t) for t, p in enumerate(params)])
        # 1 tau of B-coupled Nose-Hoover in mixed precision
        assert np.all(np.abs(h1 - h0) < 1e-3 * np.abs(h0))
        worst = 0.0
        for t, p in enumerate(params):
            case = _case_from(p, states[t])
            ea, sa = _table(p)
            st = s.download(t)
            c = dict(case)
            c.update(st)
            c["atrskin"] = s.scalars(t).atrskin
            lst, idx = s.halflist(t)
            o = po.Oracle(c, "nuc", list_cap=len(lst) + 16, eps_ab=ea, sig_ab=sa)
            o.arr["atnlist"][:len(lst)] = lst
            o.arr["atnidx"][:] = idx
            o.accatom()
            err = _force_err(st, o.arr)
            worst = max(worst, err)
            sc = s.scalars(t)
            assert abs(sc.epot - o.st.epot) < 1e-5 * abs(o.st.epot), t
            assert abs(sc.virial - o.st.virial) < 1e-5 * abs(o.st.virial), t
            if t % 8 == 0:
                # A-only Stillinger (stoddard-pbc-n-m.f) on the same snapshot
                out = s.stoddard(200, traj=t)
                o.stoddard(200)
                np.testing.assert_array_equal(out["cl"], o.arr["cl"])
                np.testing.assert_array_equal(out["cs"], o.arr["cs"])
                np.testing.assert_array_equal(out["csi"][1:], o.arr["csi"][1:])
        assert worst < 1e-5


def test_mixture_table_with_lb_values_matches_the_reference_form():
    """The mixture's table path with eps_AB = sqrt(eps_A eps_B) (the
    Lorentz-Berthelot value) against the oracle's for-a-lj-n-pbc-m.f form on
    the GPU's snapshots, 8 trajectories of 65 536 atoms, mixed and FP64."""
    params, states, meta = W.mixture64(ntraj=8)
    for p in params:
        e1, e2 = p.eps[0], p.eps[1]
        p.eps_ab[1] = p.eps_ab[fcmd.MAXS] = np.sqrt(e1 * e2)
    for prec, tol in ((fcmd.FP64, 1e-12), (fcmd.MIXED, 1e-5)):
        with fcmd.System(params, prec) as s:
            W.setup(s, params, states)
            s.run(30, use_graph=True)
            for t, p in enumerate(params):
                o, lst, idx = _oracle_on_snapshot(s, _case_from(p, states[t]), "nuc", traj=t)
                o.accatom()
                assert _force_err(s.download(t), o.arr) < tol
                sc = s.scalars(t)
                assert abs(sc.epot - o.st.epot) < 10 * tol * abs(o.st.epot)


@pytest.mark.parametrize("precision", [fcmd.FP64, fcmd.MIXED, fcmd.MIXED_KAHAN])
def test_ab_dimer_at_eps_ab_03(precision):
    """V(r) = 4 eps_AB ((s/r)^12 - (s/r)^6) and F(r) = 48 eps_AB/r
    ((s/r)^12 - (s/r)^6/2) for an A-B pair with eps_AB = 0.3 eps (not the
    Lorentz-Berthelot 0.316), sigma_AB = sigma; the mass of each atom its
    species' own."""
    tol = 1e-12 if precision == fcmd.FP64 else 1e-5
    c = cases.make_case(n=2, rho=2.0 / 15.0 ** 3, seed=3, nsorts=2, frac_a=0.5)
    L = c["box"][0]
    for r_sig in (1.0, 2 ** (1 / 6), 1.5):
        r = r_sig * SIG
        c["x"] = np.array([0.5 * L, 0.5 * L + r])
        c["y"] = np.full(2, 0.5 * L)
        c["z"] = np.full(2, 0.5 * L)
        c["mas"] = np.array([MAS, 3.0 * MAS])
        c["sig"] = np.array([SIG, SIG])
        c["eps"] = np.array([EPS, 0.1 * EPS])
        eab = [[EPS, 0.3 * EPS], [0.3 * EPS, 0.1 * EPS]]
        sab = [[SIG, SIG], [SIG, SIG]]
        with fcmd.System(fcmd.make_params(c, "nuc", eps_ab=eab, sig_ab=sab), precision) as s:
            s.upload(c)
            s.neighbor(0)
            z = np.zeros(2)
            s.upload({"ax": z, "ay": z, "az": z})
            s.accatom()
            got = s.download()
            sc = s.scalars()
        sr6 = (SIG / r) ** 6
        e_ab = 0.3 * EPS
        f = 48.0 * e_ab / r * (sr6 * sr6 - 0.5 * sr6)  # on atom 2 along +x
        fscale = 24.0 * e_ab / SIG
        assert abs(got["ax"][0] - (-f / MAS)) <= tol * fscale / MAS
        assert abs(got["ax"][1] - (f / (3.0 * MAS))) <= tol * fscale / (3.0 * MAS)
        assert abs(sc.epot - 4.0 * e_ab * (sr6 * sr6 - sr6)) <= tol * e_ab
        assert abs(sc.virial - f * r) <= tol * fscale * SIG


@pytest.mark.parametrize("precision", [fcmd.MIXED, fcmd.MIXED_KAHAN])
def test_ensemble_mixed_precision_against_the_oracle(precision):
    """64 trajectories drawn like ensemble512's, advanced in mixed precision
    (pair forces summed in FP64, or in FP32 with Kahan compensation) with
    graph replay and analyses; on each trajectory's snapshot the forces are
    within 1e-5 RMS of the oracle's FP64 AccAtom on the same list, energy and
    virial within 1e-5, and the Stillinger labels bit-exact."""
    params, states, meta = W.ensemble512(ntraj=64)
    with fcmd.System(params, precision) as s:
        W.setup(s, params, states)
        series, fpt = s.run(100, cluster_every=20, fpt_size=50, use_graph=True)
        assert series.shape == (5, 64)
        errs = []
        for t, p in enumerate(params):
            o, lst, idx = _oracle_on_snapshot(s, _case_from(p, states[t]), "nuc", traj=t)
            o.accatom()
            errs.append(_force_err(s.download(t), o.arr))
            sc = s.scalars(t)
            assert abs(sc.epot - o.st.epot) < 1e-5 * abs(o.st.epot), t
            assert abs(sc.virial - o.st.virial) < 1e-5 * abs(o.st.virial), t
            if t % 4 == 0:
                out = s.stoddard(100, traj=t)
                o.stoddard(100)
                np.testing.assert_array_equal(out["cl"], o.arr["cl"])
                np.testing.assert_array_equal(out["csi"][1:], o.arr["csi"][1:])
        assert max(errs) < 1e-5
