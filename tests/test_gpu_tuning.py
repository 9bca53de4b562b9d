"""The launch tunables (force-kernel load batch, list-search load batch,
rebuild grid) and the autotuner (needs a GPU).  No tunable may change a
single bit of the trajectory: every variant sums the same terms in the same
order."""
import numpy as np
import pytest

import fcmd
from fcmd import workloads as W

pytestmark = pytest.mark.gpu

CONFIGS = [(4, 2, 2048), (2, 4, 1024), (8, 2, 4096), (8, 4, 2048)]


def _small_ensemble():
    params, states, _ = W.ensemble512(ntraj=12, nmin=300, nmax=6000)
    return params, states


@pytest.mark.parametrize("precision", [fcmd.FP64, fcmd.MIXED])
def test_every_tuning_gives_the_same_bits(precision):
    params, states = _small_ensemble()
    outs = []
    for cfg in CONFIGS:
        with fcmd.System(params, precision) as s:
            s.set_tuning(*cfg)
            assert tuple(s.tuning().values()) == cfg
            W.setup(s, params, states)
            s.run(60, use_graph=True)
            outs.append([(s.download(t), s.scalars(t).epot, s.scalars(t).psi, s.scalars(t).nrebuild)
                         for t in range(len(params))])
    for other in outs[1:]:
        for (a, ea, pa, ra), (b, eb, pb, rb) in zip(outs[0], other):
            for k in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az"):
                np.testing.assert_array_equal(a[k], b[k])
            assert (ea, pa, ra) == (eb, pb, rb)


def test_autotune_picks_a_candidate_and_the_run_goes_on():
    params, states = _small_ensemble()
    with fcmd.System(params, fcmd.MIXED) as s:
        W.setup(s, params, states)
        before = [s.scalars(t).step for t in range(len(params))]
        got = s.autotune(rounds=2)
        assert got["force_u"] in (2, 4, 8) and got["nbr_cu"] in (2, 4)
        assert got["wl_grid"] in (1024, 2048, 4096)
        assert s.tuning() == got
        # the tuning steps are ordinary steps of every trajectory
        assert [s.scalars(t).step - b for t, b in enumerate(before)] == [12] * len(params)
        s.run(20, use_graph=True)
        with pytest.raises(fcmd.FcmdError):
            s.set_tuning(3, 2, 2048)


@pytest.mark.parametrize("precision", [fcmd.FP64, fcmd.MIXED])
def test_list_entry_width_gives_the_same_bits(precision, monkeypatch):
    """16-bit (trajectory-relative) and 32-bit list entries hold the same
    lists: states, energies, the reference-layout half list, the
    largest-cluster series and the cluster labels agree bit for bit (the
    clusters' kinetic-energy sums, accumulated with atomics, to rounding)."""
    params, states = _small_ensemble()
    outs = []
    for bits in ("16", "32"):
        if bits == "32":
            monkeypatch.setenv("FCMD_LIST_BITS", "32")
        else:
            monkeypatch.delenv("FCMD_LIST_BITS", raising=False)
        with fcmd.System(params, precision) as s:
            W.setup(s, params, states)
            series, _ = s.run(60, cluster_every=30, use_graph=True)
            halves = [s.halflist(t) for t in (0, len(params) - 1)]
            clus = [s.stoddard(0, traj=t) for t in (0, len(params) - 1)]
            outs.append(([(s.download(t), s.scalars(t).epot, s.scalars(t).psi, s.scalars(t).nrebuild)
                          for t in range(len(params))], halves, clus, series))
    (sa, ha, ca, ra_), (sb, hb, cb, rb_) = outs
    np.testing.assert_array_equal(ra_, rb_)
    for (a, ea, pa, ra), (b, eb, pb, rb) in zip(sa, sb):
        for k in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az"):
            np.testing.assert_array_equal(a[k], b[k])
        assert (ea, pa, ra) == (eb, pb, rb)
    for (la, ia), (lb, ib) in zip(ha, hb):
        np.testing.assert_array_equal(la, lb)
        np.testing.assert_array_equal(ia, ib)
    for x, y in zip(ca, cb):
        for k in x:
            u, v = np.asarray(x[k]), np.asarray(y[k])
            if u.dtype.kind == "f":  # kinetic-energy sums: atomics, order-free to rounding
                np.testing.assert_allclose(u, v, rtol=1e-12, atol=0)
            else:
                np.testing.assert_array_equal(u, v)


@pytest.mark.parametrize("precision", [fcmd.FP64, fcmd.MIXED])
@pytest.mark.parametrize("use_graph", [True, False])
def test_forces_beside_the_rebuild_give_the_same_bits(use_graph, precision, monkeypatch):
    """The step runs the rebuild of the trajectories that rebuild (and then
    their forces) on a side stream beside the forces of the trajectories
    that keep their lists; the same step on one stream (FCMD_OVERLAP=0) must
    give the same bits, graph or not."""
    params, states = _small_ensemble()
    outs = []
    for ov in ("1", "0"):
        monkeypatch.setenv("FCMD_OVERLAP", ov)
        with fcmd.System(params, precision) as s:
            W.setup(s, params, states)
            series, _ = s.run(80, cluster_every=40, use_graph=use_graph)
            outs.append(([(s.download(t), s.scalars(t).epot, s.scalars(t).psi, s.scalars(t).nrebuild)
                          for t in range(len(params))], series))
    (sa, ra_), (sb, rb_) = outs
    np.testing.assert_array_equal(ra_, rb_)
    assert sum(r for *_, r in sa) > len(params)  # the run did rebuild
    for (a, ea, pa, ra), (b, eb, pb, rb) in zip(sa, sb):
        for k in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az"):
            np.testing.assert_array_equal(a[k], b[k])
        assert (ea, pa, ra) == (eb, pb, rb)


@pytest.mark.parametrize("precision", [fcmd.FP64, fcmd.MIXED])
def test_resort_interval_keeps_the_atoms_bits(precision, monkeypatch):
    """Rebuilds that keep the atoms in their slots (a sorted view for the
    list search) and rebuilds that re-sort them by cell build the same
    lists in the same order (cells z-major, reference numbers within a
    cell), so every per-atom quantity is the same bits whatever the re-sort
    interval.  Velocity Verlet with the adaptive skin: no sum over atoms
    feeds back into the motion (the rebuild rule takes a maximum), so whole
    trajectories agree bit for bit, and so do the reference-layout half
    lists, the cluster labels and the largest-cluster series; the energies
    (sums in slot order) to rounding."""
    params, states = _small_ensemble()
    for p in params:
        p.integrator, p.ethst_nh = fcmd.INT_MAV, 0
    res = []
    for resort in ("1", "3", "8"):
        monkeypatch.setenv("FCMD_RESORT", resort)
        with fcmd.System(params, precision) as s:
            W.setup(s, params, states)
            series, _ = s.run(90, cluster_every=30, use_graph=True)
            st = [s.download(t) for t in range(len(params))]
            sc = [s.scalars(t) for t in range(len(params))]
            hl = [s.halflist(t) for t in (0, 5, 11)]
            cl = [s.stoddard(90, traj=t)["cl"] for t in (0, 5, 11)]
        res.append((st, sc, hl, cl, series))
    monkeypatch.delenv("FCMD_RESORT", raising=False)
    assert max(sc.nrebuild for sc in res[0][1]) >= 4  # several rebuilds, so both kinds ran
    for st, sc, hl, cl, series in res[1:]:
        for a, b in zip(res[0][0], st):
            for k in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az"):
                np.testing.assert_array_equal(a[k], b[k])
        for a, b in zip(res[0][1], sc):
            assert (a.nrebuild, a.npairs, a.nlist, a.atrskin) == (b.nrebuild, b.npairs, b.nlist, b.atrskin)
            assert b.epot == pytest.approx(a.epot, rel=1e-12)
        for (la, ia), (lb, ib) in zip(res[0][2], hl):
            np.testing.assert_array_equal(la, lb)
            np.testing.assert_array_equal(ia, ib)
        for a, b in zip(res[0][3], cl):
            np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(res[0][4], series)
