"""On-device nucleation statistics (needs a GPU): the largest-cluster series,
first-passage times and sizer-style cluster-size counts that fcmd_run keeps
on the device, against the same quantities taken from the per-snapshot
Stillinger histograms of an identical run (fcmd_stoddard between run
chunks), with sizer's counting rule restated from helpers/sizer.f:59-103."""
import numpy as np
import pytest

import cases
import fcmd

pytestmark = pytest.mark.gpu

EVERY, NSNAP = 50, 8


def _sizer(snapshots, nmax, warmup):
    """helpers/sizer.f:59-103, SWITCH = 1: rows (cg, cn) of each snapshot in
    increasing cg; once a row with cg >= nmax is met, that row is still
    counted (if ts >= warmup) and nothing after it."""
    count = None
    flag = False
    for ts, csi in snapshots:
        if count is None:
            count = np.zeros_like(csi)
        for cg in range(1, len(csi)):
            cn = int(csi[cg])
            if cn == 0 or flag:
                continue
            if cg >= nmax:
                flag = True
            if ts >= warmup:
                count[cg] += cn
    return count, flag


def _systems():
    specs = [dict(n=600, rho=0.05, seed=71, dense_frac=0.5, ncenters=2),
             dict(n=900, rho=0.04, seed=72, dense_frac=0.4, ncenters=3),
             dict(n=400, rho=0.03, seed=73, dense_frac=0.2, ncenters=1)]
    cs = [cases.make_case(ethst_nh=1, **k) for k in specs]
    params = [fcmd.make_params(c, "nose") for c in cs]

    def build():
        s = fcmd.System(params)
        for t, c in enumerate(cs):
            s.upload(c, traj=t)
            s.set_scalars(t, psi=float(c["psi"]), s=float(c["s"]), dmmh=c["dmmh"],
                          dmmsu=float(c["dmmsu"]), atrskin=float(c["atrskin"]))
            s.neighbor(0, traj=t)
        z = [np.zeros(c["natom"]) for c in cs]
        for t, c in enumerate(cs):
            s.upload({"ax": z[t], "ay": z[t], "az": z[t]}, traj=t)
        s.accatom()
        return s
    return cs, build


@pytest.mark.parametrize("nmax,warmup", [(6, 100), (10**6, 0)])
def test_device_statistics_match_snapshot_histograms(nmax, warmup):
    cs, build = _systems()
    a = build()
    step0 = a.scalars(0).step
    a.size_counts(nmax, warmup=step0 + warmup)
    series, fpt = a.run(EVERY * NSNAP, cluster_every=EVERY, fpt_size=8)
    b = build()
    snaps = [[] for _ in cs]
    for k in range(NSNAP):
        for t in range(len(cs)):
            out = b.stoddard(step0 + k * EVERY, traj=t)
            snaps[t].append((step0 + k * EVERY, out["csi"].copy()))
        b.run(EVERY, cluster_every=0)
    for t in range(len(cs)):
        largest = [int(np.nonzero(h[1:])[0].max()) + 1 if h[1:].any() else 0 for _, h in snaps[t]]
        assert list(series[:, t]) == largest
        first = next((ts for (ts, _), g in zip(snaps[t], largest) if g >= 8), -1)
        assert fpt[t] == first
        want, stopped = _sizer(snaps[t], nmax, step0 + warmup)
        got, got_stopped = a.get_size_counts(t)
        np.testing.assert_array_equal(got[1:], want[1:])
        assert got_stopped == stopped
    # the two runs are the same run
    for t in range(len(cs)):
        np.testing.assert_array_equal(a.download(t)["x"], b.download(t)["x"])


def test_device_statistics_match_the_oracle_run():
    """The same statistics against the oracle: each trajectory integrated by
    the C restatement of the reference (int-verlet-nh-mav.f with
    neighbor-pbc-mav.f, for-a-lj-n-pbc-m.f) from the same state, its
    Stillinger analysis (stoddard-pbc-n-m.f) taken before the steps the
    device analyses (main.f:133-164 order).  The largest-cluster series, the
    first passage to 8 atoms (helpers/passage-time.f:55-75: the first
    analysis at which the largest cluster reaches the size) and the sizer
    counts (helpers/sizer.f:59-103) must agree exactly."""
    import pyoracle as po

    cs, build = _systems()
    a = build()
    step0 = a.scalars(0).step
    nmax, warmup = 6, 100
    a.size_counts(nmax, warmup=step0 + warmup)
    series, fpt = a.run(EVERY * NSNAP, cluster_every=EVERY, fpt_size=8)
    for t, c in enumerate(cs):
        o = po.Oracle(c, "nose")
        o.neighbor(0)
        o.accatom()
        snaps, largest = [], []
        for k in range(NSNAP):
            st = step0 + k * EVERY
            o.stoddard(st)
            snaps.append((st, o.arr["csi"].copy()))
            h = o.arr["csi"][1:]
            largest.append(int(np.nonzero(h)[0].max()) + 1 if h.any() else 0)
            for i in range(st, st + EVERY):
                o.integrator(i)
        assert list(series[:, t]) == largest, t
        assert max(largest) > 1
        first = next((ts for (ts, _), g in zip(snaps, largest) if g >= 8), -1)
        assert fpt[t] == first, t
        want, stopped = _sizer(snaps, nmax, step0 + warmup)
        got, got_stopped = a.get_size_counts(t)
        np.testing.assert_array_equal(got[1:], want[1:])
        assert got_stopped == stopped
        # and the trajectories themselves still agree (summation order only)
        got_x = a.download(t)["x"]
        L = c["box"][0]
        d = np.abs(got_x - o.arr["x"])
        assert np.minimum(d, np.abs(d - L)).max() < 1e-6 * L
