"""The nucleation campaign of the ensemble workload run to its end (needs a
GPU): 512 Nose-Hoover trajectories (4.39 M atoms, N log-uniform in
500..32768, rho* U(0.01, 0.05), T* U(0.6, 0.8)) for 5e4 steps in mixed
precision, the Stillinger analysis every 100 steps and the first passage to
a 50-atom cluster recorded on the device.

Over the campaign the vapour nucleates: droplets of several hundred atoms
grow, the widest neighbour list passes 100 entries and list lengths double.
Checked:

* the run ends without an error (list width, skin, Nose-Hoover);
* every first passage is the first analysis at which that trajectory's
  largest cluster reaches 50 atoms (helpers/passage-time.f:55-75), chunk after
  chunk of the series the device recorded;
* the Stillinger analysis of the trajectory with the largest droplet equals
  the oracle's (stoddard-pbc-n-m.f) on the same snapshot and list: labels,
  sizes and the size histogram, bit for bit."""
import numpy as np
import pytest

import fcmd
from fcmd import workloads as W
from test_gpu_scale import _case_from, _oracle_on_snapshot

pytestmark = pytest.mark.gpu

STEPS, CHUNK, EVERY, FPT = 50000, 10000, 100, 50


def test_nucleation_campaign_runs_to_its_end():
    params, states, meta = W.ensemble512()
    n = len(params)
    with fcmd.System(params, fcmd.MIXED) as s:
        W.setup(s, params, states)
        step0 = [s.scalars(t).step for t in range(n)]
        rows, first = [], np.full(n, -1)
        for c in range(STEPS // CHUNK):
            series, fpt = s.run(CHUNK, cluster_every=EVERY, fpt_size=FPT)
            assert series.shape == (CHUNK // EVERY, n)
            rows.append(series)
            fresh = (first < 0) & (fpt >= 0)
            first[fresh] = fpt[fresh]
        series = np.concatenate(rows)
        sc = [s.scalars(t) for t in range(n)]
        assert all(x.err == 0 for x in sc)
        assert all(x.step == step0[t] + STEPS for t, x in enumerate(sc))
        # droplets formed: many trajectories past the passage size, dense lists
        assert (first >= 0).sum() > n // 4
        assert max(x.nlist for x in sc) > 64
        for t in range(n):
            k = np.nonzero(series[:, t] >= FPT)[0]
            want = step0[t] + int(k[0]) * EVERY if k.size else -1
            assert first[t] == want, t
        # the largest droplet's analysis against the oracle on the same snapshot
        t = int(np.argmax(series[-1]))
        assert series[-1, t] > 200
        step = sc[t].step
        o, lst, idx = _oracle_on_snapshot(s, _case_from(params[t], states[t]), "nuc", traj=t)
        out = s.stoddard(step, traj=t)
        o.stoddard(step)
        np.testing.assert_array_equal(out["cl"], o.arr["cl"])
        np.testing.assert_array_equal(out["cs"], o.arr["cs"])
        np.testing.assert_array_equal(out["csi"][1:], o.arr["csi"][1:])
        assert int(np.nonzero(out["csi"][1:])[0].max()) + 1 > 200
