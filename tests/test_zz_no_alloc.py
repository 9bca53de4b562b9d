"""fcmd_run_noalloc, the allocation-free variant of fcmd_run (needs a GPU).

Two systems start from the same state; one advances with fcmd_run (stream
launches, library-allocated series buffers), the other with
fcmd_run_noalloc (the device's fixed scratch array).  Twice in a row, the
second run smaller than the first, the states, first-passage times and
largest-cluster series must be the same bits.  A run larger than the
scratch is refused before anything runs.
"""
import ctypes as C

import numpy as np
import pytest

import fcmd
from fcmd import workloads as W

pytestmark = pytest.mark.gpu

NOALLOC_INTS = 65536  # FCMD_NOALLOC_INTS in include/fcmd.h


def run_noalloc(s, nsteps, cluster_every, fpt_size, fpt):
    nsnap = (nsteps + cluster_every - 1) // cluster_every if cluster_every else 0
    series = np.zeros((max(nsnap, 1), s.ntraj), dtype=np.int32)
    fpt = np.ascontiguousarray(fpt, dtype=np.int32).copy()
    L = fcmd.lib()
    L.fcmd_run_noalloc.restype = C.c_int
    fcmd.check(L.fcmd_run_noalloc(s.h, int(nsteps), int(cluster_every), int(fpt_size),
                                  series.ctypes.data_as(C.POINTER(C.c_int)),
                                  fpt.ctypes.data_as(C.POINTER(C.c_int))))
    return series[:nsnap], fpt


def run_alloc(s, nsteps, cluster_every, fpt_size, fpt):
    nsnap = (nsteps + cluster_every - 1) // cluster_every if cluster_every else 0
    series = np.zeros((max(nsnap, 1), s.ntraj), dtype=np.int32)
    fpt = np.ascontiguousarray(fpt, dtype=np.int32).copy()
    fcmd.check(fcmd.lib().fcmd_run(s.h, int(nsteps), int(cluster_every), int(fpt_size),
                                   series.ctypes.data_as(C.POINTER(C.c_int)),
                                   fpt.ctypes.data_as(C.POINTER(C.c_int)), 0))
    return series[:nsnap], fpt


@pytest.mark.parametrize("precision", [fcmd.FP64, fcmd.MIXED])
def test_noalloc_run_matches_the_allocating_run(precision):
    params, states, _ = W.ensemble512(ntraj=12, nmax=4096)
    a = fcmd.System(params, precision)
    b = fcmd.System(params, precision)
    for s in (a, b):
        W.setup(s, params, states)
    fpt0 = np.full(len(params), -1, dtype=np.int32)
    # a small fpt_size so that first passages happen and are recorded
    for nsteps, every in ((120, 20), (45, 10)):
        sa, fa = run_alloc(a, nsteps, every, 4, fpt0)
        sb, fb = run_noalloc(b, nsteps, every, 4, fpt0)
        assert sa.shape == sb.shape == ((nsteps + every - 1) // every, len(params))
        np.testing.assert_array_equal(sa, sb)
        np.testing.assert_array_equal(fa, fb)
        assert (sa >= 1).all()
        for t in range(len(params)):
            da, db = a.download(t), b.download(t)
            for k in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az"):
                np.testing.assert_array_equal(da[k], db[k], err_msg=f"{nsteps} {t} {k}")
            assert a.scalars(t).epot == b.scalars(t).epot
        fpt0 = fa
    a.close()
    b.close()


def test_noalloc_refuses_runs_beyond_its_scratch():
    """(analyses + 1) x trajectories ints of scratch: 255 trajectories fit
    256 analyses (65 535 ints) and not 257 (65 790)."""
    ntraj = 255
    params, states, _ = W.ensemble512(ntraj=ntraj, nmin=500, nmax=600)
    with fcmd.System(params, fcmd.FP64) as s:
        W.setup(s, params, states)
        x0 = s.download(0)["x"]
        fpt = np.full(ntraj, -1, dtype=np.int32)
        assert (257 + 1) * ntraj > NOALLOC_INTS >= (256 + 1) * ntraj
        with pytest.raises(fcmd.FcmdError) as e:
            run_noalloc(s, 257, 1, 50, fpt)
        assert e.value.code == 6
        np.testing.assert_array_equal(s.download(0)["x"], x0)  # nothing ran
        series, _ = run_noalloc(s, 256, 1, 50, fpt)
        assert series.shape == (256, ntraj) and (series >= 1).all()
