"""The package's restart writer and reader (fcmd.restart, the layout of
wrt-rst.f:30-45 / read-rst.f:37-49) without a GPU: a round trip through
Python, and the reference's own program (oracle/_ref, built from the
reference sources) reading a file the writer made and writing its own,
which the reader reads back."""
import os

import numpy as np
import pytest

import cases
import decks
import fcmd
import pyoracle as po
from fcmd import restart

BIN = os.path.join(po.REF_DIR, "bin", "mxa6000")


def _state(n, seed):
    c = cases.make_case(n=n, rho=0.02, seed=seed)
    x, y, z, vx, vy, vz = fcmd.init_random(seed, n, c["box"], 0.9 * c["sig"][0], c["att"], c["mas"],
                                           np.asarray(c["api"], dtype=np.int32))
    rng = np.random.default_rng(seed)
    a = {k: rng.normal(size=n) * 1e-6 for k in ("ax", "ay", "az")}
    return c, dict(x=x, y=y, z=z, vx=vx, vy=vy, vz=vz, **a)


def test_round_trip(tmp_path):
    c, st = _state(300, 3)
    p = str(tmp_path / "a.rst")
    restart.write_restart(p, st, mxa=500, ekin=1.25, epot=-0.5)
    back = restart.read_restart(p, 300)
    assert back["mxa"] == 500 and back["ekin"] == 1.25 and back["epot"] == -0.5
    for k in restart.KEYS:
        np.testing.assert_array_equal(back[k], st[k])
    # 9 records, each framed by its length: 4 + 40 + 32 + 6 x 3 x 500 x 8 bytes of payload
    assert os.path.getsize(p) == 9 * 8 + 4 + 40 + 32 + 6 * 3 * 500 * 8
    with pytest.raises(ValueError):
        restart.write_restart(p, st, mxa=299)


@pytest.mark.skipif(not os.path.exists(os.path.join(BIN, "cluster_mav")), reason="reference program not built")
def test_reference_program_reads_and_writes_the_layout(tmp_path):
    """cluster_mav (the reference's main.f with the mav plugins) starts from
    the writer's file (readRST) and writes wrtRST at step 0; the reader
    returns the same state: positions as neighbor(btim) wrapped them,
    velocities as stopdead(-1) left them (zero momentum and exactly T
    already: a rescale by 1 to rounding), accelerations as written."""
    c, st = _state(500, 7)
    c.update(st)
    c = decks.deck_units(c)
    st = {k: c[k] for k in restart.KEYS}
    decks.write_deck(str(tmp_path / "t.3d"), c, ntim=0, ftcsi=0, ftave=0, sctim=1, ftrst=1000)
    restart.write_restart(str(tmp_path / "start.rst"), st, mxa=6000)
    r = decks.run_program(os.path.join(BIN, "cluster_mav"), "t.3d", str(tmp_path))
    assert r.returncode == 0, r.stdout[-2000:]
    back = restart.read_restart(str(tmp_path / "test.rst"), c["natom"])
    assert back["mxa"] == 6000
    L = c["box"][0]
    for k in ("x", "y", "z"):
        assert np.abs(back[k] - st[k]).max() <= 4e-16 * L, k
    for k in ("vx", "vy", "vz"):
        assert np.abs(back[k] - st[k]).max() <= 1e-12 * np.abs(st[k]).max(), k
    for k in ("ax", "ay", "az"):
        np.testing.assert_array_equal(back[k], st[k])
