"""The drop-in: the reference's own driver (main.f, read-3d.f, read-rst.f,
stopdead, writers) linked once with the reference plugins and once with the
GPU plugins of fortran-cluster-md_amd/fortran, run on the same deck and
restart file.  Cluster-size histograms (.csi) must be identical; energies
(.ave) agree to 1e-9; the per-rebuild log lines name the same steps and pair
counts.  Both programs are built by oracle/Makefile (targets progs, dropin)
from /root/reference where it exists, and travel as binaries."""
import os
import shutil

import numpy as np
import pytest

import cases
import decks
import pyoracle as po

pytestmark = pytest.mark.gpu

BIN = os.path.join(po.REF_DIR, "bin", "mxa6000")


def _have(name):
    return os.path.exists(os.path.join(BIN, name))


def _run_pair(tmp_path, variant, case, ntim, ftcsi, ftave, ethst="-", env=None):
    o = po.Oracle(case, "mav" if variant == "mav" else "nuc")
    o.neighbor(0)
    o.accatom()
    c2 = dict(case, ax=o.arr["ax"], ay=o.arr["ay"], az=o.arr["az"])
    out = {}
    for prog in (f"cluster_{variant}", f"cluster_{variant}_hip"):
        d = tmp_path / prog
        d.mkdir()
        decks.write_deck(str(d / "t.3d"), c2, ntim=ntim, ftcsi=ftcsi, ftave=ftave, sctim=ftcsi,
                         ethst=ethst)
        decks.write_restart(str(d / "start.rst"), c2)
        r = decks.run_program(os.path.join(BIN, prog), "t.3d", str(d), env=env)
        assert r.returncode == 0, (prog, r.stdout[-2000:], r.stderr[-2000:])
        out[prog] = dict(stdout=r.stdout, csi=decks.read_csi(str(d / "test.csi")),
                         ave=decks.read_ave(str(d / "test.ave")), dir=d)
    return out[f"cluster_{variant}"], out[f"cluster_{variant}_hip"]


def _rebuild_lines(stdout):
    rows = []
    for line in stdout.splitlines():
        p = line.split()
        if len(p) == 7:
            try:
                rows.append((int(p[0]), int(p[1])))
            except ValueError:
                pass
    return rows


def _compare(ref, hip, csi_steps=None):
    steps = sorted(ref["csi"]) if csi_steps is None else csi_steps
    assert steps and set(steps) <= set(hip["csi"]), (sorted(hip["csi"]), hip["stdout"][-3000:])
    for st in steps:
        a, b = ref["csi"][st], hip["csi"][st]
        assert sorted(a) == sorted(b), st
        for size in a:
            assert a[size][0] == b[size][0], (st, size)
            assert a[size][1] == pytest.approx(b[size][1], rel=1e-9, abs=1e-14)
    assert len(ref["ave"]) == len(hip["ave"]) > 0
    for (s1, k1, p1), (s2, k2, p2) in zip(ref["ave"], hip["ave"]):
        assert s1 == s2
        assert k2 == pytest.approx(k1, rel=1e-9)
        assert p2 == pytest.approx(p1, rel=1e-9)


@pytest.mark.skipif(not (_have("cluster_mav") and _have("cluster_mav_hip")), reason="programs not built")
@pytest.mark.parametrize("resident", ["0", "1"])
def test_dropin_mav_nve(tmp_path, resident):
    case = decks.deck_units(cases.make_case(n=1000, rho=0.02, seed=21, dense_frac=0.3, ncenters=4))
    ref, hip = _run_pair(tmp_path, "mav", case, ntim=400, ftcsi=100, ftave=50,
                         env={"FCMD_RESIDENT": resident})
    _compare(ref, hip)
    assert _rebuild_lines(ref["stdout"]) == _rebuild_lines(hip["stdout"])


@pytest.mark.skipif(not (_have("cluster_nose") and _have("cluster_nose_hip")), reason="programs not built")
def test_dropin_nose_nh(tmp_path):
    case = decks.deck_units(cases.make_case(n=800, rho=0.03, seed=22, dense_frac=0.3, ethst_nh=1))
    ref, hip = _run_pair(tmp_path, "nose", case, ntim=300, ftcsi=100, ftave=50, ethst="N")
    _compare(ref, hip)


@pytest.mark.skipif(not (_have("cluster_fre") and _have("cluster_fre_hip")), reason="programs not built")
def test_dropin_fre(tmp_path):
    """stoddard-pbc-n-m-fre.f is well defined only on its first call in a
    process (its cvtl() is never zeroed), and main.f:112 already spends that
    call on stoddard(-1): the reference program's .csi is undefined from its
    first row.  So the energies are compared with the reference program over
    the whole run, and the drop-in's first histogram with the oracle's
    (= the reference routine's first call) on the same start."""
    case = decks.deck_units(cases.make_case(n=800, rho=0.04, seed=23, dense_frac=0.4, nsorts=2,
                                            frac_a=0.5))
    ref, hip = _run_pair(tmp_path, "fre", case, ntim=200, ftcsi=100, ftave=50)
    assert len(ref["ave"]) == len(hip["ave"]) > 0
    for (s1, k1, p1), (s2, k2, p2) in zip(ref["ave"], hip["ave"]):
        assert s1 == s2 and k2 == pytest.approx(k1, rel=1e-9) and p2 == pytest.approx(p1, rel=1e-9)
    o = po.Oracle(case, "fre")
    o.neighbor(0)
    o.stoddard(0)
    want = {int(k): int(o.arr["csi"][k]) for k in np.nonzero(o.arr["csi"][1:])[0] + 1}
    got = {k: v[0] for k, v in hip["csi"][0].items()}
    assert got == want
