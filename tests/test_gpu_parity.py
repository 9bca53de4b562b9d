"""HIP path vs the oracle and the reference's golden vectors (needs a GPU).

Bars (the reference's FP64 semantics):
* neighbour lists: identical to the reference's half list (same pairs, same
  ascending order) on the same positions; wrapped positions bit-exact;
* forces and energy: within 1e-12 relative after one FP64 step (sums are
  re-associated on the GPU, every pair term is the reference's);
* cluster labels / size histograms: bit-exact on identical snapshots
  (same positions, same list); KE sums within 1e-12;
* mixed precision: RMS relative force error <= 1e-5 against FP64.
"""
import numpy as np
import pytest

import cases
import fcmd
import pyoracle as po
from conftest import golden_names, load_golden

pytestmark = pytest.mark.gpu

REL = 1e-12


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    scale = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / scale)


def acc_err(got, ref):
    return max(rel_err(got[k], ref[k]) for k in ("ax", "ay", "az"))


def gpu_system(case, variant, precision=fcmd.FP64, list_width=0):
    s = fcmd.System(fcmd.make_params(case, variant), precision, list_width=list_width)
    s.upload(case)
    s.set_scalars(psi=float(case["psi"]), s=float(case["s"]), dmmh=case["dmmh"],
                  dmms=float(case["dmms"]), dmmsu=float(case["dmmsu"]), dmmax=float(case["dmmax"]),
                  dmmav=float(case["dmmav"]), dmmst=int(case["dmmst"]), dmmi=int(case["dmmi"]),
                  atrskin=float(case["atrskin"]))
    return s


def oracle_from_gpu(s, case, variant):
    """An oracle holding exactly the GPU's current snapshot: positions,
    velocities, list and skin."""
    st = s.download()
    c = dict(case)
    c.update(st)
    sc = s.scalars()
    c["atrskin"] = sc.atrskin
    o = po.Oracle(c, variant)
    lst, idx = s.halflist()
    o.arr["atnlist"][:len(lst)] = lst
    o.arr["atnidx"][:] = idx
    return o


def test_device_present():
    assert fcmd.device_count() >= 1


@pytest.mark.parametrize("name", golden_names())
def test_golden_first_step(name):
    """neighbor(0) + AccAtom on the fixture's start state vs the reference."""
    variant, case, actions, recs = load_golden(name)
    s = gpu_system(case, variant)
    assert s.neighbor(0) == recs[0]["step"]
    got = s.download()
    for k in ("x", "y", "z"):
        np.testing.assert_array_equal(got[k], recs[0][k], err_msg=f"wrapped {k}")
    lst, idx = s.halflist()
    np.testing.assert_array_equal(lst, recs[0]["atnlist"])
    np.testing.assert_array_equal(idx, recs[0]["atnidx"])
    s.upload({"ax": np.zeros(case["natom"]), "ay": np.zeros(case["natom"]), "az": np.zeros(case["natom"])})
    s.accatom()
    got = s.download()
    r = recs[1]
    assert acc_err(got, r) < REL
    assert abs(s.scalars().epot - r["epot"]) <= REL * abs(r["epot"])


@pytest.mark.parametrize("name", golden_names())
def test_golden_trajectory(name):
    """Replay the fixture's actions; integrated states stay within 1e-9 of the
    reference (summation order only), every rebuild decision matches, and each
    cluster analysis is bit-exact against the oracle on the GPU's snapshot."""
    variant, case, actions, recs = load_golden(name)
    s = gpu_system(case, variant)
    stt, ost = {}, {}
    for (code, a1, a2), r in zip(actions, recs):
        if code == 1:
            s.neighbor(a1)
        elif code == 2:
            n = case["natom"]
            s.upload({"ax": np.zeros(n), "ay": np.zeros(n), "az": np.zeros(n)})
            s.accatom()
        elif code == 3:
            for i in range(a1, a2 + 1):
                s.integrator(i)
            got = s.download()
            sc = s.scalars()
            L = case["box"][0]
            for k in ("x", "y", "z"):
                assert np.abs(got[k] - r[k]).max() < 1e-9 * L, k
            for k in ("vx", "vy", "vz"):
                assert rel_err(got[k], r[k]) < 1e-8, k
            assert abs(sc.epot - r["epot"]) < 1e-8 * abs(r["epot"])
            assert abs(sc.ekin - r["ekin"]) < 1e-8 * abs(r["ekin"])
            assert sc.atrskin == pytest.approx(r["atrskin"], rel=1e-10)
            assert (sc.dmmst, sc.dmmi) == (r["dmmst"], r["dmmi"])
            if variant == "nose":
                assert sc.psi == pytest.approx(r["psi"], rel=1e-8, abs=1e-12)
                assert sc.s == pytest.approx(r["s"], rel=1e-8, abs=1e-12)
        elif code == 4:
            o = oracle_from_gpu(s, case, variant)
            # the first-passage bookkeeping carries over between calls, as
            # COMMON /CVTPP/ carries it (stoddard-pbc-n-m-fre.f:229-258)
            for k in ("cbiggest", "cbiggestf", "clookfor", "clookforf"):
                if k in ost:
                    setattr(o.st, k, ost[k])
            for k in ("cfpt", "cfptf"):
                if k in ost:
                    o.arr[k][:] = ost[k]
            o.stoddard(a1)
            ost.update({k: getattr(o.st, k) for k in ("cbiggest", "cbiggestf", "clookfor", "clookforf")})
            ost.update(cfpt=o.arr["cfpt"].copy(), cfptf=o.arr["cfptf"].copy())
            out = s.stoddard(a1, state=stt)
            for k in ("cl", "cs", "csi"):
                got = out[k] if k != "csi" else out[k][1:]
                ref = o.arr[k] if k != "csi" else o.arr[k][1:]
                np.testing.assert_array_equal(got, ref, err_msg=k)
            assert rel_err(out["cek"], o.arr["cek"]) < REL
            assert out["cnum"] == o.st.cnum
            if variant == "fre":
                check_twf(out, o)
            # and the fixture's own labels where the snapshot is the reference's
            if r["step"] == 0 and code == 4:
                np.testing.assert_array_equal(out["cl"], r["cl"])


TWF_INT = ("clf", "ccf", "csf", "csif", "cgs", "cgsf", "cgsfa", "cfpt", "cfptf")
TWF_COUNTERS = ("cnumf", "cbiggest", "cbiggestf", "clookfor", "clookforf")


def check_twf(out, o):
    """Every ten Wolde-Frenkel output of stoddard-pbc-n-m-fre.f against the
    oracle: labels and sizes (:198-225), the carrier-gas columns cgsf/cgsfa
    and ceksf that wrt-csi-fre.f:40-55 prints, the first-passage arrays and
    counters (:229-258); kinetic-energy sums to summation order."""
    for k in TWF_INT:
        np.testing.assert_array_equal(out[k], o.arr[k], err_msg=k)
    for k in TWF_COUNTERS:
        assert out[k] == getattr(o.st, k), k
    assert rel_err(out["cekf"], o.arr["cekf"]) < REL
    assert rel_err(out["ceksf"], o.arr["ceksf"]) < REL


@pytest.mark.parametrize("n,rho,seed,variant", [
    (4000, 0.03, 7, "mav"),
    (3000, 0.05, 8, "nose"),
    (3000, 0.04, 9, "fre"),
])
def test_larger_system_one_step(n, rho, seed, variant):
    kw = dict(nsorts=2, frac_a=0.4) if variant == "fre" else {}
    case = cases.make_case(n=n, rho=rho, seed=seed, dense_frac=0.3, ncenters=6,
                           ethst_nh=int(variant == "nose"), **kw)
    o = po.Oracle(case, variant)
    o.neighbor(0)
    o.accatom()
    s = gpu_system(case, variant)
    s.neighbor(0)
    lst, idx = s.halflist()
    ref_l, ref_i = o.halflist()
    np.testing.assert_array_equal(lst, ref_l)
    s.upload({"ax": np.zeros(n), "ay": np.zeros(n), "az": np.zeros(n)})
    s.accatom()
    got = s.download()
    assert acc_err(got, o.arr) < REL
    sc = s.scalars()
    assert abs(sc.epot - o.st.epot) <= REL * abs(o.st.epot)
    assert abs(sc.virial - o.st.virial) <= REL * abs(o.st.virial)
    assert abs(sc.pressure - o.st.pressure) <= REL * abs(o.st.pressure)
    # one integrator step from the identical state
    o.integrator(1)
    s.integrator(1)
    got = s.download()
    sc = s.scalars()
    for k in ("x", "y", "z", "vx", "vy", "vz"):
        assert rel_err(got[k], o.arr[k]) < REL, k
    assert acc_err(got, o.arr) < REL
    assert abs(sc.epot - o.st.epot) <= REL * abs(o.st.epot)
    assert abs(sc.ekin - o.st.ekin) <= REL * abs(o.st.ekin)
    # cluster analysis on the identical snapshot
    o2 = oracle_from_gpu(s, case, variant)
    o2.stoddard(1)
    out = s.stoddard(1)
    np.testing.assert_array_equal(out["cl"], o2.arr["cl"])
    np.testing.assert_array_equal(out["csi"][1:], o2.arr["csi"][1:])
    if variant == "fre":
        check_twf(out, o2)


def test_list_width_is_the_widest_atom():
    """list_width bounds the neighbours of one atom: the rebuild reports the
    widest atom (nlist, against the oracle's list), runs with list_width equal
    to it, and one below stops with the reference's list-full error."""
    case = cases.make_case(n=2000, rho=0.04, seed=35, dense_frac=0.5, ncenters=3)
    o = po.Oracle(case, "mav")
    o.neighbor(0)
    deg = np.zeros(case["natom"] + 1, dtype=np.int64)
    for i, j in o.pairs():
        deg[i] += 1
        deg[j] += 1
    widest = int(deg.max())
    assert widest > 8
    s = gpu_system(case, "mav")
    s.neighbor(0)
    assert s.scalars().nlist == widest
    s = gpu_system(case, "mav", list_width=widest)
    s.neighbor(0)
    lst, idx = s.halflist()
    ref_l, ref_i = o.halflist()
    np.testing.assert_array_equal(lst, ref_l)
    s = gpu_system(case, "mav", list_width=widest - 1)
    with pytest.raises(fcmd.FcmdError, match="Neighborlist is nearly full") as e:
        s.neighbor(0)
    assert e.value.code == 2


def test_ensemble_matches_single_runs():
    """Several trajectories of different sizes in one launch give what each
    gives alone (bit for bit: same kernels, same per-trajectory order)."""
    specs = [dict(n=300, rho=0.02, seed=31), dict(n=777, rho=0.04, seed=32),
             dict(n=1500, rho=0.03, seed=33, ethst_nh=1)]
    cs = [cases.make_case(dense_frac=0.3, **k) for k in specs]
    variants = ["mav", "mav", "nose"]
    ens = fcmd.System([fcmd.make_params(c, v) for c, v in zip(cs, variants)])
    singles = []
    for t, (c, v) in enumerate(zip(cs, variants)):
        ens.upload(c, traj=t)
        ens.neighbor(0, traj=t)
        s = gpu_system(c, v)
        s.neighbor(0)
        singles.append(s)
    ens.run(60, use_graph=True)
    for s in singles:
        s.run(60, use_graph=False)
    for t, s in enumerate(singles):
        a = ens.download(t)
        b = s.download()
        for k in ("x", "vx", "ax"):
            np.testing.assert_array_equal(a[k], b[k])
        assert ens.scalars(t).epot == s.scalars().epot


@pytest.mark.parametrize("precision", [fcmd.MIXED, fcmd.MIXED_KAHAN])
def test_mixed_precision_force_error(precision):
    """Both mixed forms (FP32 pair terms summed in FP64, or in FP32 with
    Kahan compensation): RMS relative force error, energy and virial within
    1e-5 of the oracle's FP64 AccAtom on a system with liquid-like clusters."""
    case = cases.make_case(n=4000, rho=0.05, seed=41, dense_frac=0.4, ncenters=5)
    o = po.Oracle(case, "mav")
    o.neighbor(0)
    o.accatom()
    s = gpu_system(case, "mav", precision=precision)
    s.neighbor(0)
    s.upload({"ax": np.zeros(4000), "ay": np.zeros(4000), "az": np.zeros(4000)})
    s.accatom()
    got = s.download()
    num = sum(((got[k] - o.arr[k]) ** 2).sum() for k in ("ax", "ay", "az"))
    den = sum((o.arr[k] ** 2).sum() for k in ("ax", "ay", "az"))
    assert np.sqrt(num / den) < 1e-5
    sc = s.scalars()
    assert abs(sc.epot - o.st.epot) < 1e-5 * abs(o.st.epot)
    assert abs(sc.virial - o.st.virial) < 1e-5 * abs(o.st.virial)


@pytest.mark.parametrize("precision", [fcmd.MIXED, fcmd.MIXED_KAHAN])
@pytest.mark.parametrize("n,rho", [(8, 0.05), (12, 0.05), (100, 0.05)])
def test_mixed_precision_small_boxes(n, rho, precision):
    """Boxes of 1, 2 and 4 cells a side (list radius 2.8 sigma): the mixed
    path's cell-offset separations need the rounding image below 3 cells."""
    case = cases.make_case(n=n, rho=rho, seed=43, dense_frac=0.0)
    o = po.Oracle(case, "mav")
    o.neighbor(0)
    o.accatom()
    s = gpu_system(case, "mav", precision=precision)
    s.neighbor(0)
    s.upload({"ax": np.zeros(n), "ay": np.zeros(n), "az": np.zeros(n)})
    s.accatom()
    got = s.download()
    num = sum(((got[k] - o.arr[k]) ** 2).sum() for k in ("ax", "ay", "az"))
    den = sum((o.arr[k] ** 2).sum() for k in ("ax", "ay", "az"))
    assert np.sqrt(num / den) < 1e-5
    assert abs(s.scalars().epot - o.st.epot) < 1e-5 * abs(o.st.epot)


@pytest.mark.parametrize("precision", [fcmd.FP64, fcmd.MIXED])
def test_cell_grid_capped_at_255_a_side(precision):
    """A box of more than 255 list radii a side (cell coordinates are bytes):
    the grid keeps 255 wider cells a side, which only adds candidates.  The
    list is still the reference's and the forces the oracle's (FP64 to 1e-12,
    mixed to 1e-5), also after steps that rebuild on the capped grid."""
    case = cases.make_case(n=3000, rho=4e-5, seed=71, dense_frac=0.5, ncenters=8,
                           rc_sig=1.5, rskin_sig=1.6)
    assert case["box"][0] / case["atrskin"] > 255
    n = case["natom"]
    o = po.Oracle(case, "mav")
    o.neighbor(0)
    o.accatom()
    s = gpu_system(case, "mav", precision=precision)
    s.neighbor(0)
    lst, idx = s.halflist()
    ref_l, ref_i = o.halflist()
    np.testing.assert_array_equal(lst, ref_l)
    assert len(lst) > n + 100  # the clusters give pairs
    s.upload({"ax": np.zeros(n), "ay": np.zeros(n), "az": np.zeros(n)})
    s.accatom()
    got = s.download()
    tol = REL if precision == fcmd.FP64 else 1e-5
    num = sum(((got[k] - o.arr[k]) ** 2).sum() for k in ("ax", "ay", "az"))
    den = sum((o.arr[k] ** 2).sum() for k in ("ax", "ay", "az"))
    assert np.sqrt(num / den) < tol
    for i in range(1, 31):
        s.integrator(i)
    assert s.scalars().nrebuild >= 2
    # forces on the GPU's current list and positions, against the oracle's
    o2 = oracle_from_gpu(s, case, "mav")
    o2.accatom()
    s.upload({"ax": np.zeros(n), "ay": np.zeros(n), "az": np.zeros(n)})
    s.accatom()
    got = s.download()
    num = sum(((got[k] - o2.arr[k]) ** 2).sum() for k in ("ax", "ay", "az"))
    den = sum((o2.arr[k] ** 2).sum() for k in ("ax", "ay", "az"))
    assert np.sqrt(num / den) < tol
