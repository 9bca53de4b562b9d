"""Pins the C restatement (oracle/fcmd_oracle.c) before anything is checked
against it:

* golden vectors written by the compiled reference (tests/golden/*.npz):
  replayed bit for bit;
* the compiled reference itself on fresh random systems, where it is built
  (oracle/_ref, this container only);
* closed forms and invariants chosen so that a dropped term, a wrong sign, a
  wrong minimum-image shift or a transposed operand fails: the LJ dimer
  (V(2^1/6 sigma) = -eps, F(sigma) = 24 eps/sigma), force superposition and
  Newton's third law, brute-force pair sets and connected components,
  cluster chains at 1.4 / 1.6 sigma across the periodic boundary, the
  ten Wolde-Frenkel neighbour threshold, and energy conservation.
"""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components

import cases
import pyoracle as po
from conftest import golden_names, load_golden

STATE_KEYS = ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az")
SCALAR_KEYS = ("epot", "ekin", "sumv2", "psi", "s", "atrskin", "dmms", "dmmsu", "dmmax", "dmmav")
INT_KEYS = ("cl", "cs", "cc", "csi")
FRE_INT_KEYS = ("clf", "csf", "ccf", "csif", "cgs", "cgsf", "cgsfa", "cfpt", "cfptf")


def replay(o, actions):
    for code, a1, a2 in actions:
        if code == 1:
            yield o.neighbor(a1)
        elif code == 2:
            o.accatom()
            yield a1
        elif code == 3:
            s = a1
            for i in range(a1, a2 + 1):
                s = o.integrator(i)
            yield s
        elif code == 4:
            o.stoddard(a1)
            yield a1


def assert_matches(o, r, variant, what):
    for k in STATE_KEYS:
        np.testing.assert_array_equal(o.arr[k], r[k], err_msg=f"{what}: {k}")
    for k in SCALAR_KEYS:
        assert getattr(o.st, k) == r[k], (what, k, getattr(o.st, k), r[k])
    np.testing.assert_array_equal(np.array(o.st.dmmh[:]), r["dmmh"], err_msg=what)
    assert (o.st.dmmst, o.st.dmmi) == (r["dmmst"], r["dmmi"]), what
    if r["nlen"]:
        al, ai = o.halflist()
        np.testing.assert_array_equal(al, r["atnlist"], err_msg=f"{what}: atnlist")
        np.testing.assert_array_equal(ai, r["atnidx"], err_msg=f"{what}: atnidx")
    if r["code"] == 4:
        for k in INT_KEYS + (FRE_INT_KEYS if variant == "fre" else ()):
            np.testing.assert_array_equal(o.arr[k], r[k], err_msg=f"{what}: {k}")
        np.testing.assert_array_equal(o.arr["cek"], r["cek"], err_msg=f"{what}: cek")
        np.testing.assert_array_equal(o.arr["ceks"], r["ceks"], err_msg=f"{what}: ceks")
        assert o.st.cnum == r["cnum"]
        if variant == "fre":
            np.testing.assert_array_equal(o.arr["cekf"], r["cekf"])
            np.testing.assert_array_equal(o.arr["ceksf"], r["ceksf"])
            assert (o.st.cnumf, o.st.cbiggest, o.st.cbiggestf, o.st.clookfor, o.st.clookforf) == \
                (r["cnumf"], r["cbiggest"], r["cbiggestf"], r["clookfor"], r["clookforf"])


@pytest.mark.parametrize("name", golden_names())
def test_oracle_replays_golden_bit_exact(name):
    variant, case, actions, recs = load_golden(name)
    o = po.Oracle(case, variant)
    for (act, step, r) in zip(actions, replay(o, actions), recs):
        assert step == r["step"]
        assert_matches(o, r, variant, f"{name} action {act}")


@pytest.mark.skipif(not po.have_harness("mav"), reason="reference harness not built here")
@pytest.mark.parametrize("variant,kw", [
    ("mav", dict(n=700, rho=0.02, seed=101, dense_frac=0.3)),
    ("vv", dict(n=400, rho=0.05, seed=102, dense_frac=0.2)),
    ("nose", dict(n=500, rho=0.03, seed=103, dense_frac=0.4, ethst_nh=1)),
    ("nuc", dict(n=500, rho=0.04, seed=104, dense_frac=0.4, nsorts=2, frac_a=0.2)),
    ("fre", dict(n=500, rho=0.04, seed=105, dense_frac=0.5, nsorts=2, frac_a=0.5)),
])
def test_oracle_vs_compiled_reference(variant, kw):
    case = cases.make_case(**kw)
    actions = [(1, 0, 0), (2, 0, 0), (3, 1, 60), (4, 60, 0)]
    recs = po.run_harness(variant, case, actions)
    o = po.Oracle(case, variant)
    for act, step, r in zip(actions, replay(o, actions), recs):
        assert step == r["step"]
        assert_matches(o, r, variant, f"{variant} {act}")


# --------------------------------------------------------------------------
def _dimer(r_sig, box_sig=20.0, shift_sig=0.1):
    """Two atoms r apart along x, straddling the periodic boundary."""
    c = cases.make_case(n=2, rho=2.0 / box_sig ** 3, seed=1)
    L = c["box"][0]
    s = cases.SIG_AR
    c["x"] = np.array([shift_sig * s, L - (r_sig - shift_sig) * s])
    c["y"] = np.array([5.0 * s, 5.0 * s])
    c["z"] = np.array([7.0 * s, 7.0 * s])
    return c


def lj_force(r, sig, eps):
    return 48.0 * eps / r * ((sig / r) ** 12 - 0.5 * (sig / r) ** 6)


def test_dimer_closed_forms():
    s, e, m = cases.SIG_AR, cases.EPS_AR, cases.MAS_AR
    # potential minimum: V = -eps, F = 0
    o = po.Oracle(_dimer(2 ** (1 / 6)), "mav")
    o.neighbor(0)
    o.accatom()
    assert o.st.epot == pytest.approx(-e, rel=1e-12)
    assert abs(o.arr["ax"][0]) < 1e-12 * 24 * e / (s * m)
    # r = sigma: V = 0, |F| = 24 eps / sigma; atom 2 sits just below the
    # periodic face at x = L, i.e. on atom 1's -x side, so atom 1 is pushed +x
    o = po.Oracle(_dimer(1.0), "mav")
    o.neighbor(0)
    o.accatom()
    assert abs(o.st.epot) < 1e-12 * e
    assert o.arr["ax"][0] == pytest.approx(+24 * e / (s * m), rel=1e-12)
    assert o.arr["ax"][1] == pytest.approx(-24 * e / (s * m), rel=1e-12)
    # general r across the boundary
    for r in (0.95, 1.3, 1.7, 2.2, 2.49):
        o = po.Oracle(_dimer(r), "mav")
        o.neighbor(0)
        o.accatom()
        f = lj_force(r * s, s, e)
        assert o.arr["ax"][0] == pytest.approx(f / m, rel=1e-11)
        assert o.st.epot == pytest.approx(4 * e * (r ** -12 - r ** -6), rel=1e-11)
        assert o.arr["ay"][0] == 0.0 and o.arr["az"][1] == 0.0
    # beyond the cutoff but inside the list radius: in the list, no force
    o = po.Oracle(_dimer(2.6), "mav")
    o.neighbor(0)
    assert o.pairs() == [(1, 2)]
    o.accatom()
    assert o.st.epot == 0.0 and o.arr["ax"][0] == 0.0


def test_mixture_pair_parameters():
    """for-a-lj-n-pbc-m.f:61-76: Lorentz-Berthelot, eps12 = sqrt(eps1 eps2)."""
    s, e, m = cases.SIG_AR, cases.EPS_AR, cases.MAS_AR
    c = _dimer(1.2)
    c.update(nsorts=2, api=np.array([1, 2], dtype=np.int32), num=np.array([1, 1]),
             mas=np.array([m, 2 * m]), sig=np.array([s, 0.8 * s]), eps=np.array([e, 0.25 * e]))
    o = po.Oracle(c, "nuc")
    o.neighbor(0)
    o.accatom()
    s12, e12 = 0.9 * s, 0.5 * e
    f = lj_force(1.2 * s, s12, e12)
    assert o.arr["ax"][0] == pytest.approx(f / m, rel=1e-11)
    assert o.arr["ax"][1] == pytest.approx(-f / (2 * m), rel=1e-11)
    r = 1.2 * s / s12
    assert o.st.epot == pytest.approx(4 * e12 * (r ** -12 - r ** -6), rel=1e-11)


def test_virial_and_pressure_closed_forms():
    """Pair virial W = r F(r) of the dimer (for-a-lj-n-pbc.f:106,114) in the
    one-species, Lorentz-Berthelot and pair-table forms; pressure
    1e13/V (N k T - W/3) (:116-117)."""
    s, e, m = cases.SIG_AR, cases.EPS_AR, cases.MAS_AR
    for r in (0.93, 2 ** (1 / 6), 1.3, 2.2, 2.49):
        c = _dimer(r)
        o = po.Oracle(c, "mav")
        o.neighbor(0)
        o.accatom()
        w = lj_force(r * s, s, e) * r * s
        assert o.st.virial == pytest.approx(w, rel=1e-11, abs=1e-12 * 24 * e)
        V = float(np.prod(c["box"]))
        assert o.st.pressure == pytest.approx(1e13 / V * (2 * cases.CIKB * c["att"] - w / 3), rel=1e-11)
    # beyond the cutoff: no virial
    o = po.Oracle(_dimer(2.6), "mav")
    o.neighbor(0)
    o.accatom()
    assert o.st.virial == 0.0
    # A-B pair at eps_AB = 0.3, sigma_AB = 0.9 sigma through the pair table
    c = _dimer(1.2)
    c.update(nsorts=2, api=np.array([1, 2], dtype=np.int32), num=np.array([1, 1]),
             mas=np.array([m, m]), sig=np.array([s, s]), eps=np.array([e, 0.1 * e]))
    tab_e = [[e, 0.3 * e], [0.3 * e, 0.1 * e]]
    tab_s = [[s, 0.9 * s], [0.9 * s, s]]
    o = po.Oracle(c, "nuc", eps_ab=tab_e, sig_ab=tab_s)
    o.neighbor(0)
    o.accatom()
    f = lj_force(1.2 * s, 0.9 * s, 0.3 * e)
    assert o.arr["ax"][0] == pytest.approx(f / m, rel=1e-11)
    assert o.st.virial == pytest.approx(f * 1.2 * s, rel=1e-11)
    q = 1.2 / 0.9
    assert o.st.epot == pytest.approx(4 * 0.3 * e * (q ** -12 - q ** -6), rel=1e-11)


def test_virial_theorem_for_a_free_cluster():
    """For a cluster whose pairs never cross the periodic faces,
    sum over pairs r_ij . f_ij = sum over atoms r_i . F_i (F_i = m a_i): the
    virial from the pair terms equals the one from the forces."""
    c = cases.make_case(n=300, rho=0.05, seed=31, dense_frac=1.0, ncenters=1, dense_rho=0.7)
    L = c["box"][0]
    for k in ("x", "y", "z"):  # move the droplet to the box centre
        v = c[k] - c[k][0]
        v -= L * np.round(v / L)
        c[k] = v + 0.5 * L
    for variant, kw in (("mav", {}), ("nuc", dict(nsorts=2))):
        cc = dict(c)
        if kw:
            cc = cases.make_case(n=300, rho=0.05, seed=31, dense_frac=1.0, ncenters=1, dense_rho=0.7,
                                 nsorts=2, frac_a=0.5)
            for k in ("x", "y", "z"):
                cc[k] = c[k]
        o = po.Oracle(cc, variant)
        o.neighbor(0)
        o.accatom()
        pos = np.stack([o.arr["x"], o.arr["y"], o.arr["z"]], 1)
        span = pos.max(0) - pos.min(0)
        assert np.all(span < L - 2 * cc["atrskin"])  # no pair across a face
        mass = np.asarray(cc["mas"])[o.arr["api"] - 1]
        acc = np.stack([o.arr["ax"], o.arr["ay"], o.arr["az"]], 1)
        w_forces = float((pos * acc * mass[:, None]).sum())
        assert o.st.virial == pytest.approx(w_forces, rel=1e-9)
        assert abs(o.st.virial) > 0


def test_pair_table_with_lorentz_berthelot_values_is_the_m_form():
    """ORC_ACC_TABLE holding the LB values reproduces for-a-lj-n-pbc-m.f
    (forces, energy and virial) bit for bit."""
    c = cases.make_case(n=500, rho=0.05, seed=33, dense_frac=0.4, nsorts=2, frac_a=0.3)
    s1, s2 = c["sig"]
    e1, e2 = c["eps"]
    tab_e = [[e1, np.sqrt(e1 * e2)], [np.sqrt(e1 * e2), e2]]
    tab_s = [[s1, 0.5 * (s1 + s2)], [0.5 * (s1 + s2), s2]]
    out = []
    for kw in ({}, dict(eps_ab=tab_e, sig_ab=tab_s)):
        o = po.Oracle(c, "nuc", **kw)
        o.neighbor(0)
        o.accatom()
        out.append(o)
    for k in ("ax", "ay", "az"):
        np.testing.assert_array_equal(out[0].arr[k], out[1].arr[k])
    assert out[0].st.epot == out[1].st.epot
    assert out[0].st.virial == out[1].st.virial


def _brute_pairs(case, radius):
    p = np.stack([case["x"], case["y"], case["z"]], 1)
    L = case["box"]
    d = p[:, None, :] - p[None, :, :]
    d -= L * np.round(d / L)
    r2 = (d ** 2).sum(-1)
    i, j = np.nonzero(np.triu(r2 <= radius * radius, 1))
    return set(zip((i + 1).tolist(), (j + 1).tolist())), r2


def test_list_and_clusters_vs_brute_force():
    case = cases.make_case(n=400, rho=0.04, seed=21, dense_frac=0.5, ncenters=3)
    o = po.Oracle(case, "mav")
    o.neighbor(0)
    wrapped = dict(case, x=o.arr["x"], y=o.arr["y"], z=o.arr["z"])
    pairs, r2 = _brute_pairs(wrapped, case["atrskin"])
    assert set(o.pairs()) == pairs
    o.accatom()
    o.stoddard(0)
    adj = sp.csr_matrix(np.triu(r2 <= case["atskin"] ** 2, 1))
    ncomp, lab = connected_components(adj, directed=False)
    n = case["natom"]
    minidx = np.full(ncomp, n + 1)
    np.minimum.at(minidx, lab, np.arange(1, n + 1))
    np.testing.assert_array_equal(o.arr["cl"], minidx[lab])
    sizes = np.bincount(lab)
    hist = np.bincount(sizes, minlength=n + 1)
    np.testing.assert_array_equal(o.arr["csi"][1:], hist[1:])
    assert o.arr["csi"][0] == n - ncomp  # csi(0) collects the non-roots
    assert o.arr["csi"][1:].dot(np.arange(1, n + 1)) == n


@pytest.mark.parametrize("spacing,expect_one", [(1.4, True), (1.6, False)])
def test_chain_across_boundary(spacing, expect_one):
    n = 12
    c = cases.make_case(n=n, rho=0.001, seed=3)
    L = c["box"][0]
    s = cases.SIG_AR
    c["x"] = np.array([(L - 3 * s + k * spacing * s) % L for k in range(n)])
    c["y"] = np.full(n, 2.0 * s)
    c["z"] = np.full(n, 4.0 * s)
    o = po.Oracle(c, "mav")
    o.neighbor(0)
    o.accatom()
    o.stoddard(0)
    if expect_one:
        assert o.arr["csi"][n] == 1 and (o.arr["cl"] == 1).all()
        assert o.arr["cs"][0] == n
    else:
        assert o.arr["csi"][1] == n and (o.arr["cl"] == np.arange(1, n + 1)).all()


def test_ten_wolde_frenkel_threshold():
    """An A atom with exactly atnenum A neighbours within atskin is liquid-like,
    with one fewer it is not (stoddard-pbc-n-m-fre.f:148-168)."""
    s = cases.SIG_AR
    for nnb, liquid in ((5, True), (4, False)):
        n = 1 + nnb + 3
        c = cases.make_case(n=n, rho=0.0005, seed=5, nsorts=2, frac_a=(1 + nnb) / n)
        L = c["box"][0]
        dirs = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]])
        pos = [np.array([L / 2] * 3)]
        pos += [pos[0] + 1.2 * s * dirs[k] for k in range(nnb)]
        pos += [np.array([3 * s + 4 * k * s, 3 * s, 3 * s]) for k in range(3)]  # B, far away
        pos = np.array(pos)
        c["x"], c["y"], c["z"] = pos[:, 0].copy(), pos[:, 1].copy(), pos[:, 2].copy()
        o = po.Oracle(c, "fre")
        o.neighbor(0)
        o.accatom()
        o.stoddard(0)
        assert o.arr["ccf"][0] == (1 if liquid else 0)
        assert o.st.cnumf == (1 if liquid else 0)
        if liquid:
            assert o.arr["csif"][1] == 1 and o.arr["clf"][0] == 1
        assert o.arr["csif"][0] == c["num"][0] - (1 if liquid else 0)


def test_newton_third_law_and_superposition():
    case = cases.make_case(n=300, rho=0.05, seed=31, dense_frac=0.5)
    o = po.Oracle(case, "mav")
    o.neighbor(0)
    o.accatom()
    for k in ("ax", "ay", "az"):
        scale = np.abs(o.arr[k]).max()
        assert abs(o.arr[k].sum()) < 1e-12 * scale * case["natom"]
    # symmetric trimer: middle atom feels nothing
    s = cases.SIG_AR
    c = cases.make_case(n=3, rho=0.001, seed=2)
    c["x"] = np.array([10 * s, 11.15 * s, 12.3 * s])
    c["y"] = np.full(3, 5 * s)
    c["z"] = np.full(3, 5 * s)
    o = po.Oracle(c, "mav")
    o.neighbor(0)
    o.accatom()
    assert abs(o.arr["ax"][1]) < 1e-13 * abs(o.arr["ax"][0])
    assert o.arr["ax"][0] == pytest.approx(-o.arr["ax"][2], rel=1e-14)


def _pairs_within(o, case, radius):
    p = np.stack([o.arr["x"], o.arr["y"], o.arr["z"]], 1)
    L = case["box"]
    d = p[:, None] - p[None]
    d -= L * np.round(d / L)
    return int(np.triu((d ** 2).sum(-1) <= radius * radius, 1).sum())


def _drift(variant, dt_tau, nh):
    """Relative spread of the shifted-energy (extended) Hamiltonian over one
    tau.  The reference's energy is unshifted, so each pair crossing rc jumps
    by V(rc); adding back npairs * V(rc) leaves only the integrator error."""
    case = cases.make_case(n=300, rho=0.03, seed=41, dense_frac=0.2, rmin_sig=1.05,
                           dt_tau=dt_tau, ethst_nh=int(nh), q_star=10.0)
    vrc = 4 * case["eps"][0] * (2.5 ** -12 - 2.5 ** -6)
    o = po.Oracle(case, variant)
    o.neighbor(0)
    o.accatom()
    if nh:
        ekin0 = 0.5 * (case["mas"][0] * (case["vx"] ** 2 + case["vy"] ** 2 + case["vz"] ** 2)).sum()
        o.st.psi = cases.nh_initial_psi(case, ekin0)
    nkt = 3 * case["natom"] * po.CIKB * case["att"]
    h = []
    for i in range(1, int(round(1 / dt_tau)) + 1):
        o.integrator(i)
        e = o.st.ekin + o.st.epot - _pairs_within(o, case, case["atrcut"]) * vrc
        if nh:
            e += o.st.psi ** 2 * case["Q"] / 2 + nkt * o.st.s
            assert 1 <= o.st.nh_iters <= 10
        h.append(e)
    h = np.array(h)
    return np.ptp(h) / abs(h).max()


@pytest.mark.parametrize("variant,nh", [("mav", False), ("nose", True)])
def test_energy_conservation_second_order(variant, nh):
    d1 = _drift(variant, 0.005, nh)
    d2 = _drift(variant, 0.0025, nh)
    assert d1 < 1e-3
    assert 2.5 < d1 / d2 < 6.5  # velocity Verlet: error ~ dt^2
