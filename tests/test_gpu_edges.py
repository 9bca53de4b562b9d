"""GPU edge cases: the LJ dimer against its closed forms, atoms with no
neighbours, and the reference's fatal conditions forced from the host.

* Dimer (for-a-lj-n-pbc.f:95-110): V(r) = 4 eps ((s/r)^12 - (s/r)^6) and
  a = 48 eps / m * r^-2 (s^12 r^-12 - s^6 r^-6 / 2) * dx, so V(2^(1/6) s) = -eps
  with no force and V(s) = 0 with |F| = 24 eps / s, and the pair virial
  W = r F(r) = 48 eps ((s/r)^12 - (s/r)^6/2) with the pressure
  1e13/V (N k T - W/3) (:106,114-117), both on the FP64 and the mixed path,
  on grids of 1 and 5 cells a side and across the periodic face.
* No neighbours: forces and energy exactly zero.
* Pairs within 1e-5 of the cutoff on the mixed path: the reference's FP64
  test on the FP64 positions decides them (for-a-lj-n-pbc.f:89-95), so a
  rejected pair leaves its atoms exactly no force (its terms are added and
  taken out again, and the FP64 products of FP32 terms are exact) and an
  accepted one the oracle's.
* Crowded cells (27 and 100 atoms in one list cell): the half list equals
  the oracle's whichever way the cell sort orders a cell's atoms.
* neighbor-pbc*.f "Skin radius exceeded!!" when the largest displacement
  since the last build exceeds the skin; the Nose-Hoover update reports
  "did not converge" for a thermostat state with 1 + dt/2 psi <= 0, where
  v' = (v + dt/2 a) / (1 + dt/2 psi') has no meaning (int-verlet-nh-mav.f:106-166).
"""
import numpy as np
import pytest

import cases
import fcmd
from test_gpu_parity import gpu_system

pytestmark = pytest.mark.gpu

SIG = cases.SIG_AR
R_MIN = 2.0 ** (1.0 / 6.0) * SIG


def dimer_case(box_sig, r, across):
    """Two argon atoms at separation r (along x) in a cubic box of box_sig
    sigma; across: the pair straddles the periodic face x = 0."""
    L = box_sig * SIG
    c = cases.make_case(n=2, rho=2.0 / box_sig ** 3, seed=3)
    assert np.isclose(c["box"][0], L)
    mid = 0.5 * L
    if across:
        x1, x2 = 0.3 * SIG, L - (r - 0.3 * SIG)
    else:
        x1, x2 = mid, mid + r
    c["x"] = np.array([x1, x2])
    c["y"] = np.array([mid, mid])
    c["z"] = np.array([mid, mid])
    for k in ("vx", "vy", "vz"):
        c[k] = np.zeros(2)
    return c


def forces(case, precision, with_virial=False):
    s = gpu_system(case, "mav", precision=precision)
    s.neighbor(0)
    n = case["natom"]
    s.upload({"ax": np.zeros(n), "ay": np.zeros(n), "az": np.zeros(n)})
    s.accatom()
    sc = s.scalars()
    if with_virial:
        return s.download(), sc.epot, sc.virial, sc.pressure
    return s.download(), sc.epot


@pytest.mark.parametrize("precision", [fcmd.FP64, fcmd.MIXED, fcmd.MIXED_KAHAN])
@pytest.mark.parametrize("box_sig", [5.0, 15.0])
@pytest.mark.parametrize("across", [False, True])
def test_lj_dimer_closed_forms(precision, box_sig, across):
    tol = 1e-12 if precision == fcmd.FP64 else 1e-5
    eps, m = cases.EPS_AR, cases.MAS_AR
    scale = 24.0 * eps / (m * SIG)  # |a| at r = sigma
    for r in (R_MIN, SIG):
        c = dimer_case(box_sig, r, across)
        got, epot, vir, press = forces(c, precision, with_virial=True)
        L = c["box"][0]
        dx = c["x"][0] - c["x"][1]
        dx -= L * np.rint(dx / L)  # minimum image: -r, or +r across the face
        dy = c["y"][0] - c["y"][1]
        assert abs(abs(dx) - r) < 1e-9 * SIG
        r_true = np.hypot(dx, dy)
        sr6 = (SIG / r_true) ** 6
        v_ref = 4.0 * eps * (sr6 * sr6 - sr6)
        fij = r_true ** -2 * (sr6 * sr6 - 0.5 * sr6)
        ax_ref = 48.0 * eps / m * fij * dx
        if r == SIG:
            assert v_ref == pytest.approx(0.0, abs=1e-5 * eps)
            assert abs(ax_ref) == pytest.approx(scale, rel=1e-5)
        else:
            assert v_ref == pytest.approx(-eps, rel=1e-6)
        assert abs(epot - v_ref) <= tol * eps
        # pair virial W = r . f (for-a-lj-n-pbc.f:106,114): 48 eps (s6^2 - s6/2)
        w_ref = 48.0 * eps * (sr6 * sr6 - 0.5 * sr6)
        assert abs(vir - w_ref) <= tol * 24.0 * eps
        V = L ** 3
        p_ref = 1e13 / V * (2 * cases.CIKB * c["att"] - w_ref / 3)
        assert abs(press - p_ref) <= tol * 1e13 / V * 8.0 * eps + 1e-12 * abs(p_ref)
        assert abs(got["ax"][0] - ax_ref) <= tol * scale
        assert abs(got["ax"][1] + ax_ref) <= tol * scale  # Newton's third law
        assert abs(got["az"][0]) <= tol * scale and abs(got["az"][1]) <= tol * scale


@pytest.mark.parametrize("precision", [fcmd.FP64, fcmd.MIXED])
def test_atoms_without_neighbours_feel_nothing(precision):
    c = cases.make_case(n=27, rho=2e-4, seed=5)
    L = c["box"][0]
    g = (np.arange(3) + 0.5) * L / 3  # 3x3x3 lattice, spacing > 5 sigma
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    c["x"], c["y"], c["z"] = X.ravel(), Y.ravel(), Z.ravel()
    assert L / 3 > 2.8 * SIG
    got, epot = forces(c, precision)
    for k in ("ax", "ay", "az"):
        assert np.all(got[k] == 0.0), k
    assert epot == 0.0
    s = gpu_system(c, "mav", precision=precision)
    s.neighbor(0)
    assert s.scalars().npairs == 0


def test_skin_exceeded_is_fatal():
    c = cases.make_case(n=300, rho=0.02, seed=7)
    s = gpu_system(c, "mav")
    s.neighbor(0)
    skin = s.scalars().atrskin - c["atrcut"]
    s.set_scalars(dmms=1.01 * skin)
    with pytest.raises(fcmd.FcmdError, match="Skin radius exceeded") as e:
        s.neighbor(5)
    assert e.value.code == 1


def test_nose_hoover_without_a_meaningful_update_is_reported():
    c = cases.make_case(n=300, rho=0.02, seed=8)
    s = gpu_system(c, "nose")
    s.neighbor(0)
    n = c["natom"]
    s.upload({"ax": np.zeros(n), "ay": np.zeros(n), "az": np.zeros(n)})
    s.accatom()
    s.integrator(1)  # an ordinary step first
    s.set_scalars(psi=-10.0 / c["dt"])  # 1 + dt/2 psi = -4
    with pytest.raises(fcmd.FcmdError, match="did not converge") as e:
        s.integrator(2)
    assert e.value.code == 3


@pytest.mark.parametrize("precision", [fcmd.FP64, fcmd.MIXED])
@pytest.mark.parametrize("nsteps,use_graph", [(40, True), (37, True), (23, False)])
def test_run_ends_with_its_last_steps_energy(precision, nsteps, use_graph):
    """A run forms epot, the virial and the pressure only where they can be
    read: on its last step (and on the last step of each captured graph).  So
    after any run they must equal AccAtom()'s on the final state, bit for bit,
    whether the run ended on a graph replay or on single steps."""
    from fcmd import workloads as W
    params, states, meta = W.ensemble512(ntraj=6)
    with fcmd.System(params, precision) as s:
        W.setup(s, params, states)
        s.run(nsteps, cluster_every=0, use_graph=use_graph)
        got = [(x.epot, x.virial, x.pressure, x.step) for x in (s.scalars(t) for t in range(len(params)))]
        s.accatom()
        for t, (e, w, p, step) in enumerate(got):
            sc = s.scalars(t)
            assert sc.step == step
            assert (sc.epot, sc.virial, sc.pressure) == (e, w, p), t


def band_case(m=7, seed=9):
    """m^3 isolated dimers, each at a separation within 5e-6 relative of the
    cutoff, in random directions, one per cell of an m^3 grid whose spacing
    keeps every other atom beyond the list radius."""
    spacing = 2.5 + 2.8 + 0.8  # sigma: cutoff + list radius + margin
    npairs = m ** 3
    c = cases.make_case(n=2 * npairs, rho=2 * npairs / (m * spacing) ** 3, seed=seed)
    L = c["box"][0]
    sp = L / m
    rng = np.random.default_rng(seed)
    g = (np.indices((m, m, m)).reshape(3, -1).T + 0.5) * sp
    d = rng.normal(size=(npairs, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    r = c["atrcut"] * (1.0 + rng.uniform(-5e-6, 5e-6, npairs))
    p1 = g - 0.5 * r[:, None] * d
    p2 = g + 0.5 * r[:, None] * d
    pos = np.empty((2 * npairs, 3))
    pos[0::2], pos[1::2] = p1, p2
    c["x"], c["y"], c["z"] = pos[:, 0].copy(), pos[:, 1].copy(), pos[:, 2].copy()
    for k in ("vx", "vy", "vz"):
        c[k] = np.zeros(2 * npairs)
    return c


@pytest.mark.parametrize("precision", [fcmd.MIXED, fcmd.MIXED_KAHAN])
def test_mixed_cutoff_band_is_decided_in_fp64(precision):
    import pyoracle as po

    c = band_case()
    n = c["natom"]
    o = po.Oracle(c, "mav")
    o.neighbor(0)
    o.accatom()
    ref = np.stack([o.arr[k][:n] for k in ("ax", "ay", "az")], axis=1)
    taken = np.linalg.norm(ref, axis=1) > 0
    # both kinds of pair are present
    assert 0.2 * n < taken.sum() < 0.8 * n
    got, _ = forces(c, precision)
    a = np.stack([got[k] for k in ("ax", "ay", "az")], axis=1)
    assert np.all(a[~taken] == 0.0)
    err = np.linalg.norm(a[taken] - ref[taken], axis=1) / np.linalg.norm(ref[taken], axis=1)
    assert err.max() < 1e-5


def crowded_case():
    """Vapour around two lattice blobs, each inside one list cell: 27 atoms
    (3 x 3 x 3 at 0.6 sigma) and 100 atoms (5 x 5 x 4 at 0.5 sigma).  The
    cell sort gives a cell of more than 8 atoms to its wavefront, and one of
    more than 64 to the wavefront's first lane."""
    c = cases.make_case(n=600, rho=0.02, seed=11)
    L = c["box"][0]
    side = L / np.floor(L / (c["atrskin"] * (1.0 + 1e-9)))
    blobs = []
    for cell, (nx, ny, nz), a in (((2, 3, 4), (3, 3, 3), 0.6), ((7, 6, 5), (5, 5, 4), 0.5)):
        centre = (np.array(cell) + 0.5) * side
        g = np.stack(np.meshgrid(*(np.arange(k) - (k - 1) / 2 for k in (nx, ny, nz)), indexing="ij"), -1)
        blobs.append(centre + a * SIG * g.reshape(-1, 3))
    blob = np.concatenate(blobs)
    pos = np.stack([c["x"], c["y"], c["z"]], -1)
    # vapour atoms clear of the blobs' cells, the blob atoms in their place
    far = np.ones(len(pos), dtype=bool)
    for b in blobs:
        d = pos[:, None, :] - b[None, :, :]
        d -= L * np.rint(d / L)
        far &= (np.einsum("ijk,ijk->ij", d, d) > (1.5 * SIG) ** 2).all(axis=1)
    keep = np.flatnonzero(far)[: 600 - len(blob)]
    pos = np.concatenate([blob, pos[keep]])
    assert len(pos) == 600
    c["x"], c["y"], c["z"] = (np.ascontiguousarray(pos[:, k]) for k in range(3))
    return c


@pytest.mark.parametrize("precision", [fcmd.FP64, fcmd.MIXED])
def test_lists_with_crowded_cells(precision):
    """Cells of 27 and 100 atoms (sorted by reference number a wavefront a
    cell, and by one lane above 64 atoms): the half list equals the oracle's
    (neighbor-pbc-mav.f:64-95) after a re-sorting build and after a build
    that keeps the atoms in their slots."""
    import pyoracle as po

    c = crowded_case()
    o = po.Oracle(c, "mav")
    o.neighbor(0)
    ref_l, ref_i = o.halflist()
    s = gpu_system(c, "mav", precision=precision)
    for _ in range(2):
        s.neighbor(0)
        lst, idx = s.halflist()
        np.testing.assert_array_equal(lst, ref_l)
        np.testing.assert_array_equal(idx, ref_i)
