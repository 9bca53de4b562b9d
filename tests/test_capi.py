"""The C-ABI library without a GPU: it loads, exports every entry point the
header declares, reports errors with the reference's messages, and its
host-only seeded generator keeps its contract."""
import os
import re

import numpy as np
import pytest

import fcmd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "fcmd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*(fcmd_\w+)\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    L = fcmd.lib()
    names = declared_functions()
    assert len(names) >= 18
    for n in names:
        assert hasattr(L, n), n
    assert set(names) <= set(fcmd.EXPORTS) | {"fcmd_error_text", "fcmd_destroy"}


def test_fortran_interface_declares_the_entry_points():
    src = open(os.path.join(ROOT, "fortran-cluster-md_amd", "fortran", "fcmd_capi.f90")).read()
    for n in ("fcmd_create", "fcmd_upload", "fcmd_download", "fcmd_neighbor", "fcmd_accatom",
              "fcmd_integrator", "fcmd_stoddard", "fcmd_get_scalars", "fcmd_set_scalars"):
        assert f"name='{n}'" in src, n


def test_error_texts_are_the_reference_messages():
    t = lambda c: fcmd.lib().fcmd_error_text(c).decode()  # noqa: E731
    assert t(1) == "*** FATAL ERROR! ***  Skin radius exceeded!!"        # neighbor-pbc-mav.f:49
    assert t(2) == "*** FATAL ERROR *** Neighborlist is nearly full!"    # neighbor-pbc-mav.f:95
    assert t(3) == "Numerical integration did not converge!"             # int-verlet-nh-mav.f:151


def test_struct_layouts_match_header():
    import ctypes as C
    # fcmd_params: 2 ints, 3+30 doubles, 10 ints, 6 doubles, 8 ints, 200 doubles, 2 ints
    assert C.sizeof(fcmd.Params) == 8 + 24 + 240 + 40 + 48 + 32 + 1600 + 8
    # fcmd_scalars: 6 + 256 + 5 doubles, 8 ints, 2 doubles (virial, pressure)
    assert C.sizeof(fcmd.Scalars) == 6 * 8 + 256 * 8 + 5 * 8 + 8 * 4 + 2 * 8
    assert fcmd.Scalars.virial.offset == 6 * 8 + 256 * 8 + 5 * 8 + 8 * 4


@pytest.mark.skipif(fcmd.device_count() > 0, reason="checks the no-device path")
def test_create_without_device_fails_loudly():
    p = fcmd.Params()
    p.natom = 10
    with pytest.raises(fcmd.FcmdError):
        fcmd.System(p)


def test_create_rejects_an_unknown_precision():
    """fcmd_create checks its precision argument before anything else is
    set up (no device needed)."""
    import ctypes as C
    p = fcmd.Params()
    p.natom = 10
    h = C.c_void_p()
    assert fcmd.lib().fcmd_create(C.byref(h), 1, C.byref(p), 3, 0) == 10  # FCMD_ERR_ARG
    assert not h.value


def test_init_random_contract():
    box = np.array([120.0, 120.0, 120.0])
    n = 2000
    mas = np.array([66.4224, 33.2])
    api = np.concatenate([np.ones(400, dtype=np.int32), np.full(1600, 2, dtype=np.int32)])
    x, y, z, vx, vy, vz = fcmd.init_random(17, n, box, 3.0, 84.0, mas, api)
    x2, *_ = fcmd.init_random(17, n, box, 3.0, 84.0, mas, api)
    np.testing.assert_array_equal(x, x2)  # seeded
    x3, *_ = fcmd.init_random(18, n, box, 3.0, 84.0, mas, api)
    assert not np.array_equal(x, x3)
    p = np.stack([x, y, z], 1)
    assert (p >= 0).all() and (p < box).all()
    d = p[:, None] - p[None]
    d -= box * np.round(d / box)
    r2 = (d ** 2).sum(-1) + np.eye(n) * 1e9
    assert r2.min() >= 9.0 * (1 - 1e-12)
    m = mas[api - 1]
    for v in (vx, vy, vz):
        assert abs((m * v).sum()) < 1e-9 * (m * np.abs(v)).sum()
    ek = 0.5 * (m * (vx ** 2 + vy ** 2 + vz ** 2)).sum()
    assert ek == pytest.approx(1.5 * fcmd.CIKB * 84.0 * n, rel=1e-12)
