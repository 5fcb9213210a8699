"""CPU-side checks of the C-ABI library: exports, host geometry against the
oracle, error paths.  No compute call needs a GPU here."""
import math
import os
import re

import numpy as np
import pytest

from tests._common import make_pair

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"^\s*(?:int|const char\*)\s+(lsw_\w+)\s*\(", txt, re.M))
    return names


def test_library_exports_every_declared_symbol():
    import lsweeps_amd as L
    decl = _declared()
    assert len(decl) >= 25
    for name in decl:
        assert hasattr(L.lib(), name), name
    assert set(L.EXPORTED) <= decl


@pytest.mark.parametrize("nc,nel,omega", [
    ((1, 1), (30, 30), 2 * math.pi * 4.1),
    ((2, 2), (19, 19), 2 * math.pi * 4.1),
    ((3, 3), (12, 12), 2 * math.pi * 4.1),
    ((2, 4), (19, 9), 2 * math.pi * 4.1),
    ((4, 4), (99, 99), 2 * math.pi * 40),      # the 400x400, 4x4 workload
])
def test_host_geometry_matches_oracle(nc, nel, omega):
    cb, dd = make_pair(nc, nel, omega, device=False)
    assert (dd.Gx, dd.Gy) == (cb.Gx, cb.Gy)
    assert dd.mesh_size == pytest.approx(cb.mesh_size(), rel=1e-15)
    xs, ys = dd.coords()
    assert np.allclose(xs, cb.gxs, rtol=0, atol=1e-14)
    assert np.allclose(ys, cb.gys, rtol=0, atol=1e-14)
    for k, c in enumerate(cb.cells):
        inf = dd.cell_info(k)
        assert (inf["I"], inf["J"], inf["nx"], inf["ny"], inf["ox"], inf["oy"]) == \
            (c.I, c.J, c.nx, c.ny, c.ox, c.oy)
        for sh in (0, 1):
            for s in range(7):
                cols = np.unique(c.lr[bool(sh)][s] // c.ny)
                assert list(cols) == list(range(inf["vlr"][sh][s], inf["vlr"][sh][s + 1]))
                rows = np.unique(c.ud[bool(sh)][s] % c.ny)
                assert list(rows) == list(range(inf["vud"][sh][s], inf["vud"][sh][s + 1]))


def test_lens_medium_geometry_and_pml():
    from tests._common import lens_velocity
    vel = lens_velocity(40)
    cb, dd = make_pair((2, 2), (19, 19), 2 * math.pi * 2.5, velocity=vel, device=False)
    assert (dd.Gx, dd.Gy) == (cb.Gx, cb.Gy)


def test_error_paths():
    import lsweeps_amd as L
    with pytest.raises(L.LswError) as e:
        L.DDCheckerboard(L.dd_params(nels=(4, 9)), device=False)
    assert e.value.code == 1 and "Not enough elements" in str(e.value)
    dd = L.DDCheckerboard(L.dd_params(ncells=(2, 2), nels=(19, 19)), device=False)
    with pytest.raises(L.LswError) as e:
        dd.factorize()
    assert e.value.code == 3
    with pytest.raises(L.LswError) as e:
        L.DDCheckerboard(L.dd_params(order=4), device=False)
    assert e.value.code == 6
    s = L.LocalLinearSolverHIP()
    import scipy.sparse as sp
    with pytest.raises(L.LswError) as e:
        s.set_matrix(sp.random(20, 20, density=0.3, format="csc", random_state=1) + sp.eye(20))
    assert e.value.code == 6


def test_general_geometry_rectangular_domain():
    """3x1 domain, isotropic h = 1/29: 2 rows of NX=14 and 4 columns of NY=21."""
    dom = ((0.0, 3.0), (0.0, 1.0))
    cb, dd = make_pair((2, 4), (14, 21), 2 * math.pi * 2.2, device=False, dom=dom, general=True)
    assert (dd.Gx, dd.Gy) == (cb.Gx, cb.Gy)
    assert cb.h[0] == pytest.approx(cb.h[1], rel=1e-12)
    for k, c in enumerate(cb.cells):
        inf = dd.cell_info(k)
        assert (inf["nx"], inf["ny"], inf["ox"], inf["oy"]) == (c.nx, c.ny, c.ox, c.oy)
    for shift in ((False, False), (True, True)):
        owned = np.concatenate([c.gidx[c.keep[shift]] for c in cb.cells])
        assert len(np.unique(owned)) == cb.N == len(owned)


def test_reference_geometry_rejects_inconsistent_tiling():
    import lsweeps_amd as L
    with pytest.raises(L.LswError) as e:
        L.DDCheckerboard(L.dd_params(domain=((0.0, 3.0), (0.0, 1.0)), ncells=(2, 4), nels=(14, 21)),
                         device=False)
    assert e.value.code == 6
