"""Oracle: the point, element and value layout of OutputVTK::print2D.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates, for one process holding every cell (cells in mCells order):
* OutputVTK::setVals / print2D              src/OutputVTK.cpp:65-73, 82-366
* SubdomainStructured::getPhysPts           src/SubdomainStructured.cpp:508-523
* SubdomainStructured::computeVolumeElIndexSets
                                            src/SubdomainStructured.cpp:804-1000
* AssemblerHelmholtz2DFD::getNonDOFPhysPts / getElPhysPtIndices /
  expandDOFValsToPhysVals                   src/AssemblerHelmholtz2DFD.cpp:185-345

Values: setVals expands the owned values, fills the neighbours' trace lines
(communicateTR2BLTraces, src/DDCheckerboard.cpp:2813-2992: columns from the
right neighbour, then rows from the top one, which carries the diagonal
corner) and pads the non-DOF points with zero; every printed DOF point then
holds the global solution at that point.

Pinning: tests/test_host_output.py checks properties the layout must have
whatever the code says: every printed element is an hx x hy square, the
printed elements tile the global computational domain exactly once, and each
point's value is the solution at that point (zero on the Dirichlet boundary).
"""
import numpy as np


def _phys_points(c):
    """DOF points (index i*ny + j) then the 2 ex + 2 ey boundary points:
    bottom, right, top, left (getNonDOFPhysPts)."""
    ex, ey = c.nx + 1, c.ny + 1
    (x0, x1), (y0, y1) = c.comp_dom
    X, Y = np.meshgrid(c.xs, c.ys, indexing="ij")
    px, py = [X.ravel()], [Y.ravel()]
    i = np.arange(ex, dtype=np.float64)
    j = np.arange(ey, dtype=np.float64)
    px += [x0 + i * c.hx, np.full(ey, x1), x1 - i * c.hx, np.full(ey, x0)]
    py += [np.full(ex, y0), y0 + c.hy * j, np.full(ex, y1), y1 - c.hy * j]
    return np.concatenate(px), np.concatenate(py)


def _element(c, J, I):
    """getElPhysPtIndices for element (J = x index, I = y index)."""
    n0, n1 = c.nx + 1, c.ny + 1
    size = c.nx * c.ny
    if J == 0 and I == 0:
        return [size, size + 1, 0, size + 2 * n0 + 2 * n1 - 1]
    if J == n0 - 1 and I == 0:
        return [size + n0 - 1, size + n0, size + n0 + 1, (n0 - 2) * (n1 - 1)]
    if J == n0 - 1 and I == n1 - 1:
        return [(n0 - 1) * (n1 - 1) - 1, size + n0 + n1 - 1, size + n0 + n1, size + n0 + n1 + 1]
    if J == 0 and I == n1 - 1:
        return [size + 2 * n0 + n1 + 1, n1 - 2, size + 2 * n0 + n1 - 1, size + 2 * n0 + n1]
    if I == 0:
        return [size + J, size + J + 1, J * (n1 - 1), (J - 1) * (n1 - 1)]
    if J == n0 - 1:
        return [(n0 - 2) * (n1 - 1) + I - 1, size + n0 + I, size + n0 + I + 1, (n0 - 2) * (n1 - 1) + I]
    if I == n1 - 1:
        return [(J - 1) * (n1 - 1) + n1 - 2, J * (n1 - 1) + n1 - 2,
                size + 2 * n0 + n1 - 2 - (J - 1), size + 2 * n0 + n1 - 1 - (J - 1)]
    if J == 0:
        return [size + 2 * n0 + 2 * n1 - I, I - 1, I, size + 2 * n0 + 2 * n1 - 1 - I]
    return [(J - 1) * (n1 - 1) + (I - 1), (J - 1) * (n1 - 1) + I, J * (n1 - 1) + I, J * (n1 - 1) + (I - 1)]


def _classes(mid, coord, sets, lo, hi):
    """computeVolumeElIndexSets: element midpoint against the extent of
    trace0, trace1, traceN, traceNP (sets 1, 2, 4, 5 of the 7-set view)."""
    mn, mx = [], []
    for t, k in enumerate((1, 2, 4, 5)):
        v = coord[sets[k]]
        if v.size:
            mn.append(v.min())
            mx.append(v.max())
        else:
            mn.append(lo - 1.0 if t < 2 else hi + 1.0)
            mx.append(mn[-1])
    if all(sets[k].size == 0 for k in (1, 2, 4, 5)):
        return np.full(mid.shape, 4)
    edges = [mn[0], mx[0], mn[1], mx[1], mn[2], mx[2], mn[3], mx[3]]
    cls = np.full(mid.shape, 8)
    for k in range(7, -1, -1):
        cls = np.where(mid < edges[k], k, cls)
    return cls


def print2d_layout(cb, x):
    """(points (n, 2), elements (m, 4) of point indices, values (n,)) of the
    file print2D writes for the global grid vector x (index gi*Gy + gj)."""
    pts, els, vals = [], [], []
    base = 0
    for c in cb.cells:
        px, py = _phys_points(c)
        n0, n1 = c.nx + 1, c.ny + 1
        # element midpoints: the mean of the four corners' coordinates
        ex = np.array([np.mean(px[_element(c, J, min(1, n1 - 1))]) for J in range(n0)])
        ey = np.array([np.mean(py[_element(c, min(1, n0 - 1), I)]) for I in range(n1)])
        X, Y = np.meshgrid(c.xs, c.ys, indexing="ij")
        cx = _classes(ex, X.ravel(), c.lr[False], px.min(), px.max())
        cy = _classes(ey, Y.ravel(), c.ud[False], py.min(), py.max())
        cell_els = [_element(c, J, I) for J in range(n0) if 3 <= cx[J] <= 6
                    for I in range(n1) if 3 <= cy[I] <= 6]
        used = np.zeros(px.size, dtype=bool)
        for e in cell_els:
            used[e] = True
        idx = np.nonzero(used)[0]
        remap = np.full(px.size, -1)
        remap[idx] = base + np.arange(idx.size)
        base += idx.size
        pts.append(np.stack([px[idx], py[idx]], axis=1))
        els += [remap[e] for e in cell_els]
        v = np.zeros(idx.size, dtype=np.complex128)
        dof = idx < c.size
        gi = c.gidx[idx[dof]]
        v[dof] = x[gi]
        vals.append(v)
    return np.concatenate(pts), np.array(els), np.concatenate(vals)
