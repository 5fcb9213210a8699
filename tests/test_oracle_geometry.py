"""Pin the oracle's discretisation and geometry with closed forms and
invariants (the reference ships no tests or fixtures)."""
import math

import numpy as np
import pytest

from oracle.geometry import (Checkerboard, ConstantMedium, CSVMedium, assemble_helmholtz,
                             fd_weights, pml_coefficients)

OMEGA = 2 * math.pi * 4.1


def test_fornberg_order2_closed_forms():
    # Order 2 central stencils (src/FiniteDifferences.cpp:11-54)
    assert np.allclose(fd_weights(1, [0, 1, 2], 2), [1.0, -2.0, 1.0], atol=0, rtol=0)
    assert np.allclose(fd_weights(1, [0, 1, 2], 1), [-0.5, 0.0, 0.5], atol=0, rtol=0)
    # one-sided weights for a sanity check of the recursion: f'(0) from 0,1,2
    assert np.allclose(fd_weights(0, [0, 1, 2], 1), [-1.5, 2.0, -0.5])
    # Fornberg on 5 points reproduces the 4th-order second derivative
    assert np.allclose(fd_weights(2, [0, 1, 2, 3, 4], 2), [-1 / 12, 4 / 3, -5 / 2, 4 / 3, -1 / 12])


def test_pml_closed_form_alpha():
    intd = [[0.0, 1.0], [0.0, 1.0]]
    comp = [[-0.2, 1.2], [-0.2, 1.2]]
    C = math.log(OMEGA)
    x = np.array([-0.1, 0.5, 1.15])
    y = np.array([0.3, 1.05, -0.19])
    L1, L2, B1, B2 = pml_coefficients(x, y, intd, comp, C)
    for k in range(3):
        for v, lo_i, hi_i, L, B in ((x[k], 0.0, 1.0, L1[k], B1[k]), (y[k], 0.0, 1.0, L2[k], B2[k])):
            if v < lo_i:
                s = C * ((lo_i - v) / 0.2) ** 3
                ds = -3 * C * (lo_i - v) ** 2 / 0.2 ** 3
            elif v > hi_i:
                s = C * ((v - hi_i) / 0.2) ** 3
                ds = 3 * C * (v - hi_i) ** 2 / 0.2 ** 3
            else:
                s = ds = 0.0
            a = 1 / (1 + 1j * s)
            # alpha' by central differences of alpha(sigma(x))
            def alpha_of(t):
                if t < lo_i:
                    return 1 / (1 + 1j * C * ((lo_i - t) / 0.2) ** 3)
                if t > hi_i:
                    return 1 / (1 + 1j * C * ((t - hi_i) / 0.2) ** 3)
                return 1.0
            e = 1e-6
            ad = (alpha_of(v + e) - alpha_of(v - e)) / (2 * e)
            assert abs(L - a * a) < 1e-14
            assert abs(B - ad * a) < 1e-6 * max(1, abs(ad))


def test_interior_rows_are_textbook_helmholtz():
    cb = Checkerboard([[0, 1], [0, 1]], (1, 1), (30, 30), OMEGA, 2.0, ConstantMedium())
    A = cb.global_operator().tocsr()
    g = cb.gcell
    h = g.hx
    # an interior node well inside [0,1]^2: -Lap_h - w^2
    i, j = g.nx // 2, g.ny // 2
    d = i * g.ny + j
    row = A.getrow(d).toarray().ravel()
    assert abs(row[d] - (4 / h ** 2 - OMEGA ** 2)) < 1e-9 * 4 / h ** 2
    for q in (d - 1, d + 1, d - g.ny, d + g.ny):
        assert abs(row[q] + 1 / h ** 2) < 1e-9 / h ** 2
    assert np.count_nonzero(row) == 5


def test_matrix_is_complex_symmetric_only_in_interior():
    # the PML convection term makes A non-symmetric; interior block is symmetric
    cb = Checkerboard([[0, 1], [0, 1]], (1, 1), (30, 30), OMEGA, 2.0, ConstantMedium())
    A = cb.global_operator().toarray()
    assert np.abs(A - A.T).max() > 1e-3
    g = cb.gcell
    ii = [i * g.ny + j for i in range(25, 35) for j in range(25, 35)]
    S = A[np.ix_(ii, ii)]
    assert np.abs(S - S.T).max() < 1e-9


@pytest.mark.parametrize("nc,nel", [((2, 2), (19, 19)), ((3, 3), (12, 12)), ((2, 4), (19, 9))])
def test_owned_sets_partition_global_grid(nc, nel):
    cb = Checkerboard([[0, 1], [0, 1]], nc, nel, OMEGA, 2.0, ConstantMedium())
    for shift in ((False, False), (True, False), (False, True), (True, True)):
        owned = np.concatenate([c.gidx[c.keep[shift]] for c in cb.cells])
        assert len(owned) == cb.N
        assert len(np.unique(owned)) == cb.N
        assert owned.min() == 0 and owned.max() == cb.N - 1


@pytest.mark.parametrize("nc,nel", [((2, 2), (19, 19)), ((3, 3), (12, 12))])
def test_owned_local_rows_equal_global_rows(nc, nel):
    """applySystemMatrix via local matrices == global operator."""
    cb = Checkerboard([[0, 1], [0, 1]], nc, nel, OMEGA, 2.0, ConstantMedium())
    G = cb.global_operator().tocsr()
    for c in cb.cells:
        L = c.A.tocsr()
        for p in c.keep[(False, False)][:: max(1, len(c.keep[(False, False)]) // 200)]:
            lr = L.getrow(p)
            gr = G.getrow(c.gidx[p])
            lmap = {c.gidx[q]: v for q, v in zip(lr.indices, lr.data)}
            gmap = dict(zip(gr.indices, gr.data))
            assert set(lmap) == set(gmap)
            for k in gmap:
                assert abs(lmap[k] - gmap[k]) <= 1e-9 * abs(gmap[k]) + 1e-12


def test_trace_and_ltrace_sets_geometry():
    cb = Checkerboard([[0, 1], [0, 1]], (3, 3), (12, 12), OMEGA, 2.0, ConstantMedium())
    c = cb.cell(1, 1)
    h = cb.h[0]
    a = (12 + 1) * 1 * h
    for shifted, off in ((False, 0), (True, 2)):
        lr = c.lr[shifted]
        # trace0 at a-h (+2h shifted), trace1 at a, traceN at b, traceNP at b+h
        assert np.allclose(c.xs[np.unique(lr[1] // c.ny)], a - h + off * h)
        assert np.allclose(c.xs[np.unique(lr[2] // c.ny)], a + off * h)
        b = a + 12 * h
        assert np.allclose(c.xs[np.unique(lr[4] // c.ny)], b + off * h)
        assert np.allclose(c.xs[np.unique(lr[5] // c.ny)], b + h + off * h)
    # BottomLeft L-traces: outer L through (a-h, a-h), inner through (a, a)
    s0, s1 = c.ltr[("BottomLeft", False, False)]
    X, Y = c.dof_points()
    x0, y0 = X[s0], Y[s0]
    on_col = np.isclose(x0, a - h) & (Y[s0] >= a - h - 1e-12)
    on_row = np.isclose(y0, a - h) & (X[s0] >= a - h - 1e-12)
    assert np.all(on_col | on_row)
    assert len(s0) == (c.ny - np.argmin(np.abs(c.ys - (a - h)))) + (c.nx - np.argmin(np.abs(c.xs - (a - h)))) - 1
    assert len(np.intersect1d(s0, s1)) == 0


def test_csv_medium_lookup_and_clamp():
    v = np.array([[1.0, 2.0], [3.0, 4.0]], dtype=np.float32)  # row 0 = top
    m = CSVMedium(v, [[0.0, 1.0], [0.0, 1.0]])
    assert m.min_real() == pytest.approx(1 / 16)
    # top-left cell holds 1.0, bottom-right 4.0
    assert m.eval(np.array([0.25]), np.array([0.75]))[0] == pytest.approx(1.0)
    assert m.eval(np.array([0.75]), np.array([0.25]))[0] == pytest.approx(1 / 16)
    # points outside clamp to the edge cells (PML region)
    assert m.eval(np.array([-0.3]), np.array([1.4]))[0] == pytest.approx(1.0)
    assert m.eval(np.array([1.3]), np.array([-0.2]))[0] == pytest.approx(1 / 16)


def test_point_source_integrates_to_one():
    cb = Checkerboard([[0, 1], [0, 1]], (1, 1), (60, 60), OMEGA, 2.0, ConstantMedium())
    f = cb.global_rhs([(0.5, 0.5)])
    h = cb.gcell.hx
    # eps = h and the source on a node: the discrete integral is the Jacobi
    # theta sum (sum_k exp(-k^2))^2 / pi = 1.000207...
    theta = np.exp(-np.arange(-30, 31) ** 2.0).sum()
    assert abs(f.real.sum() * h * h - theta ** 2 / np.pi) < 1e-12
    assert np.all(f.imag == 0)


@pytest.mark.parametrize("general,dom,nc,nel", [
    (False, [[0.0, 1.0], [0.0, 1.0]], (3, 3), (12, 12)),
    (True, [[0.0, 3.0], [0.0, 1.0]], (2, 4), (14, 21)),
])
def test_geometry_only_mode_matches_full_checkerboard(general, dom, nc, nel):
    """The geometry-only checkerboard (no cell index sets) gives the same
    global operator, DD mesh size and global right-hand side."""
    omega = 2 * math.pi * 3.3
    full = Checkerboard(dom, nc, nel, omega, 2.0, ConstantMedium(), general=general)
    lite = Checkerboard(dom, nc, nel, omega, 2.0, ConstantMedium(), general=general, cells=False)
    assert lite.mesh_size() == full.mesh_size()
    assert (lite.Gx, lite.Gy) == (full.Gx, full.Gy)
    assert (lite.global_operator() != full.global_operator()).nnz == 0
    srcs = [(0.41 * dom[0][1], 0.37), (0.8 * dom[0][1], 0.61)]
    a, b = lite.global_rhs(srcs), full.global_rhs(srcs)
    # owned points of the cells tile the grid; values agree to coordinate rounding
    assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))


def _pml_operator_error(n):
    """max |A u - L u| over the DOFs of an n x n grid on [0,1]^2 whose outer
    0.2 on every side is cubic PML, for u = sin(pi x) sin(pi y) e^{i(2x+y)}
    (zero on the Dirichlet boundary), L u = -a1 d/dx(a1 du/dx) -
    a2 d/dy(a2 du/dy) - w^2 u with a = 1/(1 + i sigma) and sigma the cubic
    profile of src/PMLCubic2D.cpp:119-147.  The derivatives of a1 du/dx are
    taken numerically (Richardson-extrapolated central differences), not from
    the closed-form alpha' the assembly uses."""
    omega = 10.0
    C = math.log(omega)
    comp = [[0.0, 1.0], [0.0, 1.0]]
    intd = [[0.2, 0.8], [0.2, 0.8]]
    h = 1.0 / n
    xs = (np.arange(n - 1) + 1) * h
    ys = xs.copy()
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    m = np.ones(X.size)
    A = assemble_helmholtz(xs, ys, h, h, intd, comp, C, omega, m)

    def sigma(v, lo_c, lo_i, hi_i, hi_c):
        s = np.zeros_like(v)
        s = np.where(v < lo_i, C * ((lo_i - v) / (lo_i - lo_c)) ** 3, s)
        return np.where(v > hi_i, C * ((v - hi_i) / (hi_c - hi_i)) ** 3, s)

    a1 = lambda x: 1.0 / (1.0 + 1j * sigma(x, 0.0, 0.2, 0.8, 1.0))
    u = lambda x, y: np.sin(np.pi * x) * np.sin(np.pi * y) * np.exp(1j * (2 * x + y))
    ux = lambda x, y: (np.pi * np.cos(np.pi * x) + 2j * np.sin(np.pi * x)) * np.sin(np.pi * y) * np.exp(1j * (2 * x + y))
    uy = lambda x, y: np.sin(np.pi * x) * (np.pi * np.cos(np.pi * y) + 1j * np.sin(np.pi * y)) * np.exp(1j * (2 * x + y))

    def ddx(f, x, y, d=1e-3):
        c = lambda t: (f(x + t, y) - f(x - t, y)) / (2 * t)
        return (4 * c(d / 2) - c(d)) / 3

    Lu = (-a1(X) * ddx(lambda x, y: a1(x) * ux(x, y), X, Y)
          - a1(Y) * ddx(lambda y, x: a1(y) * uy(x, y), Y, X)
          - omega ** 2 * u(X, Y))
    return np.max(np.abs(A @ u(X, Y).ravel() - Lu.ravel())) / np.max(np.abs(Lu))


def test_pml_rows_converge_to_the_continuous_operator():
    """The assembled rows, PML rows included, approximate the continuous
    stretched-coordinate operator to second order: halving h divides the
    error by about 4.  A dropped convection term (alpha' alpha u_x), its sign
    flipped, or alpha used where alpha^2 belongs leaves an error of 0.37,
    0.73 or 0.084 that does not fall with h."""
    e = [_pml_operator_error(n) for n in (40, 80, 160)]
    # measured 9.9e-4, 2.5e-4, 6.3e-5
    assert e[2] < 1e-4
    for a, b in zip(e, e[1:]):
        assert 3.5 < a / b < 4.5, e


def test_csv_clamp_on_non_square_file_grids():
    """FunctionCSV::eval (src/FunctionCSV.cpp:214-235) clamps every axis from
    below at lo + mH[0]/2 and from above at hi - mH[1]/2, with mH[0] the
    x spacing and mH[1] the y spacing of the file grid.  On a non-square grid
    that moves points into other cells; the expected cells below are worked
    out by hand from those lines."""
    c = np.arange(1, 17, dtype=np.float32).reshape(4, 4)   # row 0 = top
    # 4 x 4 cells over [0,4] x [0,1]: mH = (1.0, 0.25); the y clamp from
    # below is at 0.5 (not 0.125), so the bottom row (y < 0.25) is never read
    med = CSVMedium(c, [[0.0, 4.0], [0.0, 1.0]])
    pts = [((2.5, 0.1), (2, 2)),    # y -> 0.5: row (1 - 0.5) / 0.25 = 2
           ((2.5, 0.6), (1, 2)),    # no clamp: row floor(0.4 / 0.25) = 1
           ((0.2, 0.95), (0, 0)),   # x -> 0.5, y -> 1 - 0.125
           ((3.95, 0.3), (2, 3))]   # x -> 4 - 0.125: col 3; y -> 0.5: row 2
    for (x, y), (r, col) in pts:
        assert med.eval(np.array([x]), np.array([y]))[0] == 1.0 / float(c[r, col]) ** 2, (x, y)
    # 4 x 4 cells over [0,1] x [0,4]: mH = (0.25, 1.0); the x clamp from
    # above is at 1 - 0.5, so the last column (x > 0.75) is never read
    med = CSVMedium(c, [[0.0, 1.0], [0.0, 4.0]])
    pts = [((0.9, 2.5), (1, 2)),    # x -> 0.5: col 2; row floor(1.5 / 1) = 1
           ((0.1, 0.05), (3, 0)),   # x -> 0.125: col 0; y -> 0.125: row 3
           ((0.6, 3.9), (0, 2))]    # y -> 4 - 0.5: row 0; col floor(0.6 / 0.25) = 2
    for (x, y), (r, col) in pts:
        assert med.eval(np.array([x]), np.array([y]))[0] == 1.0 / float(c[r, col]) ** 2, (x, y)
