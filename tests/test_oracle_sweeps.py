"""Pin the oracle's L-sweeps preconditioner with exactness, linearity,
re-layout and convergence properties."""
import math

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle.geometry import Checkerboard, ConstantMedium
from oracle.gmres import fgmres, gmres_reference
from oracle.sweeps import Preconditioner

OMEGA = 2 * math.pi * 4.1


def _cb(nc, nel, omega=OMEGA):
    cb = Checkerboard([[0, 1], [0, 1]], nc, nel, omega, 2.0, ConstantMedium())
    cb.factorize()
    return cb


def test_one_cell_preconditioner_is_the_inverse():
    cb = _cb((1, 1), (30, 30))
    P = Preconditioner(cb)
    rng = np.random.default_rng(0)
    x = rng.standard_normal(cb.N) + 1j * rng.standard_normal(cb.N)
    y = P.apply_global(x)
    A = cb.global_operator()
    assert np.linalg.norm(A @ y - x) / np.linalg.norm(x) < 1e-12
    assert P.nsolves == 4  # one stage-1 solve per windowed DD, no sweeps


def test_solve_count_per_application():
    cb = _cb((3, 3), (12, 12))
    P = Preconditioner(cb)
    P.apply_global(np.ones(cb.N, dtype=complex))
    n = 3
    assert P.nsolves == 4 * (3 * n - 2) ** 2


def test_preconditioner_is_linear():
    cb = _cb((3, 3), (12, 12))
    P = Preconditioner(cb)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(cb.N) + 1j * rng.standard_normal(cb.N)
    y = rng.standard_normal(cb.N) + 1j * rng.standard_normal(cb.N)
    a, b = 0.7 - 0.2j, -1.3 + 0.5j
    lhs = P.apply_global(a * x + b * y)
    rhs = a * P.apply_global(x) + b * P.apply_global(y)
    assert np.linalg.norm(lhs - rhs) / np.linalg.norm(rhs) < 1e-12


def test_shift_conversions_are_exact_relayouts():
    cb = _cb((3, 3), (12, 12))
    P = Preconditioner(cb)
    rng = np.random.default_rng(2)
    g = rng.standard_normal(cb.N) + 1j * rng.standard_normal(cb.N)
    x = cb.from_global(g)
    for conv, back, shift in ((P.nonshifted_to_shifted_lr, P.shifted_to_nonshifted_lr, (True, False)),
                              (P.nonshifted_to_shifted_ud, P.shifted_to_nonshifted_ud, (False, True))):
        xs = conv(x, False)
        # shifted layout holds the same grid function at the shifted owned sets
        for c, v in zip(cb.cells, xs):
            assert np.array_equal(v, g[c.gidx[c.keep[shift]]])
        xb = back(xs, False)
        for u, v in zip(x, xb):
            assert np.array_equal(u, v)


def _precond_error(min_wl_in_pml):
    cb = Checkerboard([[0, 1], [0, 1]], (3, 3), (12, 12), OMEGA, min_wl_in_pml, ConstantMedium())
    cb.factorize()
    P = Preconditioner(cb)
    A = cb.global_operator()
    rng = np.random.default_rng(3)
    x = rng.standard_normal(cb.N) + 1j * rng.standard_normal(cb.N)
    y = P.apply_global(A @ x)
    return np.linalg.norm(y - x) / np.linalg.norm(x)


def test_preconditioner_error_falls_with_the_pml_width():
    """The L-sweeps would invert A exactly if the subdomain PMLs absorbed
    perfectly; the only error is their reflection, so ||M^-1 A x - x|| / ||x||
    must fall as the PML widens.  Measured on this 3x3 board: 4.3e-3, 1.4e-4,
    1.5e-5, 4.9e-6 at 1, 2, 3, 4 wavelengths (9, 19, 28, 37 PML cells).  A
    skipped L-sweep leaves 0.25 / 0.18 and a flipped sign of the L
    equivalent sources 3.4 / 2.5 at 2 / 4 wavelengths: errors no PML width
    removes."""
    e = [_precond_error(w) for w in (1.0, 2.0, 4.0)]
    assert e[1] < 3e-4
    assert e[0] > 10 * e[1] and e[1] > 10 * e[2]
    assert e[2] < 1e-5


def test_windows_partition_the_input_and_avoid_each_dd_s_own_interfaces():
    """getSeparatedRHS (src/DDCheckerboard.cpp:3564-3674) splits the input
    among the four DDs.  The four windows sum to the input, and the window of
    DD (slr, sud) vanishes on the interface traces (sets 1, 2, 4, 5 on both
    axes) of that DD's own layout, so that no DD carries a source on its own
    interfaces: the reason three of the four DDs are shifted.  Giving DD2 and
    DD3 each other's window (the reference's comments on the two else-if
    branches name them the other way round) puts O(1) values there."""
    cb = Checkerboard([[0, 1], [0, 1]], (3, 3), (12, 12), OMEGA, 2.0, ConstantMedium(), assemble=False)
    P = Preconditioner(cb)
    rng = np.random.default_rng(5)
    g = rng.standard_normal(cb.N) + 1j * rng.standard_normal(cb.N)
    x = cb.from_global(g)
    total = np.zeros(cb.N, dtype=complex)
    for slr in (False, True):
        for sud in (False, True):
            w = P.separated_rhs(x, slr, sud)
            gw = np.zeros(cb.N, dtype=complex)
            for c, v in zip(cb.cells, w):
                k = c.keep[(False, False)]
                gw[c.gidx[k]] = v[k]
            total += gw
            own = np.concatenate([c.gidx[idx] for c in cb.cells for si in (1, 2, 4, 5)
                                  for idx in (c.lr[slr][si], c.ud[sud][si])])
            own = own[(own >= 0) & (own < cb.N)]
            assert own.size > 0
            assert np.all(gw[own] == 0), (slr, sud)
    assert np.array_equal(total, g)


@pytest.mark.parametrize("nc,nel", [((2, 2), (19, 19)), ((4, 4), (9, 9))])
def test_reference_gmres_converges_to_direct_solution(nc, nel):
    cb = _cb(nc, nel)
    P = Preconditioner(cb)
    A = cb.global_operator()
    b = cb.global_rhs([(0.5, 0.5)])
    x, log = gmres_reference(lambda v: A @ v, P.apply_global, b, tol=1e-8, max_iter=30)
    u = spla.spsolve(A.tocsc(), b)
    assert len(log) - 1 <= 5
    assert np.linalg.norm(x - u) / np.linalg.norm(u) < 1e-6
    xf, it, hist = fgmres(lambda v: A @ v, P.apply_global, b, tol=1e-10, restart=20, max_iter=40)
    assert np.linalg.norm(xf - u) / np.linalg.norm(u) < 1e-8
    assert it <= 8


def test_block_rhs_superposition():
    """The solution for a sum of sources equals the sum of solutions."""
    cb = _cb((2, 2), (19, 19))
    P = Preconditioner(cb)
    A = cb.global_operator()
    s1, s2 = (0.3, 0.6), (0.7, 0.2)
    u = []
    for srcs in ([s1], [s2], [s1, s2]):
        x, _, _ = fgmres(lambda v: A @ v, P.apply_global, cb.global_rhs(srcs), tol=1e-12, restart=20)
        u.append(x)
    assert np.linalg.norm(u[0] + u[1] - u[2]) / np.linalg.norm(u[2]) < 1e-9


def test_solves_on_worker_processes_give_the_same_bits():
    """The CPU baseline runs the oracle's local solves on worker processes,
    an anti-diagonal at a time (cells of one diagonal read only traces of
    earlier diagonals); the result is the sequential oracle's, bit for bit,
    for both cell-to-process mappings."""
    from oracle.parallel import CellPool
    cb = _cb((3, 3), (12, 12))
    rng = np.random.default_rng(6)
    x = rng.standard_normal(cb.N) + 1j * rng.standard_normal(cb.N)
    ref = Preconditioner(cb).apply_global(x)
    for assign in ("cells", "rows"):
        with CellPool(cb, 3, assign) as pool:
            P = Preconditioner(cb, solver=pool)
            assert np.array_equal(P.apply_global(x), ref)
            assert P.nsolves == 4 * (3 * 3 - 2) ** 2
