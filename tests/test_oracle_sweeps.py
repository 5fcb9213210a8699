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


def test_preconditioner_approximates_inverse():
    """M^{-1} A is close to the identity for the L-sweeps (few GMRES steps)."""
    cb = _cb((3, 3), (12, 12))
    P = Preconditioner(cb)
    A = cb.global_operator()
    rng = np.random.default_rng(3)
    x = rng.standard_normal(cb.N) + 1j * rng.standard_normal(cb.N)
    y = P.apply_global(A @ x)
    assert np.linalg.norm(y - x) / np.linalg.norm(x) < 0.2


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
