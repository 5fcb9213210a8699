"""Block right-hand sides: M^{-1} applied to several vectors at once equals
the single-vector application, for both precisions.

With FP64 factors the two differ only by the order of the additions (1e-12).
With FP32 factors the grouping of right-hand sides into solve tasks changes
which tree nodes a task visits and so the order of FP32 additions; through
the 21 sweep steps the two results then differ at the FP32 level of the
preconditioner itself (1e-6 .. 1e-4 relative), which FP64 GMRES absorbs.  An
indexing mistake shows as an O(1) difference."""
import math

import numpy as np
import pytest

from oracle.sweeps import Preconditioner
from tests._common import make_pair, relerr

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("nrhs", [2, 5])
def test_block_precond_equals_single(prec, nrhs):
    cb, dd = make_pair((3, 3), (12, 12), 2 * math.pi * 4.1, precision=prec)
    dd.factorize()
    rng = np.random.default_rng(11)
    X = rng.standard_normal((nrhs, dd.N)) + 1j * rng.standard_normal((nrhs, dd.N))
    Y = dd.apply_precond_block(X)
    for k in range(nrhs):
        y1 = dd.apply_precond(X[k])
        assert relerr(Y[k], y1) < (1e-12 if prec == 0 else 1e-4)


def test_block_precond_fp64_matches_oracle():
    cb, dd = make_pair((2, 4), (19, 9), 2 * math.pi * 4.1)
    cb.factorize()
    dd.factorize()
    rng = np.random.default_rng(12)
    X = rng.standard_normal((3, dd.N)) + 1j * rng.standard_normal((3, dd.N))
    Y = dd.apply_precond_block(X)
    P = Preconditioner(cb)
    for k in range(3):
        assert relerr(Y[k], P.apply_global(X[k])) < 1e-9


@pytest.mark.parametrize("prec", [0, 1])
def test_block_gmres_matches_direct_and_single(prec):
    import scipy.sparse.linalg as spla
    cb, dd = make_pair((3, 3), (19, 19), 2 * math.pi * 3.3, precision=prec)
    dd.factorize()
    srcs = [(0.5, 0.5), (0.2, 0.8), (0.7, 0.3), (0.45, 0.15), (0.9, 0.9)]
    B = np.stack([dd.assemble_rhs([s]) for s in srcs])
    X, infos = dd.gmres_block(B, tol=1e-8, max_iter=60, restart=20)
    A = cb.global_operator().tocsc()
    lu = spla.splu(A)
    for k, s in enumerate(srcs):
        u = lu.solve(B[k])
        assert infos[k]["true_residual"] < 1e-8
        assert relerr(X[k], u) < 1e-6
        x1, info1 = dd.gmres(B[k], mode=1, tol=1e-8, max_iter=60, restart=20)
        assert relerr(X[k], x1) < 1e-6
        assert abs(infos[k]["iterations"] - info1["iterations"]) <= 1
    # superposition: the summed source (the reference's single RHS) equals the sum
    xs, _ = dd.gmres(B.sum(axis=0), mode=1, tol=1e-10, max_iter=60, restart=20)
    assert relerr(X.sum(axis=0), xs) < 1e-6


@pytest.mark.parametrize("prec", [0, 1])
def test_block_precond_wide_fronts(prec):
    """Cells of 100 x 100 nodes: the separators near the root are solved by
    several wavefronts per front (single and block right-hand sides)."""
    cb, dd = make_pair((2, 2), (59, 59), 2 * math.pi * 6.0, precision=prec)
    dd.factorize()
    rng = np.random.default_rng(13)
    X = rng.standard_normal((6, dd.N)) + 1j * rng.standard_normal((6, dd.N))
    Y = dd.apply_precond_block(X)
    if prec == 0:
        cb.factorize()
        P = Preconditioner(cb)
    for k in range(6):
        y1 = dd.apply_precond(X[k])
        assert relerr(Y[k], y1) < (1e-12 if prec == 0 else 1e-4)
        if prec == 0 and k < 2:
            assert relerr(Y[k], P.apply_global(X[k])) < 1e-9


def test_block_gmres_converged_columns_leave_the_block():
    """Right-hand sides that converge at different steps (here the point
    sources take one step more than the random vectors on an 8 x 8 board at
    tol 1e-10): once a column's estimate is below tol its inner steps stop
    (its preconditioner applications equal its own iterations), the others
    go on as a smaller block, and every column still equals its single-source
    solve."""
    cb, dd = make_pair((8, 8), (12, 12), 2 * math.pi * 9.0, precision=1)
    dd.factorize()
    srcs = [(0.5, 0.5), (0.2, 0.8), (0.7, 0.3), (0.05, 0.5), (0.93, 0.07), (0.31, 0.62)]
    rng = np.random.default_rng(5)
    B = np.stack([dd.assemble_rhs([s]) for s in srcs] +
                 [rng.standard_normal(dd.N) + 1j * rng.standard_normal(dd.N) for _ in range(2)])
    X, infos = dd.gmres_block(B, tol=1e-10, max_iter=60, restart=40)
    its = [inf["iterations"] for inf in infos]
    assert len(set(its)) > 1, its
    for k, inf in enumerate(infos):
        assert inf["precond_applies"] == inf["iterations"], (k, its)
        assert inf["true_residual"] < 1e-10
        x1, info1 = dd.gmres(B[k], mode=1, tol=1e-10, max_iter=60, restart=40)
        assert relerr(X[k], x1) < 1e-8
        assert abs(inf["iterations"] - info1["iterations"]) <= 1


def test_block_gmres_returns_when_right_hand_sides_hit_the_cap():
    """Block flexible GMRES terminates when right-hand sides stop at different
    steps: an easy one converges on its estimate, a hard one runs into
    max_iter across restarts.  Each iteration count stays within max_iter."""
    cb, dd = make_pair((2, 2), (19, 19), 2 * math.pi * 3.3)
    dd.factorize()
    b1 = dd.assemble_rhs([(0.5, 0.5)])
    B = np.stack([b1, b1 * 0.0 + b1[::-1]])
    X, infos = dd.gmres_block(B, tol=1e-15, max_iter=5, restart=2)
    for inf in infos:
        assert 1 <= inf["iterations"] <= 5
    X2, infos2 = dd.gmres_block(B, tol=1e-3, max_iter=7, restart=2)
    for inf in infos2:
        assert inf["iterations"] <= 7
        assert inf["true_residual"] <= 1e-3 or inf["iterations"] == 7
