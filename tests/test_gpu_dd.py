"""The DD preconditioner, operator and GMRES on the GPU against the oracle."""
import math

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle.gmres import gmres_reference
from oracle.sweeps import Operator, Preconditioner
from tests._common import lens_velocity, make_pair, relerr

pytestmark = pytest.mark.gpu

CASES = [
    ((1, 1), (30, 30), 2 * math.pi * 4.1),
    ((2, 2), (19, 19), 2 * math.pi * 4.1),
    ((3, 3), (12, 12), 2 * math.pi * 4.1),
    ((4, 4), (9, 9), 2 * math.pi * 4.1),
    ((2, 4), (19, 9), 2 * math.pi * 4.1),
    ((2, 2), (59, 59), 2 * math.pi * 6.0),   # fronts wide enough for multi-wavefront solves
]


@pytest.mark.parametrize("ncells,nels,omega", CASES)
def test_operator_and_precond_match_oracle(ncells, nels, omega):
    cb, dd = make_pair(ncells, nels, omega)
    assert (dd.Gx, dd.Gy) == (cb.Gx, cb.Gy)
    cb.factorize()
    dd.factorize()
    rng = np.random.default_rng(7)
    x = rng.standard_normal(cb.N) + 1j * rng.standard_normal(cb.N)
    A = cb.global_operator()
    assert relerr(dd.apply_a(x), A @ x) < 1e-13
    P = Preconditioner(cb)
    ref = P.apply_global(x)
    got = dd.apply_precond(x)
    assert relerr(got, ref) < 1e-9, relerr(got, ref)


@pytest.mark.parametrize("ncells,nels,omega", [CASES[2], CASES[4]])
def test_final_gather_equals_accumulate_and_combine(ncells, nels, omega, monkeypatch):
    """The one-pass final gather (owned windows partition the grid) and
    stage 1 built and solved in F give bit-for-bit the result of the
    accumulate-then-combine path with stage 1 built in X."""
    cb, dd = make_pair(ncells, nels, omega)
    monkeypatch.setenv("LSW_FINAL_GATHER", "0")
    monkeypatch.setenv("LSW_STAGE1_IN_F", "0")
    _, dd_acc = make_pair(ncells, nels, omega)
    dd_acc.factorize()
    monkeypatch.delenv("LSW_FINAL_GATHER")
    monkeypatch.delenv("LSW_STAGE1_IN_F")
    cb.factorize()
    dd.factorize()
    rng = np.random.default_rng(11)
    x = rng.standard_normal(cb.N) + 1j * rng.standard_normal(cb.N)
    got = dd.apply_precond(x)
    acc = dd_acc.apply_precond(x)
    assert np.array_equal(got, acc)
    ref = Preconditioner(cb).apply_global(x)
    assert relerr(acc, ref) < 1e-9, relerr(acc, ref)


def test_rhs_matches_oracle():
    cb, dd = make_pair((3, 3), (12, 12), 2 * math.pi * 4.1)
    srcs = [(0.5, 0.5), (0.31, 0.72)]
    assert relerr(dd.assemble_rhs(srcs), cb.global_rhs(srcs)) < 1e-12


@pytest.mark.parametrize("ncells,nels,omega", CASES[1:4])
def test_gmres_reference_mode_matches_oracle(ncells, nels, omega):
    cb, dd = make_pair(ncells, nels, omega)
    cb.factorize()
    dd.factorize()
    b = cb.global_rhs([(0.5, 0.5)])
    P = Preconditioner(cb)
    A = cb.global_operator()
    xo, log = gmres_reference(lambda v: A @ v, P.apply_global, b, tol=1e-8, max_iter=30)
    x, info = dd.gmres(b, mode=0, tol=1e-8, max_iter=30)
    assert abs(info["iterations"] - (len(log) - 1)) <= 1
    u = spla.spsolve(A.tocsc(), b)
    assert relerr(x, xo) < 1e-6
    assert relerr(x, u) < 1e-6
    assert info["true_residual"] < 1e-6


def test_gmres_lens_medium_flexible_fp32():
    vel = lens_velocity(60)
    omega = 2 * math.pi * 0.6 * 3.0
    cb, dd = make_pair((3, 3), (19, 19), omega, velocity=vel, precision=1)
    dd.factorize()
    b = dd.assemble_rhs([(0.5, 0.1)])
    A = cb.global_operator()
    x, info = dd.gmres(b, mode=1, tol=1e-8, max_iter=60, restart=30)
    u = spla.spsolve(A.tocsc(), b)
    assert info["true_residual"] < 1e-8
    assert relerr(x, u) < 1e-6


def test_precond_fp32_factorisation_close_to_fp64():
    """M^{-1} from an FP32 factorisation (FP32 arithmetic in every front: the
    register-resident fronts up to 192 rows, LDS-resident panels and
    matrix-core updates above) against the FP64 one, on cells of 60 x 60
    points (fronts of up to ~300 rows): agreement at the FP32 level of the
    preconditioner (measured 3.9e-5).  (A panel update of the pivot rows as old + (T - E) B,
    which cancels A11^{-1}'s diagonal against 1, gave 0.11 on a 3 x 3 board.)"""
    omega = 2 * math.pi * 6.0
    _, d64 = make_pair((2, 2), (59, 59), omega, precision=0)
    _, d32 = make_pair((2, 2), (59, 59), omega, precision=1)
    d64.factorize()
    d32.factorize()
    rng = np.random.default_rng(21)
    x = rng.standard_normal(d64.N) + 1j * rng.standard_normal(d64.N)
    err = relerr(d32.apply_precond(x), d64.apply_precond(x))
    print(f"FP32 vs FP64 factorisation, M^-1 rel. error {err:.3e}", flush=True)
    assert err < 1e-3
