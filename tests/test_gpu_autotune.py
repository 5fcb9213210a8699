"""Launch-shape autotuning changes how the batched local solves are launched
(column splits, matrix cores), never what they compute: M^{-1} after tuning
equals M^{-1} before it, to the summation-order level of the precision
(1e-12 FP64; 1e-4 FP32, where a column split or the matrix cores reorder FP32
additions), and still matches the oracle in FP64."""
import math

import numpy as np
import pytest

from oracle.sweeps import Preconditioner
from tests._common import make_pair, relerr

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("nrhs", [1, 8])
def test_autotune_keeps_the_preconditioner(prec, nrhs):
    cb, dd = make_pair((3, 3), (12, 12), 2 * math.pi * 4.1, precision=prec)
    dd.factorize()
    rng = np.random.default_rng(21)
    X = rng.standard_normal((nrhs, dd.N)) + 1j * rng.standard_normal((nrhs, dd.N))
    before = dd.apply_precond_block(X) if nrhs > 1 else dd.apply_precond(X[0])[None, :]
    r = dd.autotune(nrhs)
    assert r["tried"] >= 6
    assert 0 < r["best_ms"] <= r["default_ms"] * 1.0001
    assert r["wide_cap"] in (2, 4, 8, 16)
    after = dd.apply_precond_block(X) if nrhs > 1 else dd.apply_precond(X[0])[None, :]
    for k in range(nrhs):
        assert relerr(after[k], before[k]) < (1e-12 if prec == 0 else 1e-4)


def test_autotuned_fp64_matches_oracle():
    cb, dd = make_pair((2, 4), (19, 9), 2 * math.pi * 4.1)
    cb.factorize()
    dd.factorize()
    dd.autotune(1)
    x = np.random.default_rng(22).standard_normal(dd.N) + 0j
    assert relerr(dd.apply_precond(x), Preconditioner(cb).apply_global(x)) < 1e-9
