"""lsw_lu_solve_dev_noalloc: the local solve with no device allocation gives
the bits of lsw_lu_solve_dev, over two calls in a row (the second with a
smaller matrix and fewer right-hand sides, so the shared scratch array holds
the first call's larger contents), for FP64 and FP32 factors; sizes beyond
the scratch array fail with LSW_ERR_CAPACITY; the call leaves the device's
free memory unchanged."""
import math

import numpy as np
import pytest

from tests._common import make_pair

pytestmark = pytest.mark.gpu


def _solver(L, nc, nel, cell, precision):
    cb, _ = make_pair((nc, nc), (nel, nel), 2 * math.pi * 4.1, device=False)
    s = L.LocalLinearSolverHIP(precision)
    s.set_matrix(cb.cells[cell].A.tocsc())
    s.factorize()
    return s, cb.cells[cell].size


def _both(torch, s, n, nrhs, seed):
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((nrhs, n)) + 1j * rng.standard_normal((nrhs, n))
    dB = torch.from_numpy(B).to("cuda:0")        # row r = right-hand side r (n x nrhs column-major)
    x1 = torch.zeros_like(dB)
    x2 = torch.full_like(dB, float("nan"))
    torch.cuda.synchronize()
    s.solve_dev(nrhs, dB.data_ptr(), x1.data_ptr())
    torch.cuda.synchronize()
    s.solve_dev_noalloc(nrhs, dB.data_ptr(), x2.data_ptr())
    torch.cuda.synchronize()
    return x1.cpu().numpy(), x2.cpu().numpy()


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_equals_solve_dev_twice(precision):
    import torch
    import lsweeps_amd as L
    big, n_big = _solver(L, 2, 19, 3, precision)
    small, n_small = _solver(L, 3, 12, 4, precision)
    assert n_small < n_big
    for s, n, nrhs, seed in ((big, n_big, 6, 1), (small, n_small, 2, 2)):
        ref, got = _both(torch, s, n, nrhs, seed)
        print(f"precision {precision}: n {n} nrhs {nrhs}: max |diff| {np.abs(ref - got).max():.3e}", flush=True)
        # same plan, same kernels, fixed order of additions: identical bits
        assert np.array_equal(ref, got)


def test_noalloc_allocates_nothing_and_rejects_oversize():
    import torch
    import lsweeps_amd as L
    warm, n_w = _solver(L, 2, 19, 0, 0)
    _both(torch, warm, n_w, 3, 3)
    s, n = _solver(L, 2, 19, 1, 0)
    rng = np.random.default_rng(4)
    B = torch.from_numpy(rng.standard_normal((64, n)) + 1j * rng.standard_normal((64, n))).to("cuda:0")
    X = torch.zeros_like(B)
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    s.solve_dev_noalloc(64, B.data_ptr(), X.data_ptr())
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    # solve_dev would allocate X, R (64 x n x 16 bytes = 3.6 MB each), G and the plan
    print(f"free memory before {free0}, after {free1}", flush=True)
    assert abs(free0 - free1) < (1 << 20)
    assert np.all(np.isfinite(X.cpu().numpy()))
    # a solve whose X and R buffers alone take twice the 64 MiB array:
    # rejected on the host before any launch (the input and output buffers
    # are sized for the call all the same, so no access could leave them)
    big = 2 * (2 * (64 << 20)) // (2 * 16 * n) + 1
    Bb = torch.zeros((big, n), dtype=torch.complex128, device="cuda:0")
    torch.cuda.synchronize()
    with pytest.raises(L.LswError) as e:
        s.solve_dev_noalloc(big, Bb.data_ptr(), Bb.data_ptr())
    assert e.value.code == 5 and "scratch" in str(e.value)
