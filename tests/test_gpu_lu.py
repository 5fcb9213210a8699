"""LocalLinearSolverHIP against SuperLU (scipy splu) and dense solves."""
import math

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from tests._common import make_pair, relerr

pytestmark = pytest.mark.gpu


def _local_matrix(nc=2, nel=19, omega=2 * math.pi * 4.1, cell=0):
    cb, _ = make_pair((nc, nc), (nel, nel), omega, device=False)
    return cb.cells[cell].A.tocsc()


@pytest.mark.parametrize("cell", [0, 3])
def test_lu_fp64_matches_superlu(cell):
    import lsweeps_amd as L
    A = _local_matrix(cell=cell)
    rng = np.random.default_rng(1)
    b = rng.standard_normal(A.shape[0]) + 1j * rng.standard_normal(A.shape[0])
    ref = spla.splu(A, permc_spec="COLAMD").solve(b)
    s = L.LocalLinearSolverHIP(L.LSW_PREC_FP64)
    s.set_matrix(A)
    s.factorize()
    x = s.solve(b)
    assert relerr(x, ref) < 1e-11
    # the fronts keep explicit inverses (F11^{-1}), which are backward stable
    # to a small multiple of cond(F11) eps, not to the last bit; the corner
    # cell (PML on two sides) measured 1.8e-12 with the blocked factorisation
    assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) < 1e-11


def test_lu_multi_rhs_and_fp32():
    import lsweeps_amd as L
    A = _local_matrix(nc=1, nel=30)
    rng = np.random.default_rng(2)
    B = rng.standard_normal((A.shape[0], 6)) + 1j * rng.standard_normal((A.shape[0], 6))
    ref = spla.splu(A).solve(B)
    s = L.LocalLinearSolverHIP(L.LSW_PREC_FP64)
    s.set_matrix(A)
    s.factorize()
    X = s.solve(B)
    # this all-PML 59x59 grid is ill-conditioned, and the front factors hold
    # explicit inverses (F11^{-1}), which are not backward stable to the last
    # bit: allow a backward error of 1e-9 and forward differences at the
    # conditioning level
    assert np.linalg.norm(A @ X - B) / np.linalg.norm(B) < 1e-9
    assert relerr(X, ref) < 1e-8
    # one right-hand side gives the column of the block up to the order of
    # the sums: the launch shapes (how a row's columns are split over lanes)
    # depend on the block width through the LDS budget and prefetch depth
    for k in range(B.shape[1]):
        assert relerr(s.solve(B[:, k]), X[:, k]) < 1e-12
    # FP32 factors (FP32 arithmetic in every front, FP32 storage and
    # solves): measured forward error 3.4e-5 and normwise backward error
    # 1.5e-7 on this grid
    s32 = L.LocalLinearSolverHIP(L.LSW_PREC_FP32)
    s32.set_matrix(A)
    s32.factorize()
    X32 = s32.solve(B)
    bwd = np.linalg.norm(A @ X32 - B) / (spla.norm(A) * np.linalg.norm(X32) + np.linalg.norm(B))
    print(f"FP32 all-PML grid: forward {relerr(X32, ref):.3e}, normwise backward {bwd:.3e}", flush=True)
    assert relerr(X32, ref) < 5e-4 and bwd < 2e-6


def test_lu_fp32_interior_cell():
    import lsweeps_amd as L
    A = _local_matrix(nc=3, nel=19, cell=4)
    rng = np.random.default_rng(4)
    B = rng.standard_normal((A.shape[0], 3)) + 1j * rng.standard_normal((A.shape[0], 3))
    ref = spla.splu(A).solve(B)
    s32 = L.LocalLinearSolverHIP(L.LSW_PREC_FP32)
    s32.set_matrix(A)
    s32.factorize()
    X32 = s32.solve(B)
    # FP32 factorisation, storage and solves: backward error at the FP32
    # level times the front growth
    assert np.linalg.norm(A @ X32 - B) / np.linalg.norm(B) < 1e-4
    assert relerr(X32, ref) < 1e-2


def test_lu_small_dense_bruteforce():
    import lsweeps_amd as L
    import scipy.sparse as sp
    # 5x7 grid with random complex stencil (diagonally shifted to be nonsingular)
    nx, ny = 5, 7
    N = nx * ny
    rng = np.random.default_rng(3)
    rows, cols, vals = [], [], []
    for i in range(nx):
        for j in range(ny):
            d = i * ny + j
            for di, dj in ((0, 0), (-1, 0), (1, 0), (0, -1), (0, 1)):
                ii, jj = i + di, j + dj
                if 0 <= ii < nx and 0 <= jj < ny:
                    v = rng.standard_normal() + 1j * rng.standard_normal()
                    if di == 0 and dj == 0:
                        v += 3.0
                    rows.append(d)
                    cols.append(ii * ny + jj)
                    vals.append(v)
    A = sp.csc_matrix((vals, (rows, cols)), shape=(N, N))
    b = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    ref = np.linalg.solve(A.toarray(), b)
    s = L.LocalLinearSolverHIP()
    s.set_matrix(A)
    s.factorize()
    info = s.info()
    assert (info["nx"], info["ny"]) == (nx, ny)
    assert relerr(s.solve(b), ref) < 1e-12


def test_lu_solve_dev_matches_host_and_superlu():
    """lsw_lu_solve_dev on device buffers (n x nrhs column-major, on a caller
    stream and in place) gives the host entry point's bits and SuperLU's
    solution."""
    import torch
    import lsweeps_amd as L
    A = _local_matrix(nc=2, nel=19, cell=1)
    n = A.shape[0]
    rng = np.random.default_rng(5)
    B = rng.standard_normal((n, 5)) + 1j * rng.standard_normal((n, 5))
    ref = spla.splu(A, permc_spec="COLAMD").solve(B)
    s = L.LocalLinearSolverHIP(L.LSW_PREC_FP64)
    s.set_matrix(A)
    s.factorize()
    Xh = s.solve(B)
    # column-major n x nrhs == row-major (nrhs, n)
    dB = torch.from_numpy(np.ascontiguousarray(B.T)).to("cuda:0")
    dX = torch.zeros_like(dB)
    st = torch.cuda.Stream()
    torch.cuda.synchronize()
    s.solve_dev(5, dB.data_ptr(), dX.data_ptr(), stream=st.cuda_stream)
    st.synchronize()
    Xd = dX.cpu().numpy().T
    assert np.array_equal(Xd, Xh)
    assert relerr(Xd, ref) < 1e-11
    # in place, on the solver's own stream, fewer right-hand sides
    dB2 = dB[:2].clone()
    torch.cuda.synchronize()
    s.solve_dev(2, dB2.data_ptr(), dB2.data_ptr())
    torch.cuda.synchronize()
    s_host2 = s.solve(B[:, :2])
    assert relerr(dB2.cpu().numpy().T, s_host2) < 1e-14
