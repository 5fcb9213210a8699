"""Pin the GMRES restatements on small dense systems."""
import numpy as np

from oracle.gmres import fgmres, gmres_reference


def _sys(n=40, seed=0):
    rng = np.random.default_rng(seed)
    A = np.eye(n) * 4 + (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    return A, b


def test_reference_mode_exact_preconditioner_returns_at_iteration_zero():
    A, b = _sys()
    Ainv = np.linalg.inv(A)
    x, log = gmres_reference(lambda v: A @ v, lambda v: Ainv @ v, b, tol=1e-8, max_iter=30)
    assert len(log) == 1 and log[0] < 1e-12
    assert np.linalg.norm(x - np.linalg.solve(A, b)) / np.linalg.norm(x) < 1e-12


def test_reference_mode_first_estimate_is_preconditioned_residual():
    A, b = _sys(seed=1)
    D = np.diag(1 / np.diag(A))
    x, log = gmres_reference(lambda v: A @ v, lambda v: D @ v, b, tol=1e-30, max_iter=1)
    x0 = D @ b
    r0 = D @ b - D @ (A @ x0)
    assert abs(log[0] - np.linalg.norm(r0) / np.linalg.norm(D @ b)) < 1e-14
    # after one step the estimate is |s0| * ||r0|| / ||M^-1 b|| with the
    # complex-symmetric rotation of (h00, h10)
    w = D @ (A @ (r0 / np.linalg.norm(r0)))
    q0 = r0 / np.linalg.norm(r0)
    h0 = np.vdot(w, q0)                    # conj(w) . q0
    h1 = np.linalg.norm(w - h0 * q0)
    t = np.sqrt(h0 ** 2 + h1 ** 2)
    c = abs(h0) / t
    s = c * h1 / h0
    assert abs(log[1] - abs(s) * np.linalg.norm(r0) / np.linalg.norm(D @ b)) < 1e-12


def test_flexible_gmres_solves_dense_system():
    A, b = _sys(seed=2)
    D = np.diag(1 / np.diag(A))
    x, it, hist = fgmres(lambda v: A @ v, lambda v: D @ v, b, tol=1e-12, restart=10, max_iter=200)
    assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) < 1e-11
    assert hist[-1] <= 1e-12


def test_reference_mode_skips_the_previous_rotation():
    """GlobalIterativeSolverGMRES applies only rotations 0..k-2 to column k
    (src/GlobalIterativeSolverGMRES.cpp:315-320), so at step 1 column 1 is
    rotated by nothing.  Worked by hand for A = diag(1, 2, 3, 4), M = I,
    b = (1, 1, 1, 1), where everything is real:
      x0 = b, r0 = b - A b = (0, -1, -2, -3), ||M^-1 b|| = 2;
      h00 = 25/7, h10 = sqrt(19)/7, s0 = sqrt(19/644): log[1] = s0 sqrt(14) / 2;
      q1 = (0, 11, 8, -9)/sqrt(266): h01 = 38/sqrt(3724), h11 = 758/266,
      h21^2 = 2356/266 - h01^2 - h11^2;
      log[2] = log[1] |s1| with |s1| = h21 / sqrt(h11^2 + h21^2) from the
      unrotated (h11, h21).
    Rotating column 1 by rotation 0 first (textbook GMRES) gives a log[2]
    5% larger."""
    A = np.diag([1.0, 2.0, 3.0, 4.0]).astype(complex)
    b = np.ones(4, dtype=complex)
    x, log = gmres_reference(lambda v: A @ v, lambda v: v, b, tol=1e-30, max_iter=2)
    assert abs(log[0] - np.sqrt(14.0) / 2) < 1e-15
    s0 = np.sqrt(19.0 / 644.0)
    c0 = 25.0 / np.sqrt(644.0)
    assert abs(log[1] - s0 * np.sqrt(14.0) / 2) < 1e-14
    h01 = 38.0 / np.sqrt(3724.0)
    h11 = 758.0 / 266.0
    h21 = np.sqrt(2356.0 / 266.0 - h01 ** 2 - h11 ** 2)
    expected = log[1] * h21 / np.sqrt(h11 ** 2 + h21 ** 2)
    assert abs(log[2] - expected) < 1e-13
    h11_rot = -s0 * h01 + c0 * h11
    textbook = log[1] * h21 / np.sqrt(h11_rot ** 2 + h21 ** 2)
    assert abs(textbook - expected) > 0.04 * expected
