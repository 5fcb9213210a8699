"""Oracle: GMRES for the preconditioned system.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

``gmres_reference`` restates GlobalIterativeSolverGMRES::solve/arnoldi/
apply_givens_rotation/givens_rotation (src/GlobalIterativeSolverGMRES.cpp:
65-348) including its behaviours:

* x0 = M^{-1} b (:75-78);
* left preconditioning, residual and stop test relative to ||M^{-1} b||;
* no restart, at most MaxIter steps (:134-190);
* Arnoldi coefficient h_i = sum_j conj(w_j) q_ij (Eigen ``w.dot(q_i)``, :282),
  modified Gram-Schmidt;
* at step k only the rotations 0..k-2 are applied to the new column (:315-320);
* complex-symmetric rotation t = sqrt(v1^2+v2^2), c = |v1|/t, s = c v2/v1,
  (c, s) = (0, 1) when v1 == 0 (:333-348);
* beta_{k+1} = -s_k beta_k, beta_k *= c_k, logged |beta_{k+1}|/||M^{-1}b||;
* y from the leading (k+1)x(k+1) block of the rotated H (colPivHouseholderQr
  of an upper-triangular block; here numpy lstsq, identical when nonsingular).

``fgmres`` is a textbook restarted flexible GMRES (right preconditioning,
Hermitian modified Gram-Schmidt, unitary Givens) whose stop test is the true
relative residual estimate; the mixed-precision path uses it.

Pinning: tests/test_oracle_gmres.py checks both against a dense solve on
small random systems (identity preconditioner: fgmres must reach tolerance,
the reference mode reproduces hand-computed first steps), and checks that
reference mode on a 1x1 board returns at iteration 0 with x = A^{-1} b.
"""
import numpy as np


def _givens_reference(v1, v2):
    if v1 == 0.0:
        return 0.0 + 0.0j, 1.0 + 0.0j
    t = (v1 ** 2.0 + v2 ** 2.0) ** 0.5
    c = abs(v1) / t
    s = c * v2 / v1
    return c, s


def gmres_reference(apply_A, apply_M, b, tol=1e-6, max_iter=40, callback=None):
    """Reference-mode GMRES on flat complex vectors.

    Returns (x, log) where log[k] is the logged relative residual estimate.
    """
    b = np.asarray(b, dtype=np.complex128)
    rhs_prec = apply_M(b)
    x = rhs_prec.copy()
    r = rhs_prec - apply_M(apply_A(x))
    rhs_norm = np.linalg.norm(rhs_prec)
    r_norm = np.linalg.norm(r)
    log = [r_norm / rhs_norm]
    if callback:
        callback(0, log[-1])
    if log[0] < tol or max_iter == 0:
        return x, log
    Q = [r / r_norm]
    beta = [complex(r_norm)]
    cs, sn = [], []
    Hcols = []
    lastk = 0
    for k in range(max_iter):
        w = apply_M(apply_A(Q[k]))
        h = np.zeros(k + 2, dtype=np.complex128)
        for i in range(k + 1):
            h[i] = np.vdot(w, Q[i])
            w = w - h[i] * Q[i]
        h[k + 1] = np.linalg.norm(w)
        q = w / h[k + 1]
        for i in range(1, k):
            tmp = cs[i - 1] * h[i - 1] + sn[i - 1] * h[i]
            h[i] = -sn[i - 1] * h[i - 1] + cs[i - 1] * h[i]
            h[i - 1] = tmp
        c, s = _givens_reference(h[k], h[k + 1])
        h[k] = c * h[k] + s * h[k + 1]
        h[k + 1] = 0.0
        cs.append(c)
        sn.append(s)
        Hcols.append(h)
        Q.append(q)
        beta.append(-sn[k] * beta[k])
        beta[k] = beta[k] * cs[k]
        log.append(abs(beta[-1]) / rhs_norm)
        if callback:
            callback(k + 1, log[-1])
        if log[-1] <= tol or k == max_iter - 1:
            lastk = k
            break
    k = lastk
    H = np.zeros((k + 1, k + 1), dtype=np.complex128)
    for j in range(k + 1):
        H[:j + 2 if j + 2 <= k + 1 else k + 1, j] = Hcols[j][:min(j + 2, k + 1)]
    y = np.linalg.lstsq(H, np.array(beta[:k + 1]), rcond=None)[0]
    for i in range(k + 1):
        x = x + Q[i] * y[i]
    return x, log


def fgmres(apply_A, apply_M, b, tol=1e-6, restart=30, max_iter=300, x0=None):
    """Flexible right-preconditioned GMRES(restart).

    Returns (x, iterations, history of estimated true relative residuals).
    """
    b = np.asarray(b, dtype=np.complex128)
    bn = np.linalg.norm(b)
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.complex128)
    hist = []
    it = 0
    while it < max_iter:
        r = b - apply_A(x)
        beta = np.linalg.norm(r)
        hist.append(beta / bn)
        if beta / bn <= tol:
            return x, it, hist
        m = restart
        V = [r / beta]
        Z = []
        H = np.zeros((m + 1, m), dtype=np.complex128)
        cs = np.zeros(m, dtype=np.complex128)
        sn = np.zeros(m, dtype=np.complex128)
        g = np.zeros(m + 1, dtype=np.complex128)
        g[0] = beta
        j_end = 0
        for j in range(m):
            z = apply_M(V[j])
            Z.append(z)
            w = apply_A(z)
            for i in range(j + 1):
                H[i, j] = np.vdot(V[i], w)
                w = w - H[i, j] * V[i]
            H[j + 1, j] = np.linalg.norm(w)
            V.append(w / H[j + 1, j] if H[j + 1, j] != 0 else w)
            for i in range(j):
                t = np.conj(cs[i]) * H[i, j] + np.conj(sn[i]) * H[i + 1, j]
                H[i + 1, j] = -sn[i] * H[i, j] + cs[i] * H[i + 1, j]
                H[i, j] = t
            a, bb = H[j, j], H[j + 1, j]
            den = np.sqrt(abs(a) ** 2 + abs(bb) ** 2)
            cs[j] = a / den
            sn[j] = bb / den
            H[j, j] = np.conj(cs[j]) * a + np.conj(sn[j]) * bb
            H[j + 1, j] = 0.0
            g[j + 1] = -sn[j] * g[j]
            g[j] = np.conj(cs[j]) * g[j]
            it += 1
            j_end = j + 1
            hist.append(abs(g[j + 1]) / bn)
            if abs(g[j + 1]) / bn <= tol or it >= max_iter:
                break
        y = np.linalg.solve(np.triu(H[:j_end, :j_end]), g[:j_end])
        for i in range(j_end):
            x = x + Z[i] * y[i]
    return x, it, hist
