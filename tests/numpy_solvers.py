"""Textbook dense solvers in numpy, written from the definitions, used to pin
the oracle's GAMG coarsest-level solve and its GAMG preconditioner.

  * normFactor: lduMatrixSolver.C:174-197 states it as
    sum |A psi - xRef sumA| + |b - xRef sumA| + 1e-20 with xRef = mean(psi);
  * the convergence test of SolverPerformance.C:31-92: final < tolerance, or
    relTol > 1e-20 and final < relTol * initial; singular: |value| < DBL_MIN;
  * preconditioned conjugate gradients (Hestenes-Stiefel with a
    preconditioner; the loop controls of PCG.C:65-193: convergence checked
    before the first iteration unless minIter > 0, then after each);
  * the diagonal incomplete Cholesky factor of Meijerink and van der Vorst
    (IC(0) kept to its diagonal): M = (D + L) D^-1 (D + L^T) with
    D_u = a_uu - sum_{l<u} a_lu^2 / D_l over the matrix's sparsity.
"""
import numpy as np

DBL_MIN = np.finfo(float).tiny


def norm_factor(A, psi, b):
    sumA = A.sum(axis=1)
    xRef = psi.mean()
    return np.abs(A @ psi - xRef * sumA).sum() + np.abs(b - xRef * sumA).sum() + 1e-20


def converged(final, initial, tol, relTol):
    return final < tol or (relTol > 1e-20 and final < relTol * initial)


def dic_dense(A):
    """M^-1 application of the DIC factor of the symmetric matrix A."""
    n = A.shape[0]
    D = np.diag(A).astype(float).copy()
    for u in range(n):
        for l in range(u):
            if A[l, u] != 0.0:
                D[u] -= A[l, u] ** 2 / D[l]
    L = np.tril(A, -1)
    M = (np.diag(D) + L) @ np.diag(1.0 / D) @ (np.diag(D) + L.T)
    return lambda r: np.linalg.solve(M, r)


def pcg(A, b, x0, prec, tol, relTol, maxIter=1000, minIter=0):
    """Returns (x, nIterations, residual history [initial, after each
    iteration])."""
    x = x0.astype(float).copy()
    r = b - A @ x
    nf = norm_factor(A, x, b)
    initial = np.abs(r).sum() / nf
    final = initial
    hist = [initial]
    n = 0
    if minIter > 0 or not converged(final, initial, tol, relTol):
        wr = None
        p = None
        while True:
            w = prec(r)
            wr_old, wr = wr, w @ r
            p = w.copy() if n == 0 else w + (wr / wr_old) * p
            q = A @ p
            wp = q @ p
            if abs(wp) / nf < DBL_MIN:
                break
            alpha = wr / wp
            x += alpha * p
            r -= alpha * q
            final = np.abs(r).sum() / nf
            hist.append(final)
            n += 1
            if not ((n < maxIter and not converged(final, initial, tol, relTol))
                    or n < minIter):
                break
    return x, n, np.array(hist)
