"""Direct solve of a symmetric positive definite level system (oracle; test infrastructure only).

Restates the reference's "cholmod" solver (jspsr/core/solver.py:147-157 -> _sparse_solve.solve_cusparse,
csrc/sparse_solve/solve_cusparse.cu:15-42: cuSOLVER's sparse Cholesky, which reads the lower
triangle of A): the matrix is completed from its lower triangle, factored by a dense FP64 Cholesky
and solved by two triangular substitutions. The singularity report follows cuSOLVER: the first
pivot at or below tol (relative to its diagonal here), or -1.
"""
import numpy as np
import scipy.linalg as sla


def lower_symmetric(a):
    """Dense FP64 matrix built from the lower triangle of a (scipy sparse or dense)."""
    d = a.toarray() if hasattr(a, "toarray") else np.asarray(a, np.float64)
    low = np.tril(d)
    return low + np.tril(low, -1).T


def solve_spd(a, b, tol=1e-6):
    """x with A x = b (A from its lower triangle) and the first singular pivot (-1: none)."""
    A = lower_symmetric(a)
    n = A.shape[0]
    L = np.zeros_like(A)
    sing = -1
    for j in range(n):
        d = A[j, j] - L[j, :j] @ L[j, :j]
        if not d > tol * abs(A[j, j]):
            if sing < 0:
                sing = j
            d = abs(A[j, j]) if A[j, j] != 0 else 1.0
        L[j, j] = np.sqrt(d)
        L[j + 1:, j] = (A[j + 1:, j] - L[j + 1:, :j] @ L[j, :j]) / L[j, j]
    y = sla.solve_triangular(L, np.asarray(b, np.float64), lower=True)
    return sla.solve_triangular(L.T, y, lower=False), sing
