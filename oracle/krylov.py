"""
Oracle (test infrastructure): Hermitian Lanczos, Krylov eigensolver and
Krylov matrix exponential. Restates pytenet/krylov.py:12-57,110-150
(the Arnoldi branch, krylov.py:60-107,137-139, is never reached by the sweep
drivers, which always pass hermitian=True).
"""

import warnings
import numpy as np


def lanczos_iteration(afunc, vstart, numiter):
    """Plain three-term Lanczos without reorthogonalisation (krylov.py:12-57).
    Breakdown when beta_j < 100 * n * eps: warn and truncate to j+1 vectors.
    Returns (alpha, beta, V) with V of shape (n, numiter)."""
    vstart = np.asarray(vstart)
    nrmv = np.linalg.norm(vstart)
    assert nrmv > 0
    vstart = vstart / nrmv
    alpha = np.zeros(numiter)
    beta = np.zeros(numiter - 1)
    v = np.zeros((numiter, len(vstart)), dtype=vstart.dtype)
    v[0] = vstart
    for j in range(numiter - 1):
        w = afunc(v[j])
        alpha[j] = np.vdot(w, v[j]).real
        w -= alpha[j] * v[j] + (beta[j-1] * v[j-1] if j > 0 else 0)
        beta[j] = np.linalg.norm(w)
        if beta[j] < 100 * len(vstart) * np.finfo(float).eps:
            warnings.warn(f"beta[{j}] ~= 0 encountered during Lanczos iteration.",
                          RuntimeWarning)
            return alpha[:j+1], beta[:j], v[:j+1, :].T
        v[j+1] = w / beta[j]
    j = numiter - 1
    w = afunc(v[j])
    alpha[j] = np.vdot(w, v[j]).real
    return alpha, beta, v.T


def eigh_tridiag(d, e):
    """Dense eigh of the symmetric tridiagonal matrix (krylov.py:142-150)."""
    t = np.diag(d) + np.diag(e, k=1) + np.diag(e, k=-1)
    return np.linalg.eigh(t)


def eigh_krylov(afunc, vstart, numiter, numeig):
    """Lowest `numeig` Ritz pairs (krylov.py:110-119)."""
    alpha, beta, v = lanczos_iteration(afunc, vstart, numiter)
    w, u = eigh_tridiag(alpha, beta)
    return w[:numeig], v @ u[:, :numeig]


def expm_krylov(afunc, vec, dt, numiter, hermitian=True):
    """expm(dt*A) vec from the Lanczos decomposition (krylov.py:122-136):
    V (U (|vec| exp(dt lambda) * U[0,:]))."""
    if not hermitian:
        raise NotImplementedError("the sweep drivers only use the Hermitian branch")
    alpha, beta, v = lanczos_iteration(afunc, vec, numiter)
    w, u = eigh_tridiag(alpha, beta)
    return v @ (u @ (np.linalg.norm(vec) * np.exp(dt * w) * u[0]))
