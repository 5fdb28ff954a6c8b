"""
Krylov subspace algorithms (reference pytenet/krylov.py).

Two layers:
  * the sweep drivers call the fused device Lanczos of the library
    (``backend.krylov_local``): matvec, recurrence, tridiagonal eigensolver and
    the final combination run as gfx950 kernels with FP64 scalars in device
    memory, so a local solve needs no host round trip;
  * the functions below keep the reference's callback API
    (``afunc(x) -> A x``) for arbitrary Hermitian operators. Their Krylov
    vectors live on the GPU; `afunc` receives and returns the kind of array
    the caller passed (NumPy arrays are exchanged with the host around each
    call). Arnoldi (the non-Hermitian branch) is not on the sweep path and is
    not provided.
"""

import warnings
import numpy as np
import torch

from . import backend

__all__ = ["lanczos_iteration", "eigh_krylov", "expm_krylov", "eigh_tridiag"]


def _apply(afunc, x, host):
    y = afunc(backend.to_host(x) if host else x)
    return backend.as_dev(y).reshape(-1)


def lanczos_iteration(afunc, vstart, numiter):
    """Matrix-free Lanczos without reorthogonalisation (krylov.py:12-57).
    Returns (alpha, beta, v) with v of shape (n, numiter); a breakdown
    (beta_j < 100 n eps) warns and truncates the Krylov space."""
    host = isinstance(vstart, np.ndarray) or not isinstance(vstart, torch.Tensor)
    v0 = backend.as_dev(vstart).reshape(-1)
    n = v0.numel()
    nrmv = float(torch.linalg.vector_norm(v0).item())
    assert nrmv > 0
    alpha = np.zeros(numiter)
    beta = np.zeros(numiter - 1)
    v = torch.zeros((numiter, n), dtype=torch.complex128, device=v0.device)
    v[0] = v0 / nrmv
    for j in range(numiter - 1):
        w = _apply(afunc, v[j], host)
        alpha[j] = float(torch.vdot(w, v[j]).real.item())
        w = w - (alpha[j] * v[j] + (beta[j - 1] * v[j - 1] if j > 0 else 0))
        beta[j] = float(torch.linalg.vector_norm(w).item())
        if beta[j] < 100 * n * np.finfo(float).eps:
            warnings.warn(f"beta[{j}] ~= 0 encountered during Lanczos iteration.",
                          RuntimeWarning)
            m = j + 1
            vt = v[:m, :].T
            return alpha[:m], beta[:m - 1], (backend.to_host(vt) if host else vt)
        v[j + 1] = w / beta[j]
    j = numiter - 1
    w = _apply(afunc, v[j], host)
    alpha[j] = float(torch.vdot(w, v[j]).real.item())
    vt = v.T
    return alpha, beta, (backend.to_host(vt) if host else vt)


def eigh_tridiag(d, e):
    """Eigendecomposition of the small symmetric tridiagonal matrix (krylov.py:142-150).
    The sweep drivers use the device QL solver instead; this helper serves the
    callback API."""
    t = np.diag(d) + np.diag(e, k=1) + np.diag(e, k=-1)
    return np.linalg.eigh(t)


def _combine(v, c):
    if isinstance(v, np.ndarray):
        return v @ c
    ct = torch.as_tensor(np.asarray(c, dtype=np.complex128), device=v.device)
    return v @ ct


def eigh_krylov(afunc, vstart, numiter, numeig):
    """Lowest Ritz pairs (krylov.py:110-119)."""
    alpha, beta, v = lanczos_iteration(afunc, vstart, numiter)
    w, u = eigh_tridiag(alpha, beta)
    return w[:numeig], _combine(v, u[:, :numeig])


def expm_krylov(afunc, vec, dt, numiter, hermitian=False):
    """Krylov approximation of expm(dt*A) vec for Hermitian A (krylov.py:122-136).
    The Arnoldi branch of the reference (hermitian=False, krylov.py:137-139)
    is not on the sweep path (the drivers always pass hermitian=True) and is
    not provided."""
    if not hermitian:
        raise NotImplementedError("expm_krylov: only the Hermitian (Lanczos) branch is provided")
    nrm = float(np.linalg.norm(backend.to_host(vec)))
    alpha, beta, v = lanczos_iteration(afunc, vec, numiter)
    w, u = eigh_tridiag(alpha, beta)
    return _combine(v, u @ (nrm * np.exp(dt * w) * u[0]))
