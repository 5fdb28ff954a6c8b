"""
Many independent chains evolved together on one GPU.

The reference evolves one chain per call; a study of many chains (disorder
realisations, parameter scans) calls its ``tdvp_singlesite``
(pytenet/tdvp.py:26-118) in a loop. ``tdvp_singlesite_batched`` runs the same
algorithm, step for step, on all chains at once: chains whose bond and MPO
dimensions agree form a bucket, their tensors are stacked along a leading
batch axis, and every local operation of the sweep (Krylov exponential of the
one-site and bond operators, QR, environment steps, bond absorption) is a
single batched launch sequence for the whole bucket. Each chain gets exactly
the result of its own tdvp_singlesite call (up to rounding).
"""

import numpy as np
import torch

from . import backend, _lib

__all__ = ["tdvp_singlesite_batched"]


def _flip(a):
    """Swap the virtual bonds of stacked MPS tensors (B, Dl, d, Dr) -> (B, Dr, d, Dl)."""
    return a.permute(0, 3, 2, 1).contiguous()


def _absorb_left(c, a):
    """c (B, k, D) x a (B, D, d, Dr) -> (B, k, d, Dr) (tdvp.py:84)."""
    nb, D, d, Dr = a.shape
    return backend.matmul_batched(c, a.reshape(nb, D, d * Dr)).reshape(nb, c.shape[1], d, Dr)


def _absorb_right(a, c):
    """a (B, Dl, d, D) x c (B, D, k) -> (B, Dl, d, k) (tdvp.py:112)."""
    nb, Dl, d, D = a.shape
    return backend.matmul_batched(a.reshape(nb, Dl * d, D), c).reshape(nb, Dl, d, c.shape[2])


class _Bucket:
    """Stacked device state of chains with identical shapes."""

    def __init__(self, hamiltonians, psis, numiter, precision):
        self.psis = psis
        self.n = psis[0].nsites
        self.nb = len(psis)
        self.numiter = int(numiter)
        self.precision = precision
        self.W = [torch.stack([backend.as_dev(H.a[i]) for H in hamiltonians])
                  for i in range(self.n)]
        self.status = []

    def site_expm(self, a, l, w, r, dt):
        out, _, info = backend.krylov_local_batched(_lib.PTN_OP_HEFF, a, l, w, r, self.numiter,
                                                    "expm", dt, self.precision)
        self.status.append(info)
        return out

    def bond_expm(self, c, l, r, dt):
        out, _, info = backend.krylov_local_batched(_lib.PTN_OP_BOND, c, l, None, r,
                                                    self.numiter, "expm", dt, self.precision)
        self.status.append(info)
        return out

    def flush(self):
        if self.status:
            rows = torch.cat(self.status).cpu().numpy()
            self.status = []
            backend.report_krylov_status(rows)


def _orthonormalize_right(A):
    """MPS.orthonormalize(mode="right") (mps.py:162-178) of stacked chains:
    batched QR sweep from the right; returns the per-chain norms (sign >= 0)."""
    n = len(A)
    for i in reversed(range(n)):
        af = _flip(A[i])
        nb, s0, s1, s2 = af.shape
        q, r = backend.qr_batched(af.reshape(nb, s0 * s1, s2))
        A[i] = _flip(q.reshape(nb, s0, s1, q.shape[2]))
        if i > 0:
            A[i - 1] = _absorb_right(A[i - 1], r.transpose(1, 2).contiguous())
    assert tuple(r.shape[1:]) == (1, 1)
    nrm = r[:, 0, 0].real.cpu().numpy()
    sign = torch.as_tensor(np.where(nrm < 0, -1.0, 1.0), dtype=torch.complex128,
                           device=A[0].device)
    A[0] = A[0] * sign.reshape(-1, 1, 1, 1)
    return [float(abs(x)) for x in nrm]


def _run_bucket(bk, dt, numsteps):
    """tdvp.py:26-118 on a bucket of stacked chains; returns the initial norms."""
    n, nb, W, prec = bk.n, bk.nb, bk.W, bk.precision
    A = [torch.stack([p.a[i] for p in bk.psis]) for i in range(n)]
    nrms = _orthonormalize_right(A)
    one = torch.ones((nb, 1, 1, 1), dtype=torch.complex128, device=A[0].device)
    rblocks = [None] * n
    rblocks[n - 1] = one
    for i in reversed(range(n - 1)):
        rblocks[i] = backend.env_right_batched(A[i + 1], W[i + 1], rblocks[i + 1], prec)
    lblocks = [None] * n
    lblocks[0] = one
    for _ in range(numsteps):
        # left to right
        for i in range(n - 1):
            A[i] = bk.site_expm(A[i], lblocks[i], W[i], rblocks[i], -0.5 * dt)
            _, Dl, d, Dr = A[i].shape
            q, c = backend.qr_batched(A[i].reshape(nb, Dl * d, Dr))
            A[i] = q.reshape(nb, Dl, d, q.shape[2])
            lblocks[i + 1] = backend.env_left_batched(A[i], W[i], lblocks[i], prec)
            c = bk.bond_expm(c, lblocks[i + 1], rblocks[i], 0.5 * dt)
            A[i + 1] = _absorb_left(c, A[i + 1])
        i = n - 1
        A[i] = bk.site_expm(A[i], lblocks[i], W[i], rblocks[i], -dt)
        # right to left
        for i in reversed(range(1, n)):
            af = _flip(A[i])
            _, s0, s1, s2 = af.shape
            q, c = backend.qr_batched(af.reshape(nb, s0 * s1, s2))
            A[i] = _flip(q.reshape(nb, s0, s1, q.shape[2]))
            rblocks[i - 1] = backend.env_right_batched(A[i], W[i], rblocks[i], prec)
            c = c.transpose(1, 2).contiguous()
            c = bk.bond_expm(c, lblocks[i], rblocks[i - 1], 0.5 * dt)
            A[i - 1] = _absorb_right(A[i - 1], c)
            A[i - 1] = bk.site_expm(A[i - 1], lblocks[i - 1], W[i - 1], rblocks[i - 1], -0.5 * dt)
        bk.flush()
    for b, psi in enumerate(bk.psis):
        psi.a = [A[i][b].contiguous() for i in range(n)]
    return nrms


def tdvp_singlesite_batched(hamiltonians, psis, dt, numsteps: int, numiter_lanczos: int = 25,
                            precision: str = "fp64"):
    """Symmetric single-site TDVP (tdvp.py:26-118) of many independent chains.

    `hamiltonians` and `psis` are equally long lists of MPO / MPS; chain k is
    evolved under hamiltonians[k] exactly as tdvp_singlesite(hamiltonians[k],
    psis[k], dt, numsteps, numiter_lanczos) would, and psis[k] is overwritten.
    Chains must carry zero quantum numbers (dense tensors). Returns the list of
    initial norms (the return value of each per-chain call)."""
    assert len(hamiltonians) == len(psis)
    backend.precision_code(precision)
    buckets = {}
    for k, (H, psi) in enumerate(zip(hamiltonians, psis)):
        assert H.nsites == psi.nsites
        assert not np.any(np.asarray(psi.qsite)) and all(not np.any(np.asarray(q)) for q in psi.qbonds), \
            "tdvp_singlesite_batched requires zero quantum numbers"
        key = (tuple(psi.bond_dims), tuple(H.bond_dims), len(psi.qsite))
        buckets.setdefault(key, []).append(k)
    nrms = [None] * len(psis)
    host = [not psi.on_device for psi in psis]
    for psi in psis:
        psi.to_device()
    for idx in buckets.values():
        bk = _Bucket([hamiltonians[k] for k in idx], [psis[k] for k in idx], numiter_lanczos,
                     precision)
        for k, nrm in zip(idx, _run_bucket(bk, dt, numsteps)):
            nrms[k] = nrm
    for k, psi in enumerate(psis):
        if host[k]:
            psi.to_host()
    return nrms
