"""
Many independent chains evolved together on one GPU.

The reference evolves one chain per call; a study of many chains (disorder
realisations, parameter scans) calls its ``tdvp_singlesite``
(pytenet/tdvp.py:26-118) in a loop. ``tdvp_singlesite_batched`` runs the same
algorithm, step for step, on all chains at once: chains whose bond and MPO
dimensions agree form a bucket, their tensors are stacked along a leading
batch axis, and every local operation of the sweep (Krylov exponential of the
one-site and bond operators, QR, environment steps, bond absorption) is a
single batched launch sequence for the whole bucket. The buckets run on
separate streams and advance together, one site step at a time, so their
launch sequences overlap on the GPU (the local solves of small buckets are
replayed as captured graphs). Each chain gets exactly the result of its own
tdvp_singlesite call (up to rounding).
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
    """Stacked device state of chains with identical shapes, with its own HIP
    stream: the buckets of a call advance site step by site step together,
    their launch sequences overlapping on the GPU."""

    def __init__(self, hamiltonians, psis, numiter, precision):
        self.psis = psis
        self.n = psis[0].nsites
        self.nb = len(psis)
        self.numiter = int(numiter)
        self.precision = precision
        self.stream = torch.cuda.Stream()      # default priority
        self.stream.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(self.stream):
            for p in psis:
                for a in p.a:
                    a.record_stream(self.stream)
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

    # tdvp.py:26-118, split into the site steps of one time step
    def setup(self):
        n, nb, W, prec = self.n, self.nb, self.W, self.precision
        A = [torch.stack([p.a[i] for p in self.psis]) for i in range(n)]
        self.nrms = _orthonormalize_right(A)
        one = torch.ones((nb, 1, 1, 1), dtype=torch.complex128, device=A[0].device)
        self.rblocks = [None] * n
        self.rblocks[n - 1] = one
        for i in reversed(range(n - 1)):
            self.rblocks[i] = backend.env_right_batched(A[i + 1], W[i + 1], self.rblocks[i + 1],
                                                        prec)
        self.lblocks = [None] * n
        self.lblocks[0] = one
        self.A = A

    def left_step(self, i, dt):
        A, W, nb = self.A, self.W, self.nb
        lb, rb = self.lblocks, self.rblocks
        A[i] = self.site_expm(A[i], lb[i], W[i], rb[i], -0.5 * dt)
        _, Dl, d, Dr = A[i].shape
        q, c = backend.qr_batched(A[i].reshape(nb, Dl * d, Dr))
        A[i] = q.reshape(nb, Dl, d, q.shape[2])
        lb[i + 1] = backend.env_left_batched(A[i], W[i], lb[i], self.precision)
        c = self.bond_expm(c, lb[i + 1], rb[i], 0.5 * dt)
        A[i + 1] = _absorb_left(c, A[i + 1])

    def last_step(self, dt):
        i = self.n - 1
        self.A[i] = self.site_expm(self.A[i], self.lblocks[i], self.W[i], self.rblocks[i], -dt)

    def right_step(self, i, dt):
        A, W, nb = self.A, self.W, self.nb
        lb, rb = self.lblocks, self.rblocks
        af = _flip(A[i])
        _, s0, s1, s2 = af.shape
        q, c = backend.qr_batched(af.reshape(nb, s0 * s1, s2))
        A[i] = _flip(q.reshape(nb, s0, s1, q.shape[2]))
        rb[i - 1] = backend.env_right_batched(A[i], W[i], rb[i], self.precision)
        c = c.transpose(1, 2).contiguous()
        c = self.bond_expm(c, lb[i], rb[i - 1], 0.5 * dt)
        A[i - 1] = _absorb_right(A[i - 1], c)
        A[i - 1] = self.site_expm(A[i - 1], lb[i - 1], W[i - 1], rb[i - 1], -0.5 * dt)

    def finish(self):
        for b, psi in enumerate(self.psis):
            psi.a = [self.A[i][b].contiguous() for i in range(self.n)]


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


def _run_group(bks, dt, numsteps):
    """tdvp.py:26-118 on buckets of chains of one length: every site step is
    issued for all buckets (each on its own stream) before the next one."""
    def on(bk, fn, *args):
        with torch.cuda.stream(bk.stream):
            fn(*args)
    for bk in bks:
        on(bk, bk.setup)
    n = bks[0].n
    for _ in range(numsteps):
        for i in range(n - 1):
            for bk in bks:
                on(bk, bk.left_step, i, dt)
        for bk in bks:
            on(bk, bk.last_step, dt)
        for i in reversed(range(1, n)):
            for bk in bks:
                on(bk, bk.right_step, i, dt)
        for bk in bks:
            on(bk, bk.flush)
    main = torch.cuda.current_stream()
    for bk in bks:
        on(bk, bk.finish)
        main.wait_stream(bk.stream)


def tdvp_singlesite_batched(hamiltonians, psis, dt, numsteps: int, numiter_lanczos: int = 25,
                            precision: str = "fp64", autotune: bool = True):
    """Symmetric single-site TDVP (tdvp.py:26-118) of many independent chains.

    `hamiltonians` and `psis` are equally long lists of MPO / MPS; chain k is
    evolved under hamiltonians[k] exactly as tdvp_singlesite(hamiltonians[k],
    psis[k], dt, numsteps, numiter_lanczos) would, and psis[k] is overwritten.
    Chains with equal bond and MPO dimensions form a bucket (stacked tensors,
    one batched launch sequence per local operation); the buckets of chains of
    one length run on separate streams and advance one site step at a time
    together. With `autotune` (default) the tile autotuner of the complex
    products picks the tile shape per product shape, i.e. per (D, d, M)
    bucket. Chains must carry zero quantum numbers (dense tensors). Returns
    the list of initial norms (the return value of each per-chain call)."""
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
    prev = backend.set_autotune(autotune)
    try:
        groups = {}
        for idx in buckets.values():
            bk = _Bucket([hamiltonians[k] for k in idx], [psis[k] for k in idx], numiter_lanczos,
                         precision)
            bk.idx = idx
            groups.setdefault(bk.n, []).append(bk)
        for bks in groups.values():
            _run_group(bks, dt, numsteps)
            for bk in bks:
                for k, nrm in zip(bk.idx, bk.nrms):
                    nrms[k] = nrm
    finally:
        backend.set_autotune(prev)
    for k, psi in enumerate(psis):
        if host[k]:
            psi.to_host()
    return nrms
