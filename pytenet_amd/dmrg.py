"""
DMRG ground-state search (reference pytenet/dmrg.py:22-189), with the site
tensors, environment blocks and the local Lanczos eigensolves resident on the GPU.

Same signatures as the reference plus the keyword extensions `precision`
("fp64" default, or "fp32" dense contractions) and `max_bond` (two-site bond
cap, None = no cap as in the reference).

Reference:
    Ulrich Schollwoeck
    The density-matrix renormalization group in the age of matrix product states
    Ann. Phys. (N. Y.) 326, 96-192 (2011)
"""

import numpy as np
import torch

from .mps import (MPS, mps_local_orthonormalize_left_qr, mps_local_orthonormalize_right_qr,
                  mps_merge_tensor_pair, mps_split_tensor_svd)
from .mpo import MPO, mpo_merge_tensor_pair
from . import backend
from .sweep import SweepContext

__all__ = ["dmrg_singlesite", "dmrg_twosite"]


def _energy(en):
    return 0.0 if en is None else float(en.item())


def dmrg_singlesite(hamiltonian: MPO, psi: MPS, numsweeps: int, numiter_lanczos: int = 25,
                    precision: str = "fp64"):
    """Single-site DMRG (dmrg.py:22-93): returns the energy after each sweep
    (the last local Ritz value); `psi` is overwritten and its bond dimensions
    cannot grow."""
    nsites = hamiltonian.nsites
    assert nsites == psi.nsites
    ctx = SweepContext(hamiltonian, psi, numiter_lanczos, precision)
    W = ctx.W
    psi.orthonormalize(mode="right")
    lblocks, rblocks = ctx.init_blocks(psi)
    en_min = np.zeros(numsweeps)
    for n in range(numsweeps):
        en = None
        for i in range(nsites - 1):
            en, psi.a[i] = ctx.site_eigh(psi.a[i], lblocks[i], W[i], rblocks[i], ctx.q1(psi, i))
            psi.a[i], psi.a[i + 1], psi.qbonds[i + 1] = mps_local_orthonormalize_left_qr(
                psi.a[i], psi.a[i + 1], psi.qsite, psi.qbonds[i:i + 2])
            lblocks[i + 1] = ctx.env_left(psi.a[i], W[i], lblocks[i],
                                          ctx.q1(psi, i))
        for i in reversed(range(1, nsites)):
            en, psi.a[i] = ctx.site_eigh(psi.a[i], lblocks[i], W[i], rblocks[i], ctx.q1(psi, i))
            psi.a[i], psi.a[i - 1], psi.qbonds[i] = mps_local_orthonormalize_right_qr(
                psi.a[i], psi.a[i - 1], psi.qsite, psi.qbonds[i:i + 2])
            rblocks[i - 1] = ctx.env_right(psi.a[i], W[i], rblocks[i],
                                           ctx.q1(psi, i))
        psi.a[0], _, psi.qbonds[0] = mps_local_orthonormalize_right_qr(
            psi.a[0], ctx.one(), psi.qsite, psi.qbonds[:2])
        ctx.flush()
        en_min[n] = _energy(en)
    ctx.finish(psi)
    return en_min


def dmrg_twosite(hamiltonian: MPO, psi: MPS, numsweeps: int, numiter_lanczos: int = 25,
                 tol_split: float = 0, precision: str = "fp64", max_bond=None,
                 svd_precond32=None):
    """Two-site DMRG (dmrg.py:96-178): returns the energy after each sweep.
    svd_precond32: split SVD with FP32 Jacobi preconditioning sweeps (FP64
    output accuracy); None = on when precision is not "fp64"."""
    pre32 = (precision != "fp64") if svd_precond32 is None else bool(svd_precond32)
    nsites = hamiltonian.nsites
    assert nsites == psi.nsites
    ctx = SweepContext(hamiltonian, psi, numiter_lanczos, precision)
    W = ctx.W
    psi.orthonormalize(mode="right")
    lblocks, rblocks = ctx.init_blocks(psi)
    en_min = np.zeros(numsweeps)
    h2 = [backend.as_dev(mpo_merge_tensor_pair(hamiltonian.a[i], hamiltonian.a[i + 1]))
          for i in range(nsites - 1)]
    qs = psi.qsite
    for n in range(numsweeps):
        en = None
        for i in range(nsites - 2):
            a2 = mps_merge_tensor_pair(psi.a[i], psi.a[i + 1])
            en, a2 = ctx.site_eigh(a2, lblocks[i], h2[i], rblocks[i + 1], ctx.q2(psi, i))
            ctx.flush()   # a failed local solve raises here, before its split
            psi.a[i], psi.a[i + 1], psi.qbonds[i + 1] = mps_split_tensor_svd(
                a2, qs, qs, [psi.qbonds[i], psi.qbonds[i + 2]], "right", tol=tol_split,
                max_bond=max_bond, svd_precond32=pre32)
            lblocks[i + 1] = ctx.env_left(psi.a[i], W[i], lblocks[i],
                                          ctx.q1(psi, i))
        for i in reversed(range(nsites - 1)):
            a2 = mps_merge_tensor_pair(psi.a[i], psi.a[i + 1])
            en, a2 = ctx.site_eigh(a2, lblocks[i], h2[i], rblocks[i + 1], ctx.q2(psi, i))
            ctx.flush()   # a failed local solve raises here, before its split
            psi.a[i], psi.a[i + 1], psi.qbonds[i + 1] = mps_split_tensor_svd(
                a2, qs, qs, [psi.qbonds[i], psi.qbonds[i + 2]], "left", tol=tol_split,
                max_bond=max_bond, svd_precond32=pre32)
            rblocks[i] = ctx.env_right(psi.a[i + 1], W[i + 1], rblocks[i + 1],
                                       ctx.q1(psi, i + 1))
        psi.a[0], _, psi.qbonds[0] = mps_local_orthonormalize_right_qr(
            psi.a[0], ctx.one(), psi.qsite, psi.qbonds[:2])
        ctx.flush()
        en_min[n] = _energy(en)
    ctx.finish(psi)
    return en_min
