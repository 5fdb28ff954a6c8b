"""
TDVP time evolution (reference pytenet/tdvp.py:26-238), with the site tensors,
environment blocks and all local Krylov solves resident on the GPU.

Same signatures as the reference plus keyword extensions with
reference-preserving defaults:
  precision: "fp64" (default; the parity mode), "fp32", "bf16x6" (FP32
             accuracy on the bf16 MFMA) or "bf16x3" for the dense
             contractions (Krylov scalars, QR and SVD stay FP64);
  max_bond:  bond-dimension cap for the two-site split (None = no cap, as
             in the reference);
  svd_precond32: two-site split SVD with FP32 Jacobi preconditioning sweeps
             finished in FP64 (same output accuracy); None = on when
             precision is not "fp64".

Reference:
    J. Haegeman, C. Lubich, I. Oseledets, B. Vandereycken, F. Verstraete
    Unifying time evolution and optimization with matrix product states
    Phys. Rev. B 94, 165116 (2016) (arXiv:1408.5056)
"""

from .mps import MPS, mps_merge_tensor_pair, mps_split_tensor_svd
from .mpo import MPO, mpo_merge_tensor_pair
from .block_sparse_util import qnumber_flatten, block_sparse_qr
from . import backend
from .sweep import SweepContext, absorb_left, absorb_right, flip

__all__ = ["tdvp_singlesite", "tdvp_twosite", "TwoSiteTDVP"]


def tdvp_singlesite(hamiltonian: MPO, psi: MPS, dt, numsteps: int, numiter_lanczos: int = 25,
                    precision: str = "fp64"):
    """Symmetric single-site TDVP (tdvp.py:26-118). `psi` is overwritten with the
    evolved state; returns the norm of the initial psi."""
    nsites = hamiltonian.nsites
    assert nsites == psi.nsites
    ctx = SweepContext(hamiltonian, psi, numiter_lanczos, precision)
    W = ctx.W
    nrm = psi.orthonormalize(mode="right")
    lblocks, rblocks = ctx.init_blocks(psi)
    for _ in range(numsteps):
        # left to right
        for i in range(nsites - 1):
            psi.a[i] = ctx.site_expm(psi.a[i], lblocks[i], W[i], rblocks[i], -0.5 * dt,
                                     ctx.q1(psi, i))
            s = psi.a[i].shape
            q, c, psi.qbonds[i + 1] = block_sparse_qr(
                psi.a[i].reshape(s[0] * s[1], s[2]),
                qnumber_flatten((psi.qbonds[i], psi.qsite)), psi.qbonds[i + 1])
            psi.a[i] = q.reshape(s[0], s[1], q.shape[1])
            lblocks[i + 1] = ctx.env_left(psi.a[i], W[i], lblocks[i],
                                          ctx.q1(psi, i))
            c = ctx.bond_expm(c, lblocks[i + 1], rblocks[i], 0.5 * dt)
            psi.a[i + 1] = absorb_left(c, psi.a[i + 1])
        i = nsites - 1
        psi.a[i] = ctx.site_expm(psi.a[i], lblocks[i], W[i], rblocks[i], -dt, ctx.q1(psi, i))
        # right to left
        for i in reversed(range(1, nsites)):
            psi.a[i] = flip(psi.a[i])
            s = psi.a[i].shape
            q, c, qbond = block_sparse_qr(
                psi.a[i].reshape(s[0] * s[1], s[2]),
                qnumber_flatten((-psi.qbonds[i + 1], psi.qsite)), -psi.qbonds[i])
            psi.qbonds[i] = -qbond
            psi.a[i] = flip(q.reshape(s[0], s[1], q.shape[1]))
            rblocks[i - 1] = ctx.env_right(psi.a[i], W[i], rblocks[i],
                                           ctx.q1(psi, i))
            c = c.transpose(0, 1).contiguous()
            c = ctx.bond_expm(c, lblocks[i], rblocks[i - 1], 0.5 * dt)
            psi.a[i - 1] = absorb_right(psi.a[i - 1], c)
            psi.a[i - 1] = ctx.site_expm(psi.a[i - 1], lblocks[i - 1], W[i - 1], rblocks[i - 1],
                                         -0.5 * dt, ctx.q1(psi, i - 1))
        ctx.flush()
    ctx.finish(psi)
    return nrm


class TwoSiteTDVP:
    """State of a two-site TDVP run on the GPU (tdvp.py:121-220), split into
    setup() (right-orthonormalisation, environments, merged MPO pairs) and
    sweep() (one symmetric time step), so that sweeps can be timed on their own."""

    def __init__(self, hamiltonian: MPO, psi: MPS, dt, numiter_lanczos: int = 25, tol_split=0,
                 precision: str = "fp64", max_bond=None, svd_precond32=None):
        self.nsites = hamiltonian.nsites
        assert self.nsites == psi.nsites
        assert self.nsites >= 2
        self.H, self.psi, self.dt = hamiltonian, psi, dt
        self.tol, self.max_bond = tol_split, max_bond
        self.pre32 = (precision != "fp64") if svd_precond32 is None else bool(svd_precond32)
        self.ctx = SweepContext(hamiltonian, psi, numiter_lanczos, precision)

    def setup(self):
        psi, H = self.psi, self.H
        self.nrm = psi.orthonormalize(mode="right")
        self.lblocks, self.rblocks = self.ctx.init_blocks(psi)
        self.h2 = [backend.as_dev(mpo_merge_tensor_pair(H.a[i], H.a[i + 1]))
                   for i in range(self.nsites - 1)]
        return self.nrm

    def sweep(self):
        psi, ctx, W, h2 = self.psi, self.ctx, self.ctx.W, self.h2
        lblocks, rblocks = self.lblocks, self.rblocks
        dt, qs, n = self.dt, psi.qsite, self.nsites
        for i in range(n - 2):
            a2 = mps_merge_tensor_pair(psi.a[i], psi.a[i + 1])
            a2 = ctx.site_expm(a2, lblocks[i], h2[i], rblocks[i + 1], -0.5 * dt, ctx.q2(psi, i))
            ctx.flush()   # a failed local solve raises here, before its split
            psi.a[i], psi.a[i + 1], psi.qbonds[i + 1] = mps_split_tensor_svd(
                a2, qs, qs, (psi.qbonds[i], psi.qbonds[i + 2]), "right", tol=self.tol,
                max_bond=self.max_bond, svd_precond32=self.pre32)
            lblocks[i + 1] = ctx.env_left(psi.a[i], W[i], lblocks[i],
                                          ctx.q1(psi, i))
            psi.a[i + 1] = ctx.site_expm(psi.a[i + 1], lblocks[i + 1], W[i + 1], rblocks[i + 1],
                                         0.5 * dt, ctx.q1(psi, i + 1))
        i = n - 2
        a2 = mps_merge_tensor_pair(psi.a[i], psi.a[i + 1])
        a2 = ctx.site_expm(a2, lblocks[i], h2[i], rblocks[i + 1], -dt, ctx.q2(psi, i))
        ctx.flush()   # a failed local solve raises here, before its split
        psi.a[i], psi.a[i + 1], psi.qbonds[i + 1] = mps_split_tensor_svd(
            a2, qs, qs, [psi.qbonds[i], psi.qbonds[i + 2]], "left", tol=self.tol,
            max_bond=self.max_bond, svd_precond32=self.pre32)
        rblocks[i] = ctx.env_right(psi.a[i + 1], W[i + 1], rblocks[i + 1],
                                   ctx.q1(psi, i + 1))
        for i in reversed(range(n - 2)):
            psi.a[i + 1] = ctx.site_expm(psi.a[i + 1], lblocks[i + 1], W[i + 1], rblocks[i + 1],
                                         0.5 * dt, ctx.q1(psi, i + 1))
            a2 = mps_merge_tensor_pair(psi.a[i], psi.a[i + 1])
            a2 = ctx.site_expm(a2, lblocks[i], h2[i], rblocks[i + 1], -0.5 * dt, ctx.q2(psi, i))
            ctx.flush()   # a failed local solve raises here, before its split
            psi.a[i], psi.a[i + 1], psi.qbonds[i + 1] = mps_split_tensor_svd(
                a2, qs, qs, [psi.qbonds[i], psi.qbonds[i + 2]], "left", tol=self.tol,
                max_bond=self.max_bond, svd_precond32=self.pre32)
            rblocks[i] = ctx.env_right(psi.a[i + 1], W[i + 1], rblocks[i + 1],
                                       ctx.q1(psi, i + 1))
        ctx.flush()

    def finish(self):
        self.ctx.finish(self.psi)


def tdvp_twosite(hamiltonian: MPO, psi: MPS, dt, numsteps: int, numiter_lanczos: int = 25,
                 tol_split=0, precision: str = "fp64", max_bond=None, svd_precond32=None):
    """Symmetric two-site TDVP (tdvp.py:121-220). `psi` is overwritten with the
    evolved state; returns the norm of the initial psi."""
    run = TwoSiteTDVP(hamiltonian, psi, dt, numiter_lanczos, tol_split, precision, max_bond,
                      svd_precond32)
    nrm = run.setup()
    for _ in range(numsteps):
        run.sweep()
    run.finish()
    return nrm
