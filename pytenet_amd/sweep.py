"""
Shared state of the sweep drivers: device copies of the MPO tensors, the
precision mode, the Krylov calls with their status words, and the
environment-block initialisation (reference pytenet/tdvp.py:52-63,
pytenet/dmrg.py:45-56).
"""

import numpy as np
import torch

from . import backend, _lib
from .block_sparse_util import is_qsparse, qnumber_flatten


class SweepContext:
    """Per-call device state of a sweep driver."""

    def __init__(self, hamiltonian, psi, numiter, precision="fp64"):
        assert hamiltonian.nsites == psi.nsites
        self.dev = backend.device()
        backend.precision_code(precision)
        self.precision = precision
        self.numiter = int(numiter)
        self.H = hamiltonian
        self.W = [backend.as_dev(w) for w in hamiltonian.a]
        self.host_input = not psi.on_device
        psi.to_device()
        self._status = []
        # nontrivial quantum numbers make every tensor of the sweep block sparse
        # (the reference keeps them as dense arrays with zero blocks): let the
        # contractions skip the zero tiles
        self.block_sparse = bool(np.any(np.asarray(psi.qsite) != 0) or
                                 any(np.any(np.asarray(q) != 0) for q in psi.qbonds))

    def one(self):
        return torch.ones((1, 1, 1), dtype=torch.complex128, device=self.dev)

    def q1(self, psi, i):
        """U(1) labels of the one-site H_eff at site i (left bond, physical, right
        bond, MPO bonds), or None without quantum numbers."""
        if not self.block_sparse:
            return None
        return (psi.qbonds[i], psi.qsite, psi.qbonds[i + 1], self.H.qbonds[i],
                self.H.qbonds[i + 1])

    def q2(self, psi, i):
        """U(1) labels of the two-site H_eff on sites (i, i+1): the physical
        labels are the flattened pair sums of the merged pair (mps.py:528-535)."""
        if not self.block_sparse:
            return None
        return (psi.qbonds[i], qnumber_flatten((psi.qsite, psi.qsite)), psi.qbonds[i + 2],
                self.H.qbonds[i], self.H.qbonds[i + 2])

    def site_expm(self, a, l, w, r, dt, qnums=None):
        """expm(dt * H_eff) a (the reference passes -dt to expm_krylov, tdvp.py:227);
        `qnums` (q1 / q2) lets the matvecs run over the sector blocks only."""
        out, _, info = backend.krylov_local(_lib.PTN_OP_HEFF, a, l, w, r, self.numiter, "expm",
                                            dt, self.precision, block_sparse=self.block_sparse,
                                            qnums=qnums)
        self._status.append(info)
        return out

    def bond_expm(self, c, l, r, dt):
        """expm(dt * H_bond) c (tdvp.py:232-238)."""
        out, _, info = backend.krylov_local(_lib.PTN_OP_BOND, c, l, None, r, self.numiter,
                                            "expm", dt, self.precision,
                                            block_sparse=self.block_sparse)
        self._status.append(info)
        return out

    def site_eigh(self, a, l, w, r, qnums=None):
        """Lowest Ritz pair of H_eff started from a (dmrg.py:181-189);
        the energy stays on the device."""
        out, en, info = backend.krylov_local(_lib.PTN_OP_HEFF, a, l, w, r, self.numiter, "eigh",
                                             0.0, self.precision, block_sparse=self.block_sparse,
                                             qnums=qnums)
        self._status.append(info)
        return en, out

    def env_left(self, a, w, l, qnums=None):
        return backend.env_left(a, a, w, l, self.precision, self.block_sparse, qnums)

    def env_right(self, a, w, r, qnums=None):
        return backend.env_right(a, a, w, r, self.precision, self.block_sparse, qnums)

    def flush(self):
        """Read the Krylov status words of the calls so far (one synchronisation)
        and raise / warn as the reference does."""
        if self._status:
            rows = torch.stack(self._status).cpu().numpy()
            self._status = []
            backend.report_krylov_status(rows)

    def init_blocks(self, psi):
        """Right environments of a right-orthonormal psi and the trivial left block,
        with the reference's quantum-number consistency check (tdvp.py:56-63)."""
        n = psi.nsites
        rblocks = [None] * n
        rblocks[n - 1] = self.one()
        for i in reversed(range(n - 1)):
            rblocks[i] = self.env_right(psi.a[i + 1], self.W[i + 1], rblocks[i + 1],
                                        self.q1(psi, i + 1))
        lblocks = [None] * n
        lblocks[0] = self.one()
        for i, rb in enumerate(rblocks):
            assert is_qsparse(rb, [psi.qbonds[i + 1], self.H.qbonds[i + 1], -psi.qbonds[i + 1]]), \
                "sparsity pattern of operator blocks must match quantum numbers"
        return lblocks, rblocks

    def finish(self, psi):
        self.flush()
        if self.host_input:
            psi.to_host()


def absorb_left(c, a):
    """tensordot(c, a, (1, 0)): (k, D) x (D, d, Dr) -> (k, d, Dr) (tdvp.py:84)."""
    D, d, Dr = a.shape
    return backend.matmul(c, a.reshape(D, d * Dr)).reshape(c.shape[0], d, Dr)


def absorb_right(a, c):
    """tensordot(a, c, (2, 0)): (Dl, d, D) x (D, k) -> (Dl, d, k) (tdvp.py:112)."""
    Dl, d, D = a.shape
    return backend.matmul(a.reshape(Dl * d, D), c).reshape(Dl, d, c.shape[1])


def flip(a):
    """Swap the virtual bonds of an MPS tensor: transpose((2, 1, 0)), made contiguous."""
    return a.permute(2, 1, 0).contiguous()
