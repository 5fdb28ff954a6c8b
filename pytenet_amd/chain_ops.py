"""
Local tensor-network contractions on a chain (reference pytenet/chain_ops.py:16-113,151-189,237-317),
each one a call into the GPU library.

Index conventions (as in the reference): A (Dl, d, Dr); W (Ml, d_out, d_in, Mr);
L (Dl_ket, Ml, Dl_bra); R (Dr_ket, Mr, Dr_bra). Functions accept NumPy arrays
or device tensors and return the same kind. `precision` selects the
arithmetic of the dense contraction ("fp64" reproduces the reference;
"fp32" runs complex FP32 MFMA products).
"""

import numpy as np
import torch

from . import backend
from .mps import MPS
from .mpo import MPO

__all__ = ["contraction_operator_step_right", "contraction_operator_step_left",
           "compute_right_operator_blocks", "mpo_average", "mpo_inner_product",
           "apply_local_hamiltonian", "apply_local_bond_contraction"]


def _host(*xs):
    return any(isinstance(x, np.ndarray) for x in xs[:1])


def _ret(x, host):
    return backend.to_host(x) if host else x


def contraction_operator_step_right(a, b, w, r, precision="fp64"):
    """R'[l, m, lb] = sum A[l,s,r] W[m,sb,s,m'] R[r,m',rb] conj(B[lb,sb,rb])
    (chain_ops.py:16-57)."""
    assert a.ndim == 3 and b.ndim == 3 and w.ndim == 4 and r.ndim == 3
    return _ret(backend.env_right(a, b, w, r, precision), _host(a))


def contraction_operator_step_left(a, b, w, l, precision="fp64"):
    """L'[r, m', rb] = sum A[l,s,r] W[m,sb,s,m'] L[l,m,lb] conj(B[lb,sb,rb])
    (chain_ops.py:60-99)."""
    assert a.ndim == 3 and b.ndim == 3 and w.ndim == 4 and l.ndim == 3
    return _ret(backend.env_left(a, b, w, l, precision), _host(a))


def compute_right_operator_blocks(psi: MPS, op: MPO, precision="fp64"):
    """All right environments from the right end, R[L-1] = [[[1]]]
    (chain_ops.py:102-113). Returns device tensors."""
    n = psi.nsites
    assert n == op.nsites
    dev = backend.device()
    blocks = [None] * n
    blocks[n - 1] = torch.ones((1, 1, 1), dtype=torch.complex128, device=dev)
    for i in reversed(range(n - 1)):
        blocks[i] = backend.env_right(psi.a[i + 1], psi.a[i + 1], op.a[i + 1], blocks[i + 1],
                                      precision)
    return blocks


def mpo_inner_product(chi: MPS, op: MPO, psi: MPS):
    """<chi|op|psi> as a chain of right environment steps (chain_ops.py:165-189)."""
    assert chi.nsites == op.nsites and psi.nsites == op.nsites
    if psi.nsites == 0:
        return 0
    assert chi.a[-1].shape[2] == psi.a[-1].shape[2]
    n = psi.a[-1].shape[2]
    dev = backend.device()
    t = torch.eye(n, dtype=torch.complex128, device=dev).reshape(n, 1, n)
    for i in reversed(range(psi.nsites)):
        t = backend.env_right(psi.a[i], chi.a[i], op.a[i], t)
    assert tuple(t.shape) == (1, 1, 1)
    return complex(t[0, 0, 0].item())


def mpo_average(psi: MPS, op: MPO):
    """<psi|op|psi> (chain_ops.py:151-162)."""
    return mpo_inner_product(psi, op, psi)


def apply_local_hamiltonian(a, w, l, r, precision="fp64"):
    """Out[lb, sb, rb] = sum L[l,m,lb] W[m,sb,s,m'] A[l,s,r] R[r,m',rb]
    (chain_ops.py:237-279)."""
    assert a.ndim == 3 and w.ndim == 4 and l.ndim == 3 and r.ndim == 3
    return _ret(backend.heff_apply(a, w, l, r, precision), _host(a))


def apply_local_bond_contraction(c, l, r, precision="fp64"):
    """Out[lb, rb] = sum L[l,m,lb] C[l,r] R[r,m,rb] (chain_ops.py:282-317)."""
    assert c.ndim == 2 and l.ndim == 3 and r.ndim == 3
    return _ret(backend.bond_apply(c, l, r, precision), _host(c))
