"""
Oracle (test infrastructure): the local tensor-network contractions.

Restates pytenet/chain_ops.py:16-113,151-189,237-317 with the same
pairwise contraction order (tensordot chains), hence the same BLAS calls.
Index conventions: A (Dl, d, Dr); W (Ml, d_out, d_in, Mr); L (Dl_ket, Ml, Dl_bra);
R (Dr_ket, Mr, Dr_bra).
"""

import numpy as np


def contraction_operator_step_right(a, b, w, r):
    """R'[l, m, lb] = sum A[l,s,r] W[m,sb,s,m'] R[r,m',rb] conj(B[lb,sb,rb])
    (chain_ops.py:16-57)."""
    assert a.ndim == 3 and b.ndim == 3 and w.ndim == 4 and r.ndim == 3
    t = np.tensordot(a, r, 1)                           # [l, s, m', rb]
    t = np.tensordot(w, t, axes=((2, 3), (1, 2)))       # [m, sb, l, rb]
    t = t.transpose((2, 0, 1, 3))                       # [l, m, sb, rb]
    return np.tensordot(t, b.conj(), axes=((2, 3), (1, 2)))


def contraction_operator_step_left(a, b, w, l):
    """L'[r, m', rb] = sum A[l,s,r] W[m,sb,s,m'] L[l,m,lb] conj(B[lb,sb,rb])
    (chain_ops.py:60-99)."""
    assert a.ndim == 3 and b.ndim == 3 and w.ndim == 4 and l.ndim == 3
    t = np.tensordot(l, b.conj(), axes=(2, 0))          # [l, m, sb, rb]
    t = np.tensordot(w, t, axes=((0, 1), (1, 2)))       # [s, m', l, rb]
    return np.tensordot(a, t, axes=((0, 1), (2, 0)))


def compute_right_operator_blocks(psi, op):
    """All right environments, R[L-1] = [[[1]]] (chain_ops.py:102-113)."""
    n = psi.nsites
    assert n == op.nsites
    blocks = [None] * n
    blocks[n - 1] = np.array([[[1]]])
    for i in reversed(range(n - 1)):
        blocks[i] = contraction_operator_step_right(psi.a[i+1], psi.a[i+1], op.a[i+1],
                                                    blocks[i+1])
    return blocks


def mpo_inner_product(chi, op, psi):
    """<chi|op|psi> as a right-environment chain (chain_ops.py:165-189)."""
    assert chi.nsites == op.nsites and psi.nsites == op.nsites
    if psi.nsites == 0:
        return 0
    assert chi.a[-1].shape[2] == psi.a[-1].shape[2]
    n = psi.a[-1].shape[2]
    t = np.identity(n, dtype=psi.a[-1].dtype).reshape(n, 1, n)
    for i in reversed(range(psi.nsites)):
        t = contraction_operator_step_right(psi.a[i], chi.a[i], op.a[i], t)
    assert t.shape == (1, 1, 1)
    return t[0, 0, 0]


def mpo_average(psi, op):
    """<psi|op|psi> (chain_ops.py:151-162)."""
    return mpo_inner_product(psi, op, psi)


def apply_local_hamiltonian(a, w, l, r):
    """Out[lb, sb, rb] = sum L[l,m,lb] W[m,sb,s,m'] A[l,s,r] R[r,m',rb]
    (chain_ops.py:237-279)."""
    assert a.ndim == 3 and w.ndim == 4 and l.ndim == 3 and r.ndim == 3
    t = np.tensordot(a, r, 1)                           # [l, s, m', rb]
    t = np.tensordot(w, t, axes=((2, 3), (1, 2)))       # [m, sb, l, rb]
    return np.tensordot(l, t, axes=((0, 1), (2, 0)))


def apply_local_bond_contraction(c, l, r):
    """Out[lb, rb] = sum L[l,m,lb] C[l,r] R[r,m,rb] (chain_ops.py:282-317)."""
    assert c.ndim == 2 and l.ndim == 3 and r.ndim == 3
    t = np.tensordot(c, r, 1)                           # [l, m, rb]
    return np.tensordot(l, t, axes=((0, 1), (0, 1)))
