"""
Bond truncation (reference pytenet/bond_ops.py:23-54).

The truncation decision is integer bookkeeping on at most a few thousand
singular values that the host needs anyway to size the new bond; it follows
the reference's rule exactly (same NumPy sort), so the GPU path keeps the same
singular values as the reference whenever the singular values agree.
"""

import numpy as np
from .block_sparse_util import block_sparse_svd

__all__ = ["retained_bond_indices", "split_block_sparse_matrix_svd", "cap_bond_indices"]


def retained_bond_indices(s, tol):
    """Indices (ascending) of the singular values kept by the discarded-weight
    rule (bond_ops.py:23-38): normalise by the 2-norm, square, accumulate from
    the smallest upwards, keep the entries whose running sum exceeds `tol`."""
    s = np.asarray(s, dtype=float)
    w = np.linalg.norm(s)
    if w == 0:
        return np.array([], dtype=int)
    p = (s / w)**2
    order = np.argsort(p)
    p[order] = np.cumsum(p[order])
    return np.where(p > tol)[0]


def cap_bond_indices(s, idx, max_bond):
    """Bond-dimension cap (extension; the reference truncates by weight only):
    of the retained indices keep the `max_bond` largest singular values, ties
    resolved by position, returned in ascending index order."""
    if max_bond is None or len(idx) <= max_bond:
        return idx
    order = np.argsort(-np.asarray(s)[idx], kind="stable")
    return np.sort(idx[order[:max_bond]])


def split_block_sparse_matrix_svd(a, q0, q1, tol, max_bond=None, svd_precond32=False):
    """Block-sparse SVD followed by truncation (bond_ops.py:41-54);
    returns (u, s, v, q) with u, v device tensors and s, q host arrays."""
    u, s, v, q = block_sparse_svd(a, q0, q1, svd_precond32)
    idx = cap_bond_indices(s, retained_bond_indices(s, tol), max_bond)
    import torch
    it = torch.as_tensor(idx, device=u.device)
    return u.index_select(1, it), s[idx], v.index_select(0, it), q[idx]
