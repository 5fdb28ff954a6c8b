"""
Oracle (test infrastructure): quantum-number bookkeeping, block-sparse QR/SVD
and the truncation rule.

Restates pytenet/block_sparse_util.py:13-180,244-319 and pytenet/bond_ops.py:23-54.
Dense storage with zeros, exactly as the reference: quantum numbers only select
the blocks that are factorised.
"""

import numpy as np


def qnumber_outer_sum(qnums):
    """Tensor of all sums q0[i0] + q1[i1] + ... (block_sparse_util.py:13-30)."""
    if len(qnums) == 0:
        return np.array(0)
    acc = np.asarray(qnums[0])
    for q in qnums[1:]:
        acc = np.add.outer(acc, np.asarray(q))
    return acc


def common_qnumbers(qnums0, qnums1):
    """Sorted unique quantum numbers present in both lists (block_sparse_util.py:33-37)."""
    return np.intersect1d(qnums0, qnums1)


def qnumber_flatten(qnums):
    """Outer sum flattened in row-major order (block_sparse_util.py:40-44)."""
    return qnumber_outer_sum(qnums).reshape(-1)


def is_qsparse(a, qnums):
    """True when every entry whose quantum numbers do not sum to zero vanishes
    (block_sparse_util.py:47-53)."""
    mask = qnumber_outer_sum(qnums)
    return not np.any(np.where(mask == 0, 0, a))


def enforce_qsparsity(a, qnums):
    """Zero, in place, every entry whose quantum numbers do not sum to zero.
    The reference loops element by element (block_sparse_util.py:56-64);
    the boolean-mask assignment sets exactly the same entries."""
    mask = qnumber_outer_sum(qnums)
    a[np.broadcast_to(mask != 0, a.shape)] = 0


def sort_by_qnumbers(a, q0, q1):
    """Stable (merge) sort of rows and columns by quantum number
    (block_sparse_util.py:67-92). Returns the permuted matrix and, per axis,
    the sorting permutation and the sorted quantum numbers."""
    q0 = np.asarray(q0)
    q1 = np.asarray(q1)
    p0 = np.argsort(q0, kind="mergesort")
    p1 = np.argsort(q1, kind="mergesort")
    if np.any(p0 != np.arange(len(p0))):
        q0 = q0[p0]
        a = a[p0, :]
    if np.any(p1 != np.arange(len(p1))):
        q1 = q1[p1]
        a = a[:, p1]
    return a, (p0, q0), (p1, q1)


def slice_with_qnumber(qn, qnums):
    """Contiguous range of `qn` inside sorted `qnums` (block_sparse_util.py:95-103)."""
    hit = np.flatnonzero(qnums == qn)
    return slice(hit[0], hit[-1] + 1)


def _unsort(q_or_u, perm0, r_or_v, perm1):
    if np.any(perm0 != np.arange(len(perm0))):
        q_or_u = q_or_u[np.argsort(perm0), :]
    if np.any(perm1 != np.arange(len(perm1))):
        r_or_v = r_or_v[:, np.argsort(perm1)]
    return q_or_u, r_or_v


def block_sparse_qr(a, q0, q1):
    """Per-sector QR (block_sparse_util.py:106-180): a[i, j] may be non-zero only
    if q0[i] == q1[j]. Each common sector is factorised with LAPACK's reduced
    Householder QR; Q/R are assembled in sector order and unsorted."""
    a = np.asarray(a)
    assert a.ndim == 2
    assert len(q0) == a.shape[0] and len(q1) == a.shape[1]
    assert is_qsparse(a, [q0, -np.asarray(q1)])
    sectors = common_qnumbers(q0, q1)
    if len(sectors) == 0:
        # no common sector: a dummy bond of dimension 1 (block_sparse_util.py:124-134)
        assert np.linalg.norm(a) == 0
        q = np.zeros((a.shape[0], 1), dtype=a.dtype)
        q[0, 0] = 1
        r = np.zeros((1, a.shape[1]), dtype=a.dtype)
        return q, r, np.asarray(q0)[:1]
    a, (p0, q0), (p1, q1) = sort_by_qnumbers(a, q0, q1)
    kmax = min(a.shape)
    q = np.zeros((a.shape[0], kmax), dtype=a.dtype)
    r = np.zeros((kmax, a.shape[1]), dtype=a.dtype)
    qi = np.zeros(kmax, dtype=q0.dtype)
    k = 0
    for qn in sectors:
        rows = slice_with_qnumber(qn, q0)
        cols = slice_with_qnumber(qn, q1)
        qs, rs = np.linalg.qr(a[rows, cols], mode="reduced")
        q[rows, k:k + qs.shape[1]] = qs
        r[k:k + qs.shape[1], cols] = rs
        qi[k:k + qs.shape[1]] = qn
        k += qs.shape[1]
    assert k <= kmax
    q, r = _unsort(q[:, :k], p0, r[:k, :], p1)
    return q, r, qi[:k]


def block_sparse_svd(a, q0, q1):
    """Per-sector SVD (block_sparse_util.py:244-319) with LAPACK gesdd,
    singular values of each sector in descending order, sectors ascending."""
    a = np.asarray(a)
    assert a.ndim == 2
    assert len(q0) == a.shape[0] and len(q1) == a.shape[1]
    assert is_qsparse(a, [q0, -np.asarray(q1)])
    sectors = common_qnumbers(q0, q1)
    if len(sectors) == 0:
        # dummy bond of dimension 1 (block_sparse_util.py:257-270)
        assert np.linalg.norm(a) == 0
        u = np.zeros((a.shape[0], 1), dtype=a.dtype)
        if a.shape[0] > 0:
            u[0, 0] = 1
        v = np.zeros((1, a.shape[1]), dtype=a.dtype)
        return u, np.zeros(1), v, np.asarray(q0)[:1]
    a, (p0, q0), (p1, q1) = sort_by_qnumbers(a, q0, q1)
    kmax = min(a.shape)
    u = np.zeros((a.shape[0], kmax), dtype=a.dtype)
    v = np.zeros((kmax, a.shape[1]), dtype=a.dtype)
    s = np.zeros(kmax)
    qi = np.zeros(kmax, dtype=q0.dtype)
    k = 0
    for qn in sectors:
        rows = slice_with_qnumber(qn, q0)
        cols = slice_with_qnumber(qn, q1)
        us, ss, vs = np.linalg.svd(a[rows, cols], full_matrices=False)
        u[rows, k:k + len(ss)] = us
        v[k:k + len(ss), cols] = vs
        s[k:k + len(ss)] = ss
        qi[k:k + len(ss)] = qn
        k += len(ss)
    assert k <= kmax
    u, v = _unsort(u[:, :k], p0, v[:k, :], p1)
    return u, s[:k], v, qi[:k]


def retained_bond_indices(s, tol):
    """Indices (ascending) of singular values kept by the discarded-weight rule
    (bond_ops.py:23-38): normalise by the 2-norm, square, accumulate from the
    smallest upwards (NumPy's default argsort), keep entries whose running sum
    exceeds `tol`."""
    w = np.linalg.norm(s)
    if w == 0:
        return np.array([], dtype=int)
    p = (s / w)**2
    order = np.argsort(p)
    p[order] = np.cumsum(p[order])
    return np.where(p > tol)[0]


def cap_bond_indices(s, idx, max_bond):
    """Extension (no counterpart in the reference, which has no bond cap):
    of the retained indices keep the `max_bond` with the largest singular
    values (stable on ties), returned in ascending index order."""
    if max_bond is None or len(idx) <= max_bond:
        return idx
    order = np.argsort(-s[idx], kind="stable")
    return np.sort(idx[order[:max_bond]])


def split_block_sparse_matrix_svd(a, q0, q1, tol, max_bond=None):
    """SVD split with truncation (bond_ops.py:41-54); `max_bond` is the
    bond-cap extension (None reproduces the reference)."""
    u, s, v, q = block_sparse_svd(a, q0, q1)
    idx = cap_bond_indices(s, retained_bond_indices(s, tol), max_bond)
    return u[:, idx], s[idx], v[idx, :], q[idx]
