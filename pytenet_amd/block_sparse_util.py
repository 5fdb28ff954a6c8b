"""
Quantum-number bookkeeping and block-sparse factorisations
(reference pytenet/block_sparse_util.py:13-180,244-319).

Storage is dense with zeros, exactly as in the reference: quantum numbers only
select which row/column blocks are factorised. Integer bookkeeping (outer sums,
stable sorts, sector slices) stays on the host; the factorisations of each
sector (Householder QR, one-sided Jacobi SVD) and all dense data movement run
on the GPU through the C-ABI library.
"""

import numpy as np

__all__ = [
    "qnumber_outer_sum", "common_qnumbers", "qnumber_flatten", "is_qsparse",
    "enforce_qsparsity", "sort_by_qnumbers", "slice_with_qnumber",
    "block_sparse_qr", "block_sparse_svd"]


def qnumber_outer_sum(qnums):
    """All sums q0[i0] + q1[i1] + ... as a tensor (block_sparse_util.py:13-30).

    >>> qnumber_outer_sum([[0, 2, -1], [0, 1]]).tolist()
    [[0, 1], [2, 3], [-1, 0]]
    """
    if len(qnums) == 0:
        return np.array(0)
    acc = np.asarray(qnums[0])
    for q in qnums[1:]:
        acc = np.add.outer(acc, np.asarray(q))
    return acc


def common_qnumbers(qnums0, qnums1):
    """Sorted quantum numbers present in both lists (block_sparse_util.py:33-37)."""
    return np.intersect1d(qnums0, qnums1)


def qnumber_flatten(qnums):
    """Outer sum flattened in row-major order (block_sparse_util.py:40-44)."""
    return qnumber_outer_sum(qnums).reshape(-1)


def _all_zero(qnums):
    return all(not np.any(np.asarray(q)) for q in qnums)


def is_qsparse(a, qnums):
    """True if every entry whose quantum numbers do not sum to zero vanishes
    (block_sparse_util.py:47-53). Accepts NumPy arrays and device tensors;
    device tensors are checked on the GPU without copying them back."""
    if _all_zero(qnums):
        return True
    if not isinstance(a, np.ndarray):
        from . import backend
        return backend.qsparse_violation(a, qnums) == 0
    mask = qnumber_outer_sum(qnums)
    return not np.any(np.where(mask == 0, 0, a))


def enforce_qsparsity(a, qnums):
    """Zero, in place, every entry whose quantum numbers do not sum to zero
    (block_sparse_util.py:56-64). The reference visits entries one by one; a
    boolean mask (host) or a device kernel sets exactly the same entries."""
    if _all_zero(qnums):
        return
    if not isinstance(a, np.ndarray):
        from . import backend
        backend.enforce_qsparsity(a, qnums)
        return
    mask = qnumber_outer_sum(qnums)
    a[np.broadcast_to(mask != 0, a.shape)] = 0


def sort_by_qnumbers(a, q0, q1):
    """Stable sort of rows and columns by quantum number (block_sparse_util.py:67-92).
    Works for host arrays and device tensors (the permutation is applied by
    indexing on the device)."""
    q0 = np.asarray(q0)
    q1 = np.asarray(q1)
    p0 = np.argsort(q0, kind="mergesort")
    p1 = np.argsort(q1, kind="mergesort")
    if np.any(p0 != np.arange(len(p0))):
        q0 = q0[p0]
        a = a[_index(a, p0), :]
    if np.any(p1 != np.arange(len(p1))):
        q1 = q1[p1]
        a = a[:, _index(a, p1)]
    return a, (p0, q0), (p1, q1)


def _index(a, perm):
    if isinstance(a, np.ndarray):
        return perm
    import torch
    return torch.as_tensor(perm, device=a.device)


def slice_with_qnumber(qn, qnums):
    """First-to-last range of `qn` in sorted `qnums` (block_sparse_util.py:95-103)."""
    hit = np.flatnonzero(np.asarray(qnums) == qn)
    return slice(int(hit[0]), int(hit[-1]) + 1)


def block_sparse_qr(a, q0, q1):
    """Block-wise QR (block_sparse_util.py:106-180): a[i, j] may be non-zero only
    if q0[i] == q1[j]. Returns (q, r, qinterm); each common sector is factorised
    by the GPU Householder QR, which follows LAPACK's zgeqrf/zungqr conventions.
    `a` is a device tensor (complex128)."""
    from . import backend
    return backend.block_sparse_qr(a, np.asarray(q0), np.asarray(q1))


def block_sparse_svd(a, q0, q1, precond32=False):
    """Block-wise SVD (block_sparse_util.py:244-319): per common sector, singular
    values in descending order, sectors in ascending quantum number.
    `a` is a device tensor (complex128). precond32 (extension) starts the Jacobi
    iteration with FP32 sweeps; the result keeps FP64 accuracy."""
    from . import backend
    return backend.block_sparse_svd(a, np.asarray(q0), np.asarray(q1), precond32)
