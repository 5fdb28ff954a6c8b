"""
Matrix product operator container (reference pytenet/mpo.py:15-73,161-259,314-322).

MPO tensors are tiny (bond dimension 3..5, d = 2) and are built once on the
host; the sweep drivers upload them to the GPU once per call. Layout of
tensor i: (b[i], d_out, d_in, b[i+1]), row-major.
"""

import numpy as np
from .block_sparse_util import enforce_qsparsity, is_qsparse
from .util import crandn

__all__ = ["MPO", "mpo_merge_tensor_pair"]


class MPO:
    """Matrix product operator (mpo.py:15-73). `qsite` are the physical quantum
    numbers, `qbonds` the virtual-bond quantum numbers; entry
    a[i][m, s, t, n] may be non-zero only if
    qbonds[i][m] + qsite[s] - qsite[t] - qbonds[i+1][n] == 0."""

    def __init__(self, qsite, qbonds, fill=0.0, rng=None):
        self.qsite = np.array(qsite)
        self.qbonds = [np.array(qb) for qb in qbonds]
        d = len(qsite)
        b = [len(qb) for qb in qbonds]
        shapes = [(b[i], d, d, b[i + 1]) for i in range(len(b) - 1)]
        if isinstance(fill, (int, float, complex)):
            self.a = [np.full(s, fill) for s in shapes]
        elif fill == "random":
            if rng is None:
                rng = np.random.default_rng()
            self.a = [crandn(s, rng) / np.sqrt(s[0] * s[1] * s[3]) for s in shapes]
        elif fill == "random real":
            if rng is None:
                rng = np.random.default_rng()
            self.a = [rng.normal(size=s) / np.sqrt(s[0] * s[1] * s[3]) for s in shapes]
        elif fill == "postpone":
            self.a = [None] * len(shapes)
        else:
            raise ValueError(f'`fill` = {fill} invalid; must be a number, '
                             f'"random", "random real" or "postpone".')
        if fill != "postpone":
            for i, ai in enumerate(self.a):
                enforce_qsparsity(ai, (self.qbonds[i], self.qsite, -self.qsite,
                                       -self.qbonds[i + 1]))
        self.nid_map = None

    @classmethod
    def from_tensors(cls, qsite, qbonds, a):
        """Wrap explicit tensors; checks the quantum-number sparsity pattern."""
        op = cls(qsite, qbonds, fill="postpone")
        op.a = [np.asarray(x) for x in a]
        for i, ai in enumerate(op.a):
            assert is_qsparse(ai, (op.qbonds[i], op.qsite, -op.qsite, -op.qbonds[i + 1])), \
                "sparsity pattern of MPO tensor does not match quantum numbers"
        return op

    @classmethod
    def identity(cls, qsite, nsites: int, scale: float = 1, dtype=float):
        """Identity operator (mpo.py:75-84)."""
        d = len(qsite)
        op = cls(qsite, (nsites + 1) * [[0]])
        for i in range(nsites):
            op.a[i] = scale * np.identity(d, dtype=dtype).reshape((1, d, d, 1))
        return op

    @property
    def nsites(self) -> int:
        return len(self.a)

    @property
    def bond_dims(self) -> list:
        if len(self.a) == 0:
            return []
        return [ai.shape[0] for ai in self.a] + [self.a[-1].shape[3]]

    def zero_qnumbers(self):
        """Set all quantum numbers to zero; returns self (mpo.py:177-185)."""
        self.qsite.fill(0)
        for qb in self.qbonds:
            qb.fill(0)
        return self

    def to_matrix(self):
        """Dense matrix on the full Hilbert space (mpo.py:225-236). Host-side,
        for small chains (tests and measurement only)."""
        op = self.a[0]
        for ai in self.a[1:]:
            op = mpo_merge_tensor_pair(op, ai)
        assert op.ndim == 4 and op.shape[0] == 1 and op.shape[3] == 1
        return op.reshape((op.shape[1], op.shape[2]))


def mpo_merge_tensor_pair(a0, a1) -> np.ndarray:
    """Two-site MPO tensor (Ml, d0*d1, d0*d1, Mr) with combined output index
    s0o*d1 + s1o and input index s0i*d1 + s1i (mpo.py:314-322). Host-side: it
    is built once per driver call from tensors with at most a few hundred entries."""
    a0 = np.asarray(a0)
    a1 = np.asarray(a1)
    t = np.einsum("aste,eulb->asutlb", a0, a1)
    s = t.shape
    return t.reshape((s[0], s[1] * s[2], s[3] * s[4], s[5]))
