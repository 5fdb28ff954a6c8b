"""
Oracle (test infrastructure): MPS/MPO containers and the tensor-pair and
orthonormalisation operations of the sweep.

Restates pytenet/mps.py:14-178,292-301,352-491,528-568 and pytenet/mpo.py:15-73,225-259,314-322.
"""

import numpy as np
from .util import crandn
from .qnum import (qnumber_outer_sum, qnumber_flatten, enforce_qsparsity, block_sparse_qr,
                   split_block_sparse_matrix_svd)


class MPS:
    """Matrix product state with tensors a[i] of shape (b[i], d, b[i+1]) and
    additive integer quantum numbers (mps.py:14-71)."""

    def __init__(self, qsite, qbonds, fill=0.0, rng=None):
        self.qsite = np.asarray(qsite)
        self.qbonds = [np.asarray(qb) for qb in qbonds]
        d = len(qsite)
        b = [len(qb) for qb in qbonds]
        assert b[0] == 1 and b[-1] == 1
        shapes = [(b[i], d, b[i+1]) for i in range(len(b) - 1)]
        if isinstance(fill, (int, float, complex)):
            self.a = [np.full(s, fill) for s in shapes]
        elif fill == "random":
            if rng is None:
                rng = np.random.default_rng()
            # one crandn call per site, in site order (mps.py:51-56)
            self.a = [crandn(s, rng) / np.sqrt(s[0] * s[1] * s[2]) for s in shapes]
        elif fill == "random real":
            if rng is None:
                rng = np.random.default_rng()
            self.a = [rng.normal(size=s) / np.sqrt(s[0] * s[1] * s[2]) for s in shapes]
        elif fill == "postpone":
            self.a = [None] * len(shapes)
        else:
            raise ValueError(f"`fill` = {fill} invalid")
        if fill != "postpone":
            for i, ai in enumerate(self.a):
                enforce_qsparsity(ai, [self.qbonds[i], self.qsite, -self.qbonds[i+1]])

    @classmethod
    def construct_random(cls, nsites, qsite, qnum_sector, max_vdim=256, dtype="complex",
                         rng=None):
        """Random MPS in a fixed quantum-number sector (mps.py:73-113): bond
        quantum numbers are all combinations from the left (resp. right) end,
        randomly subsampled (without replacement) to `max_vdim`."""
        assert nsites > 0
        qsite = np.asarray(qsite)
        if rng is None:
            rng = np.random.default_rng()
        qbonds = [None] * (nsites + 1)
        qbonds[0] = [0]
        qbonds[-1] = [qnum_sector]
        for l in range(1, (nsites + 1) // 2):
            full = qnumber_outer_sum([qbonds[l - 1], qsite]).reshape(-1)
            qbonds[l] = full if len(full) <= max_vdim else \
                full[rng.choice(len(full), size=max_vdim, replace=False)]
        for l in reversed(range((nsites + 1) // 2, nsites)):
            full = qnumber_outer_sum([qbonds[l + 1], -qsite]).reshape(-1)
            qbonds[l] = full if len(full) <= max_vdim else \
                full[rng.choice(len(full), size=max_vdim, replace=False)]
        fill = "random" if dtype in (complex, "complex") else "random real"
        return cls(qsite, qbonds, fill=fill, rng=rng)

    @property
    def nsites(self):
        return len(self.a)

    @property
    def bond_dims(self):
        if not self.a:
            return []
        return [ai.shape[0] for ai in self.a] + [self.a[-1].shape[2]]

    def zero_qnumbers(self):
        self.qsite.fill(0)
        for qb in self.qbonds:
            qb.fill(0)
        return self

    def orthonormalize(self, mode="left"):
        """QR sweep (mps.py:141-178); returns the norm, sign fixed to >= 0."""
        if not self.a:
            return 1
        if mode == "left":
            for i in range(len(self.a) - 1):
                self.a[i], self.a[i+1], self.qbonds[i+1] = mps_local_orthonormalize_left_qr(
                    self.a[i], self.a[i+1], self.qsite, self.qbonds[i:i+2])
            self.a[-1], t, self.qbonds[-1] = mps_local_orthonormalize_left_qr(
                self.a[-1], np.ones((1, 1, 1)), self.qsite, self.qbonds[-2:])
            assert t.shape == (1, 1, 1)
            nrm = t[0, 0, 0].real
            if nrm < 0:
                self.a[-1] = -self.a[-1]
                nrm = -nrm
            return nrm
        if mode == "right":
            for i in reversed(range(1, len(self.a))):
                self.a[i], self.a[i-1], self.qbonds[i] = mps_local_orthonormalize_right_qr(
                    self.a[i], self.a[i-1], self.qsite, self.qbonds[i:i+2])
            self.a[0], t, self.qbonds[0] = mps_local_orthonormalize_right_qr(
                self.a[0], np.ones((1, 1, 1)), self.qsite, self.qbonds[:2])
            assert t.shape == (1, 1, 1)
            nrm = t[0, 0, 0].real
            if nrm < 0:
                self.a[0] = -self.a[0]
                nrm = -nrm
            return nrm
        raise ValueError(f"`mode` = {mode} invalid; must be \"left\" or \"right\".")

    def to_vector(self):
        """Full Hilbert-space vector by successive pair merges (mps.py:292-301)."""
        v = self.a[0]
        for ai in self.a[1:]:
            v = mps_merge_tensor_pair(v, ai)
        assert v.ndim == 3 and v.shape[0] == 1 and v.shape[2] == 1
        return v.reshape(-1)


class MPO:
    """Matrix product operator with tensors (b[i], d, d, b[i+1]) (mpo.py:15-73);
    the oracle only needs the container."""

    def __init__(self, qsite, qbonds, a):
        self.qsite = np.array(qsite)
        self.qbonds = [np.array(qb) for qb in qbonds]
        self.a = [np.asarray(x) for x in a]

    @property
    def nsites(self):
        return len(self.a)

    @property
    def bond_dims(self):
        return [ai.shape[0] for ai in self.a] + [self.a[-1].shape[3]]

    def to_matrix(self):
        """Dense matrix on the full Hilbert space by pair merges (mpo.py:225-236)."""
        op = self.a[0]
        for ai in self.a[1:]:
            op = mpo_merge_tensor_pair(op, ai)
        assert op.ndim == 4 and op.shape[0] == 1 and op.shape[3] == 1
        return op.reshape(op.shape[1], op.shape[2])


def _mps_step_right(a, b, r):
    """<b|a> transfer step from the right (mps.py:352-381)."""
    t = np.tensordot(a, r, 1)
    return np.tensordot(t, b.conj(), axes=((1, 2), (1, 2)))


def mps_vdot(chi, psi):
    """<chi|psi>, conjugating chi (mps.py:430-450)."""
    assert psi.nsites == chi.nsites
    if psi.nsites == 0:
        return 0
    t = np.identity(psi.a[-1].shape[2], dtype=psi.a[-1].dtype)
    for i in reversed(range(psi.nsites)):
        t = _mps_step_right(psi.a[i], chi.a[i], t)
    assert t.shape == (1, 1)
    return t[0, 0]


def mps_norm(psi):
    """sqrt(Re <psi|psi>) (mps.py:453-457)."""
    return np.sqrt(mps_vdot(psi, psi).real)


def mps_local_orthonormalize_left_qr(a, a_next, qsite, qbonds):
    """QR of a reshaped to (b0*d, b1); R is absorbed into a_next (mps.py:460-473)."""
    s = a.shape
    assert len(s) == 3
    q, r, qbond = block_sparse_qr(a.reshape(s[0] * s[1], s[2]),
                                  qnumber_flatten((qbonds[0], qsite)), qbonds[1])
    return q.reshape(s[0], s[1], q.shape[1]), np.tensordot(r, a_next, (1, 0)), qbond


def mps_local_orthonormalize_right_qr(a, a_prev, qsite, qbonds):
    """QR of the bond-flipped tensor; R^T is absorbed into a_prev (mps.py:476-491)."""
    a = a.transpose((2, 1, 0))
    s = a.shape
    assert len(s) == 3
    q, r, qbond = block_sparse_qr(a.reshape(s[0] * s[1], s[2]),
                                  qnumber_flatten((-qbonds[1], qsite)), -qbonds[0])
    a = q.reshape(s[0], s[1], q.shape[1]).transpose((2, 1, 0))
    return a, np.tensordot(a_prev, r, (2, 1)), -qbond


def mps_merge_tensor_pair(a0, a1):
    """Theta[l, s0*d1+s1, r] = sum_k a0[l,s0,k] a1[k,s1,r] (mps.py:528-535)."""
    t = np.einsum(a0, (0, 1, 4), a1, (4, 2, 3), (0, 1, 2, 3), optimize=True)
    return t.reshape(t.shape[0], t.shape[1] * t.shape[2], t.shape[3])


def mps_split_tensor_svd(a, qsite0, qsite1, qbonds_outer, svd_distr, tol=0, max_bond=None):
    """SVD split of (b0, d0*d1, b2) into (b0, d0, k) and (k, d1, b2), singular
    values multiplied into the left or right factor (mps.py:538-568).
    `max_bond` is the bond-cap extension (None reproduces the reference)."""
    assert a.ndim == 3
    qsite0 = np.asarray(qsite0)
    qsite1 = np.asarray(qsite1)
    d0, d1 = len(qsite0), len(qsite1)
    assert d0 * d1 == a.shape[1]
    b0, b2 = a.shape[0], a.shape[2]
    q0 = qnumber_flatten([qbonds_outer[0], qsite0])
    q1 = qnumber_flatten([-qsite1, qbonds_outer[1]])
    u, sigma, v, qbond = split_block_sparse_matrix_svd(
        a.reshape(b0 * d0, d1 * b2), q0, q1, tol, max_bond)
    a0 = u.reshape(b0, d0, len(sigma))
    a1 = v.reshape(len(sigma), d1, b2)
    if svd_distr == "left":
        a0 = a0 * sigma
    elif svd_distr == "right":
        a1 = a1 * sigma[:, None, None]
    elif svd_distr == "sqrt":
        sq = np.sqrt(sigma)
        a0 = a0 * sq
        a1 = a1 * sq[:, None, None]
    else:
        raise ValueError('`svd_distr` parameter must be "left", "right" or "sqrt".')
    return a0, a1, qbond


def mpo_merge_tensor_pair(a0, a1):
    """Two-site MPO tensor (Ml, d0*d1, d0*d1, Mr), output pair (s0o*d1+s1o),
    input pair (s0i*d1+s1i) (mpo.py:314-322)."""
    t = np.einsum(a0, (0, 1, 3, 6), a1, (6, 2, 4, 5), (0, 1, 2, 3, 4, 5), optimize=True)
    s = t.shape
    return t.reshape(s[0], s[1] * s[2], s[3] * s[4], s[5])
