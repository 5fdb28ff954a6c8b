"""
Matrix product state container and the tensor-pair / orthonormalisation
operations of the sweep (reference pytenet/mps.py:14-178,292-301,430-491,528-568).

Site tensors a[i] have shape (b[i], d, b[i+1]) in row-major layout. They may
be NumPy arrays (host) or complex128 torch tensors on the GPU; ``to_device()``
moves them once, and the sweep drivers keep them resident for a whole call.
Every numerical operation below runs on the GPU; a function given NumPy
arrays uploads them and returns NumPy arrays, so reference-style code keeps
working unchanged.
"""

import numpy as np
import torch

from . import backend
from .block_sparse_util import (qnumber_outer_sum, qnumber_flatten, enforce_qsparsity,
                                block_sparse_qr, block_sparse_svd)
from .bond_ops import retained_bond_indices, cap_bond_indices
from .util import crandn

__all__ = ["MPS", "mps_vdot", "mps_norm", "mps_merge_tensor_pair", "mps_split_tensor_svd",
           "mps_local_orthonormalize_left_qr", "mps_local_orthonormalize_right_qr"]


def _is_host(x):
    return isinstance(x, np.ndarray) or not isinstance(x, torch.Tensor)


def _like(x, host):
    return backend.to_host(x) if host else x


def _dev_real(v, dev):
    return torch.as_tensor(np.ascontiguousarray(v, dtype=np.float64), device=dev)


def _negate(t):
    """-t for a complex128 device tensor (sign fix of the norm), on the GPU."""
    t = t.contiguous().clone()
    backend.scale_axis(t, _dev_real(-np.ones(t.shape[0]), t.device), 0)
    return t


class MPS:
    """Matrix product state (mps.py:14-71). Quantum numbers are additive
    integers: entry a[i][l, s, r] may be non-zero only if
    qbonds[i][l] + qsite[s] - qbonds[i+1][r] == 0."""

    def __init__(self, qsite, qbonds, fill=0.0, rng=None):
        self.qsite = np.asarray(qsite)
        self.qbonds = [np.asarray(qb) for qb in qbonds]
        d = len(qsite)
        b = [len(qb) for qb in qbonds]
        assert b[0] == 1 and b[-1] == 1
        shapes = [(b[i], d, b[i + 1]) for i in range(len(b) - 1)]
        if isinstance(fill, (int, float, complex)):
            self.a = [np.full(s, fill) for s in shapes]
        elif fill == "random":
            # one complex normal draw per site, in site order (mps.py:51-56),
            # on the host so that a seed reproduces the reference bit for bit
            if rng is None:
                rng = np.random.default_rng()
            self.a = [crandn(s, rng) / np.sqrt(s[0] * s[1] * s[2]) for s in shapes]
        elif fill == "random real":
            if rng is None:
                rng = np.random.default_rng()
            self.a = [rng.normal(size=s) / np.sqrt(s[0] * s[1] * s[2]) for s in shapes]
        elif fill == "postpone":
            self.a = [None] * len(shapes)
        else:
            raise ValueError(f'`fill` = {fill} invalid; must be a number, '
                             f'"random", "random real" or "postpone".')
        if fill != "postpone":
            for i, ai in enumerate(self.a):
                enforce_qsparsity(ai, [self.qbonds[i], self.qsite, -self.qbonds[i + 1]])

    @classmethod
    def construct_random(cls, nsites: int, qsite, qnum_sector: int, max_vdim: int = 256,
                         dtype="complex", rng: np.random.Generator = None):
        """Random MPS in a quantum-number sector (mps.py:73-113)."""
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

    @classmethod
    def padded_product_state(cls, qsite, nsites: int, max_vdim: int,
                             rng: np.random.Generator = None):
        """Random product state embedded in an MPS of bond dimensions
        min(max_vdim, d^i, d^(L-i)) (zero quantum numbers): site i holds one
        complex normal vector of length d, normalised, at a[i][0, :, 0]; every
        other entry is zero. The single-site sweeps of such a state reach the
        whole bond dimension through the QR completion of its rank-deficient
        tensors. Draws one complex vector per site in site order."""
        qsite = np.asarray(qsite)
        d = len(qsite)
        if rng is None:
            rng = np.random.default_rng()
        bonds = [min(max_vdim, d**min(i, nsites - i, 40)) for i in range(nsites + 1)]
        psi = cls(np.zeros_like(qsite), [np.zeros(b, dtype=int) for b in bonds], fill=0.0)
        for i in range(nsites):
            v = crandn(d, rng)
            psi.a[i] = np.zeros((bonds[i], d, bonds[i + 1]), dtype=complex)
            psi.a[i][0, :, 0] = v / np.linalg.norm(v)
        return psi

    @property
    def nsites(self) -> int:
        return len(self.a)

    @property
    def bond_dims(self) -> list:
        if len(self.a) == 0:
            return []
        return [int(ai.shape[0]) for ai in self.a] + [int(self.a[-1].shape[2])]

    @property
    def on_device(self) -> bool:
        return len(self.a) > 0 and isinstance(self.a[0], torch.Tensor)

    def to_device(self):
        """Upload all site tensors (complex128) to the current GPU; returns self."""
        self.a = [backend.as_dev(ai) for ai in self.a]
        return self

    def to_host(self):
        """Copy all site tensors back to NumPy arrays; returns self."""
        self.a = [backend.to_host(ai) for ai in self.a]
        return self

    def zero_qnumbers(self):
        """Set all quantum numbers to zero; returns self (mps.py:131-139)."""
        self.qsite.fill(0)
        for qb in self.qbonds:
            qb.fill(0)
        return self

    def orthonormalize(self, mode="left"):
        """Left- or right-orthonormalise by QR sweeps; returns the norm with the
        sign fixed to >= 0 (mps.py:141-178)."""
        if len(self.a) == 0:
            return 1
        host = _is_host(self.a[0])
        if host:
            self.to_device()
        one = torch.ones((1, 1, 1), dtype=torch.complex128, device=self.a[0].device)
        if mode == "left":
            for i in range(len(self.a) - 1):
                self.a[i], self.a[i + 1], self.qbonds[i + 1] = mps_local_orthonormalize_left_qr(
                    self.a[i], self.a[i + 1], self.qsite, self.qbonds[i:i + 2])
            self.a[-1], t, self.qbonds[-1] = mps_local_orthonormalize_left_qr(
                self.a[-1], one, self.qsite, self.qbonds[-2:])
            assert tuple(t.shape) == (1, 1, 1)
            nrm = float(t[0, 0, 0].real.item())
            if nrm < 0:
                self.a[-1] = _negate(self.a[-1])
                nrm = -nrm
        elif mode == "right":
            for i in reversed(range(1, len(self.a))):
                self.a[i], self.a[i - 1], self.qbonds[i] = mps_local_orthonormalize_right_qr(
                    self.a[i], self.a[i - 1], self.qsite, self.qbonds[i:i + 2])
            self.a[0], t, self.qbonds[0] = mps_local_orthonormalize_right_qr(
                self.a[0], one, self.qsite, self.qbonds[:2])
            assert tuple(t.shape) == (1, 1, 1)
            nrm = float(t[0, 0, 0].real.item())
            if nrm < 0:
                self.a[0] = _negate(self.a[0])
                nrm = -nrm
        else:
            raise ValueError(f'`mode` = {mode} invalid; must be "left" or "right".')
        if host:
            self.to_host()
        return nrm

    def to_vector(self):
        """Vector on the full Hilbert space (mps.py:292-301); for small chains."""
        host = _is_host(self.a[0])
        v = backend.as_dev(self.a[0])
        for ai in self.a[1:]:
            v = backend.merge_pair(v, ai)
        assert v.dim() == 3 and v.shape[0] == 1 and v.shape[2] == 1
        return _like(v.reshape(-1), host)


def _step_right(a, b, r):
    """<b|a> transfer step from the right (mps.py:352-381)."""
    Dl, d, Dr = a.shape
    t = backend.matmul(a.reshape(Dl * d, Dr), r).reshape(Dl, d * r.shape[1])
    bl = b.shape[0]
    return backend.matmul(t, b.reshape(bl, -1).transpose(0, 1), conjB=True)


def mps_vdot(chi: MPS, psi: MPS):
    """<chi|psi>, conjugating chi (mps.py:430-450)."""
    assert psi.nsites == chi.nsites
    if psi.nsites == 0:
        return 0
    a = [backend.as_dev(x) for x in psi.a]
    b = [backend.as_dev(x) for x in chi.a]
    n = a[-1].shape[2]
    t = torch.eye(n, dtype=torch.complex128, device=a[-1].device)
    for i in reversed(range(psi.nsites)):
        t = _step_right(a[i], b[i], t)
    assert tuple(t.shape) == (1, 1)
    return complex(t[0, 0].item())


def mps_norm(psi: MPS):
    """sqrt(Re <psi|psi>) (mps.py:453-457)."""
    return float(np.sqrt(mps_vdot(psi, psi).real))


def mps_local_orthonormalize_left_qr(a, a_next, qsite, qbonds):
    """QR of a reshaped to (b0*d, b1); R is multiplied into a_next (mps.py:460-473)."""
    host = _is_host(a)
    a = backend.as_dev(a)
    a_next = backend.as_dev(a_next)
    s = a.shape
    assert len(s) == 3
    q0 = qnumber_flatten((qbonds[0], qsite))
    q, r, qbond = block_sparse_qr(a.reshape(s[0] * s[1], s[2]), q0, qbonds[1])
    a = q.reshape(s[0], s[1], q.shape[1])
    sn = a_next.shape
    a_next = backend.matmul(r, a_next.reshape(sn[0], sn[1] * sn[2])).reshape(r.shape[0], sn[1], sn[2])
    return _like(a, host), _like(a_next, host), qbond


def mps_local_orthonormalize_right_qr(a, a_prev, qsite, qbonds):
    """QR of the bond-flipped tensor; R^T is multiplied into a_prev (mps.py:476-491)."""
    host = _is_host(a)
    a = backend.as_dev(a).permute(2, 1, 0).contiguous()
    a_prev = backend.as_dev(a_prev)
    s = a.shape
    q0 = qnumber_flatten((-np.asarray(qbonds[1]), qsite))
    q, r, qbond = block_sparse_qr(a.reshape(s[0] * s[1], s[2]), q0, -np.asarray(qbonds[0]))
    a = q.reshape(s[0], s[1], q.shape[1]).permute(2, 1, 0).contiguous()
    sp = a_prev.shape
    a_prev = backend.matmul(a_prev.reshape(sp[0] * sp[1], sp[2]), r.transpose(0, 1))
    a_prev = a_prev.reshape(sp[0], sp[1], r.shape[0])
    return _like(a, host), _like(a_prev, host), -qbond


def mps_merge_tensor_pair(a0, a1):
    """Theta[l, s0*d1+s1, r] = sum_k a0[l,s0,k] a1[k,s1,r] (mps.py:528-535)."""
    host = _is_host(a0)
    return _like(backend.merge_pair(a0, a1), host)


def mps_split_tensor_svd(a, qsite0, qsite1, qbonds_outer, svd_distr: str, tol=0, max_bond=None,
                         svd_precond32=False):
    """Split (b0, d0*d1, b2) into (b0, d0, k) and (k, d1, b2) by a block-sparse
    SVD with discarded-weight truncation; sigma multiplied into the left or right
    factor, or its square root into both (mps.py:538-568). Extensions: `max_bond`
    caps the new bond (None reproduces the reference); `svd_precond32` starts the
    Jacobi SVD with FP32 sweeps (FP64 output accuracy)."""
    host = _is_host(a)
    a = backend.as_dev(a)
    assert a.dim() == 3
    qsite0 = np.asarray(qsite0)
    qsite1 = np.asarray(qsite1)
    d0, d1 = len(qsite0), len(qsite1)
    assert d0 * d1 == a.shape[1], "physical dimension of MPS tensor must be equal to d0 * d1"
    if svd_distr not in ("left", "right", "sqrt"):
        raise ValueError('`svd_distr` parameter must be "left", "right" or "sqrt".')
    b0, b2 = a.shape[0], a.shape[2]
    q0 = qnumber_flatten([qbonds_outer[0], qsite0])
    q1 = qnumber_flatten([-qsite1, qbonds_outer[1]])
    if svd_distr != "sqrt" and not np.any(q0) and not np.any(q1):
        # one sector: the truncating split computes only the isometric factor's
        # singular vectors and projects onto them (same kept subspace and weights)
        a0, a1, s = backend.svd_split(a.reshape(b0 * d0, d1 * b2), svd_distr, tol, max_bond,
                                      svd_precond32)
        k = a0.shape[1]
        q = np.zeros(k, dtype=np.asarray(q0).dtype)
        return (_like(a0.reshape(b0, d0, k), host), _like(a1.reshape(k, d1, b2), host), q)
    u, s, v, q = block_sparse_svd(a.reshape(b0 * d0, d1 * b2), q0, q1, svd_precond32)
    idx = cap_bond_indices(s, retained_bond_indices(s, tol), max_bond)
    k = len(idx)
    if k != len(s):
        it = torch.as_tensor(idx, device=a.device)
        u = u.index_select(1, it)
        v = v.index_select(0, it)
    s = s[idx]
    q = q[idx]
    a0 = u.contiguous().reshape(b0, d0, k)
    a1 = v.contiguous().reshape(k, d1, b2)
    if svd_distr == "left":
        backend.scale_axis(a0, _dev_real(s, a.device), 2)
    elif svd_distr == "right":
        backend.scale_axis(a1, _dev_real(s, a.device), 0)
    else:
        sq = _dev_real(np.sqrt(s), a.device)
        backend.scale_axis(a0, sq, 2)
        backend.scale_axis(a1, sq, 0)
    return _like(a0, host), _like(a1, host), q
