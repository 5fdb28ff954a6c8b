"""
Parity of the truncating two-site split (ptn_svd_split, called through the C
ABI) with the oracle's mps_split_tensor_svd (pytenet/mps.py:538-568 with
bond_ops.py:23-54): the kept count, the singular values against LAPACK, the
truncated two-site tensor a0 . a1 (gauge invariant), and the isometry of the
factor that carries no singular values. Shapes cover tall, wide and square
sector matrices, both distributions, the tolerance rule and the bond cap, the
FP32-preconditioned Jacobi iteration, inputs scaled to 1e-250 / 1e250, and the
production size of the two-site TDVP at D = 1024 (a 2048 x 2048 split with a
spectrum graded over nine decades).
"""

import numpy as np
import pytest

import oracle as orc
import pytenet_amd as ptn
from pytenet_amd import backend

pytestmark = pytest.mark.gpu


def _graded(m, n, decades, rng, k=None):
    """Random m x n complex matrix with singular values 10^(-decades * i / (k-1))."""
    q = min(m, n) if k is None else k
    u, _ = np.linalg.qr(orc.crandn((m, q), rng))
    v, _ = np.linalg.qr(orc.crandn((n, q), rng))
    s = 10.0 ** (-decades * np.arange(q) / max(q - 1, 1))
    return (u * s) @ v.conj().T, s


def _check_split(theta, distr, tol, max_bond, precond32, s_rtol=1e-9, prod_rtol=1e-11):
    m, n = theta.shape
    a0, a1, s = backend.svd_split(backend.as_dev(theta), distr, tol, max_bond, precond32)
    a0, a1 = backend.to_host(a0), backend.to_host(a1)
    s_ref = np.linalg.svd(theta, compute_uv=False)
    # every singular value within the backward-error bound of a stable SVD
    # (4 q eps sigma_max absolute: LAPACK's own values carry errors of that
    # order), the kept ones (below) also to s_rtol relative
    q = min(m, n)
    assert np.all(np.abs(s - s_ref) <= s_rtol * s_ref + 4 * q * np.finfo(float).eps * s_ref[0])
    # the same kept set as the oracle's split (retained_bond_indices + cap)
    qs0, qs1 = np.zeros(1, int), np.zeros(1, int)
    r0, r1, _ = orc.mps_split_tensor_svd(theta.reshape(m, 1, n), qs0, qs1,
                                         (np.zeros(m, int), np.zeros(n, int)), distr, tol=tol,
                                         max_bond=max_bond)
    k = r0.shape[2]
    assert a0.shape == (m, k) and a1.shape == (k, n)
    assert np.all(np.abs(s[:k] - s_ref[:k]) <= s_rtol * s_ref[:k])
    ref = r0.reshape(m, k) @ r1.reshape(k, n)
    got = a0 @ a1
    assert np.abs(got - ref).max() <= prod_rtol * np.abs(theta).max()
    iso = a0 if distr == "right" else a1.conj().T
    assert np.abs(iso.conj().T @ iso - np.eye(k)).max() < 1e-12
    return k


@pytest.mark.parametrize("m,n", [(40, 24), (24, 40), (64, 64), (130, 96), (96, 130), (1, 7), (7, 1)])
@pytest.mark.parametrize("distr", ["left", "right"])
def test_split_matches_oracle_small(m, n, distr):
    rng = np.random.default_rng(m * 100 + n)
    theta, _ = _graded(m, n, 6, rng)
    _check_split(theta, distr, 1e-10, None, False)


@pytest.mark.parametrize("distr", ["left", "right"])
@pytest.mark.parametrize("precond32", [False, True])
def test_split_tolerance_and_cap(distr, precond32):
    rng = np.random.default_rng(5)
    theta, _ = _graded(300, 260, 8, rng)
    k_tol = _check_split(theta, distr, 1e-6, None, precond32)
    assert 0 < k_tol < 260
    assert _check_split(theta, distr, 1e-6, 50, precond32) == 50
    assert _check_split(theta, distr, 0.0, 300, precond32) == 260


@pytest.mark.parametrize("scale", [1e-250, 1e250])
@pytest.mark.parametrize("m,n", [(80, 48), (48, 80)])
def test_split_extreme_scales(scale, m, n):
    """Singular values of inputs scaled to 1e-250 / 1e250 keep their relative
    accuracy (power-of-two input scaling of the QR). The kept count follows the
    reference's rule literally: the 2-norm of s underflows to 0 at 1e-250 (no
    value kept) and overflows at 1e250 (every normalised weight is 0, so only a
    negative tol keeps values), as the oracle's retained_bond_indices gives."""
    rng = np.random.default_rng(9)
    theta, _ = _graded(m, n, 4, rng)
    theta = theta * scale
    s_ref = np.linalg.svd(theta, compute_uv=False)
    qs = np.zeros(1, int)
    for distr in ("left", "right"):
        for tol in (1e-10, -1.0):
            a0, a1, s = backend.svd_split(backend.as_dev(theta), distr, tol, None, False)
            assert np.all(np.abs(s - s_ref) <= 1e-9 * s_ref)
            r0, _, _ = orc.mps_split_tensor_svd(theta.reshape(m, 1, n), qs, qs,
                                                (np.zeros(m, int), np.zeros(n, int)), distr,
                                                tol=tol)
            k = r0.shape[2]
            assert a0.shape == (m, k) and a1.shape == (k, n)
            if k == min(m, n):
                got = backend.to_host(a0) @ backend.to_host(a1)
                assert np.abs(got - theta).max() <= 1e-11 * np.abs(theta).max()
    assert (scale < 1 and k == 0) or (scale > 1 and k == min(m, n))


def test_split_of_zero_keeps_nothing():
    a0, a1, s = backend.svd_split(backend.as_dev(np.zeros((12, 8), complex)), "right", 1e-10)
    assert a0.shape == (12, 0) and a1.shape == (0, 8) and np.all(s == 0)


def test_split_nonfinite_raises():
    theta = np.ones((20, 16), complex)
    theta[3, 4] = np.nan
    with pytest.raises(FloatingPointError):
        backend.svd_split(backend.as_dev(theta), "right", 1e-10)


@pytest.mark.parametrize("precond32", [False, True])
@pytest.mark.parametrize("distr", ["left", "right"])
def test_split_production_size_2048(precond32, distr):
    """The split of the two-site TDVP at D = 1024: a 2048 x 2048 matrix with
    a spectrum graded over nine decades, tol 1e-10 and the cap 1024; the kept
    singular values against LAPACK to 1e-9 relative (all of them to the
    backward-error bound), orthogonality 1e-12, and the truncated product
    against the oracle's split."""
    rng = np.random.default_rng(2048 + precond32)
    theta, _ = _graded(2048, 2048, 9, rng)
    k = _check_split(theta, distr, 1e-10, 1024, precond32, s_rtol=1e-9, prod_rtol=1e-11)
    assert k == 1024


def test_mps_split_dense_uses_split_and_matches_block_sparse_path():
    """mps_split_tensor_svd with zero quantum numbers (the truncating split)
    equals the oracle's split of the same two-site tensor."""
    rng = np.random.default_rng(3)
    a = orc.crandn((12, 4, 10), rng)
    qs = np.zeros(2, int)
    qb = (np.zeros(12, int), np.zeros(10, int))
    for distr in ("left", "right"):
        g0, g1, gq = ptn.mps_split_tensor_svd(a, qs, qs, qb, distr, tol=1e-3, max_bond=9)
        r0, r1, rq = orc.mps_split_tensor_svd(a, qs, qs, qb, distr, tol=1e-3, max_bond=9)
        assert g0.shape == r0.shape and g1.shape == r1.shape and np.array_equal(gq, rq)
        got = np.einsum("lsk,ktr->lstr", g0, g1)
        ref = np.einsum("lsk,ktr->lstr", r0, r1)
        assert np.abs(got - ref).max() < 1e-12 * np.abs(a).max()
