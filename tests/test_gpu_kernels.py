"""
Parity of the individual GPU kernels (called through the C ABI) with the
oracle on the same seeded inputs. FP64 mode: relative error <= 1e-12 of the
result's max norm; FP32 mode: <= 2e-5; split-bf16 mode (three bf16 products
per real product, about 16 significant bits): <= 2e-4, which a plain bf16
product (8 bits, errors of order 4e-3) would not meet.
"""

import numpy as np
import pytest
import torch

import oracle as orc
import pytenet_amd as ptn
from conftest import load_fixture
from pytenet_amd import backend, _lib
from pytenet_amd.hamiltonian import heisenberg_xxz_1d_mpo

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 2e-5
BF16X3_TOL = 2e-4


def relerr(got, ref):
    got = backend.to_host(got)
    ref = np.asarray(ref)
    return np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-300)


def dev(x):
    return backend.as_dev(x)


@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (7, 5, 3), (33, 65, 17), (130, 70, 9), (64, 128, 64),
                                   (257, 129, 200)])
@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
def test_zgemm_strides_and_ragged_shapes(M, N, K, ta, tb):
    rng = np.random.default_rng(M * 1000 + N + K)
    A = orc.crandn((K, M) if ta else (M, K), rng)
    B = orc.crandn((N, K) if tb else (K, N), rng)
    Ad, Bd = dev(A), dev(B)
    opA = Ad.transpose(0, 1) if ta else Ad
    opB = Bd.transpose(0, 1) if tb else Bd
    refA = A.T if ta else A
    refB = B.T if tb else B
    assert relerr(backend.matmul(opA, opB), refA @ refB) < FP64_TOL
    assert relerr(backend.matmul(opA, opB, conjA=True, conjB=True),
                  refA.conj() @ refB.conj()) < FP64_TOL


@pytest.mark.parametrize("M,N,K", [(300, 200, 64), (1024, 768, 256)])
def test_autotuned_products_match(M, N, K):
    """With the tile autotuner on, products (first call: tuning; second call:
    the cached choice) equal those of the fixed tile choice."""
    rng = np.random.default_rng(M + N)
    A, B = dev(orc.crandn((M, K), rng)), dev(orc.crandn((K, N), rng))
    a, w, l, r = _heff_problem(3, Dl=40, d=4, Dr=36, M=5)
    ref = backend.to_host(backend.matmul(A, B))
    ref32 = backend.to_host(backend.heff_apply(a, w, l, r, "fp32"))
    prev = backend.set_autotune(True)
    try:
        for _ in range(2):
            assert relerr(backend.matmul(A, B), ref) < 1e-14
            assert relerr(backend.heff_apply(a, w, l, r, "fp32"), ref32) < 1e-6
    finally:
        backend.set_autotune(prev)
    assert backend.set_autotune(prev) == prev


def test_zgemm_accumulate():
    rng = np.random.default_rng(2)
    A, B, C = orc.crandn((40, 30), rng), orc.crandn((30, 50), rng), orc.crandn((40, 50), rng)
    Cd = dev(C)
    backend.matmul(dev(A), dev(B), out=Cd, accumulate=True)
    assert relerr(Cd, C + A @ B) < FP64_TOL


def test_local_kernels_match_reference_fixture():
    z = load_fixture("kernels_dense")
    a, b, w, l, r, c, rb = (z[k] for k in ("a", "b", "w", "l", "r", "c", "rb"))
    assert relerr(backend.heff_apply(a, w, l, r), z["heff"]) < FP64_TOL
    assert relerr(backend.bond_apply(c, l, rb), z["bond"]) < FP64_TOL
    assert relerr(backend.env_left(a, b, w, l), z["env_left"]) < FP64_TOL
    assert relerr(backend.env_right(a, b, w, r), z["env_right"]) < FP64_TOL


SHAPES = [(1, 2, 1, 1, 4), (3, 2, 5, 4, 5), (17, 2, 23, 5, 5), (40, 4, 36, 5, 5),
          (64, 2, 64, 3, 3), (100, 4, 90, 5, 4)]


@pytest.mark.parametrize("Dl,d,Dr,Ml,Mr", SHAPES)
@pytest.mark.parametrize("precision,tol", [("fp64", FP64_TOL), ("fp32", FP32_TOL),
                                           ("bf16x3", BF16X3_TOL), ("bf16x6", FP32_TOL)])
def test_local_kernels_random(Dl, d, Dr, Ml, Mr, precision, tol):
    rng = np.random.default_rng(Dl * 7 + Dr)
    a = orc.crandn((Dl, d, Dr), rng)
    b = orc.crandn((Dl, d, Dr), rng)
    w = orc.crandn((Ml, d, d, Mr), rng)
    l = orc.crandn((Dl, Ml, Dl), rng)
    r = orc.crandn((Dr, Mr, Dr), rng)
    c = orc.crandn((Dl, Dr), rng)
    rm = orc.crandn((Dr, Ml, Dr), rng)
    assert relerr(backend.heff_apply(a, w, l, r, precision),
                  orc.apply_local_hamiltonian(a, w, l, r)) < tol
    assert relerr(backend.env_left(a, b, w, l, precision),
                  orc.contraction_operator_step_left(a, b, w, l)) < tol
    assert relerr(backend.env_right(a, b, w, r, precision),
                  orc.contraction_operator_step_right(a, b, w, r)) < tol
    assert relerr(backend.bond_apply(c, l, rm, precision),
                  orc.apply_local_bond_contraction(c, l, rm)) < tol


def _hermitian_env(rng, D, M):
    """Environment blocks of a physical chain are Hermitian in (ket, bra) for each m
    when the MPO is Hermitian; random Hermitian blocks make H_eff Hermitian."""
    x = orc.crandn((D, M, D), rng)
    return 0.5 * (x + x.transpose(2, 1, 0).conj())


def _heff_problem(seed, Dl=12, d=2, Dr=10, M=3):
    rng = np.random.default_rng(seed)
    w = orc.crandn((M, d, d, M), rng)
    # Hermitian physical blocks: W[m, s', s, m'] = conj(W[m, s, s', m'])
    w = 0.5 * (w + w.transpose(0, 2, 1, 3).conj())
    l = _hermitian_env(rng, Dl, M)
    r = _hermitian_env(rng, Dr, M)
    a = orc.crandn((Dl, d, Dr), rng)
    return a, w, l, r


def test_heff_is_hermitian_for_hermitian_blocks():
    a, w, l, r = _heff_problem(1)
    rng = np.random.default_rng(9)
    b = orc.crandn(a.shape, rng)
    ha = backend.to_host(backend.heff_apply(a, w, l, r))
    hb = backend.to_host(backend.heff_apply(b, w, l, r))
    assert abs(np.vdot(b, ha) - np.vdot(hb, a)) < 1e-12 * abs(np.vdot(b, ha))


@pytest.mark.parametrize("numiter", [1, 4, 8, 25])
def test_krylov_expm_matches_oracle(numiter):
    a, w, l, r = _heff_problem(numiter)
    dt = -0.05j + 0.01
    out, _, info = backend.krylov_local(_lib.PTN_OP_HEFF, a, l, w, r, numiter, "expm", dt)
    ref = orc.expm_krylov(lambda x: orc.apply_local_hamiltonian(x.reshape(a.shape), w, l, r).reshape(-1),
                          a.reshape(-1), dt, numiter, hermitian=True).reshape(a.shape)
    assert relerr(out, ref) < 1e-11
    inf = info.cpu().numpy()
    assert inf[0] == numiter and inf[1] == 0 and inf[2] == 0


def test_krylov_eigh_matches_oracle():
    a, w, l, r = _heff_problem(5)
    out, en, info = backend.krylov_local(_lib.PTN_OP_HEFF, a, l, w, r, 12, "eigh")
    ev, u = orc.eigh_krylov(lambda x: orc.apply_local_hamiltonian(x.reshape(a.shape), w, l, r).reshape(-1),
                            a.reshape(-1), 12, 1)
    assert abs(float(en.item()) - ev[0]) < 1e-11 * abs(ev[0])
    got = backend.to_host(out).reshape(-1)
    ref = u[:, 0]
    # eigenvector sign is a convention of the tridiagonal solver
    phase = np.vdot(ref, got)
    assert abs(abs(phase) - 1) < 1e-10
    assert np.abs(got - phase * ref).max() < 1e-9


def test_krylov_bond_expm_matches_oracle():
    rng = np.random.default_rng(31)
    D1, D2, M = 9, 11, 4
    l = _hermitian_env(rng, D1, M)
    r = _hermitian_env(rng, D2, M)
    c = orc.crandn((D1, D2), rng)
    out, _, _ = backend.krylov_local(_lib.PTN_OP_BOND, c, l, None, r, 6, "expm", 0.025j)
    ref = orc.expm_krylov(lambda x: orc.apply_local_bond_contraction(x.reshape(c.shape), l, r).reshape(-1),
                          c.reshape(-1), 0.025j, 6, hermitian=True).reshape(c.shape)
    assert relerr(out, ref) < 1e-11


def test_krylov_breakdown_on_invariant_subspace():
    """A start vector in a 2-dimensional invariant subspace of H_eff stops after 2 vectors."""
    Dl, d, Dr = 1, 4, 1
    w = np.zeros((1, d, d, 1), dtype=complex)
    w[0, :, :, 0] = np.diag([1.0, 2.0, 3.0, 4.0])
    one = np.ones((1, 1, 1))
    a = np.array([1.0, 1.0, 0.0, 0.0], dtype=complex).reshape(Dl, d, Dr)
    out, _, info = backend.krylov_local(_lib.PTN_OP_HEFF, a, one, w, one, 4, "expm", -0.3j)
    inf = info.cpu().numpy()
    assert inf[0] == 2 and inf[1] == 2
    with pytest.warns(RuntimeWarning):
        ref = orc.expm_krylov(lambda x: (np.diag([1., 2, 3, 4]) @ x), a.reshape(-1), -0.3j, 4)
    assert relerr(out.reshape(-1), ref) < 1e-12
    with pytest.warns(RuntimeWarning, match="beta\\[1\\]"):
        backend.report_krylov_status(inf)


def test_krylov_zero_start_vector_raises():
    a, w, l, r = _heff_problem(3)
    _, _, info = backend.krylov_local(_lib.PTN_OP_HEFF, np.zeros_like(a), l, w, r, 4, "expm", 0.1)
    with pytest.raises(AssertionError):
        backend.report_krylov_status(info.cpu().numpy())


def test_krylov_nonfinite_input_raises():
    a, w, l, r = _heff_problem(4)
    a = a.copy()
    a[0, 0, 0] = np.nan
    _, _, info = backend.krylov_local(_lib.PTN_OP_HEFF, a, l, w, r, 4, "expm", 0.1)
    with pytest.raises(FloatingPointError):
        backend.report_krylov_status(info.cpu().numpy())


@pytest.mark.parametrize("m,n", [(1, 1), (5, 3), (3, 5), (16, 8), (33, 33), (96, 40), (40, 96)])
def test_qr_matches_lapack(m, n):
    rng = np.random.default_rng(m * 100 + n)
    a = orc.crandn((m, n), rng)
    q, r = backend.qr(dev(a))
    qr_, rr_ = np.linalg.qr(a)
    assert relerr(q, qr_) < 1e-12
    assert relerr(r, rr_) < 1e-12


@pytest.mark.parametrize("scale", [1e-250, 1e250])
def test_qr_and_svd_of_extreme_scales_match_lapack(scale):
    """Entries whose squares under- or overflow: the factorisations scale the
    input by a power of two first, as LAPACK's scaled norms would."""
    rng = np.random.default_rng(3)
    a = orc.crandn((40, 24), rng) * scale
    q, r = backend.qr(dev(a))
    q0, r0 = np.linalg.qr(a)
    assert relerr(q, q0) < 1e-12 and relerr(r, r0) < 1e-12
    qb, rb = backend.qr_batched(np.stack([a, a / scale]))
    assert relerr(qb[0], q0) < 1e-12 and relerr(rb[0], r0) < 1e-12
    assert relerr(qb[1], q0) < 1e-12 and relerr(rb[1], r0 / scale) < 1e-12
    for precond32 in (False, True):
        b = orc.crandn((160, 150), rng) * scale
        u, s, vh = backend.svd(dev(b), precond32=precond32)
        s = s.cpu().numpy()
        s0 = np.linalg.svd(b, compute_uv=False)
        np.testing.assert_allclose(s, s0, rtol=1e-11)
        uh, vhh = backend.to_host(u), backend.to_host(vh)
        np.testing.assert_allclose(uh.conj().T @ uh, np.identity(150), atol=1e-12)
        assert np.abs((uh * s) @ vhh - b).max() < 1e-12 * s0[0]


def test_orthonormalize_of_a_tiny_state_matches_oracle():
    """A chain whose norm (about 1e-288) squares to below the double range."""
    rng = np.random.default_rng(8)
    L, D = 24, 6
    qs = np.array([0, 0])
    qb = [np.zeros(min(D, 2**i, 2**(L - i)), dtype=int) for i in range(L + 1)]
    ref = orc.MPS(qs, qb, fill="random", rng=rng)
    ref.a = [x * 1e-12 for x in ref.a]
    psi = ptn.MPS(qs, qb, fill=0.0)
    psi.a = [x.copy() for x in ref.a]
    nrm = psi.orthonormalize("right")
    nrm0 = ref.orthonormalize(mode="right")
    assert 0 < nrm0 < 1e-250 and abs(nrm - nrm0) < 1e-12 * nrm0
    for x, y in zip(psi.a, ref.a):
        assert np.abs(backend.to_host(x) - y).max() < 1e-12


@pytest.mark.parametrize("precond32", [False, True])
def test_svd_nonfinite_input_raises(precond32):
    """A NaN or an infinity in the input is reported (FloatingPointError), in
    the plain and the FP32-preconditioned Jacobi SVD and in a batch."""
    rng = np.random.default_rng(4)
    for bad in (np.nan, np.inf):
        a = orc.crandn((150, 140), rng)
        a[7, 9] = bad
        with pytest.raises(FloatingPointError):
            backend.svd(dev(a), precond32=precond32)
        b = np.stack([orc.crandn((150, 140), rng), a])
        with pytest.raises(FloatingPointError):
            backend.svd_batched(b, precond32=precond32)
    # the library stays usable afterwards
    c = orc.crandn((150, 140), rng)
    _, s, _ = backend.svd(dev(c), precond32=precond32)
    np.testing.assert_allclose(s.cpu().numpy(), np.linalg.svd(c, compute_uv=False), rtol=1e-11)


def test_qr_rank_deficient_matches_lapack():
    rng = np.random.default_rng(77)
    a = orc.crandn((12, 8), rng)
    a[:, 3] = 0
    a[:, 6] = 0
    a[9:, :] = 0
    q, r = backend.qr(dev(a))
    qr_, rr_ = np.linalg.qr(a)
    assert relerr(q, qr_) < 1e-12
    assert relerr(r, rr_) < 1e-12


@pytest.mark.parametrize("m,n", [(1, 1), (6, 4), (4, 6), (24, 24), (64, 30), (30, 64), (130, 128)])
def test_svd_matches_lapack(m, n):
    rng = np.random.default_rng(m * 31 + n)
    a = orc.crandn((m, n), rng)
    # graded singular values down to 1e-9 relative
    u0, _, v0 = np.linalg.svd(a, full_matrices=False)
    k = min(m, n)
    s0 = np.logspace(0, -9, k)
    a = (u0 * s0) @ v0
    u, s, vh = backend.svd(dev(a))
    s = s.cpu().numpy()
    np.testing.assert_allclose(s, s0, rtol=1e-9, atol=1e-14)
    uh = backend.to_host(u)
    vhh = backend.to_host(vh)
    np.testing.assert_allclose(uh.conj().T @ uh, np.identity(k), atol=1e-12)
    np.testing.assert_allclose(vhh @ vhh.conj().T, np.identity(k), atol=1e-12)
    np.testing.assert_allclose((uh * s) @ vhh, a, atol=1e-13)


def test_block_sparse_factorisations_match_reference_fixture():
    z = load_fixture("kernels_blocksparse")
    q, r, qi = backend.block_sparse_qr(z["qr_a"], z["qr_q0"], z["qr_q1"])
    assert relerr(q, z["qr_q"]) < 1e-12 and relerr(r, z["qr_r"]) < 1e-12
    assert np.array_equal(qi, z["qr_qi"])
    u, s, v, qs = backend.block_sparse_svd(z["svd_a"], z["svd_q0"], z["svd_q1"])
    np.testing.assert_allclose(s, z["svd_s"], rtol=1e-12, atol=1e-14)
    assert np.array_equal(qs, z["svd_q"])
    uh, vh = backend.to_host(u), backend.to_host(v)
    np.testing.assert_allclose((uh * s) @ vh, z["svd_a"], atol=1e-13)


@pytest.mark.parametrize("m,n", [(40, 24), (24, 40), (70, 70), (33, 1)])
def test_svd_batched_matches_lapack(m, n):
    rng = np.random.default_rng(m * 7 + n)
    k = min(m, n)
    mats, svals = [], []
    for b in range(5):
        u0, _, v0 = np.linalg.svd(orc.crandn((m, n), rng), full_matrices=False)
        s0 = np.logspace(0, -(3 + 2 * b), k)
        mats.append((u0 * s0) @ v0)
        svals.append(s0)
    u, s, vh = backend.svd_batched(np.stack(mats))
    u, s, vh = backend.to_host(u), s.cpu().numpy(), backend.to_host(vh)
    for b in range(5):
        np.testing.assert_allclose(s[b], svals[b], rtol=1e-9, atol=1e-14)
        np.testing.assert_allclose(u[b].conj().T @ u[b], np.identity(k), atol=1e-12)
        np.testing.assert_allclose(vh[b] @ vh[b].conj().T, np.identity(k), atol=1e-12)
        np.testing.assert_allclose((u[b] * s[b]) @ vh[b], mats[b], atol=1e-13)


@pytest.mark.parametrize("m,n", [(130, 128), (300, 200), (200, 300), (256, 256)])
def test_svd_precond32_matches_lapack(m, n):
    """FP32 preconditioning sweeps, FP64 finish: the FP64 path's tolerances on a
    spectrum graded to 1e-9 and on an MPS-like exponentially decaying one."""
    rng = np.random.default_rng(m * 3 + n)
    k = min(m, n)
    u0, _, v0 = np.linalg.svd(orc.crandn((m, n), rng), full_matrices=False)
    for s0 in (np.logspace(0, -9, k), np.exp(-np.arange(k) / (k / 10))):
        a = (u0 * s0) @ v0
        u, s, vh, inf = backend.svd(dev(a), return_info=True, precond32=True)
        assert inf[3] > 0 and inf[1] == 0
        s = s.cpu().numpy()
        np.testing.assert_allclose(s, s0, rtol=1e-9, atol=1e-14)
        uh, vhh = backend.to_host(u), backend.to_host(vh)
        np.testing.assert_allclose(uh.conj().T @ uh, np.identity(k), atol=1e-12)
        np.testing.assert_allclose(vhh @ vhh.conj().T, np.identity(k), atol=1e-12)
        np.testing.assert_allclose((uh * s) @ vhh, a, atol=1e-13)


@pytest.mark.parametrize("precond32", [False, True])
def test_svd_batched_zero_padding_stays_exact(precond32):
    """A block zero-padded into a larger shape: the first min(m_s, n_s) triplets
    are the block's own, with exactly zero entries in the padding (also for a
    rank-deficient and an all-zero block)."""
    rng = np.random.default_rng(12)
    if precond32:
        P, Q = 200, 160
        shapes = [(200, 160), (150, 140), (90, 60), (190, 30), (100, 80), (9, 5)]
    else:
        P, Q = 50, 30
        shapes = [(50, 30), (20, 12), (7, 7), (45, 3), (20, 12), (9, 5)]
    blocks = [orc.crandn(sh, rng) for sh in shapes]
    blocks[4] = orc.crandn((shapes[4][0], 2), rng) @ orc.crandn((2, shapes[4][1]), rng)   # rank 2
    blocks[5] = np.zeros((9, 5), dtype=complex)
    pad = np.zeros((len(shapes), P, Q), dtype=complex)
    for b, (ms, ns) in enumerate(shapes):
        pad[b, :ms, :ns] = blocks[b]
    u, s, vh, inf = backend.svd_batched(pad, return_info=True, precond32=precond32)
    assert (inf[3] > 0) == precond32
    u, s, vh = backend.to_host(u), s.cpu().numpy(), backend.to_host(vh)
    for b, (ms, ns) in enumerate(shapes):
        kk = min(ms, ns)
        assert np.all(u[b, ms:, :kk] == 0) and np.all(vh[b, :kk, ns:] == 0)
        s_ref = np.linalg.svd(blocks[b], compute_uv=False)
        np.testing.assert_allclose(s[b, :kk], s_ref, rtol=1e-11, atol=1e-13)
        rec = (u[b, :ms, :kk] * s[b, :kk]) @ vh[b, :kk, :ns]
        np.testing.assert_allclose(rec, blocks[b], atol=1e-13)
        nz = s[b, :kk] > 1e-12
        uu, vv = u[b, :ms, :kk][:, nz], vh[b, :kk, :ns][nz]
        np.testing.assert_allclose(uu.conj().T @ uu, np.identity(nz.sum()), atol=1e-12)
        np.testing.assert_allclose(vv @ vv.conj().T, np.identity(nz.sum()), atol=1e-12)


def test_block_sparse_svd_many_sectors_matches_oracle():
    """Sectors of many sizes, tall and wide (several batched size classes)."""
    rng = np.random.default_rng(21)
    q0 = rng.binomial(12, 0.5, size=300) - 6
    q1 = rng.binomial(10, 0.45, size=240) - 5
    a = orc.crandn((300, 240), rng) * (q0[:, None] == q1[None, :])
    u, s, v, qs = backend.block_sparse_svd(a, q0, q1)
    u0, s0, v0, qs0 = orc.block_sparse_svd(a, q0, q1)
    assert np.array_equal(qs, qs0)
    np.testing.assert_allclose(s, s0, rtol=1e-11, atol=1e-13)
    uh, vh = backend.to_host(u), backend.to_host(v)
    mask = q0[:, None] == qs[None, :]
    assert np.all(uh[~mask] == 0) and np.all(vh[(q1[None, :] != qs[:, None])] == 0)
    np.testing.assert_allclose(uh.conj().T @ uh, np.identity(len(s)), atol=1e-12)
    np.testing.assert_allclose(vh @ vh.conj().T, np.identity(len(s)), atol=1e-12)
    np.testing.assert_allclose((uh * s) @ vh, a, atol=1e-12)


def test_qsparse_kernels():
    rng = np.random.default_rng(6)
    q0, q1, q2 = rng.integers(-2, 3, 5), np.array([1, -1]), rng.integers(-2, 3, 7)
    a = orc.crandn((5, 2, 7), rng)
    mask = np.add.outer(np.add.outer(q0, q1), q2) != 0
    t = dev(a)
    assert backend.qsparse_violation(t, [q0, q1, q2]) == pytest.approx(
        np.max(np.maximum(np.abs(a.real), np.abs(a.imag))[mask]))
    backend.enforce_qsparsity(t, [q0, q1, q2])
    ref = a.copy()
    ref[mask] = 0
    assert np.array_equal(backend.to_host(t), ref)
    assert backend.qsparse_violation(t, [q0, q1, q2]) == 0


def _sorted_bond(rng, D, spread):
    """Quantum numbers of a bond grouped into contiguous ascending sectors
    (as the bond-splitting SVD produces them), drawn around zero."""
    return np.sort(rng.binomial(2 * spread, 0.5, size=D) - spread)


def _qsparse(rng, shape, qnums):
    x = orc.crandn(shape, rng)
    orc.enforce_qsparsity(x, qnums)
    return x


@pytest.mark.parametrize("Dl,Dr,two_site", [(200, 180, False), (260, 300, True), (37, 41, True)])
@pytest.mark.parametrize("precision,tol", [("fp64", FP64_TOL), ("fp32", FP32_TOL),
                                           ("bf16x3", BF16X3_TOL), ("bf16x6", FP32_TOL)])
def test_block_sparse_contractions_match_dense(Dl, Dr, two_site, precision, tol):
    """U(1) block-sparse operands (XXZ MPO with quantum numbers, sector-sorted
    bonds): the tile-skipping contractions equal the dense ones and the oracle."""
    rng = np.random.default_rng(Dl + Dr)
    H = heisenberg_xxz_1d_mpo(4, 1.0, 0.8, -0.1)
    qd = np.asarray(H.qsite)
    w = np.asarray(H.a[1], dtype=complex)
    qm_l, qm_r = np.asarray(H.qbonds[1]), np.asarray(H.qbonds[2])
    qs = qd
    if two_site:
        w = orc.mpo_merge_tensor_pair(w, np.asarray(H.a[2], dtype=complex))
        qm_r = np.asarray(H.qbonds[3])
        qs = orc.qnumber_flatten([qd, qd])
    ql, qr = _sorted_bond(rng, Dl, 4), _sorted_bond(rng, Dr, 4)
    d = len(qs)
    a = _qsparse(rng, (Dl, d, Dr), [ql, qs, -qr])
    l = _qsparse(rng, (Dl, len(qm_l), Dl), [ql, qm_l, -ql])
    r = _qsparse(rng, (Dr, len(qm_r), Dr), [qr, qm_r, -qr])
    assert np.count_nonzero(a) < a.size // 3       # genuinely sparse operands
    dense = backend.to_host(backend.heff_apply(a, w, l, r, precision))
    sparse = backend.to_host(backend.heff_apply(a, w, l, r, precision, block_sparse=True))
    ref = orc.apply_local_hamiltonian(a, w, l, r)
    assert relerr(sparse, ref) < tol
    assert np.abs(sparse - dense).max() <= 1e-13 * np.abs(dense).max()
    assert relerr(backend.env_left(a, a, w, l, precision, block_sparse=True),
                  orc.contraction_operator_step_left(a, a, w, l)) < tol
    assert relerr(backend.env_right(a, a, w, r, precision, block_sparse=True),
                  orc.contraction_operator_step_right(a, a, w, r)) < tol
    c = _qsparse(rng, (Dl, Dl), [ql, -ql])
    rb = _qsparse(rng, (Dl, len(qm_l), Dl), [ql, qm_l, -ql])
    assert relerr(backend.bond_apply(c, l, rb, precision, block_sparse=True),
                  orc.apply_local_bond_contraction(c, l, rb)) < tol


def _chain_pair(L, D, seeds):
    out = []
    for seed in seeds:
        bonds = [min(D, 2**min(i, 30), 2**min(L - i, 30)) for i in range(L + 1)]
        out.append(ptn.MPS(np.array([0, 0]), [np.zeros(b, dtype=int) for b in bonds],
                           fill="random", rng=np.random.default_rng(seed)))
    return out


@pytest.mark.parametrize("L,D", [(8, 8), (16, 96)])
def test_mps_vdot_and_norm_match_oracle(L, D):
    """ptn.mps_vdot / mps_norm (the fidelity measure of the drivers and the
    bench) against the oracle (mps.py:430-457) at D >= 64, chi != psi, and
    against np.vdot of the full vectors where those fit."""
    chi, psi = _chain_pair(L, D, (21, 22))
    ochi = orc.MPS(chi.qsite, chi.qbonds, fill="postpone")
    ochi.a = [a.copy() for a in chi.a]
    opsi = orc.MPS(psi.qsite, psi.qbonds, fill="postpone")
    opsi.a = [a.copy() for a in psi.a]
    ref = orc.mps_vdot(ochi, opsi)
    scale = orc.mps_norm(ochi) * orc.mps_norm(opsi)
    gchi, gpsi = ptn.MPS(chi.qsite, chi.qbonds, fill="postpone"), ptn.MPS(psi.qsite, psi.qbonds,
                                                                          fill="postpone")
    gchi.a, gpsi.a = [a.copy() for a in chi.a], [a.copy() for a in psi.a]
    gchi.to_device(); gpsi.to_device()
    got = ptn.mps_vdot(gchi, gpsi)
    assert abs(got - ref) <= 1e-12 * scale
    assert abs(ptn.mps_vdot(gpsi, gchi) - np.conj(ref)) <= 1e-12 * scale
    assert abs(ptn.mps_norm(gpsi) - orc.mps_norm(opsi)) <= 1e-12 * orc.mps_norm(opsi)
    if L <= 10:
        x, y = ochi.to_vector(), opsi.to_vector()
        assert abs(got - np.vdot(x, y)) <= 1e-12 * np.linalg.norm(x) * np.linalg.norm(y)


def test_live_launch_timing_counts_products():
    """ptn_timing_begin/end: every complex-product launch between them is
    counted once per kernel instance with its algorithmic flops."""
    rng = np.random.default_rng(3)
    A, B = dev(orc.crandn((300, 70), rng)), dev(orc.crandn((70, 90), rng))
    backend.timing_begin()
    for _ in range(3):
        backend.matmul(A, B)
    rows = backend.timing_end()
    assert sum(r["launches"] for r in rows) == 3
    assert abs(sum(r["flop"] for r in rows) - 3 * 8.0 * 300 * 90 * 70) < 1
    assert all(r["total_ms"] > 0 and r["name"].startswith("zgemm_kernel<double, c128") for r in rows)
    # off again: further launches are not recorded
    backend.matmul(A, B)
    backend.timing_begin()
    assert backend.timing_end() == []


def test_callback_krylov_api_matches_oracle():
    """The callback Krylov API (lanczos_iteration / eigh_krylov / expm_krylov
    of pytenet/krylov.py:12-57,110-136) with a user matvec: equal to the
    oracle on the same operator, exact on the full space (numiter = n) against
    scipy's expm / numpy's eigh, for NumPy and device inputs."""
    from scipy.linalg import expm
    rng = np.random.default_rng(17)
    n = 12
    a = orc.crandn((n, n), rng)
    a = 0.5 * (a + a.conj().T)
    v = orc.crandn(n, rng)
    # Lanczos without reorthogonalisation amplifies rounding differences in the
    # late vectors of a full-space run, hence the looser bound there
    for numiter, vtol in ((5, 1e-12), (n, 1e-8)):
        al, be, V = ptn.lanczos_iteration(lambda x: a @ x, v, numiter)
        ral, rbe, rV = orc.lanczos_iteration(lambda x: a @ x, v, numiter)
        assert np.allclose(al, ral, rtol=vtol, atol=vtol) and np.allclose(be, rbe, rtol=vtol)
        assert np.abs(V - rV).max() < vtol
    y = ptn.expm_krylov(lambda x: a @ x, v, -0.3j, n, hermitian=True)
    assert np.abs(y - expm(-0.3j * a) @ v).max() < 1e-11
    w, u = ptn.eigh_krylov(lambda x: a @ x, v, n, 1)
    assert abs(w[0] - np.linalg.eigvalsh(a)[0]) < 1e-11
    assert abs(abs(np.vdot(u[:, 0], np.linalg.eigh(a)[1][:, 0])) - 1) < 1e-10
    ad = dev(a)
    yd = ptn.expm_krylov(lambda x: ad @ x, dev(v), -0.3j, n, hermitian=True)
    assert isinstance(yd, torch.Tensor) and np.abs(backend.to_host(yd) - expm(-0.3j * a) @ v).max() < 1e-11
    with pytest.raises(NotImplementedError):
        ptn.expm_krylov(lambda x: a @ x, v, 0.1, 4, hermitian=False)


def test_zero_width_bond_after_a_zero_split_matches_reference_semantics():
    """A split of an all-zero two-site tensor keeps no singular value
    (pytenet/bond_ops.py:28-29): the new bond has width 0, the environment
    step of the empty tensor gives the empty / zero blocks NumPy's tensordot
    gives (as the oracle), and a local Krylov step on the empty tensor raises
    AssertionError as the reference's expm_krylov does (krylov.py:30)."""
    rng = np.random.default_rng(8)
    qs = np.zeros(2, dtype=int)
    theta = np.zeros((4, 4, 6), dtype=complex)
    a0, a1, q = ptn.mps_split_tensor_svd(theta, qs, qs, (np.zeros(4, int), np.zeros(6, int)),
                                          "right", tol=1e-10)
    r0, r1, rq = orc.mps_split_tensor_svd(theta, qs, qs, (np.zeros(4, int), np.zeros(6, int)),
                                          "right", tol=1e-10)
    assert a0.shape == r0.shape == (4, 2, 0) and a1.shape == r1.shape == (0, 2, 6)
    assert len(q) == len(rq) == 0
    w = heisenberg_xxz_1d_mpo(4, 1.0, 0.8, -0.1).a[1]
    l = orc.crandn((4, w.shape[0], 4), rng)
    got = backend.env_left(a0, a0, w, l)
    ref = orc.contraction_operator_step_left(r0, r0, w, l)
    assert tuple(got.shape) == ref.shape == (0, w.shape[3], 0)
    rr = orc.crandn((6, w.shape[3], 6), rng)
    got = backend.env_right(a1, a1, w, rr)
    ref = orc.contraction_operator_step_right(r1, r1, w, rr)
    assert tuple(got.shape) == ref.shape == (0, w.shape[0], 0)
    # an empty sum: the left step through the zero-width bond gives zeros
    l0 = np.zeros((0, w.shape[0], 0), dtype=complex)
    got = backend.env_left(a1, a1, w, l0)
    ref = orc.contraction_operator_step_left(r1, r1, w, l0)
    assert tuple(got.shape) == ref.shape == (6, w.shape[3], 6)
    assert np.abs(backend.to_host(got)).max() == 0.0 and np.abs(ref).max() == 0.0
    with pytest.raises(AssertionError):
        backend.krylov_local(_lib.PTN_OP_HEFF, a1, dev(np.zeros((0, w.shape[0], 0))), dev(w), dev(rr),
                             4, "expm", 0.1j)


def test_graph_captured_solves_equal_direct_launches():
    """hipGraph capture of short local solves (ptn_set_graphs): a sequence of
    solves of one shape bucket with different operands and start vectors,
    interleaved with eager work and synchronisations, replays the cached graph
    with updated pointers; every replay equals the directly launched solve bit
    for bit, including a solve that breaks down (predicated iterations) and
    eigh solves, and the oracle to 1e-12."""
    prev = backend.set_graphs(True)
    try:
        n0 = backend.graph_launches()
        results = []
        for seed in range(6):
            a, w, l, r = _heff_problem(100 + seed, Dl=10, d=4, Dr=12, M=3)
            la, wa, ra, va = dev(l), dev(w), dev(r), dev(a)
            g_out, _, g_info = backend.krylov_local(_lib.PTN_OP_HEFF, va, la, wa, ra, 6, "expm",
                                                    -0.2j)
            g_eig, g_en, _ = backend.krylov_local(_lib.PTN_OP_HEFF, va, la, wa, ra, 6, "eigh")
            # eager work and a synchronisation between the replays
            backend.heff_apply(va, wa, la, ra)
            torch.cuda.synchronize()
            results.append((seed, a, w, l, r, backend.to_host(g_out), backend.to_host(g_info),
                            backend.to_host(g_eig), float(g_en.item())))
        assert backend.graph_launches() - n0 == 12
        # invariant-subspace breakdown through the same capture path
        wb = np.zeros((1, 4, 4, 1), dtype=complex)
        wb[0, :, :, 0] = np.diag([1.0, 2.0, 3.0, 4.0])
        one = np.ones((1, 1, 1))
        ab = np.array([1.0, 1.0, 0.0, 0.0], dtype=complex).reshape(1, 4, 1)
        gb, _, gbi = backend.krylov_local(_lib.PTN_OP_HEFF, ab, one, wb, one, 4, "expm", -0.3j)
        gb, gbi = backend.to_host(gb), backend.to_host(gbi)
        backend.set_graphs(False)
        n1 = backend.graph_launches()
        for seed, a, w, l, r, out, info, eig, en in results:
            d_out, _, d_info = backend.krylov_local(_lib.PTN_OP_HEFF, a, l, w, r, 6, "expm", -0.2j)
            d_eig, d_en, _ = backend.krylov_local(_lib.PTN_OP_HEFF, a, l, w, r, 6, "eigh")
            assert np.array_equal(out, backend.to_host(d_out)), seed
            assert np.array_equal(info, backend.to_host(d_info)), seed
            assert np.array_equal(eig, backend.to_host(d_eig)) and en == float(d_en.item()), seed
            ref = orc.expm_krylov(lambda x: orc.apply_local_hamiltonian(x.reshape(a.shape), w, l, r)
                                  .reshape(-1), a.reshape(-1), -0.2j, 6)
            assert relerr(out.reshape(-1), ref) < 1e-12
        db, _, dbi = backend.krylov_local(_lib.PTN_OP_HEFF, ab, one, wb, one, 4, "expm", -0.3j)
        assert np.array_equal(gb, backend.to_host(db)) and np.array_equal(gbi, backend.to_host(dbi))
        assert gbi[0] == 2 and gbi[1] == 2
        assert backend.graph_launches() == n1
    finally:
        backend.set_graphs(prev)


@pytest.mark.parametrize("precision", ["fp32", "bf16x3", "bf16x6"])
def test_low_precision_lanczos_breaks_down_on_invariant_subspace(precision):
    """A start vector in a 2-dimensional invariant subspace of a block-diagonal
    H_eff: in FP64 the residual after two steps is at rounding level and the
    reference's rule stops the iteration; with complex64 / split-bf16 matvecs
    the residual is the matvec's rounding noise, which the precision-aware
    rule also treats as a breakdown (instead of continuing with normalised
    noise), so the result stays within the mode's accuracy of the FP64 one."""
    rng = np.random.default_rng(21)
    blk = [orc.crandn((2, 2), rng) for _ in range(2)]
    blk = [0.5 * (b + b.conj().T) for b in blk]
    w = np.zeros((1, 4, 4, 1), dtype=complex)
    w[0, :2, :2, 0], w[0, 2:, 2:, 0] = blk
    one = np.ones((1, 1, 1))
    a = np.zeros((1, 4, 1), dtype=complex)
    a[0, :2, 0] = orc.crandn(2, rng)
    o64, _, i64 = backend.krylov_local(_lib.PTN_OP_HEFF, a, one, w, one, 4, "expm", -0.7j)
    olp, _, ilp = backend.krylov_local(_lib.PTN_OP_HEFF, a, one, w, one, 4, "expm", -0.7j,
                                       precision)
    i64, ilp = backend.to_host(i64), backend.to_host(ilp)
    assert i64[0] == 2 and i64[1] == 2
    assert ilp[0] == 2 and ilp[1] == 2
    tol = 1e-3 if precision == "bf16x3" else 1e-6
    assert relerr(olp, backend.to_host(o64)) < tol


@pytest.mark.parametrize("Dl,Dr", [(515, 509), (512, 512)])
def test_presplit_bf16x3_products_match_oracle(Dl, Dr):
    """The pre-split split-bf16 products (operands split once into bf16 hi / lo
    planes, the real embedding of B built in that pass, 256 x 256 MFMA tiles)
    at ragged and aligned shapes large enough to select them: H_eff, both
    environment steps and the bond operator against the oracle at the split-bf16
    bound; the launched kernels include the pre-split product."""
    a, w, l, r = _heff_problem(Dl + Dr, Dl=Dl, d=4, Dr=Dr, M=5)
    backend.timing_begin()
    got = backend.heff_apply(a, w, l, r, "bf16x3")
    names = [t["name"] for t in backend.timing_end()]
    assert any(n.startswith("zgemm_bf16x3_v2<") for n in names), names
    assert relerr(got, orc.apply_local_hamiltonian(a, w, l, r)) < BF16X3_TOL
    a2 = a[:, :2, :] + 0.0
    w2 = w[:, :2, :2, :] + 0.0
    assert relerr(backend.env_left(a2, a2, w2, l, "bf16x3"),
                  orc.contraction_operator_step_left(a2, a2, w2, l)) < BF16X3_TOL
    assert relerr(backend.env_right(a2, a2, w2, r, "bf16x3"),
                  orc.contraction_operator_step_right(a2, a2, w2, r)) < BF16X3_TOL
    rng = np.random.default_rng(4)
    c = orc.crandn((Dl, Dr), rng)
    assert relerr(backend.bond_apply(c, l, r, "bf16x3"),
                  orc.apply_local_bond_contraction(c, l, r)) < BF16X3_TOL


@pytest.mark.parametrize("Dl,Dr", [(515, 509), (512, 512)])
def test_presplit_bf16x6_products_have_fp32_accuracy(Dl, Dr):
    """The three-term split-bf16 products (hi / mid / lo terms interleaved per
    16 real k, six MFMA products, 256 x 256 tiles, k-chunked where the tiles
    do not fill the chip) at ragged and aligned shapes: H_eff, both environment
    steps and the bond operator against the oracle at the FP32 bound. The
    error is within twice the FP32 product's on the same operands (a dropped or
    misplaced third term would leave the two-term split's error, several times
    larger, which the comparison with bf16x3 rules out); the launched kernels
    include the three-term product."""
    a, w, l, r = _heff_problem(Dl + Dr, Dl=Dl, d=4, Dr=Dr, M=5)
    ref = orc.apply_local_hamiltonian(a, w, l, r)
    backend.timing_begin()
    got = backend.heff_apply(a, w, l, r, "bf16x6")
    names = [t["name"] for t in backend.timing_end()]
    assert any(n.startswith("zgemm_bf16x6_v3<") for n in names), names
    e6 = relerr(got, ref)
    e32 = relerr(backend.heff_apply(a, w, l, r, "fp32"), ref)
    e3 = relerr(backend.heff_apply(a, w, l, r, "bf16x3"), ref)
    assert e6 < FP32_TOL and e6 < 2 * e32 and e6 < e3 / 3, (e6, e32, e3)
    a2 = a[:, :2, :] + 0.0
    w2 = w[:, :2, :2, :] + 0.0
    assert relerr(backend.env_left(a2, a2, w2, l, "bf16x6"),
                  orc.contraction_operator_step_left(a2, a2, w2, l)) < FP32_TOL
    assert relerr(backend.env_right(a2, a2, w2, r, "bf16x6"),
                  orc.contraction_operator_step_right(a2, a2, w2, r)) < FP32_TOL
    rng = np.random.default_rng(4)
    c = orc.crandn((Dl, Dr), rng)
    assert relerr(backend.bond_apply(c, l, r, "bf16x6"),
                  orc.apply_local_bond_contraction(c, l, r)) < FP32_TOL


def test_presplit_bf16x3_inside_lanczos_matches_fp64():
    """A local solve whose matvec runs the k-chunked pre-split split-bf16
    product (one-site H_eff at D = 512: 80 output tiles, two k-chunks) under
    the Lanczos predication: finite and within the split-bf16 accuracy of the
    FP64 solve (the predication flag belongs to the problem, not to a k-chunk)."""
    a, w, l, r = _heff_problem(77, Dl=512, d=2, Dr=512, M=5)
    a = a / np.linalg.norm(a)
    prev = backend.set_graphs(False)    # direct launches, so that the launch timing sees them
    try:
        backend.timing_begin()
        o16, _, i16 = backend.krylov_local(_lib.PTN_OP_HEFF, a, l, w, r, 6, "expm", -0.02j,
                                           "bf16x3")
        names = [t["name"] for t in backend.timing_end()]
    finally:
        backend.set_graphs(prev)
    g_o16, _, _ = backend.krylov_local(_lib.PTN_OP_HEFF, a, l, w, r, 6, "expm", -0.02j,
                                       "bf16x3")
    assert torch.equal(g_o16, o16)     # the captured graph computes the same
    o64, _, i64 = backend.krylov_local(_lib.PTN_OP_HEFF, a, l, w, r, 6, "expm", -0.02j)
    assert any(n.startswith("zgemm_bf16x3_v2<") for n in names), names
    assert np.all(backend.to_host(i16) == backend.to_host(i64))
    g16 = backend.to_host(o16)
    assert np.all(np.isfinite(g16))
    assert relerr(g16, backend.to_host(o64)) < 1e-3
