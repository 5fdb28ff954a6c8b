"""
Parity of the sector-grouped H_eff and environment steps (one grouped launch
per product over the admissible U(1) sector blocks; ptn_local_apply,
ptn_env_apply and the local solves with quantum numbers, called through the
C ABI) with the oracle's apply_local_hamiltonian and environment steps
(pytenet/chain_ops.py:16-99, 237-279), with the dense and the tile-mask
products, and of the solves and a U(1) two-site DMRG run through them with the
oracle. FP64: <= 1e-12 of the oracle's max norm and <= 1e-13 of
the dense product's; FP32: <= 2e-5. Operators are the XXZ MPO with its quantum
numbers (one and two sites), bonds sector-sorted as the splits leave them;
unsorted bonds fall back to the tile-mask products.
"""

import warnings

import numpy as np
import pytest

import oracle as orc
import pytenet_amd as ptn
from conftest import load_fixture, mps_arrays, mpo_arrays
from pytenet_amd import backend, _lib
from pytenet_amd.hamiltonian import heisenberg_xxz_1d_mpo

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-12, "fp32": 2e-5, "bf16x6": 2e-5}


def relerr(got, ref):
    got = backend.to_host(got)
    return np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-300)


def _sorted_bond(rng, D, spread, parity=0):
    """Labels of a bond in contiguous ascending sectors (as the splits leave
    them), even or odd as the XXZ labels (steps of 2) require."""
    return 2 * np.sort(rng.binomial(2 * spread, 0.5, size=D) - spread) + parity


def _qsparse(rng, shape, qnums):
    x = orc.crandn(shape, rng)
    orc.enforce_qsparsity(x, qnums)
    return x


def _operator(rng, Dl, Dr, two_site, ql=None, qr=None, spread=4):
    H = heisenberg_xxz_1d_mpo(4, 1.0, 0.8, -0.1)
    qd = np.asarray(H.qsite)
    w = np.asarray(H.a[1], dtype=complex)
    qm_l, qm_r = np.asarray(H.qbonds[1]), np.asarray(H.qbonds[2])
    qs = qd
    if two_site:
        w = orc.mpo_merge_tensor_pair(w, np.asarray(H.a[2], dtype=complex))
        qm_r = np.asarray(H.qbonds[3])
        qs = orc.qnumber_flatten([qd, qd])
    if ql is None:
        ql = _sorted_bond(rng, Dl, spread)
    if qr is None:
        qr = _sorted_bond(rng, Dr, spread, parity=0 if two_site else 1)
    a = _qsparse(rng, (Dl, len(qs), Dr), [ql, qs, -qr])
    l = _qsparse(rng, (Dl, len(qm_l), Dl), [ql, qm_l, -ql])
    r = _qsparse(rng, (Dr, len(qm_r), Dr), [qr, qm_r, -qr])
    return a, w, l, r, (ql, qs, qr, qm_l, qm_r)


@pytest.mark.parametrize("Dl,Dr,two_site", [(200, 180, False), (260, 300, True), (37, 41, True),
                                            (1, 6, True), (64, 1, False)])
@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16x6"])
def test_grouped_heff_matches_oracle_and_dense(Dl, Dr, two_site, precision):
    rng = np.random.default_rng(Dl * 7 + Dr)
    if Dl == 1:
        a, w, l, r, q = _operator(rng, Dl, Dr, two_site, ql=np.zeros(1, int))
    elif Dr == 1:
        a, w, l, r, q = _operator(rng, Dl, Dr, two_site, qr=np.ones(1, int))
    else:
        a, w, l, r, q = _operator(rng, Dl, Dr, two_site)
    out, grouped = backend.heff_apply_sectors(a, w, l, r, q, precision)
    assert grouped
    ref = orc.apply_local_hamiltonian(a, w, l, r)
    assert relerr(out, ref) < TOL[precision]
    dense = backend.to_host(backend.heff_apply(a, w, l, r, precision))
    if precision == "fp64":
        assert np.abs(backend.to_host(out) - dense).max() <= 1e-13 * np.abs(dense).max()
    # entries outside the admissible blocks are exactly zero
    ql, qs, qr = q[0], q[1], q[2]
    off = np.add.outer(np.add.outer(ql, qs), -qr) != 0
    assert np.all(backend.to_host(out)[off] == 0)


@pytest.mark.parametrize("Dl,Dr,two_site", [(200, 180, False), (260, 300, True), (37, 41, True),
                                            (1, 6, True), (64, 1, False)])
@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16x6"])
def test_grouped_environment_steps_match_oracle_and_dense(Dl, Dr, two_site, precision):
    """ptn_env_apply (both sides) with labels: grouped products, equal to the
    oracle's contraction_operator_step_left / _right (pytenet/chain_ops.py:16-99)
    and to the dense products, with a bra different from the ket."""
    rng = np.random.default_rng(Dl * 5 + Dr + 1)
    kw = {"ql": np.zeros(1, int)} if Dl == 1 else ({"qr": np.ones(1, int)} if Dr == 1 else {})
    a, w, l, r, q = _operator(rng, Dl, Dr, two_site, **kw)
    b = _qsparse(rng, a.shape, [q[0], q[1], -q[2]])
    for side, blk, ref, dense_fn in (
            (_lib.PTN_ENV_LEFT, l, orc.contraction_operator_step_left(a, b, w, l), backend.env_left),
            (_lib.PTN_ENV_RIGHT, r, orc.contraction_operator_step_right(a, b, w, r),
             backend.env_right)):
        out, grouped = backend._env_sectors(side, backend.as_dev(a), backend.as_dev(b),
                                            backend.as_dev(w), backend.as_dev(blk), q, precision,
                                            ref.shape)
        assert grouped
        assert relerr(out, ref) < TOL[precision]
        if precision == "fp64":
            dense = backend.to_host(dense_fn(a, b, w, blk, precision))
            assert np.abs(backend.to_host(out) - dense).max() <= 1e-13 * np.abs(dense).max()
        # the public wrappers route labelled calls the same way
        again = dense_fn(a, b, w, blk, precision, block_sparse=True, qnums=q)
        assert np.array_equal(backend.to_host(again), backend.to_host(out))


def test_grouped_heff_large_sectors_select_the_large_tile():
    """Two-site operator at D = 2048 with 8 sectors of 256: the first product
    (about 100 blocks of 256 x 256) runs on 128 x 64 tiles; the result equals
    the dense product and, at sampled entries, the defining sum."""
    rng = np.random.default_rng(2048)
    q = 2 * (np.repeat(np.arange(8), 256) - 4)
    a, w, l, r, qn = _operator(rng, 2048, 2048, True, ql=q, qr=q)
    backend.timing_begin()
    out, grouped = backend.heff_apply_sectors(a, w, l, r, qn, "fp64")
    rows = backend.timing_end()
    assert grouped
    names = [x["name"] for x in rows if x["name"].startswith("zgemm_grouped_kernel")]
    assert any(", 4, 4, " in n for n in names), names
    out = backend.to_host(out)
    dense = backend.to_host(backend.heff_apply(a, w, l, r, "fp64"))
    assert np.abs(out - dense).max() <= 1e-13 * np.abs(dense).max()
    for lb, sb, rb in [(5, 1, 130), (1000, 2, 1100), (2047, 3, 2047), (700, 0, 650)]:
        # out[lb,sb,rb] = sum L[l,m,lb] W[m,sb,s,m'] A[l,s,r] R[r,m',rb]
        t = np.tensordot(a, r[:, :, rb], axes=(2, 0))            # (l, s, m')
        val = np.einsum("lm,msn,lsn->", l[:, :, lb], w[:, sb, :, :], t)
        assert abs(out[lb, sb, rb] - val) <= 1e-12 * np.abs(dense).max()


def test_unsorted_bonds_use_the_tile_mask_products():
    rng = np.random.default_rng(11)
    a, w, l, r, q = _operator(rng, 90, 70, True)
    perm = rng.permutation(90)
    ql = q[0][perm]
    a, l = a[perm], l[perm][:, :, perm]
    out, grouped = backend.heff_apply_sectors(a, w, l, r, (ql,) + q[1:], "fp64")
    assert not grouped
    assert relerr(out, orc.apply_local_hamiltonian(a, w, l, r)) < TOL["fp64"]


def _chain_operator(rng, two_site, L=8, i=3, D=48):
    """H_eff at site i (or sites i, i+1) of a random U(1) MPS with sorted bond
    labels, the environments contracted by the oracle from the chain ends
    (Hermitian, as in a sweep)."""
    H = heisenberg_xxz_1d_mpo(L, 1.0, 0.8, -0.1)
    qsite = np.asarray(H.qsite)
    qbonds = [np.zeros(1, int)]
    for k in range(1, L):
        reach = np.arange(-min(k, L - k), min(k, L - k) + 1, 2)
        qbonds.append(np.sort(rng.choice(reach, size=min(D, 2 ** min(k, L - k)))))
    qbonds.append(np.zeros(1, int))
    psi = orc.MPS(qsite, qbonds, fill="random", rng=rng)
    mpo = [np.asarray(x, dtype=complex) for x in H.a]
    l = np.ones((1, 1, 1), dtype=complex)
    for j in range(i):
        l = orc.contraction_operator_step_left(psi.a[j], psi.a[j], mpo[j], l)
    last = i + 1 if two_site else i
    r = np.ones((1, 1, 1), dtype=complex)
    for j in reversed(range(last + 1, L)):
        r = orc.contraction_operator_step_right(psi.a[j], psi.a[j], mpo[j], r)
    if two_site:
        a = orc.mps_merge_tensor_pair(psi.a[i], psi.a[i + 1])
        w = orc.mpo_merge_tensor_pair(mpo[i], mpo[i + 1])
        qs = orc.qnumber_flatten([qsite, qsite])
    else:
        a, w, qs = psi.a[i], mpo[i], qsite
    q = (qbonds[i], qs, qbonds[last + 1], np.asarray(H.qbonds[i]), np.asarray(H.qbonds[last + 1]))
    return a / np.linalg.norm(a), w, l, r, q


@pytest.mark.parametrize("two_site", [False, True])
def test_grouped_lanczos_matches_oracle(two_site):
    """expm and lowest Ritz value with the matvecs through the grouped
    products, against the oracle's Lanczos and against the tile-mask solve."""
    rng = np.random.default_rng(5 + two_site)
    a, w, l, r, q = _chain_operator(rng, two_site)
    assert a.shape[0] > 1 and a.shape[2] > 1
    hfun = lambda x: orc.apply_local_hamiltonian(x.reshape(a.shape), w, l, r).reshape(-1)
    dt = 0.1 - 0.3j
    got, _, _ = backend.krylov_local(_lib.PTN_OP_HEFF, a, l, w, r, 10, "expm", dt,
                                     block_sparse=True, qnums=q)
    masks, _, _ = backend.krylov_local(_lib.PTN_OP_HEFF, a, l, w, r, 10, "expm", dt,
                                       block_sparse=True)
    ref = orc.expm_krylov(hfun, a.reshape(-1), dt, 10, hermitian=True).reshape(a.shape)
    assert relerr(got, ref) < 1e-11
    assert np.abs(backend.to_host(got) - backend.to_host(masks)).max() < 1e-12 * np.abs(ref).max()
    _, en, _ = backend.krylov_local(_lib.PTN_OP_HEFF, a, l, w, r, 12, "eigh",
                                    block_sparse=True, qnums=q)
    ev, _ = orc.eigh_krylov(hfun, a.reshape(-1), 12, 1)
    assert abs(float(backend.to_host(en)[0]) - ev[0]) < 1e-11 * abs(ev[0])


def test_u1_dmrg_runs_the_grouped_products_and_matches_the_reference_trajectory():
    """The reference's two-site DMRG trajectory with quantum numbers (XXZ
    L=10, tests/golden): energies as the reference's, with the local solves
    captured into graphs (the default) and launched directly; in the direct
    run the launch records show that the matvecs ran through the grouped
    products once the splits had sorted the bonds."""
    z = load_fixture("traj_dmrg2_xxz_L10_qn")
    H = ptn.MPO.from_tensors(*mpo_arrays(z))
    for graphs in (True, False):
        qsite, qb, a = mps_arrays(z, "init_")
        psi = ptn.MPS(qsite, qb, fill="postpone")
        psi.a = [x.copy() for x in a]
        prev = backend.set_graphs(graphs)
        backend.timing_begin()
        try:
            with warnings.catch_warnings():
                warnings.simplefilter("ignore", RuntimeWarning)
                en = ptn.dmrg_twosite(H, psi, int(z["numsweeps"]),
                                      numiter_lanczos=int(z["numiter"]))
        finally:
            rows = backend.timing_end()
            backend.set_graphs(prev)
        np.testing.assert_allclose(en, z["en_min"], rtol=1e-10)
        if not graphs:
            names = [x["name"] for x in rows]
            assert any(n.startswith("zgemm_grouped_kernel") for n in names), names


@pytest.mark.parametrize("precision", ["fp32", "bf16x6"])
def test_grouped_lanczos_fp32_matches_oracle(precision):
    """FP32 matvecs through the grouped products (complex64 environment blocks,
    FP64 Lanczos scalars; the three-term split-bf16 mode takes the same FP32
    grouped products): the expm result against the oracle's FP64 Lanczos
    within the FP32 bound."""
    rng = np.random.default_rng(17)
    a, w, l, r, q = _chain_operator(rng, True)
    hfun = lambda x: orc.apply_local_hamiltonian(x.reshape(a.shape), w, l, r).reshape(-1)
    dt = -0.05j
    got, _, info = backend.krylov_local(_lib.PTN_OP_HEFF, a, l, w, r, 8, "expm", dt,
                                        precision=precision, block_sparse=True, qnums=q)
    ref = orc.expm_krylov(hfun, a.reshape(-1), dt, 8, hermitian=True).reshape(a.shape)
    assert relerr(got, ref) < TOL["fp32"]


def test_environment_step_with_unsorted_bonds_falls_back():
    rng = np.random.default_rng(23)
    a, w, l, r, q = _operator(rng, 70, 90, True)
    perm = rng.permutation(90)
    qr = q[2][perm]
    a, r = a[:, :, perm], r[perm][:, :, perm]
    b = a.copy()
    out, grouped = backend._env_sectors(_lib.PTN_ENV_RIGHT, backend.as_dev(a), backend.as_dev(b),
                                        backend.as_dev(w), backend.as_dev(r),
                                        (q[0], q[1], qr, q[3], q[4]), "fp64", (70, w.shape[0], 70))
    assert not grouped
    assert relerr(out, orc.contraction_operator_step_right(a, b, w, r)) < TOL["fp64"]


def test_plan_cache_eviction_keeps_results_exact():
    """More distinct label sets than the plan cache holds (1024): the cache
    drops its older half on the way, and every application still equals the
    oracle's (released plans are rebuilt on demand)."""
    rng = np.random.default_rng(29)
    H = heisenberg_xxz_1d_mpo(4, 1.0, 0.8, -0.1)
    w = np.asarray(H.a[1], dtype=complex)
    qd, qm_l, qm_r = np.asarray(H.qsite), np.asarray(H.qbonds[1]), np.asarray(H.qbonds[2])
    for k in range(1100):
        Dl = 6 + k % 5
        ql = np.sort(rng.integers(-2, 3, Dl)) * 2
        qr = np.sort(rng.integers(-2, 3, 7)) * 2 + 1
        if k % 100 == 0:
            a = _qsparse(rng, (Dl, 2, 7), [ql, qd, -qr])
            l = _qsparse(rng, (Dl, len(qm_l), Dl), [ql, qm_l, -ql])
            r = _qsparse(rng, (7, len(qm_r), 7), [qr, qm_r, -qr])
            out, _ = backend.heff_apply_sectors(a, w, l, r, (ql, qd, qr, qm_l, qm_r))
            assert relerr(out, orc.apply_local_hamiltonian(a, w, l, r)) < TOL["fp64"]
        else:
            # a new label set per call on zero operands: only the plan matters
            a0 = np.zeros((Dl, 2, 7), complex)
            l0 = np.zeros((Dl, len(qm_l), Dl), complex)
            r0 = np.zeros((7, len(qm_r), 7), complex)
            out, _ = backend.heff_apply_sectors(a0, w, l0, r0, (ql, qd, qr, qm_l, qm_r))
            assert np.all(backend.to_host(out) == 0)
