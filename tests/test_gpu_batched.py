"""
Batched kernels and the batched single-site TDVP (many independent chains in
one launch sequence) against the single-problem GPU path and the oracle.
"""

import copy

import numpy as np
import pytest
import torch

import oracle as orc
import pytenet_amd as ptn
from pytenet_amd import backend, _lib

pytestmark = pytest.mark.gpu


def _herm_env(rng, D, M):
    x = orc.crandn((D, M, D), rng)
    return 0.5 * (x + x.transpose(2, 1, 0).conj())


def _herm_w(rng, M, d):
    w = orc.crandn((M, d, d, M), rng)
    return 0.5 * (w + w.transpose(0, 2, 1, 3).conj())


@pytest.mark.parametrize("nb,m,n", [(5, 40, 24), (3, 70, 70), (4, 33, 65), (64, 16, 8)])
def test_qr_batched_matches_single(nb, m, n):
    rng = np.random.default_rng(nb * 100 + m + n)
    a = np.stack([orc.crandn((m, n), rng) for _ in range(nb)])
    q, r = backend.qr_batched(a)
    q, r = backend.to_host(q), backend.to_host(r)
    for b in range(nb):
        q0, r0 = np.linalg.qr(a[b])
        assert np.abs(q[b] - q0).max() < 1e-12
        assert np.abs(r[b] - r0).max() < 1e-12 * max(1.0, np.abs(r0).max())


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-11), ("fp32", 2e-5), ("bf16x6", 2e-5)])
def test_krylov_batched_matches_single(precision, tol):
    rng = np.random.default_rng(11)
    nb, Dl, d, Dr, M = 6, 10, 2, 12, 3
    a = np.stack([orc.crandn((Dl, d, Dr), rng) for _ in range(nb)])
    w = np.stack([_herm_w(rng, M, d) for _ in range(nb)])
    l = np.stack([_herm_env(rng, Dl, M) for _ in range(nb)])
    r = np.stack([_herm_env(rng, Dr, M) for _ in range(nb)])
    out, _, info = backend.krylov_local_batched(_lib.PTN_OP_HEFF, a, l, w, r, 6, "expm",
                                                -0.05j + 0.01, precision)
    out = backend.to_host(out)
    assert np.all(info.cpu().numpy()[:, 0] == 6)
    for b in range(nb):
        ref = orc.expm_krylov(
            lambda x: orc.apply_local_hamiltonian(x.reshape(Dl, d, Dr), w[b], l[b], r[b]).reshape(-1),
            a[b].reshape(-1), -0.05j + 0.01, 6, hermitian=True).reshape(Dl, d, Dr)
        assert np.abs(out[b] - ref).max() < tol * np.abs(ref).max()
    # bond operator and the eigensolver
    c = np.stack([orc.crandn((Dl, Dr), rng) for _ in range(nb)])
    rb = np.stack([_herm_env(rng, Dr, M) for _ in range(nb)])
    outb, _, _ = backend.krylov_local_batched(_lib.PTN_OP_BOND, c, l, None, rb, 5, "expm", 0.2j,
                                              precision)
    outb = backend.to_host(outb)
    _, en, _ = backend.krylov_local_batched(_lib.PTN_OP_HEFF, a, l, w, r, 8, "eigh", 0.0, precision)
    en = en.cpu().numpy()
    for b in range(nb):
        ref = orc.expm_krylov(
            lambda x: orc.apply_local_bond_contraction(x.reshape(Dl, Dr), l[b], rb[b]).reshape(-1),
            c[b].reshape(-1), 0.2j, 5, hermitian=True).reshape(Dl, Dr)
        assert np.abs(outb[b] - ref).max() < tol * np.abs(ref).max()
        ev, _ = orc.eigh_krylov(
            lambda x: orc.apply_local_hamiltonian(x.reshape(Dl, d, Dr), w[b], l[b], r[b]).reshape(-1),
            a[b].reshape(-1), 8, 1)
        assert abs(en[b] - ev[0]) < 10 * tol * abs(ev[0])


def test_env_batched_matches_single():
    rng = np.random.default_rng(5)
    nb, Dl, d, Dr, Ml, Mr = 4, 9, 2, 7, 3, 4
    a = np.stack([orc.crandn((Dl, d, Dr), rng) for _ in range(nb)])
    w = np.stack([orc.crandn((Ml, d, d, Mr), rng) for _ in range(nb)])
    l = np.stack([orc.crandn((Dl, Ml, Dl), rng) for _ in range(nb)])
    r = np.stack([orc.crandn((Dr, Mr, Dr), rng) for _ in range(nb)])
    el = backend.to_host(backend.env_left_batched(a, w, l))
    er = backend.to_host(backend.env_right_batched(a, w, r))
    for b in range(nb):
        assert np.abs(el[b] - orc.contraction_operator_step_left(a[b], a[b], w[b], l[b])).max() < 1e-12 * np.abs(el[b]).max()
        assert np.abs(er[b] - orc.contraction_operator_step_right(a[b], a[b], w[b], r[b])).max() < 1e-12 * np.abs(er[b]).max()


def test_env_batched_three_term_split_bf16_matches_oracle():
    """Batched environment steps large enough for the pre-split three-term
    split-bf16 product (two problems of D = 512: the problem index on the grid's
    third axis, the term planes of each problem at their own stride), per
    problem against the oracle at the FP32 bound."""
    rng = np.random.default_rng(8)
    nb, D, d, M = 2, 512, 2, 5
    a = np.stack([orc.crandn((D, d, D), rng) / D for _ in range(nb)])
    w = np.stack([orc.crandn((M, d, d, M), rng) for _ in range(nb)])
    l = np.stack([orc.crandn((D, M, D), rng) / D for _ in range(nb)])
    r = np.stack([orc.crandn((D, M, D), rng) / D for _ in range(nb)])
    backend.timing_begin()
    el = backend.to_host(backend.env_left_batched(a, w, l, "bf16x6"))
    er = backend.to_host(backend.env_right_batched(a, w, r, "bf16x6"))
    names = [t["name"] for t in backend.timing_end()]
    assert any(n.startswith("zgemm_bf16x6_v3<") for n in names), names
    for b in range(nb):
        refl = orc.contraction_operator_step_left(a[b], a[b], w[b], l[b])
        refr = orc.contraction_operator_step_right(a[b], a[b], w[b], r[b])
        assert np.abs(el[b] - refl).max() < 2e-5 * np.abs(refl).max()
        assert np.abs(er[b] - refr).max() < 2e-5 * np.abs(refr).max()


def _ising_chains(nchains, L, Ds, seed):
    rng = np.random.default_rng(seed)
    Hs, psis = [], []
    for k in range(nchains):
        g = rng.uniform(0.5, 1.5)
        H = ptn.ising_1d_mpo(L, 1.0, 0.0, g)
        D = Ds[k % len(Ds)]
        bonds = [min(D, 2**i, 2**(L - i)) for i in range(L + 1)]
        psi = ptn.MPS(H.qsite, [np.zeros(b, dtype=int) for b in bonds], fill="random", rng=rng)
        Hs.append(H)
        psis.append(psi)
    return Hs, psis


def _to_oracle(psi):
    p = orc.MPS(psi.qsite, psi.qbonds)
    p.a = [backend.to_host(a) if isinstance(a, torch.Tensor) else np.array(a) for a in psi.a]
    return p


def _vec(psi):
    v = psi.a[0]
    for a in psi.a[1:]:
        v = np.tensordot(v, a, axes=(-1, 0))
    return v.reshape(-1)


@pytest.mark.parametrize("dt", [0.05, 0.1j])
def test_tdvp_singlesite_batched_matches_oracle_per_chain(dt):
    """Chains of two bond-dimension buckets, imaginary and real time: every chain
    equals the oracle's tdvp_singlesite on that chain."""
    Hs, psis = _ising_chains(5, 6, [4, 8], seed=3)
    refs = [_to_oracle(copy.deepcopy(p)) for p in psis]
    nrm = ptn.tdvp_singlesite_batched(Hs, psis, dt, 3, numiter_lanczos=6)
    for k in range(len(psis)):
        Ho = orc.MPO(Hs[k].qsite, Hs[k].qbonds, Hs[k].a)
        nrm_ref = orc.tdvp_singlesite(Ho, refs[k], dt, 3, numiter_lanczos=6)
        assert abs(nrm[k] - nrm_ref) < 1e-12 * abs(nrm_ref)
        got, ref = _vec(_to_oracle(psis[k])), _vec(refs[k])
        assert np.abs(got - ref).max() < 1e-10 * np.abs(ref).max()


def test_tdvp_singlesite_batched_fp32_energies():
    """complex64 contractions with FP64 Krylov scalars: per-chain energy close to FP64."""
    Hs, psis = _ising_chains(8, 8, [8, 16], seed=9)
    p64 = copy.deepcopy(psis)
    ptn.tdvp_singlesite_batched(Hs, p64, 0.05, 4, numiter_lanczos=8)
    ptn.tdvp_singlesite_batched(Hs, psis, 0.05, 4, numiter_lanczos=8, precision="fp32")
    for k in range(len(psis)):
        e64 = ptn.mpo_average(p64[k], Hs[k]).real / ptn.mps_norm(p64[k])**2
        e32 = ptn.mpo_average(psis[k], Hs[k]).real / ptn.mps_norm(psis[k])**2
        assert abs(e32 - e64) < 1e-5 * abs(e64)


def _padded_chains(nchains, L, Ds, seed):
    """The many-chain workload's inputs at a reduced size: transverse-field
    Ising chains (J = 1, h = 0, g uniform in [0.5, 1.5]) starting from random
    product states padded to their own bond dimension D_i."""
    rng = np.random.default_rng(seed)
    Hs, psis = [], []
    for k in range(nchains):
        g = rng.uniform(0.5, 1.5)
        Hs.append(ptn.ising_1d_mpo(L, 1.0, 0.0, g))
        psis.append(ptn.MPS.padded_product_state(Hs[-1].qsite, L, Ds[k % len(Ds)], rng))
    return Hs, psis


def test_tdvp_batched_padded_product_states_four_buckets_match_oracle():
    """Padded product states in four bond-dimension buckets (D = 2, 4, 8, 16),
    imaginary-time single-site TDVP dt = 0.05 (the buckets advance together on
    their own streams, the autotuner on): every chain equals the oracle's
    tdvp_singlesite of that chain, whose QR completion of the zero padding is
    LAPACK's."""
    Hs, psis = _padded_chains(8, 8, [2, 4, 8, 16], seed=12)
    refs = [_to_oracle(copy.deepcopy(p)) for p in psis]
    nrm = ptn.tdvp_singlesite_batched(Hs, psis, 0.05, 3, numiter_lanczos=6)
    assert len({tuple(p.bond_dims) for p in psis}) == 4
    for k in range(len(psis)):
        Ho = orc.MPO(Hs[k].qsite, Hs[k].qbonds, Hs[k].a)
        nrm_ref = orc.tdvp_singlesite(Ho, refs[k], 0.05, 3, numiter_lanczos=6)
        assert abs(nrm[k] - nrm_ref) < 1e-12 * abs(nrm_ref)
        got, ref = _vec(_to_oracle(psis[k])), _vec(refs[k])
        # the rank-deficient start makes the local Krylov spaces break down
        # (invariant subspaces of the padded product state), which amplifies
        # rounding differences; 1e-9 relative still pins the QR completion
        assert np.abs(got - ref).max() < 1e-9 * np.abs(ref).max(), k


def test_tdvp_batched_buckets_on_streams_equal_one_bucket_at_a_time():
    """The interleaved buckets (separate streams, graph-replayed small solves)
    give each chain the result of running its bucket alone."""
    Hs, psis = _padded_chains(6, 10, [4, 8, 16], seed=13)
    alone = [copy.deepcopy(p) for p in psis]
    ptn.tdvp_singlesite_batched(Hs, psis, 0.05, 2, numiter_lanczos=6, precision="fp32")
    for D in (4, 8, 16):
        idx = [k for k in range(6) if max(alone[k].bond_dims) == D]
        sub = [alone[k] for k in idx]
        ptn.tdvp_singlesite_batched([Hs[k] for k in idx], sub, 0.05, 2, numiter_lanczos=6,
                                    precision="fp32")
        for j, k in enumerate(idx):
            assert np.abs(_vec(_to_oracle(psis[k])) - _vec(_to_oracle(sub[j]))).max() < 1e-12
