"""
Parity at the production sizes of the headline workload (two-site TDVP on the
XXZ chain at D = 1024, M = 5): the contractions that run there, in the tile
shapes the launch rule picks at that size (checked from the live launch
timing, which names every launched product kernel), against the oracle
(pytenet/chain_ops.py:16-99,237-317); the tile-skipping products of U(1)
sector-sorted operands at D = 2048 against the dense products; and the
single-site TDVP workload of the reference's small configuration (L = 10,
bonds clamped to 8, dt = 0.01 - 0.05j, numiter 5) for its full 100 steps
against the oracle.
"""

import numpy as np
import pytest

import oracle as orc
import pytenet_amd as ptn
from pytenet_amd import backend
from pytenet_amd.hamiltonian import heisenberg_xxz_1d_mpo

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-12, "fp32": 2e-5, "bf16x3": 2e-4, "bf16x6": 2e-5}
# product-kernel name prefixes of the large tiles the launch rule selects at D = 1024
BIG_TILE = {"fp64": lambda n: n.startswith("zgemm_kernel<double,") and ", 4, 4, " in n,
            "fp32": lambda n: n.startswith("zgemm_kernel<float,") and ", 4, 8, " in n,
            "bf16x3": lambda n: n.startswith("zgemm_bf16x3_v2<"),
            "bf16x6": lambda n: n.startswith("zgemm_bf16x6_v3<")}


def relerr(got, ref):
    got = backend.to_host(got)
    return np.abs(got - ref).max() / np.abs(ref).max()


def _herm_env(rng, D, M):
    x = orc.crandn((D, M, D), rng) / D
    return 0.5 * (x + x.transpose(2, 1, 0).conj())


@pytest.fixture(scope="module")
def problem_d1024():
    rng = np.random.default_rng(1024)
    D = 1024
    op = heisenberg_xxz_1d_mpo(10, 1.0, 0.8, -0.1)
    w1 = np.asarray(op.a[4], dtype=complex)
    w2 = orc.mpo_merge_tensor_pair(w1, np.asarray(op.a[5], dtype=complex))
    l, r = _herm_env(rng, D, 5), _herm_env(rng, D, 5)
    a2 = orc.crandn((D, 4, D), rng) / D
    a1 = orc.crandn((D, 2, D), rng) / D
    c = orc.crandn((D, D), rng) / D
    ref = {"heff2": orc.apply_local_hamiltonian(a2, w2, l, r),
           "env": orc.contraction_operator_step_left(a1, a1, w1, l),
           "bond": orc.apply_local_bond_contraction(c, l, r)}
    dev = {k: backend.as_dev(v) for k, v in dict(a2=a2, a1=a1, c=c, w1=w1, w2=w2, l=l, r=r).items()}
    return dev, ref


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16x3", "bf16x6"])
def test_contractions_at_d1024_match_oracle_in_production_tiles(problem_d1024, precision):
    dv, ref = problem_d1024
    prev = backend.set_autotune(False)   # the fixed launch rule of the sweeps
    try:
        backend.timing_begin()
        heff = backend.heff_apply(dv["a2"], dv["w2"], dv["l"], dv["r"], precision)
        env = backend.env_left(dv["a1"], dv["a1"], dv["w1"], dv["l"], precision)
        bond = backend.bond_apply(dv["c"], dv["l"], dv["r"], precision)
        rows = backend.timing_end()
    finally:
        backend.set_autotune(prev)
    assert relerr(heff, ref["heff2"]) < TOL[precision]
    assert relerr(env, ref["env"]) < TOL[precision]
    assert relerr(bond, ref["bond"]) < TOL[precision]
    names = [r["name"] for r in rows]
    big = [n for n in names if BIG_TILE[precision](n)]
    assert big, names


def _sorted_bond(rng, D, spread):
    return np.sort(rng.binomial(2 * spread, 0.5, size=D) - spread)


def _qsparse(rng, shape, qnums):
    x = orc.crandn(shape, rng)
    orc.enforce_qsparsity(x, qnums)
    return x


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_tile_skipping_at_d2048_equals_dense(precision):
    """Two-site H_eff and both environment steps with U(1) sector-sorted bonds
    of dimension 2048 (about 20 sectors): the tile-skipping products equal the
    dense products of the same precision to rounding (the skipped tiles are
    exactly zero)."""
    rng = np.random.default_rng(2048)
    D = 2048
    H = heisenberg_xxz_1d_mpo(4, 1.0, 0.8, -0.1)
    qd = np.asarray(H.qsite)
    w1 = np.asarray(H.a[1], dtype=complex)
    w2 = orc.mpo_merge_tensor_pair(w1, np.asarray(H.a[2], dtype=complex))
    qm_l, qm_r1, qm_r2 = (np.asarray(H.qbonds[1]), np.asarray(H.qbonds[2]),
                          np.asarray(H.qbonds[3]))
    q2 = orc.qnumber_flatten([qd, qd])
    ql, qr = _sorted_bond(rng, D, 10), _sorted_bond(rng, D, 10)
    a2 = _qsparse(rng, (D, 4, D), [ql, q2, -qr])
    a1 = _qsparse(rng, (D, 2, D), [ql, qd, -qr])
    l = _qsparse(rng, (D, len(qm_l), D), [ql, qm_l, -ql])
    r2 = _qsparse(rng, (D, len(qm_r2), D), [qr, qm_r2, -qr])
    r1 = _qsparse(rng, (D, len(qm_r1), D), [qr, qm_r1, -qr])
    dv = {k: backend.as_dev(v) for k, v in dict(a2=a2, a1=a1, w1=w1, w2=w2, l=l, r1=r1, r2=r2).items()}
    del a2, a1, l, r1, r2
    dense = backend.heff_apply(dv["a2"], dv["w2"], dv["l"], dv["r2"], precision)
    sparse = backend.heff_apply(dv["a2"], dv["w2"], dv["l"], dv["r2"], precision, block_sparse=True)
    assert float((sparse - dense).abs().max()) <= 1e-13 * float(dense.abs().max())
    for step, blk in ((backend.env_left, "l"), (backend.env_right, "r1")):
        d0 = step(dv["a1"], dv["a1"], dv["w1"], dv[blk], precision)
        s0 = step(dv["a1"], dv["a1"], dv["w1"], dv[blk], precision, block_sparse=True)
        assert float((s0 - d0).abs().max()) <= 1e-13 * float(d0.abs().max())


def test_tdvp_singlesite_small_configuration_full_100_steps():
    """The reference's CPU-runnable configuration in full: single-site TDVP on
    XXZ L = 10 (J = 1, Delta = 0.8, h = -0.1, zero quantum numbers), seeded
    random MPS with bonds [1,2,4,8,16,28,16,8,4,2,1] clamped to 8 and
    left-orthonormalised, dt = 0.01 - 0.05j, numiter 5, 100 steps; the GPU
    state vector equals the oracle's to 1e-10 (relative to its norm), as do
    the returned norm and the energy."""
    L = 10
    H = heisenberg_xxz_1d_mpo(L, 1.0, 0.8, -0.1).zero_qnumbers()
    bonds = [1, 2, 4, 8, 16, 28, 16, 8, 4, 2, 1]
    psi = ptn.MPS(H.qsite, [np.zeros(b, dtype=int) for b in bonds], fill="random",
                  rng=np.random.default_rng(42))
    for i in range(L):
        psi.a[i][8:, :, :] = 0
        psi.a[i][:, :, 8:] = 0
    ref = orc.MPS(H.qsite, [q.copy() for q in psi.qbonds], fill="postpone")
    ref.a = [a.copy() for a in psi.a]
    psi.orthonormalize(mode="left")
    orc_nrm0 = ref.orthonormalize(mode="left")
    assert orc_nrm0 > 0
    Ho = orc.MPO(H.qsite, H.qbonds, H.a)
    nrm_ref = orc.tdvp_singlesite(Ho, ref, 0.01 - 0.05j, 100, numiter_lanczos=5)
    nrm = ptn.tdvp_singlesite(H, psi, 0.01 - 0.05j, 100, numiter_lanczos=5)
    assert abs(nrm - nrm_ref) < 1e-12
    x, y = psi.to_vector(), ref.to_vector()
    assert np.abs(x - y).max() <= 1e-10 * np.linalg.norm(y)
    e, e_ref = ptn.mpo_average(psi, H), orc.mpo_average(ref, Ho)
    assert abs(e - e_ref) <= 1e-10 * abs(e_ref)
