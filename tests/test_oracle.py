"""
Pins the oracle (CPU restatement of the reference) and the host-side MPO
builders: against the reference's own outputs (tests/golden, written by
tests/golden/make_fixtures.py from cmendl/pytenet 1.3.0) and against closed
forms that do not share the oracle's code path (Kronecker-product
Hamiltonians, dense expm/eigh, brute-force operators, invariants).
"""

import copy
import numpy as np
import pytest
from scipy.linalg import expm

import oracle as orc
from conftest import load_fixture, mps_arrays, mpo_arrays
from pytenet_amd.hamiltonian import heisenberg_xxz_1d_mpo, ising_1d_mpo
import pytenet_amd as ptn

MPO_FIXTURES = ["mpo_xxz_L2", "mpo_xxz_L3", "mpo_xxz_L5", "mpo_xxz_L10", "mpo_xxz_h0_L2",
                "mpo_xxz_h0_L3", "mpo_xxz_h0_L6", "mpo_ising_L1", "mpo_ising_L2", "mpo_ising_L4",
                "mpo_ising_L5"]


def _oracle_mps(qsite, qbonds, a):
    psi = orc.MPS(qsite, qbonds, fill="postpone")
    psi.a = [x.copy() for x in a]
    return psi


def _oracle_mpo(z, prefix="h_"):
    qs, qb, a = mpo_arrays(z, prefix)
    return orc.MPO(qs, qb, a)


# ---------------------------------------------------------------------------
# closed-form Hamiltonians (independent of any MPO code)

SUP = np.array([[0., 1.], [0., 0.]])
SDN = SUP.T.copy()
SZ = np.diag([0.5, -0.5])


def _site_op(op, i, L):
    return np.kron(np.kron(np.identity(2**i), op), np.identity(2**(L - i - 1)))


def xxz_kron(L, J, D, h):
    H = np.zeros((2**L, 2**L))
    for i in range(L - 1):
        H += 0.5 * J * (_site_op(SUP, i, L) @ _site_op(SDN, i + 1, L)
                        + _site_op(SDN, i, L) @ _site_op(SUP, i + 1, L))
        H += D * _site_op(SZ, i, L) @ _site_op(SZ, i + 1, L)
    for i in range(L):
        H -= h * _site_op(SZ, i, L)
    return H


def ising_kron(L, J, h, g):
    X = np.array([[0., 1.], [1., 0.]])
    Z = np.diag([1., -1.])
    H = np.zeros((2**L, 2**L))
    for i in range(L - 1):
        H += J * _site_op(Z, i, L) @ _site_op(Z, i + 1, L)
    for i in range(L):
        H += h * _site_op(Z, i, L) + g * _site_op(X, i, L)
    return H


# ---------------------------------------------------------------------------
# MPO builders

@pytest.mark.parametrize("name", MPO_FIXTURES)
def test_mpo_builder_matches_reference_layout(name):
    z = load_fixture(name)
    L = int(z["nsites"])
    if "ising" in name:
        op = ising_1d_mpo(L, float(z["J"]), float(z["h"]), float(z["g"]))
    else:
        op = heisenberg_xxz_1d_mpo(L, float(z["J"]), float(z["D"]), float(z["h"]))
    assert np.array_equal(op.qsite, z["qsite"])
    for i in range(L):
        assert op.a[i].shape == z[f"a{i}"].shape
        assert np.array_equal(op.a[i], z[f"a{i}"]), f"site {i}"
    for i in range(L + 1):
        assert np.array_equal(op.qbonds[i], z[f"q{i}"])


@pytest.mark.parametrize("L,J,D,h", [(2, 1.0, 0.8, -0.1), (3, 0.7, 1.3, 0.0), (5, 4/3, 5/13, -2/7),
                                     (6, -1.1, 0.25, 0.6)])
def test_xxz_mpo_equals_kronecker_hamiltonian(L, J, D, h):
    op = heisenberg_xxz_1d_mpo(L, J, D, h)
    assert np.allclose(op.to_matrix(), xxz_kron(L, J, D, h), atol=1e-13)


@pytest.mark.parametrize("L,J,h,g", [(1, 1.0, 0.3, 0.7), (2, 1.0, 0.3, 0.7), (5, 0.6, 0.0, 1.2)])
def test_ising_mpo_equals_kronecker_hamiltonian(L, J, h, g):
    op = ising_1d_mpo(L, J, h, g)
    assert np.allclose(op.to_matrix(), ising_kron(L, J, h, g), atol=1e-13)


# ---------------------------------------------------------------------------
# local contractions

def test_kernels_match_reference_fixture():
    z = load_fixture("kernels_dense")
    a, b, w, l, r, c, rb = (z[k] for k in ("a", "b", "w", "l", "r", "c", "rb"))
    np.testing.assert_allclose(orc.apply_local_hamiltonian(a, w, l, r), z["heff"], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(orc.apply_local_bond_contraction(c, l, rb), z["bond"], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(orc.contraction_operator_step_left(a, b, w, l), z["env_left"],
                               rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(orc.contraction_operator_step_right(a, b, w, r), z["env_right"],
                               rtol=1e-13, atol=1e-13)


def test_heff_brute_force_dense_operator():
    """The tensordot chain equals the dense operator assembled index by index."""
    rng = np.random.default_rng(3)
    Dl, d, Dr, Ml, Mr = 3, 2, 4, 2, 3
    a = orc.crandn((Dl, d, Dr), rng)
    w = orc.crandn((Ml, d, d, Mr), rng)
    l = orc.crandn((Dl, Ml, Dl), rng)
    r = orc.crandn((Dr, Mr, Dr), rng)
    H = np.zeros((Dl, d, Dr, Dl, d, Dr), dtype=complex)
    for lb in range(Dl):
        for sb in range(d):
            for rb_ in range(Dr):
                for li in range(Dl):
                    for s in range(d):
                        for ri in range(Dr):
                            H[lb, sb, rb_, li, s, ri] = sum(
                                l[li, m, lb] * w[m, sb, s, mp] * r[ri, mp, rb_]
                                for m in range(Ml) for mp in range(Mr))
    ref = H.reshape(Dl * d * Dr, -1) @ a.reshape(-1)
    np.testing.assert_allclose(orc.apply_local_hamiltonian(a, w, l, r).reshape(-1), ref,
                               rtol=1e-13, atol=1e-13)
    # bond operator: same construction without W
    c = orc.crandn((Dl, Dr), rng)
    r2 = orc.crandn((Dr, Ml, Dr), rng)
    Hb = np.einsum("lmx,rmy->xylr", l, r2).reshape(Dl * Dr, -1)
    np.testing.assert_allclose(orc.apply_local_bond_contraction(c, l, r2).reshape(-1),
                               Hb @ c.reshape(-1), rtol=1e-13, atol=1e-13)


def test_two_site_heff_with_trivial_blocks_is_kronecker_hamiltonian():
    """For a 2-site chain with trivial environments the two-site operator is H itself."""
    L, J, D, h = 2, 0.9, 1.7, -0.3
    op = heisenberg_xxz_1d_mpo(L, J, D, h)
    h2 = orc.mpo_merge_tensor_pair(op.a[0], op.a[1])
    one = np.ones((1, 1, 1))
    Hk = xxz_kron(L, J, D, h)
    rng = np.random.default_rng(5)
    x = orc.crandn((1, 4, 1), rng)
    np.testing.assert_allclose(orc.apply_local_hamiltonian(x, h2, one, one).reshape(-1),
                               Hk @ x.reshape(-1), atol=1e-14)


def test_environment_chain_gives_expectation_value():
    """mpo_average (a chain of right environment steps) equals <x|H|x> on the full space."""
    rng = np.random.default_rng(8)
    L = 6
    op = heisenberg_xxz_1d_mpo(L, 1.0, 0.8, -0.1).zero_qnumbers()
    bonds = [1, 2, 4, 5, 4, 2, 1]
    psi = orc.MPS(op.qsite, [np.zeros(b, dtype=int) for b in bonds], fill="random", rng=rng)
    x = psi.to_vector()
    H = orc.MPO(op.qsite, op.qbonds, op.a)
    ref = np.vdot(x, xxz_kron(L, 1.0, 0.8, -0.1) @ x)
    assert abs(orc.mpo_average(psi, H) - ref) < 1e-12 * abs(ref)
    # left environments: sandwich the full chain from the left
    t = np.ones((1, 1, 1))
    for i in range(L):
        t = orc.contraction_operator_step_left(psi.a[i], psi.a[i], H.a[i], t)
    assert abs(t[0, 0, 0] - ref) < 1e-12 * abs(ref)


# ---------------------------------------------------------------------------
# Krylov

def _herm(n, rng):
    a = orc.crandn((n, n), rng) / np.sqrt(n)
    return 0.5 * (a + a.conj().T)


def test_lanczos_full_space_is_exact():
    """With numiter equal to the dimension the Krylov results are exact."""
    rng = np.random.default_rng(12)
    n = 12
    A = _herm(n, rng)
    v = orc.crandn(n, rng)
    dt = 0.3 - 0.7j
    got = orc.expm_krylov(lambda x: A @ x, v, dt, n, hermitian=True)
    np.testing.assert_allclose(got, expm(dt * A) @ v, rtol=1e-10, atol=1e-12)
    w, u = orc.eigh_krylov(lambda x: A @ x, v, n, 1)
    ev, evec = np.linalg.eigh(A)
    assert abs(w[0] - ev[0]) < 1e-12
    assert abs(abs(np.vdot(u[:, 0], evec[:, 0])) - 1) < 1e-10


def test_lanczos_tridiagonalises():
    rng = np.random.default_rng(13)
    n, k = 64, 10
    A = _herm(n, rng)
    alpha, beta, v = orc.lanczos_iteration(lambda x: A @ x, orc.crandn(n, rng), k)
    np.testing.assert_allclose(v.conj().T @ v, np.identity(k), atol=1e-12)
    T = np.diag(alpha) + np.diag(beta, 1) + np.diag(beta, -1)
    np.testing.assert_allclose(v.conj().T @ A @ v, T, atol=1e-12)


def test_lanczos_breakdown_warns_and_truncates():
    """An invariant subspace of dimension 2 stops the recurrence after 2 vectors."""
    A = np.diag([1.0, 2.0, 3.0, 4.0]).astype(complex)
    v = np.array([1.0, 1.0, 0.0, 0.0], dtype=complex)
    with pytest.warns(RuntimeWarning, match="beta"):
        alpha, beta, V = orc.lanczos_iteration(lambda x: A @ x, v, 4)
    assert len(alpha) == 2 and len(beta) == 1 and V.shape == (4, 2)


# ---------------------------------------------------------------------------
# QR / SVD / truncation

def test_block_sparse_factorisations_match_reference_fixture():
    z = load_fixture("kernels_blocksparse")
    q, r, qi = orc.block_sparse_qr(z["qr_a"], z["qr_q0"], z["qr_q1"])
    np.testing.assert_allclose(q, z["qr_q"], atol=1e-13)
    np.testing.assert_allclose(r, z["qr_r"], atol=1e-13)
    assert np.array_equal(qi, z["qr_qi"])
    u, s, v, qs = orc.block_sparse_svd(z["svd_a"], z["svd_q0"], z["svd_q1"])
    np.testing.assert_allclose(s, z["svd_s"], atol=1e-13)
    np.testing.assert_allclose(u @ np.diag(s) @ v, z["svd_a"], atol=1e-13)
    assert np.array_equal(qs, z["svd_q"])
    ut, st, vt, qt = orc.split_block_sparse_matrix_svd(z["svd_a"], z["svd_q0"], z["svd_q1"],
                                                       float(z["split_tol"]))
    np.testing.assert_allclose(st, z["split_s"], atol=1e-13)
    assert np.array_equal(qt, z["split_q"])
    assert np.array_equal(orc.retained_bond_indices(z["ret_s"], 1e-6), z["ret_idx_1e_6"])
    assert np.array_equal(orc.retained_bond_indices(z["ret_s"], 1e-2), z["ret_idx_1e_2"])


def test_truncation_discarded_weight_identity():
    rng = np.random.default_rng(21)
    s = np.sort(rng.uniform(0, 1, 40) ** 6)[::-1]
    for tol in (1e-8, 1e-4, 1e-2):
        idx = orc.retained_bond_indices(s, tol)
        kept = np.zeros(len(s), bool)
        kept[idx] = True
        p = (s / np.linalg.norm(s))**2
        assert p[~kept].sum() <= tol + 1e-15
        # keeping one fewer would violate the tolerance
        assert p[~kept].sum() + p[kept].min() > tol
    idx = orc.retained_bond_indices(s, 1e-8)
    assert np.array_equal(orc.qnum.cap_bond_indices(s, idx, 5), np.arange(5))


def test_qr_conventions_on_rank_deficient_matrix():
    """Exact-zero columns give tau = 0: R has a zero diagonal entry and Q stays orthonormal."""
    rng = np.random.default_rng(4)
    a = orc.crandn((6, 4), rng)
    a[:, 2] = 0
    q, r = np.linalg.qr(a)
    assert abs(r[2, 2]) == 0
    np.testing.assert_allclose(q.conj().T @ q, np.identity(4), atol=1e-14)
    np.testing.assert_allclose(q @ r, a, atol=1e-14)


# ---------------------------------------------------------------------------
# seeded initial states and end-to-end trajectories

def test_seeded_random_mps_matches_reference():
    z = load_fixture("traj_dmrg1_xxz_L16")
    qsite, qb, a = mps_arrays(z, "init_")
    L = len(a)
    psi = orc.MPS(qsite, [np.zeros(len(q), dtype=int) for q in qb], fill="random",
                  rng=np.random.default_rng(int(z["seed"])))
    for i in range(L):
        assert np.array_equal(psi.a[i], a[i])


def _compare_mps(psi, qsite, qb, a, rtol=1e-10):
    for i in range(len(a)):
        assert psi.a[i].shape == a[i].shape
        np.testing.assert_allclose(psi.a[i], a[i], rtol=rtol, atol=rtol * np.abs(a[i]).max())
    for i in range(len(qb)):
        assert np.array_equal(psi.qbonds[i], qb[i])


def test_trajectory_tdvp_singlesite():
    z = load_fixture("traj_tdvp1_xxz_L10")
    psi = _oracle_mps(*mps_arrays(z, "init_"))
    nrm = orc.tdvp_singlesite(_oracle_mpo(z), psi, complex(z["dt"]), int(z["numsteps"]),
                              int(z["numiter"]))
    assert abs(nrm - float(z["nrm"])) < 1e-13
    _compare_mps(psi, *mps_arrays(z, "final_"))


def test_trajectory_tdvp_twosite():
    z = load_fixture("traj_tdvp2_xxz_L12")
    psi = _oracle_mps(*mps_arrays(z, "init_"))
    H = _oracle_mpo(z)
    orc.tdvp_twosite(H, psi, complex(z["dt"]), int(z["numsteps"]), int(z["numiter"]),
                     float(z["tol"]))
    _compare_mps(psi, *mps_arrays(z, "final_"))
    assert abs(orc.mpo_average(psi, H) - complex(z["energy"])) < 1e-11


def test_trajectory_dmrg_singlesite():
    z = load_fixture("traj_dmrg1_xxz_L16")
    psi = _oracle_mps(*mps_arrays(z, "init_"))
    en = orc.dmrg_singlesite(_oracle_mpo(z), psi, int(z["numsweeps"]), int(z["numiter"]))
    np.testing.assert_allclose(en, z["en_min"], rtol=1e-12)
    _compare_mps(psi, *mps_arrays(z, "final_"))


def test_trajectory_dmrg_twosite_quantum_numbers():
    z = load_fixture("traj_dmrg2_xxz_L10_qn")
    psi = _oracle_mps(*mps_arrays(z, "init_"))
    with _nowarn():
        en = orc.dmrg_twosite(_oracle_mpo(z), psi, int(z["numsweeps"]), int(z["numiter"]))
    np.testing.assert_allclose(en, z["en_min"], rtol=1e-12)
    _compare_mps(psi, *mps_arrays(z, "final_"))


class _nowarn:
    def __enter__(self):
        import warnings
        self._c = warnings.catch_warnings()
        self._c.__enter__()
        warnings.simplefilter("ignore", RuntimeWarning)

    def __exit__(self, *a):
        self._c.__exit__(*a)


def test_trajectory_tdvp_quantum_numbers():
    z = load_fixture("traj_tdvp_xxz_L8_qn")
    H = _oracle_mpo(z)
    psi = _oracle_mps(*mps_arrays(z, "init_"))
    p1, p2 = copy.deepcopy(psi), copy.deepcopy(psi)
    with _nowarn():
        orc.tdvp_singlesite(H, p1, complex(z["dt"]), int(z["numsteps"]), 5)
        orc.tdvp_twosite(H, p2, complex(z["dt"]), int(z["numsteps"]), 10)
    _compare_mps(p1, *mps_arrays(z, "one_"))
    _compare_mps(p2, *mps_arrays(z, "two_"))


def test_tdvp_twosite_matches_exact_evolution():
    """Closed form (as test/test_tdvp.py:7-75): 2-site TDVP without truncation and
    with a Krylov space covering the local problem equals expm(-dt N H) psi."""
    rng = np.random.default_rng(17)
    L, J, D, h = 6, 4/3, 5/13, -2/7
    op = heisenberg_xxz_1d_mpo(L, J, D, h).zero_qnumbers()
    H = orc.MPO(op.qsite, op.qbonds, op.a)
    bonds = [1, 2, 4, 8, 4, 2, 1]
    psi = orc.MPS(op.qsite, [np.zeros(b, dtype=int) for b in bonds], fill="random", rng=rng)
    psi.orthonormalize("left")
    dt, n = 0.02 - 0.05j, 4
    ref = expm(-dt * n * xxz_kron(L, J, D, h)) @ psi.to_vector()
    orc.tdvp_twosite(H, psi, dt, n, numiter_lanczos=10)
    np.testing.assert_allclose(psi.to_vector(), ref, atol=1e-10)


def test_dmrg_matches_exact_diagonalisation():
    rng = np.random.default_rng(19)
    L, J, D, h = 8, 4/5, 8/3, -2/7
    op = heisenberg_xxz_1d_mpo(L, J, D, h)
    H = orc.MPO(op.qsite, op.qbonds, op.a)
    psi = orc.MPS.construct_random(L, op.qsite, 0, max_vdim=16, dtype="real", rng=rng)
    with _nowarn():
        en = orc.dmrg_singlesite(H, psi, 4)
    e_exact = np.linalg.eigvalsh(xxz_kron(L, J, D, h))[0]
    assert abs(en[-1] - e_exact) < 1e-12



# ---------------------------------------------------------------------------
# the fidelity measure and the seeded constructors of the product package

def _random_chain(L, D, seed):
    rng = np.random.default_rng(seed)
    bonds = [min(D, 2**i, 2**(L - i)) for i in range(L + 1)]
    return orc.MPS(np.array([0, 0]), [np.zeros(b, dtype=int) for b in bonds], fill="random",
                   rng=rng)


def test_oracle_vdot_and_norm_equal_dense_vectors():
    """<chi|psi> and |psi| of the oracle (mps.py:430-457) equal np.vdot of the
    full state vectors, for chi != psi and for the norm."""
    chi, psi = _random_chain(8, 6, 1), _random_chain(8, 6, 2)
    x, y = chi.to_vector(), psi.to_vector()
    assert abs(orc.mps_vdot(chi, psi) - np.vdot(x, y)) < 1e-13 * np.linalg.norm(x) * np.linalg.norm(y)
    assert abs(orc.mps_vdot(psi, chi) - np.vdot(y, x)) < 1e-13 * np.linalg.norm(x) * np.linalg.norm(y)
    assert abs(orc.mps_norm(psi) - np.linalg.norm(y)) < 1e-13 * np.linalg.norm(y)
    # conjugate symmetry and a planted mistake (no conjugation of chi) that the check catches
    assert abs(orc.mps_vdot(chi, psi) - np.conj(orc.mps_vdot(psi, chi))) < 1e-13
    assert abs(np.sum(x * y) - np.vdot(x, y)) > 1e-3 * abs(np.vdot(x, y))


def test_product_seeded_mps_matches_reference_fixture():
    """pytenet_amd.MPS(fill="random", rng) builds the reference's seeded state
    bit for bit (pytenet/mps.py:28-71, util.py:10-17), on the host."""
    z = load_fixture("traj_dmrg1_xxz_L16")
    qsite, qb, a = mps_arrays(z, "init_")
    psi = ptn.MPS(qsite, [np.zeros(len(q), dtype=int) for q in qb], fill="random",
                  rng=np.random.default_rng(int(z["seed"])))
    assert not psi.on_device
    for i in range(len(a)):
        assert psi.a[i].dtype == a[i].dtype and np.array_equal(psi.a[i], a[i])


def test_product_construct_random_matches_reference_fixture():
    """MPS.construct_random with U(1) quantum numbers (pytenet/mps.py:73-113):
    the same bond labels and bit-identical entries, incl. the zeroed entries of
    the quantum-number sparsity pattern."""
    z = load_fixture("traj_dmrg2_xxz_L10_qn")
    qsite, qb, a = mps_arrays(z, "init_")
    psi = ptn.MPS.construct_random(len(a), qsite, 4, max_vdim=18, dtype="real",
                                   rng=np.random.default_rng(int(z["seed"])))
    for i in range(len(qb)):
        assert np.array_equal(np.asarray(psi.qbonds[i]), qb[i])
    for i in range(len(a)):
        assert np.array_equal(psi.a[i], a[i])
