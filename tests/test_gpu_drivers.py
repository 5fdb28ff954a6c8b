"""
End-to-end parity of the GPU sweep drivers (FP64 mode) with the reference
trajectories (tests/golden) and with closed forms. MPS tensors are compared
through gauge-invariant quantities: the full state vector (the SVD/QR gauge
of individual tensors is a convention), norms, energies and bond dimensions.
"""

import copy
import warnings

import numpy as np
import pytest

import oracle as orc
from conftest import load_fixture, mps_arrays, mpo_arrays
import pytenet_amd as ptn
from pytenet_amd import backend

pytestmark = pytest.mark.gpu


def _mps(qsite, qb, a):
    psi = ptn.MPS(qsite, qb, fill="postpone")
    psi.a = [x.copy() for x in a]
    return psi


def _mpo(z):
    qs, qb, a = mpo_arrays(z)
    return ptn.MPO.from_tensors(qs, qb, a)


def _vector(qsite, qb, a):
    psi = orc.MPS(qsite, qb, fill="postpone")
    psi.a = a
    return psi.to_vector()


def _phase_aligned(x, ref):
    z = np.vdot(ref, x)
    return x * (abs(z) / z)


def test_tdvp_singlesite_reference_trajectory():
    z = load_fixture("traj_tdvp1_xxz_L10")
    psi = _mps(*mps_arrays(z, "init_"))
    nrm = ptn.tdvp_singlesite(_mpo(z), psi, complex(z["dt"]), int(z["numsteps"]),
                              numiter_lanczos=int(z["numiter"]))
    assert abs(nrm - float(z["nrm"])) < 1e-12
    ref = _vector(*mps_arrays(z, "final_"))
    assert isinstance(psi.a[0], np.ndarray)
    np.testing.assert_allclose(psi.to_vector(), ref, atol=1e-10)
    assert psi.bond_dims == [a.shape[0] for a in mps_arrays(z, "final_")[2]] + [1]


def test_tdvp_twosite_reference_trajectory():
    z = load_fixture("traj_tdvp2_xxz_L12")
    psi = _mps(*mps_arrays(z, "init_"))
    H = _mpo(z)
    ptn.tdvp_twosite(H, psi, complex(z["dt"]), int(z["numsteps"]),
                     numiter_lanczos=int(z["numiter"]), tol_split=float(z["tol"]))
    qsite, qb, a = mps_arrays(z, "final_")
    assert psi.bond_dims == [x.shape[0] for x in a] + [1]
    np.testing.assert_allclose(psi.to_vector(), _vector(qsite, qb, a), atol=1e-10)
    assert abs(ptn.mpo_average(psi, H) - complex(z["energy"])) < 1e-10


def test_dmrg_singlesite_reference_trajectory():
    z = load_fixture("traj_dmrg1_xxz_L16")
    psi = _mps(*mps_arrays(z, "init_"))
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        en = ptn.dmrg_singlesite(_mpo(z), psi, int(z["numsweeps"]),
                                 numiter_lanczos=int(z["numiter"]))
    np.testing.assert_allclose(en, z["en_min"], rtol=1e-10)
    ref = _vector(*mps_arrays(z, "final_"))
    np.testing.assert_allclose(_phase_aligned(psi.to_vector(), ref), ref, atol=1e-8)


def test_dmrg_twosite_reference_trajectory_quantum_numbers():
    z = load_fixture("traj_dmrg2_xxz_L10_qn")
    psi = _mps(*mps_arrays(z, "init_"))
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        en = ptn.dmrg_twosite(_mpo(z), psi, int(z["numsweeps"]), numiter_lanczos=int(z["numiter"]))
    np.testing.assert_allclose(en, z["en_min"], rtol=1e-10)
    qsite, qb, a = mps_arrays(z, "final_")
    for i in range(len(qb)):
        assert np.array_equal(np.sort(psi.qbonds[i]), np.sort(qb[i]))
    ref = _vector(qsite, qb, a)
    np.testing.assert_allclose(_phase_aligned(psi.to_vector(), ref), ref, atol=1e-8)


@pytest.mark.parametrize("precision,rtol", [("fp32", 1e-6), ("bf16x3", 1e-4), ("bf16x6", 1e-6)])
def test_dmrg_twosite_quantum_numbers_low_precision(precision, rtol):
    """U(1) sectors with FP32 / split-bf16 tile-skipping products: the energies
    of the reference trajectory to the precision of the contractions (split
    bf16 keeps about 16 bits per product; measured 1.7e-5 here), and the same
    sector labels."""
    z = load_fixture("traj_dmrg2_xxz_L10_qn")
    psi = _mps(*mps_arrays(z, "init_"))
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        en = ptn.dmrg_twosite(_mpo(z), psi, int(z["numsweeps"]), numiter_lanczos=int(z["numiter"]),
                              precision=precision)
    np.testing.assert_allclose(en, z["en_min"], rtol=rtol)
    _, qb, _ = mps_arrays(z, "final_")
    for i in range(len(qb)):
        assert np.array_equal(np.sort(psi.qbonds[i]), np.sort(qb[i]))


def test_tdvp_quantum_numbers_reference_trajectory():
    z = load_fixture("traj_tdvp_xxz_L8_qn")
    H = _mpo(z)
    psi = _mps(*mps_arrays(z, "init_"))
    p1, p2 = copy.deepcopy(psi), copy.deepcopy(psi)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        ptn.tdvp_singlesite(H, p1, complex(z["dt"]), int(z["numsteps"]), numiter_lanczos=5)
        ptn.tdvp_twosite(H, p2, complex(z["dt"]), int(z["numsteps"]), numiter_lanczos=10)
    np.testing.assert_allclose(p1.to_vector(), _vector(*mps_arrays(z, "one_")), atol=1e-10)
    np.testing.assert_allclose(p2.to_vector(), _vector(*mps_arrays(z, "two_")), atol=1e-10)


def test_tdvp_twosite_matches_exact_evolution():
    from scipy.linalg import expm
    rng = np.random.default_rng(17)
    L, J, D, h = 8, 4/3, 5/13, -2/7
    H = ptn.heisenberg_xxz_1d_mpo(L, J, D, h).zero_qnumbers()
    bonds = [1, 2, 4, 8, 16, 8, 4, 2, 1]
    psi = ptn.MPS(H.qsite, [np.zeros(b, dtype=int) for b in bonds], fill="random", rng=rng)
    psi.orthonormalize("left")
    dt, n = 0.02 - 0.05j, 6
    ref = expm(-dt * n * H.to_matrix()) @ psi.to_vector()
    ptn.tdvp_twosite(H, psi, dt, n, numiter_lanczos=10)
    np.testing.assert_allclose(psi.to_vector(), ref, atol=1e-10)


def test_dmrg_twosite_matches_exact_diagonalisation():
    rng = np.random.default_rng(23)
    L = 10
    H = ptn.heisenberg_xxz_1d_mpo(L, -4/3, 2/8, 5/7)
    psi = ptn.MPS.construct_random(L, H.qsite, 4, max_vdim=18, dtype="real", rng=rng)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        en = ptn.dmrg_twosite(H, psi, 2)
    e_exact = np.linalg.eigvalsh(H.to_matrix())[0]
    assert abs(en[-1] - e_exact) < 1e-12


def test_tdvp_time_reversal_symmetry():
    rng = np.random.default_rng(29)
    L = 8
    H = ptn.heisenberg_xxz_1d_mpo(L, 4/3, 5/13, -2/7).zero_qnumbers()
    qb = [np.array([0])] + [np.zeros(5, dtype=int) for _ in range(L - 1)] + [np.array([0])]
    psi = ptn.MPS(H.qsite, qb, fill="random", rng=rng)
    psi.orthonormalize("left")
    ref = psi.to_vector()
    p1 = copy.deepcopy(psi)
    ptn.tdvp_singlesite(H, p1, 0.5j, 1, numiter_lanczos=10)
    ptn.tdvp_singlesite(H, p1, -0.5j, 1, numiter_lanczos=10)
    np.testing.assert_allclose(p1.to_vector(), ref, atol=1e-10)


def test_fp32_contractions_keep_fidelity():
    """Mixed mode: FP32 MFMA contractions, FP64 Krylov/QR/SVD; small 2-site run."""
    rng = np.random.default_rng(41)
    L = 10
    H = ptn.heisenberg_xxz_1d_mpo(L, 1.0, 0.8, -0.1).zero_qnumbers()
    bonds = [min(16, 2**i, 2**(L - i)) for i in range(L + 1)]
    psi = ptn.MPS(H.qsite, [np.zeros(b, dtype=int) for b in bonds], fill="random", rng=rng)
    psi.orthonormalize("left")
    p64, p32 = copy.deepcopy(psi), copy.deepcopy(psi)
    ptn.tdvp_twosite(H, p64, -0.05j, 2, numiter_lanczos=8, tol_split=1e-10)
    ptn.tdvp_twosite(H, p32, -0.05j, 2, numiter_lanczos=8, tol_split=1e-10, precision="fp32")
    f = abs(ptn.mps_vdot(p64, p32))**2 / (ptn.mps_norm(p64)**2 * ptn.mps_norm(p32)**2)
    assert abs(1 - f) < 1e-8
    e64, e32 = ptn.mpo_average(p64, H).real, ptn.mpo_average(p32, H).real
    assert abs(e32 - e64) < 1e-6 * abs(e64)


def test_tdvp_twosite_precond32_svd_matches_fp64_svd():
    """Splits of 128 columns (D=64) use the FP32-preconditioned Jacobi SVD when
    asked: with FP64 contractions the state equals the plain-SVD run to
    rounding, and the default mixed mode keeps the fidelity bound."""
    rng = np.random.default_rng(43)
    L = 12
    H = ptn.heisenberg_xxz_1d_mpo(L, 1.0, 0.8, -0.1).zero_qnumbers()
    bonds = [min(64, 2**i, 2**(L - i)) for i in range(L + 1)]
    psi = ptn.MPS(H.qsite, [np.zeros(b, dtype=int) for b in bonds], fill="random", rng=rng)
    psi.orthonormalize("left")
    runs = {}
    for key, prec, pre in (("plain", "fp64", False), ("pre", "fp64", True), ("mixed", "fp32", None)):
        p = copy.deepcopy(psi)
        ptn.tdvp_twosite(H, p, -0.05j, 2, numiter_lanczos=8, tol_split=0, precision=prec,
                         max_bond=64, svd_precond32=pre)
        runs[key] = p
    def infid(a, b):
        return abs(1 - abs(ptn.mps_vdot(a, b))**2 / (ptn.mps_norm(a)**2 * ptn.mps_norm(b)**2))
    assert max(runs["pre"].bond_dims) == 64
    assert infid(runs["plain"], runs["pre"]) < 1e-12
    e0, e1 = ptn.mpo_average(runs["plain"], H).real, ptn.mpo_average(runs["pre"], H).real
    assert abs(e1 - e0) < 1e-11 * abs(e0)
    assert infid(runs["plain"], runs["mixed"]) < 1e-8


def test_tdvp_checkpoint_resume(tmp_path):
    """Two real-time steps in one run equal one step, a checkpoint, a resumed run
    and one more step (the resumed run re-normalises a state of norm one)."""
    rng = np.random.default_rng(17)
    L = 8
    H = ptn.heisenberg_xxz_1d_mpo(L, 1.0, 0.8, -0.1).zero_qnumbers()
    bonds = [min(8, 2**i, 2**(L - i)) for i in range(L + 1)]
    psi = ptn.MPS(H.qsite, [np.zeros(b, dtype=int) for b in bonds], fill="random", rng=rng)
    psi.orthonormalize("left")
    psi0 = copy.deepcopy(psi)
    ref = copy.deepcopy(psi)
    ptn.tdvp_twosite(H, ref, -0.05j, 2, numiter_lanczos=8, tol_split=0)
    run = ptn.TwoSiteTDVP(H, psi, -0.05j, 8, 0)
    run.setup()
    run.sweep()
    path = tmp_path / "ckpt.npz"
    ptn.tdvp_checkpoint(run, path, 1)
    run2, done = ptn.tdvp_resume(path, H, -0.05j, 8, 0)
    assert done == 1
    run2.setup()
    run2.sweep()
    run2.finish()
    got = run2.psi
    f = abs(ptn.mps_vdot(ref, got))**2 / (ptn.mps_norm(ref)**2 * ptn.mps_norm(got)**2)
    assert abs(1 - f) < 1e-12
    # and the resumed GPU run against the oracle's uninterrupted two steps
    orc_psi = orc.MPS(psi0.qsite, [q.copy() for q in psi0.qbonds], fill="postpone")
    orc_psi.a = [backend.to_host(a).copy() for a in psi0.a]
    orc.tdvp_twosite(orc.MPO(H.qsite, H.qbonds, H.a), orc_psi, -0.05j, 2, numiter_lanczos=8,
                     tol_split=0)
    x, y = got.to_vector(), orc_psi.to_vector()
    x = backend.to_host(x)
    y = y * np.linalg.norm(x) / np.linalg.norm(y)    # the resumed run restarts from norm one
    assert np.abs(_phase_aligned(x, y) - y).max() < 1e-10 * np.linalg.norm(y)


@pytest.mark.parametrize("L,D,steps", [(10, 16, 2), (20, 512, 1)])
def test_bf16x3_contractions_keep_fidelity(L, D, steps):
    """Split-bf16 contractions in the two-site TDVP (FP64 Krylov scalars, QR
    and SVD): fidelity and energy against the all-FP64 run of the same state.
    At D = 512 the two-site products take the pre-split split-bf16 path."""
    rng = np.random.default_rng(43)
    H = ptn.heisenberg_xxz_1d_mpo(L, 1.0, 0.8, -0.1).zero_qnumbers()
    bonds = [min(D, 2**min(i, L - i)) for i in range(L + 1)]
    psi = ptn.MPS(H.qsite, [np.zeros(b, dtype=int) for b in bonds], fill="random", rng=rng)
    psi.orthonormalize("left")
    p64, p16 = copy.deepcopy(psi), copy.deepcopy(psi)
    ptn.tdvp_twosite(H, p64, -0.05j, steps, numiter_lanczos=8, tol_split=1e-10, max_bond=D)
    ptn.tdvp_twosite(H, p16, -0.05j, steps, numiter_lanczos=8, tol_split=1e-10, max_bond=D,
                     precision="bf16x3")
    f = abs(ptn.mps_vdot(p64, p16))**2 / (ptn.mps_norm(p64)**2 * ptn.mps_norm(p16)**2)
    assert abs(1 - f) < 1e-6
    e64, e16 = ptn.mpo_average(p64, H).real, ptn.mpo_average(p16, H).real
    assert abs(e16 - e64) < 1e-4 * abs(e64)


def test_bf16x6_contractions_keep_fp32_fidelity():
    """Three-term split-bf16 contractions (six bf16 MFMA products per operand
    pair, FP32 accuracy) in the two-site TDVP at D = 512, where the two-site
    products take the pre-split three-term path: fidelity and energy against
    the all-FP64 run within the FP32 mode's bounds."""
    rng = np.random.default_rng(43)
    L, D = 20, 512
    H = ptn.heisenberg_xxz_1d_mpo(L, 1.0, 0.8, -0.1).zero_qnumbers()
    bonds = [min(D, 2**min(i, L - i)) for i in range(L + 1)]
    psi = ptn.MPS(H.qsite, [np.zeros(b, dtype=int) for b in bonds], fill="random", rng=rng)
    psi.orthonormalize("left")
    p64, p6 = copy.deepcopy(psi), copy.deepcopy(psi)
    ptn.tdvp_twosite(H, p64, -0.05j, 1, numiter_lanczos=8, tol_split=1e-10, max_bond=D)
    backend.timing_begin()
    ptn.tdvp_twosite(H, p6, -0.05j, 1, numiter_lanczos=8, tol_split=1e-10, max_bond=D,
                     precision="bf16x6")
    names = [t["name"] for t in backend.timing_end()]
    assert any(n.startswith("zgemm_bf16x6_v3<") for n in names), names
    f = abs(ptn.mps_vdot(p64, p6))**2 / (ptn.mps_norm(p64)**2 * ptn.mps_norm(p6)**2)
    assert abs(1 - f) < 1e-8
    e64, e6 = ptn.mpo_average(p64, H).real, ptn.mpo_average(p6, H).real
    assert abs(e6 - e64) < 1e-6 * abs(e64)
