"""
Writes the golden fixtures in this folder by running the reference package
(cmendl/pytenet 1.3.0, imported from /root/reference) on small seeded inputs.

Run it in a container that holds /root/reference:

    python tests/golden/make_fixtures.py

Only numeric arrays are stored (``.npz``, loadable with ``allow_pickle=False``).
Each fixture records its own inputs (seed, sizes, parameters) next to the
reference's outputs, so the oracle and the GPU path can rebuild the same inputs.

What the fixtures pin (nothing else pins these):
  * the exact MPO tensor layout (bond order and gauge) produced by the reference's
    operator-graph construction for the XXZ and Ising chains
    (pytenet/hamiltonian/heisenberg.py:14-54, pytenet/hamiltonian/ising.py:14-69);
  * end-to-end trajectories of tdvp_singlesite / tdvp_twosite /
    dmrg_singlesite / dmrg_twosite on small seeded states
    (pytenet/tdvp.py:26-220, pytenet/dmrg.py:22-178);
  * the local contraction kernels and block-sparse QR/SVD on seeded inputs
    (pytenet/chain_ops.py:16-113,237-317, pytenet/block_sparse_util.py:106-180,244-319,
    pytenet/bond_ops.py:23-54).
"""

import os
import sys
import numpy as np

REF = "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))

sys.path.insert(0, REF)
import pytenet as ptn  # noqa: E402


def _save(name, **arrays):
    path = os.path.join(OUT, name + ".npz")
    np.savez_compressed(path, **arrays)
    print("wrote", path, os.path.getsize(path), "bytes")


def _mps_arrays(prefix, psi):
    out = {prefix + "nsites": np.array(psi.nsites), prefix + "qsite": np.asarray(psi.qsite)}
    for i, a in enumerate(psi.a):
        out[f"{prefix}a{i}"] = a
    for i, q in enumerate(psi.qbonds):
        out[f"{prefix}q{i}"] = np.asarray(q)
    return out


def _mpo_arrays(prefix, op):
    out = {prefix + "nsites": np.array(op.nsites), prefix + "qsite": np.asarray(op.qsite)}
    for i, a in enumerate(op.a):
        out[f"{prefix}a{i}"] = a
    for i, q in enumerate(op.qbonds):
        out[f"{prefix}q{i}"] = np.asarray(q)
    return out


def tapered_bonds(nsites, dmax):
    return [min(dmax, 2**i, 2**(nsites - i)) for i in range(nsites + 1)]


def fixtures_mpo():
    for (L, J, D, h) in [(2, 1.0, 0.8, -0.1), (3, 1.0, 0.8, -0.1), (5, 1.0, 0.8, -0.1),
                         (10, 4.0/3, 5.0/13, -2.0/7)]:
        op = ptn.heisenberg_xxz_1d_mpo(L, J, D, h)
        _save(f"mpo_xxz_L{L}", J=np.array(J), D=np.array(D), h=np.array(h), **_mpo_arrays("", op))
    # zero field: the reference drops zero-coefficient operator chains, which changes the layout
    for (L, J, D, h) in [(2, 1.0, 1.0, 0.0), (3, 0.7, 1.3, 0.0), (6, 1.0, 1.0, 0.0)]:
        op = ptn.heisenberg_xxz_1d_mpo(L, J, D, h)
        _save(f"mpo_xxz_h0_L{L}", J=np.array(J), D=np.array(D), h=np.array(h),
              **_mpo_arrays("", op))
    for (L, J, h, g) in [(1, 1.0, 0.3, 0.7), (2, 1.0, 0.3, 0.7), (4, 0.6, 0.3, 0.7),
                         (5, 1.0, 0.0, 1.2)]:
        op = ptn.ising_1d_mpo(L, J, h, g)
        _save(f"mpo_ising_L{L}", J=np.array(J), h=np.array(h), g=np.array(g), **_mpo_arrays("", op))


def fixtures_kernels():
    rng = np.random.default_rng(1234)
    # apply_local_hamiltonian / bond / env steps on random dense tensors (no quantum numbers)
    Dl, d, Dr, Ml, Mr = 6, 2, 5, 3, 4
    a = ptn.crandn((Dl, d, Dr), rng)
    b = ptn.crandn((Dl, d, Dr), rng)
    w = ptn.crandn((Ml, d, d, Mr), rng)
    l = ptn.crandn((Dl, Ml, Dl), rng)
    r = ptn.crandn((Dr, Mr, Dr), rng)
    c = ptn.crandn((Dl, Dr), rng)
    rb = ptn.crandn((Dr, Ml, Dr), rng)
    _save("kernels_dense",
          a=a, b=b, w=w, l=l, r=r, c=c, rb=rb,
          heff=ptn.apply_local_hamiltonian(a, w, l, r),
          bond=ptn.apply_local_bond_contraction(c, l, rb),
          env_left=ptn.contraction_operator_step_left(a, b, w, l),
          env_right=ptn.contraction_operator_step_right(a, b, w, r))

    # block-sparse QR / SVD with random quantum numbers
    # (same construction as test/test_block_sparse_util.py)
    q0 = rng.integers(-2, 3, size=23)
    q1 = rng.integers(-2, 3, size=15)
    m = ptn.crandn((23, 15), rng)
    ptn.enforce_qsparsity(m, [q0, -q1])
    qq, rr, qi = ptn.block_sparse_qr(m, q0, q1)
    q0s = rng.integers(-2, 3, size=17)
    q1s = rng.integers(-2, 3, size=26)
    ms = ptn.crandn((17, 26), rng)
    ptn.enforce_qsparsity(ms, [q0s, -q1s])
    us, ss, vs, qs = ptn.block_sparse_svd(ms, q0s, q1s)
    tol = 0.05
    ut, st, vt, qt = ptn.split_block_sparse_matrix_svd(ms, q0s, q1s, tol)
    sv = np.array([0.5, 1e-3, 2.0, 1e-6, 0.7, 0.0, 1e-3])
    _save("kernels_blocksparse",
          qr_a=m, qr_q0=q0, qr_q1=q1, qr_q=qq, qr_r=rr, qr_qi=qi,
          svd_a=ms, svd_q0=q0s, svd_q1=q1s, svd_u=us, svd_s=ss, svd_v=vs, svd_q=qs,
          split_tol=np.array(tol), split_u=ut, split_s=st, split_v=vt, split_q=qt,
          ret_s=sv, ret_idx_1e_6=ptn.retained_bond_indices(sv, 1e-6),
          ret_idx_1e_2=ptn.retained_bond_indices(sv, 1e-2))


def fixtures_trajectories():
    # single-site TDVP, XXZ L=10, J=1, Delta=0.8, h=-0.1, zeroed quantum numbers,
    # bond dims [1,2,4,8,16,28,16,8,4,2,1] clamped to 8, left-orthonormalized,
    # dt = 0.01-0.05j, numiter_lanczos = 5 (10 steps here; 100 in the full workload)
    seed, L = 42, 10
    hmpo = ptn.heisenberg_xxz_1d_mpo(L, 1.0, 0.8, -0.1).zero_qnumbers()
    bonds = [1, 2, 4, 8, 16, 28, 16, 8, 4, 2, 1]
    psi = ptn.MPS(hmpo.qsite, [np.zeros(b, dtype=int) for b in bonds], fill="random",
                  rng=np.random.default_rng(seed))
    for i in range(L):
        psi.a[i][8:, :, :] = 0
        psi.a[i][:, :, 8:] = 0
    psi.orthonormalize(mode="left")
    init = _mps_arrays("init_", psi)
    nrm = ptn.tdvp_singlesite(hmpo, psi, 0.01 - 0.05j, 10, numiter_lanczos=5)
    _save("traj_tdvp1_xxz_L10", seed=np.array(seed), dt=np.array(0.01 - 0.05j),
          numsteps=np.array(10), numiter=np.array(5), nrm=np.array(nrm),
          **_mpo_arrays("h_", hmpo), **init, **_mps_arrays("final_", psi))

    # two-site TDVP, XXZ L=12, tapered bond 16, dt=-0.05j, 2 steps, numiter 8, tol 1e-10
    seed, L = 7, 12
    hmpo = ptn.heisenberg_xxz_1d_mpo(L, 1.0, 0.8, -0.1).zero_qnumbers()
    psi = ptn.MPS(hmpo.qsite, [np.zeros(b, dtype=int) for b in tapered_bonds(L, 16)],
                  fill="random", rng=np.random.default_rng(seed))
    psi.orthonormalize(mode="left")
    init = _mps_arrays("init_", psi)
    nrm = ptn.tdvp_twosite(hmpo, psi, -0.05j, 2, numiter_lanczos=8, tol_split=1e-10)
    _save("traj_tdvp2_xxz_L12", seed=np.array(seed), dt=np.array(-0.05j),
          numsteps=np.array(2), numiter=np.array(8), tol=np.array(1e-10), nrm=np.array(nrm),
          energy=np.array(ptn.mpo_average(psi, hmpo)),
          **_mpo_arrays("h_", hmpo), **init, **_mps_arrays("final_", psi))

    # single-site DMRG, XXZ L=16, J=1, Delta=1, h=0, tapered bond 32, 2 sweeps, numiter 10
    seed, L = 11, 16
    hmpo = ptn.heisenberg_xxz_1d_mpo(L, 1.0, 1.0, 0.0).zero_qnumbers()
    psi = ptn.MPS(hmpo.qsite, [np.zeros(b, dtype=int) for b in tapered_bonds(L, 32)],
                  fill="random", rng=np.random.default_rng(seed))
    init = _mps_arrays("init_", psi)
    en = ptn.dmrg_singlesite(hmpo, psi, 2, numiter_lanczos=10)
    _save("traj_dmrg1_xxz_L16", seed=np.array(seed), numsweeps=np.array(2),
          numiter=np.array(10), en_min=en, **_mpo_arrays("h_", hmpo), **init,
          **_mps_arrays("final_", psi))

    # two-site DMRG with U(1) quantum numbers (as test/test_dmrg.py:52-95), seeded
    seed, L = 5, 10
    hmpo = ptn.heisenberg_xxz_1d_mpo(L, -4/3, 2/8, 5/7)
    psi = ptn.MPS.construct_random(L, hmpo.qsite, 4, max_vdim=18, dtype="real",
                                   rng=np.random.default_rng(seed))
    init = _mps_arrays("init_", psi)
    en = ptn.dmrg_twosite(hmpo, psi, 2, numiter_lanczos=25)
    _save("traj_dmrg2_xxz_L10_qn", seed=np.array(seed), numsweeps=np.array(2),
          numiter=np.array(25), en_min=en, **_mpo_arrays("h_", hmpo), **init,
          **_mps_arrays("final_", psi))

    # TDVP with U(1) quantum numbers (as test/test_tdvp.py:7-75), seeded, L=8, 3 steps
    seed, L = 3, 8
    hmpo = ptn.heisenberg_xxz_1d_mpo(L, 4.0/3, 5.0/13, -2.0/7)
    qbonds = [np.array([0])]
    for _ in range(L - 1):
        qbonds.append(np.sort(np.array([q + hmpo.qsite for q in qbonds[-1]]).reshape(-1)))
    qbonds.append(np.array([4]))
    psi = ptn.MPS(hmpo.qsite, qbonds, fill="random", rng=np.random.default_rng(seed))
    psi.orthonormalize(mode="left")
    psi.orthonormalize(mode="right")
    for i in range(L):
        psi.a[i][6:, :, :] = 0
        psi.a[i][:, :, 6:] = 0
    psi.orthonormalize(mode="left")
    init = _mps_arrays("init_", psi)
    import copy
    psi1 = copy.deepcopy(psi)
    psi2 = copy.deepcopy(psi)
    dt = 0.02 - 0.05j
    nrm1 = ptn.tdvp_singlesite(hmpo, psi1, dt, 3, numiter_lanczos=5)
    nrm2 = ptn.tdvp_twosite(hmpo, psi2, dt, 3, numiter_lanczos=10)
    _save("traj_tdvp_xxz_L8_qn", seed=np.array(seed), dt=np.array(dt), numsteps=np.array(3),
          nrm1=np.array(nrm1), nrm2=np.array(nrm2), **_mpo_arrays("h_", hmpo), **init,
          **_mps_arrays("one_", psi1), **_mps_arrays("two_", psi2))


if __name__ == "__main__":
    fixtures_mpo()
    fixtures_kernels()
    fixtures_trajectories()
