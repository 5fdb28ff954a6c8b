"""
Oracle (test infrastructure): the four sweep drivers.

Restates pytenet/tdvp.py:26-238 and pytenet/dmrg.py:22-189 step for step.
`max_bond` (two-site drivers) is the bond-cap extension; None reproduces
the reference, which truncates by discarded weight only.
"""

import numpy as np
from .qnum import qnumber_flatten, is_qsparse, block_sparse_qr
from .mps import (mps_merge_tensor_pair, mps_split_tensor_svd, mpo_merge_tensor_pair,
                  mps_local_orthonormalize_left_qr, mps_local_orthonormalize_right_qr)
from .chain_ops import (contraction_operator_step_right, contraction_operator_step_left,
                        compute_right_operator_blocks, apply_local_hamiltonian,
                        apply_local_bond_contraction)
from .krylov import expm_krylov, eigh_krylov


def _local_hamiltonian_step(l, r, w, a, dt, numiter):
    """expm(-dt H_eff) a by Lanczos (tdvp.py:223-229)."""
    return expm_krylov(lambda x: apply_local_hamiltonian(x.reshape(a.shape), w, l, r).reshape(-1),
                       a.reshape(-1), -dt, numiter, hermitian=True).reshape(a.shape)


def _local_bond_step(l, r, c, dt, numiter):
    """expm(-dt H_bond) c by Lanczos (tdvp.py:232-238)."""
    return expm_krylov(lambda x: apply_local_bond_contraction(x.reshape(c.shape), l, r).reshape(-1),
                       c.reshape(-1), -dt, numiter, hermitian=True).reshape(c.shape)


def _minimize_local_energy(w, l, r, a_start, numiter):
    """Lowest Ritz pair (dmrg.py:181-189)."""
    en, u = eigh_krylov(lambda x: apply_local_hamiltonian(x.reshape(a_start.shape), w, l, r).reshape(-1),
                        a_start.reshape(-1), numiter, 1)
    return en[0], u[:, 0].reshape(a_start.shape)


def _init_blocks(hamiltonian, psi):
    rblocks = compute_right_operator_blocks(psi, hamiltonian)
    lblocks = [None] * psi.nsites
    lblocks[0] = np.array([[[1]]], dtype=rblocks[0].dtype)
    for i, rb in enumerate(rblocks):
        assert is_qsparse(rb, [psi.qbonds[i+1], hamiltonian.qbonds[i+1], -psi.qbonds[i+1]]), \
            "sparsity pattern of operator blocks must match quantum numbers"
    return lblocks, rblocks


def tdvp_singlesite(hamiltonian, psi, dt, numsteps, numiter_lanczos=25):
    """Symmetric single-site TDVP (tdvp.py:26-118). Returns the initial norm."""
    n = hamiltonian.nsites
    assert n == psi.nsites
    nrm = psi.orthonormalize(mode="right")
    lblocks, rblocks = _init_blocks(hamiltonian, psi)
    for _ in range(numsteps):
        for i in range(n - 1):
            psi.a[i] = _local_hamiltonian_step(lblocks[i], rblocks[i], hamiltonian.a[i],
                                               psi.a[i], 0.5 * dt, numiter_lanczos)
            s = psi.a[i].shape
            q, c, psi.qbonds[i+1] = block_sparse_qr(
                psi.a[i].reshape(s[0] * s[1], s[2]),
                qnumber_flatten((psi.qbonds[i], psi.qsite)), psi.qbonds[i+1])
            psi.a[i] = q.reshape(s[0], s[1], q.shape[1])
            lblocks[i+1] = contraction_operator_step_left(psi.a[i], psi.a[i], hamiltonian.a[i],
                                                          lblocks[i])
            c = _local_bond_step(lblocks[i+1], rblocks[i], c, -0.5 * dt, numiter_lanczos)
            psi.a[i+1] = np.tensordot(c, psi.a[i+1], (1, 0))
        i = n - 1
        psi.a[i] = _local_hamiltonian_step(lblocks[i], rblocks[i], hamiltonian.a[i], psi.a[i],
                                           dt, numiter_lanczos)
        for i in reversed(range(1, n)):
            psi.a[i] = psi.a[i].transpose((2, 1, 0))
            s = psi.a[i].shape
            q, c, qbond = block_sparse_qr(
                psi.a[i].reshape(s[0] * s[1], s[2]),
                qnumber_flatten((-psi.qbonds[i+1], psi.qsite)), -psi.qbonds[i])
            psi.qbonds[i] = -qbond
            psi.a[i] = q.reshape(s[0], s[1], q.shape[1]).transpose((2, 1, 0))
            rblocks[i-1] = contraction_operator_step_right(psi.a[i], psi.a[i], hamiltonian.a[i],
                                                           rblocks[i])
            c = np.transpose(c)
            c = _local_bond_step(lblocks[i], rblocks[i-1], c, -0.5 * dt, numiter_lanczos)
            psi.a[i-1] = np.tensordot(psi.a[i-1], c, (2, 0))
            psi.a[i-1] = _local_hamiltonian_step(lblocks[i-1], rblocks[i-1], hamiltonian.a[i-1],
                                                 psi.a[i-1], 0.5 * dt, numiter_lanczos)
    return nrm


def tdvp_twosite(hamiltonian, psi, dt, numsteps, numiter_lanczos=25, tol_split=0,
                 max_bond=None):
    """Symmetric two-site TDVP (tdvp.py:121-220). Returns the initial norm."""
    n = hamiltonian.nsites
    assert n == psi.nsites
    assert n >= 2
    nrm = psi.orthonormalize(mode="right")
    lblocks, rblocks = _init_blocks(hamiltonian, psi)
    h2 = [mpo_merge_tensor_pair(hamiltonian.a[i], hamiltonian.a[i+1]) for i in range(n - 1)]
    for _ in range(numsteps):
        for i in range(n - 2):
            a2 = mps_merge_tensor_pair(psi.a[i], psi.a[i+1])
            a2 = _local_hamiltonian_step(lblocks[i], rblocks[i+1], h2[i], a2, 0.5 * dt,
                                         numiter_lanczos)
            psi.a[i], psi.a[i+1], psi.qbonds[i+1] = mps_split_tensor_svd(
                a2, psi.qsite, psi.qsite, (psi.qbonds[i], psi.qbonds[i+2]), "right",
                tol=tol_split, max_bond=max_bond)
            lblocks[i+1] = contraction_operator_step_left(psi.a[i], psi.a[i], hamiltonian.a[i],
                                                          lblocks[i])
            psi.a[i+1] = _local_hamiltonian_step(lblocks[i+1], rblocks[i+1], hamiltonian.a[i+1],
                                                 psi.a[i+1], -0.5 * dt, numiter_lanczos)
        i = n - 2
        a2 = mps_merge_tensor_pair(psi.a[i], psi.a[i+1])
        a2 = _local_hamiltonian_step(lblocks[i], rblocks[i+1], h2[i], a2, dt, numiter_lanczos)
        psi.a[i], psi.a[i+1], psi.qbonds[i+1] = mps_split_tensor_svd(
            a2, psi.qsite, psi.qsite, [psi.qbonds[i], psi.qbonds[i+2]], "left",
            tol=tol_split, max_bond=max_bond)
        rblocks[i] = contraction_operator_step_right(psi.a[i+1], psi.a[i+1], hamiltonian.a[i+1],
                                                     rblocks[i+1])
        for i in reversed(range(n - 2)):
            psi.a[i+1] = _local_hamiltonian_step(lblocks[i+1], rblocks[i+1], hamiltonian.a[i+1],
                                                 psi.a[i+1], -0.5 * dt, numiter_lanczos)
            a2 = mps_merge_tensor_pair(psi.a[i], psi.a[i+1])
            a2 = _local_hamiltonian_step(lblocks[i], rblocks[i+1], h2[i], a2, 0.5 * dt,
                                         numiter_lanczos)
            psi.a[i], psi.a[i+1], psi.qbonds[i+1] = mps_split_tensor_svd(
                a2, psi.qsite, psi.qsite, [psi.qbonds[i], psi.qbonds[i+2]], "left",
                tol=tol_split, max_bond=max_bond)
            rblocks[i] = contraction_operator_step_right(psi.a[i+1], psi.a[i+1],
                                                         hamiltonian.a[i+1], rblocks[i+1])
    return nrm


def dmrg_singlesite(hamiltonian, psi, numsweeps, numiter_lanczos=25):
    """Single-site DMRG (dmrg.py:22-93). Returns the energy after each sweep."""
    n = hamiltonian.nsites
    assert n == psi.nsites
    psi.orthonormalize(mode="right")
    lblocks, rblocks = _init_blocks(hamiltonian, psi)
    en_min = np.zeros(numsweeps)
    for k in range(numsweeps):
        en = 0
        for i in range(n - 1):
            en, psi.a[i] = _minimize_local_energy(hamiltonian.a[i], lblocks[i], rblocks[i],
                                                  psi.a[i], numiter_lanczos)
            psi.a[i], psi.a[i+1], psi.qbonds[i+1] = mps_local_orthonormalize_left_qr(
                psi.a[i], psi.a[i+1], psi.qsite, psi.qbonds[i:i+2])
            lblocks[i+1] = contraction_operator_step_left(psi.a[i], psi.a[i], hamiltonian.a[i],
                                                          lblocks[i])
        for i in reversed(range(1, n)):
            en, psi.a[i] = _minimize_local_energy(hamiltonian.a[i], lblocks[i], rblocks[i],
                                                  psi.a[i], numiter_lanczos)
            psi.a[i], psi.a[i-1], psi.qbonds[i] = mps_local_orthonormalize_right_qr(
                psi.a[i], psi.a[i-1], psi.qsite, psi.qbonds[i:i+2])
            rblocks[i-1] = contraction_operator_step_right(psi.a[i], psi.a[i], hamiltonian.a[i],
                                                           rblocks[i])
        psi.a[0], _, psi.qbonds[0] = mps_local_orthonormalize_right_qr(
            psi.a[0], np.array([[[1]]]), psi.qsite, psi.qbonds[:2])
        en_min[k] = en
    return en_min


def dmrg_twosite(hamiltonian, psi, numsweeps, numiter_lanczos=25, tol_split=0, max_bond=None):
    """Two-site DMRG (dmrg.py:96-178). Returns the energy after each sweep."""
    n = hamiltonian.nsites
    assert n == psi.nsites
    psi.orthonormalize(mode="right")
    lblocks, rblocks = _init_blocks(hamiltonian, psi)
    en_min = np.zeros(numsweeps)
    h2 = [mpo_merge_tensor_pair(hamiltonian.a[i], hamiltonian.a[i+1]) for i in range(n - 1)]
    for k in range(numsweeps):
        en = 0
        for i in range(n - 2):
            a2 = mps_merge_tensor_pair(psi.a[i], psi.a[i+1])
            en, a2 = _minimize_local_energy(h2[i], lblocks[i], rblocks[i+1], a2, numiter_lanczos)
            psi.a[i], psi.a[i+1], psi.qbonds[i+1] = mps_split_tensor_svd(
                a2, psi.qsite, psi.qsite, [psi.qbonds[i], psi.qbonds[i+2]], "right",
                tol=tol_split, max_bond=max_bond)
            lblocks[i+1] = contraction_operator_step_left(psi.a[i], psi.a[i], hamiltonian.a[i],
                                                          lblocks[i])
        for i in reversed(range(n - 1)):
            a2 = mps_merge_tensor_pair(psi.a[i], psi.a[i+1])
            en, a2 = _minimize_local_energy(h2[i], lblocks[i], rblocks[i+1], a2, numiter_lanczos)
            psi.a[i], psi.a[i+1], psi.qbonds[i+1] = mps_split_tensor_svd(
                a2, psi.qsite, psi.qsite, [psi.qbonds[i], psi.qbonds[i+2]], "left",
                tol=tol_split, max_bond=max_bond)
            rblocks[i] = contraction_operator_step_right(psi.a[i+1], psi.a[i+1],
                                                         hamiltonian.a[i+1], rblocks[i+1])
        psi.a[0], _, psi.qbonds[0] = mps_local_orthonormalize_right_qr(
            psi.a[0], np.array([[[1]]]), psi.qsite, psi.qbonds[:2])
        en_min[k] = en
    return en_min
