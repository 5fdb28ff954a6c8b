"""
Oracle: a CPU restatement (NumPy/SciPy, complex128) of the reference's
TDVP/DMRG sweep hot path, cmendl/pytenet 1.3.0.

THIS PACKAGE IS TEST INFRASTRUCTURE. Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import it, and only as the
checker or the timed CPU baseline. The product package ``pytenet_amd`` never
imports it and has no CPU fallback.

Every function follows the operation order of the reference function it
restates (cited as ``file:line`` relative to the reference tree), so that the
oracle reproduces the reference to rounding on the same inputs.

How each part is pinned (the tests in ``tests/test_oracle_*.py``):
  * golden fixtures written by ``tests/golden/make_fixtures.py`` from the
    reference itself: local contractions, block-sparse QR/SVD, truncation,
    and seeded end-to-end trajectories of all four drivers;
  * closed forms: dense effective-Hamiltonian matrices built by brute force,
    ``expm``/``eigh`` of the full 2^L Hilbert-space Hamiltonian (Kronecker
    construction), exact Krylov results when the Krylov space is the whole
    space, and invariants (isometry, reconstruction, quantum-number sparsity,
    discarded-weight identity, time-reversal symmetry).
Functions without such a check: none at present.
"""

from .util import crandn
from .qnum import (qnumber_outer_sum, common_qnumbers, qnumber_flatten, is_qsparse,
                   enforce_qsparsity, sort_by_qnumbers, slice_with_qnumber,
                   block_sparse_qr, block_sparse_svd, retained_bond_indices,
                   split_block_sparse_matrix_svd)
from .mps import (MPS, MPO, mps_vdot, mps_norm, mps_merge_tensor_pair, mps_split_tensor_svd,
                  mps_local_orthonormalize_left_qr, mps_local_orthonormalize_right_qr,
                  mpo_merge_tensor_pair)
from .chain_ops import (contraction_operator_step_right, contraction_operator_step_left,
                        compute_right_operator_blocks, mpo_average, mpo_inner_product,
                        apply_local_hamiltonian, apply_local_bond_contraction)
from .krylov import lanczos_iteration, eigh_krylov, expm_krylov, eigh_tridiag
from .drivers import tdvp_singlesite, tdvp_twosite, dmrg_singlesite, dmrg_twosite
