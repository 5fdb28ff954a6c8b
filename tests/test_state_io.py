"""
State I/O (npz) of the host-side MPS container; runs without a GPU.
"""

import numpy as np


def test_mps_save_load_roundtrip(tmp_path):
    """State I/O: tensors and quantum numbers survive an .npz round trip exactly."""
    from pytenet_amd.state_io import mps_save, mps_load
    from pytenet_amd.mps import MPS
    rng = np.random.default_rng(3)
    qsite = [1, -1]
    qbonds = [np.array([0]), np.array([1, -1]), np.array([0, 2, 0, -2]), np.array([1, -1]),
              np.array([0])]
    psi = MPS(qsite, qbonds, fill="random", rng=rng)
    path = tmp_path / "psi.npz"
    mps_save(psi, path)
    phi = mps_load(path)
    assert phi.nsites == psi.nsites
    assert np.array_equal(phi.qsite, psi.qsite)
    for qa, qb in zip(phi.qbonds, psi.qbonds):
        assert np.array_equal(qa, qb)
    for a, b in zip(phi.a, psi.a):
        assert a.dtype == np.complex128 and np.array_equal(a, b)
