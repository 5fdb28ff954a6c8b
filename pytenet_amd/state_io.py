"""
State I/O: an MPS (site tensors and quantum numbers) to and from a NumPy
``.npz`` file, and checkpoint / resume of a two-site TDVP run between time
steps. The reference keeps states in memory only (pytenet/mps.py:14-71); the
file holds exactly those fields: ``qsite``, ``qbond_<i>`` and ``a_<i>``
(complex128, the reference's (Dl, d, Dr) layout), plus ``nsites``. Files are
read with ``allow_pickle=False``.
"""

import numpy as np

from . import backend
from .mps import MPS

__all__ = ["mps_save", "mps_load", "tdvp_checkpoint", "tdvp_resume"]


def mps_save(psi: MPS, path):
    """Write psi (host or device tensors) to `path` (.npz)."""
    arrays = {"nsites": np.asarray(psi.nsites), "qsite": np.asarray(psi.qsite)}
    for i, q in enumerate(psi.qbonds):
        arrays[f"qbond_{i}"] = np.asarray(q)
    for i, a in enumerate(psi.a):
        arrays[f"a_{i}"] = np.ascontiguousarray(backend.to_host(a), dtype=np.complex128)
    np.savez(path, **arrays)


def mps_load(path, device=False) -> MPS:
    """Read an MPS written by mps_save; on the GPU when `device`."""
    with np.load(path, allow_pickle=False) as z:
        n = int(z["nsites"])
        qsite = z["qsite"]
        qbonds = [z[f"qbond_{i}"] for i in range(n + 1)]
        a = [z[f"a_{i}"] for i in range(n)]
    psi = MPS(qsite, qbonds, fill="postpone")
    psi.a = a
    for i in range(n):
        assert psi.a[i].shape == (len(qbonds[i]), len(qsite), len(qbonds[i + 1]))
    if device:
        psi.to_device()
    return psi


def tdvp_checkpoint(run, path, step: int):
    """Save the state of a TwoSiteTDVP run after `step` completed time steps.
    The environment blocks are not stored: they follow from the state."""
    run.ctx.flush()
    mps_save(run.psi, path)
    with np.load(path, allow_pickle=False) as z:
        arrays = {k: z[k] for k in z.files}
    arrays["step"] = np.asarray(step)
    np.savez(path, **arrays)


def tdvp_resume(path, hamiltonian, dt, numiter_lanczos=25, tol_split=0, precision="fp64",
                max_bond=None):
    """A TwoSiteTDVP run continued from a checkpoint: returns (run, completed
    steps). The environments are rebuilt from the stored state by setup()."""
    from .tdvp import TwoSiteTDVP
    psi = mps_load(path, device=True)
    with np.load(path, allow_pickle=False) as z:
        step = int(z["step"])
    run = TwoSiteTDVP(hamiltonian, psi, dt, numiter_lanczos, tol_split, precision, max_bond)
    return run, step
