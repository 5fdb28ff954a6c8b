"""
pytenet_amd: the TDVP/DMRG sweep hot path of PyTeNet (cmendl/pytenet 1.3.0)
re-built for AMD MI355X (gfx950).

The public API mirrors the reference's ``pytenet`` namespace for this path
(pytenet/__init__.py:10-36): the MPS/MPO containers, the XXZ and Ising MPO
builders, the chain contractions, the Krylov solvers and the four sweep
drivers keep their names and signatures. Site tensors and environment
blocks live on the GPU during a sweep; every contraction, Lanczos step, QR
and SVD runs in the hand-written HIP kernels of ``libptn_hip.so`` (C ABI in
include/ptn.h). There is no CPU fallback.
"""

from .util import *
from .block_sparse_util import *
from .bond_ops import *
from .mps import *
from .mpo import *
from .hamiltonian import *
from .chain_ops import *
from .krylov import *
from .tdvp import *
from .dmrg import *

__version__ = "0.1.0"
from .batched import *
from .state_io import *
