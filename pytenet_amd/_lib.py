"""
ctypes binding of the C ABI in include/ptn.h (libptn_hip.so, built by
pytenet_amd/build.py). Loading fails loudly when the library is missing:
there is no CPU fallback.
"""

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libptn_hip.so")

PTN_OK = 0
PTN_ERR_ARG = 1
PTN_ERR_WORKSPACE = 2
PTN_ERR_HIP = 3
PTN_ERR_UNSUPPORTED = 4
PTN_ERR_NONFINITE = 5

PTN_FP64 = 0
PTN_FP32 = 1
PTN_BF16X3 = 2
PTN_BF16X6 = 3
PTN_BLOCK_SPARSE = 0x100

PTN_OP_HEFF = 0
PTN_OP_BOND = 1

PTN_SPLIT_LEFT = 0
PTN_SPLIT_RIGHT = 1
PTN_ENV_LEFT = 0
PTN_ENV_RIGHT = 1

PTN_NOALLOC_SCRATCH_BYTES = 64 << 20

c_i64 = ctypes.c_int64
c_p = ctypes.c_void_p
c_sz = ctypes.c_size_t
c_int = ctypes.c_int
c_dbl = ctypes.c_double


class LocalOp(ctypes.Structure):
    """ptn_local_op (include/ptn.h)."""
    _fields_ = [("kind", c_int), ("precision", c_int),
                ("Dl", c_i64), ("d", c_i64), ("Dr", c_i64), ("Ml", c_i64), ("Mr", c_i64),
                ("l", c_p), ("w", c_p), ("r", c_p),
                ("q_l", c_p), ("q_s", c_p), ("q_r", c_p), ("q_ml", c_p), ("q_mr", c_p)]


class KernelTiming(ctypes.Structure):
    """ptn_kernel_timing (include/ptn.h)."""
    _fields_ = [("name", ctypes.c_char * 160), ("launches", c_i64), ("total_ms", c_dbl),
                ("flop", c_dbl), ("bytes", c_dbl)]


# (name, restype, argtypes) for every exported symbol of include/ptn.h
SIGNATURES = [
    ("ptn_version", c_int, []),
    ("ptn_status_string", ctypes.c_char_p, [c_int]),
    ("ptn_last_hip_error", ctypes.c_char_p, []),
    ("ptn_heff_workspace", c_sz, [c_int, c_i64, c_i64, c_i64, c_i64, c_i64]),
    ("ptn_heff_apply", c_int, [c_int, c_i64, c_i64, c_i64, c_i64, c_i64,
                               c_p, c_p, c_p, c_p, c_p, c_p, c_sz, c_p]),
    ("ptn_bond_workspace", c_sz, [c_int, c_i64, c_i64, c_i64]),
    ("ptn_bond_apply", c_int, [c_int, c_i64, c_i64, c_i64, c_p, c_p, c_p, c_p, c_p, c_sz, c_p]),
    ("ptn_local_apply", c_int, [ctypes.POINTER(LocalOp), c_p, c_p, c_p, c_p, c_sz, c_p]),
    ("ptn_env_apply", c_int, [ctypes.POINTER(LocalOp), c_int, c_p, c_p, c_p, c_p, c_p, c_sz, c_p]),
    ("ptn_env_workspace", c_sz, [c_int, c_i64, c_i64, c_i64, c_i64, c_i64]),
    ("ptn_env_left", c_int, [c_int, c_i64, c_i64, c_i64, c_i64, c_i64,
                             c_p, c_p, c_p, c_p, c_p, c_p, c_sz, c_p]),
    ("ptn_env_right", c_int, [c_int, c_i64, c_i64, c_i64, c_i64, c_i64,
                              c_p, c_p, c_p, c_p, c_p, c_p, c_sz, c_p]),
    ("ptn_zgemm_workspace", c_sz, [c_i64, c_i64, c_i64]),
    ("ptn_zgemm", c_int, [c_i64, c_i64, c_i64, c_p, c_i64, c_i64, c_int,
                          c_p, c_i64, c_i64, c_int, c_p, c_i64, c_i64, c_int, c_p, c_sz, c_p]),
    ("ptn_krylov_workspace", c_sz, [ctypes.POINTER(LocalOp), c_int]),
    ("ptn_krylov_expm", c_int, [ctypes.POINTER(LocalOp), c_p, c_dbl, c_dbl, c_int, c_p, c_p, c_p,
                                c_p, c_sz, c_p]),
    ("ptn_krylov_eigh", c_int, [ctypes.POINTER(LocalOp), c_p, c_int, c_p, c_p, c_p, c_p,
                                c_p, c_sz, c_p]),
    ("ptn_qr_workspace", c_sz, [c_i64, c_i64]),
    ("ptn_qr", c_int, [c_i64, c_i64, c_p, c_i64, c_p, c_i64, c_p, c_i64, c_p, c_sz, c_p]),
    ("ptn_svd_workspace", c_sz, [c_i64, c_i64]),
    ("ptn_svd", c_int, [c_i64, c_i64, c_p, c_i64, c_p, c_i64, c_p, c_p, c_i64, c_p,
                        c_p, c_sz, c_p]),
    ("ptn_scale_axis", c_int, [c_i64, c_i64, c_i64, c_p, c_p, c_int, c_p]),
    ("ptn_qsparse3", c_int, [c_i64, c_i64, c_i64, c_p, c_p, c_p, c_p, c_int, c_p, c_p]),
    ("ptn_zgemm_batched", c_int, [c_i64, c_i64, c_i64, c_i64, c_p, c_i64, c_i64, c_i64, c_int,
                                  c_p, c_i64, c_i64, c_i64, c_int, c_p, c_i64, c_i64, c_i64,
                                  c_int, c_p]),
    ("ptn_krylov_workspace_batched", c_sz, [ctypes.POINTER(LocalOp), c_int, c_i64]),
    ("ptn_krylov_expm_batched", c_int, [ctypes.POINTER(LocalOp), c_i64, c_p, c_dbl, c_dbl, c_int,
                                        c_p, c_p, c_p, c_sz, c_p]),
    ("ptn_krylov_eigh_batched", c_int, [ctypes.POINTER(LocalOp), c_i64, c_p, c_int, c_p, c_p, c_p,
                                        c_p, c_sz, c_p]),
    ("ptn_env_workspace_batched", c_sz, [c_int, c_i64, c_i64, c_i64, c_i64, c_i64, c_i64]),
    ("ptn_env_left_batched", c_int, [c_int, c_i64, c_i64, c_i64, c_i64, c_i64, c_i64,
                                     c_p, c_p, c_p, c_p, c_p, c_p, c_sz, c_p]),
    ("ptn_env_right_batched", c_int, [c_int, c_i64, c_i64, c_i64, c_i64, c_i64, c_i64,
                                      c_p, c_p, c_p, c_p, c_p, c_p, c_sz, c_p]),
    ("ptn_qr_workspace_batched", c_sz, [c_i64, c_i64, c_i64]),
    ("ptn_qr_batched", c_int, [c_i64, c_i64, c_i64, c_p, c_p, c_p, c_p, c_sz, c_p]),
    ("ptn_svd_workspace_batched", c_sz, [c_i64, c_i64, c_i64]),
    ("ptn_svd_batched", c_int, [c_i64, c_i64, c_i64, c_p, c_p, c_p, c_p, c_p, c_p, c_sz, c_p]),
    ("ptn_set_autotune", c_int, [c_int]),
    ("ptn_svd_workspace_precond32", c_sz, [c_i64, c_i64, c_i64]),
    ("ptn_svd_precond32", c_int, [c_i64, c_i64, c_i64, c_p, c_p, c_p, c_p, c_p, c_p, c_sz, c_p]),
    ("ptn_svd_split_workspace", c_sz, [c_i64, c_i64, c_int, c_int]),
    ("ptn_svd_split", c_int, [c_i64, c_i64, c_p, c_i64, c_int, c_dbl, c_i64, c_int, c_p, c_p,
                              c_p, ctypes.POINTER(c_i64), c_p, c_p, c_sz, c_p]),
    ("ptn_krylov_expm_noalloc", c_int, [ctypes.POINTER(LocalOp), c_p, c_dbl, c_dbl, c_int, c_p,
                                        c_p, c_p, c_p]),
    ("ptn_set_graphs", c_int, [c_int]),
    ("ptn_graph_launches", c_i64, []),
    ("ptn_timing_begin", c_int, []),
    ("ptn_timing_end", c_int, [ctypes.POINTER(KernelTiming), c_int, ctypes.POINTER(c_int)]),
]

_lib = None


def load():
    """Load libptn_hip.so and declare its prototypes; raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"pytenet_amd: {LIB_PATH} is missing; build it with `python -m pytenet_amd.build` "
                "(there is no CPU fallback)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib
