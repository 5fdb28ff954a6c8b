"""The C-ABI library loads without a GPU and exports every symbol declared in include/ptn.h."""

import ctypes
import os
import re

from pytenet_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "ptn.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ptn_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = _declared()
    for required in ("ptn_heff_apply", "ptn_bond_apply", "ptn_env_left", "ptn_env_right",
                     "ptn_zgemm", "ptn_krylov_expm", "ptn_krylov_eigh", "ptn_qr", "ptn_svd"):
        assert required in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    bound = {name for name, _, _ in _lib.SIGNATURES}
    assert set(_declared()) <= bound


def test_version_and_status_strings():
    lib = _lib.load()
    assert lib.ptn_version() >= 100
    assert lib.ptn_status_string(_lib.PTN_ERR_WORKSPACE) == b"workspace too small"
    # workspace queries are pure host arithmetic
    assert lib.ptn_heff_workspace(_lib.PTN_FP64, 16, 2, 16, 5, 5) >= 2 * 16 * 2 * 5 * 16 * 16


def test_zero_width_bonds_are_valid_shapes():
    """Zero-width virtual bonds (a split that keeps no singular value,
    pytenet/bond_ops.py:28-29) are accepted by the contraction entry points;
    with an empty output nothing is launched, so this runs without a GPU.
    Negative sizes and mode mixes beyond d*M = 32 are rejected."""
    lib = _lib.load()
    nul = ctypes.c_void_p(0)
    fp64 = _lib.PTN_FP64
    ws_h = lib.ptn_heff_workspace(fp64, 0, 2, 6, 5, 5)
    buf = ctypes.create_string_buffer(max(int(ws_h), 1))
    ws = ctypes.cast(buf, ctypes.c_void_p)
    for Dl, Dr in ((0, 6), (6, 0), (0, 0)):
        n = lib.ptn_heff_workspace(fp64, Dl, 2, Dr, 5, 5)
        b = ctypes.create_string_buffer(max(int(n), 1))
        assert lib.ptn_heff_apply(fp64, Dl, 2, Dr, 5, 5, nul, nul, nul, nul, nul,
                                  ctypes.cast(b, ctypes.c_void_p), n, nul) == _lib.PTN_OK
        n = lib.ptn_bond_workspace(fp64, Dl, Dr, 5)
        b = ctypes.create_string_buffer(max(int(n), 1))
        assert lib.ptn_bond_apply(fp64, Dl, Dr, 5, nul, nul, nul, nul,
                                  ctypes.cast(b, ctypes.c_void_p), n, nul) == _lib.PTN_OK
    n = lib.ptn_env_workspace(fp64, 0, 2, 6, 5, 5)
    b = ctypes.create_string_buffer(max(int(n), 1))
    assert lib.ptn_env_right(fp64, 0, 2, 6, 5, 5, nul, nul, nul, nul, nul,
                             ctypes.cast(b, ctypes.c_void_p), n, nul) == _lib.PTN_OK
    assert lib.ptn_heff_apply(fp64, -1, 2, 6, 5, 5, nul, nul, nul, nul, nul, ws, ws_h,
                              nul) == _lib.PTN_ERR_ARG
    n = lib.ptn_heff_workspace(fp64, 4, 8, 4, 5, 5)
    b = ctypes.create_string_buffer(max(int(n), 1))
    assert lib.ptn_heff_apply(fp64, 4, 8, 4, 5, 5, nul, nul, nul, nul, nul,
                              ctypes.cast(b, ctypes.c_void_p), n, nul) == _lib.PTN_ERR_ARG
