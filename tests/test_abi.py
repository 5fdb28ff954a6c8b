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
