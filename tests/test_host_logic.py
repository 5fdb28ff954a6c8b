"""
Host-side logic of the GPU path that runs without a GPU: the grouping of
quantum-number sectors into batched SVDs, and the sector bookkeeping that
block_sparse_svd/qr share with the reference (block_sparse_util.py:67-103).
"""

import numpy as np

from pytenet_amd import backend


def test_svd_classes_partition_the_sectors():
    rng = np.random.default_rng(5)
    shapes = [(int(m), int(n)) for m, n in rng.integers(1, 1500, size=(60, 2))]
    groups = backend._svd_classes(shapes)
    flat = sorted(i for g in groups for i in g)
    assert flat == list(range(len(shapes)))
    small = [i for i, (m, n) in enumerate(shapes) if min(m, n) <= 512]
    # every latency-bound sector (smaller dimension <= 512) shares one batch
    assert any(sorted(g) == small for g in groups)
    for g in groups:
        if sorted(g) == small:
            continue
        # larger sectors: one half-octave of the smaller dimension per batch
        ks = {int(np.ceil(2 * np.log2(min(shapes[i])))) for i in g}
        assert len(ks) == 1 and all(min(shapes[i]) > 512 for i in g)


def test_svd_classes_threshold_is_inclusive():
    groups = backend._svd_classes([(600, 512), (513, 700), (40, 3)])
    assert sorted(map(sorted, groups)) == [[0, 2], [1]]


def test_sectors_follow_the_stable_sort_of_the_quantum_numbers():
    q0 = np.array([1, -1, 0, 1, 0, 2])
    q1 = np.array([0, 1, 1, -1, 0])
    secs = backend._sectors(q0, q1)
    # common quantum numbers in ascending order; rows and columns of a sector
    # in their original order
    assert [int(qn) for qn, _, _ in secs] == [-1, 0, 1]
    assert [list(r) for _, r, _ in secs] == [[1], [2, 4], [0, 3]]
    assert [list(c) for _, _, c in secs] == [[3], [0, 4], [1, 2]]


def test_krylov_status_words_map_to_the_reference_errors():
    """Each status bit of the device Lanczos (include/ptn.h) maps to the
    reference's behaviour: breakdown -> RuntimeWarning (krylov.py:44-50), zero
    start -> AssertionError (krylov.py:30), no convergence of the tridiagonal
    eigensolver -> LinAlgError (np.linalg.eigh at krylov.py:150), non-finite
    value -> FloatingPointError."""
    import warnings
    import pytest
    ok = np.array([[8, 0, 0, 0]], dtype=np.int32)
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        backend.report_krylov_status(ok)
    with pytest.warns(RuntimeWarning, match=r"beta\[3\] ~= 0"):
        backend.report_krylov_status(np.array([[4, 4, 0, 0]], dtype=np.int32))
    with pytest.raises(AssertionError):
        backend.report_krylov_status(np.array([[8, 0, 0, 1]], dtype=np.int32))
    with pytest.raises(FloatingPointError):
        backend.report_krylov_status(np.array([[8, 0, 1, 0]], dtype=np.int32))
    with pytest.raises(np.linalg.LinAlgError):
        backend.report_krylov_status(np.array([[8, 0, 2, 0]], dtype=np.int32))
    # every row of a batch is checked
    with pytest.raises(np.linalg.LinAlgError):
        backend.report_krylov_status(np.array([[8, 0, 0, 0], [8, 0, 2, 0]], dtype=np.int32))


def test_padded_product_state_is_the_product_of_its_site_vectors():
    """MPS.padded_product_state: tapered bonds min(D, d^i, d^(L-i)), zeros
    outside a[i][0, :, 0], unit site vectors, and a state vector equal to the
    Kronecker product of the site vectors (computed on the host)."""
    import pytenet_amd as ptn
    psi = ptn.MPS.padded_product_state([0, 0], 7, 4, np.random.default_rng(3))
    assert psi.bond_dims == [1, 2, 4, 4, 4, 4, 2, 1]
    vecs = []
    for a in psi.a:
        v = a[0, :, 0].copy()
        assert abs(np.linalg.norm(v) - 1) < 1e-15
        b = a.copy()
        b[0, :, 0] = 0
        assert not np.any(b)
        vecs.append(v)
    full = vecs[0]
    for v in vecs[1:]:
        full = np.kron(full, v)
    t = psi.a[0]
    for a in psi.a[1:]:
        t = np.tensordot(t, a, axes=(-1, 0))
    assert np.allclose(t.reshape(-1), full, atol=1e-15)
    # seeded: the same draws give the same state
    again = ptn.MPS.padded_product_state([0, 0], 7, 4, np.random.default_rng(3))
    assert all(np.array_equal(x, y) for x, y in zip(psi.a, again.a))


def test_local_op_carries_the_site_labels_only_when_they_fit():
    """The ptn_local_op built for a labelled operator points at the host label
    arrays (int64, contiguous) in the order q_l, q_s, q_r, q_ml, q_mr; labels of
    the wrong length are rejected before any call; without block sparsity or for
    the bond operator no labels are passed (NULL: the dense / tile-mask path)."""
    import ctypes
    import pytest
    import torch
    from pytenet_amd import _lib
    Dl, d, Dr, Ml, Mr = 5, 2, 7, 3, 4
    l = torch.zeros((Dl, Ml, Dl), dtype=torch.complex128)
    r = torch.zeros((Dr, Mr, Dr), dtype=torch.complex128)
    w = torch.zeros((Ml, d, d, Mr), dtype=torch.complex128)
    q = (np.arange(Dl), np.array([1, -1]), np.arange(Dr) - 3, np.array([0, 2, -2]),
         np.array([2, 0, -2, 0]))
    op, keep = backend._local_op(_lib.PTN_OP_HEFF, "fp64", l, w, r, Dl, d, Dr, True, q)
    assert op.precision == _lib.PTN_FP64 | _lib.PTN_BLOCK_SPARSE
    for name, k in zip(("q_l", "q_s", "q_r", "q_ml", "q_mr"), range(5)):
        ptr = getattr(op, name)
        got = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctypes.c_int64)),
                                    shape=(len(q[k]),))
        assert np.array_equal(got, q[k])
    with pytest.raises(AssertionError):
        backend._local_op(_lib.PTN_OP_HEFF, "fp64", l, w, r, Dl, d, Dr, True,
                          (q[0][:-1],) + q[1:])
    op, keep = backend._local_op(_lib.PTN_OP_HEFF, "fp64", l, w, r, Dl, d, Dr, False, q)
    assert keep is None and not op.q_l
    op, keep = backend._local_op(_lib.PTN_OP_BOND, "fp64", l, None, l, Dl, 1, Dl, True, q)
    assert keep is None and not op.q_l
    # an environment step describes its site with only one of the two blocks
    op, keep = backend._local_op(_lib.PTN_OP_HEFF, "fp32", l, w, None, Dl, d, Dr, True, q)
    assert (op.Ml, op.Mr) == (Ml, Mr) and not op.r and op.q_mr


def _bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even), returned as float32."""
    b = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32)


def test_split_bf16_terms_are_exact_and_bound_the_dropped_products():
    """The operand splits of the split-bf16 products (gemm_bf16.hip), emulated
    with exact bf16 rounding: hi + mid + lo reproduces every FP32 value
    exactly (each difference is exact in FP32), the six-term product drops at
    most 2^-23 of a b (|mid| <= 2^-8 |x|, |lo| <= 2^-16 |x|), and the two-term
    split, which drops a_lo b_lo and the representation residues, is about a
    thousand times less accurate at the median."""
    rng = np.random.default_rng(5)
    n = 200_000
    x = (rng.standard_normal(n) * np.exp(rng.uniform(-20, 20, n))).astype(np.float32)
    y = (rng.standard_normal(n) * np.exp(rng.uniform(-20, 20, n))).astype(np.float32)

    def split3(v):
        h = _bf16(v)
        r = (v - h).astype(np.float32)
        m = _bf16(r)
        return h, m, _bf16((r - m).astype(np.float32))

    hx, mx, lx = [t.astype(np.float64) for t in split3(x)]
    hy, my, ly = [t.astype(np.float64) for t in split3(y)]
    assert np.array_equal(hx + mx + lx, x.astype(np.float64))
    assert np.all(np.abs(mx) <= 2.0**-8 * np.abs(x)) and np.all(np.abs(lx) <= 2.0**-16 * np.abs(x))
    ab = x.astype(np.float64) * y
    six = lx * hy + hx * ly + mx * my + mx * hy + hx * my + hx * hy
    rel6 = np.abs(ab - six) / np.abs(ab)
    assert rel6.max() <= 2.0**-23
    h2x = _bf16(x).astype(np.float64)
    l2x = _bf16((x - _bf16(x)).astype(np.float32)).astype(np.float64)
    h2y = _bf16(y).astype(np.float64)
    l2y = _bf16((y - _bf16(y)).astype(np.float32)).astype(np.float64)
    rel3 = np.abs(ab - (h2x * h2y + h2x * l2y + l2x * h2y)) / np.abs(ab)
    assert np.median(rel3) > 500 * np.median(rel6)
