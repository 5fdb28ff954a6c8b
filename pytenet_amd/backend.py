"""
Device layer: wraps the C ABI of libptn_hip.so for torch device tensors.

PyTorch is used for device memory (its caching allocator owns every buffer),
for the current HIP stream, and for plain data movement (gathers/scatters of
quantum-number sectors). All arithmetic of the sweep runs in the library's
own gfx950 kernels. Every tensor handed to the library is complex128,
contiguous and resident on the current GPU.
"""

import ctypes
import warnings
import numpy as np
import torch

from . import _lib
from ._lib import c_p

PRECISIONS = {"fp64": _lib.PTN_FP64, "fp32": _lib.PTN_FP32, "bf16x3": _lib.PTN_BF16X3,
              "bf16x6": _lib.PTN_BF16X6}


class PtnError(RuntimeError):
    """A call into libptn_hip.so failed."""


def _check(status, what):
    if status != _lib.PTN_OK:
        lib = _lib.load()
        msg = lib.ptn_status_string(status).decode()
        if status == _lib.PTN_ERR_HIP:
            msg += ": " + lib.ptn_last_hip_error().decode()
        raise PtnError(f"{what} failed ({status}: {msg})")


def precision_code(precision, block_sparse=False):
    """Precision code of the C ABI; `block_sparse` adds the flag that makes the
    contractions skip the identically-zero tiles of U(1) block-sparse operands."""
    if precision not in PRECISIONS:
        raise ValueError(f"precision must be one of {sorted(PRECISIONS)}, got {precision!r}")
    return PRECISIONS[precision] | (_lib.PTN_BLOCK_SPARSE if block_sparse else 0)


def device():
    """The current GPU; raises when none is visible (no CPU fallback)."""
    if not torch.cuda.is_available():
        raise RuntimeError("pytenet_amd needs a ROCm GPU; none is visible")
    _lib.load()
    return torch.device("cuda", torch.cuda.current_device())


def stream():
    return c_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    return c_p(t.data_ptr())


class _Workspace:
    """Grow-only scratch buffer per (device, stream), reused across calls: calls
    on one stream are ordered, calls on different streams (the concurrent
    buckets of the many-chain driver) get separate buffers."""

    def __init__(self):
        self.buf = {}

    def get(self, nbytes):
        dev = torch.cuda.current_device()
        key = (dev, torch.cuda.current_stream().cuda_stream)
        cur = self.buf.get(key)
        if cur is None or cur.numel() < nbytes:
            nb = max(int(nbytes), 1 << 20)
            if cur is not None:
                nb = max(nb, int(cur.numel() * 1.25))
            cur = torch.empty(nb, dtype=torch.uint8, device=torch.device("cuda", dev))
            self.buf[key] = cur
        return cur


_ws = _Workspace()


def workspace(nbytes):
    return _ws.get(nbytes)


def as_dev(x):
    """complex128, contiguous, on the current GPU (uploads NumPy arrays)."""
    dev = device()
    if isinstance(x, np.ndarray) or np.isscalar(x):
        return torch.as_tensor(np.ascontiguousarray(x, dtype=np.complex128), device=dev)
    if not isinstance(x, torch.Tensor):
        x = torch.as_tensor(np.asarray(x))
    if x.device != dev:
        x = x.to(dev)
    if x.dtype != torch.complex128:
        x = x.to(torch.complex128)
    return x.contiguous()


def to_host(x):
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy()
    return np.asarray(x)


# ---------------------------------------------------------------------------
# local contractions (K1-K4)

def heff_apply(a, w, l, r, precision="fp64", block_sparse=False):
    """out[lb,sb,rb] = sum L[l,m,lb] W[m,sb,s,m'] A[l,s,r] R[r,m',rb]."""
    a, w, l, r = as_dev(a), as_dev(w), as_dev(l), as_dev(r)
    Dl, d, Dr = a.shape
    Ml, Mr = w.shape[0], w.shape[3]
    _shape_check(w.shape == (Ml, d, d, Mr) and l.shape == (Dl, Ml, Dl) and r.shape == (Dr, Mr, Dr),
                 "apply_local_hamiltonian", a, w, l, r)
    prec = precision_code(precision, block_sparse)
    lib = _lib.load()
    out = torch.empty((Dl, d, Dr), dtype=torch.complex128, device=a.device)
    nb = lib.ptn_heff_workspace(prec, Dl, d, Dr, Ml, Mr)
    ws = workspace(nb)
    _check(lib.ptn_heff_apply(prec, Dl, d, Dr, Ml, Mr, ptr(l), ptr(w), ptr(r), ptr(a), ptr(out),
                              ptr(ws), ws.numel(), stream()), "ptn_heff_apply")
    return out


def heff_apply_sectors(a, w, l, r, qnums, precision="fp64"):
    """H_eff applied once through ptn_local_apply with U(1) quantum numbers
    qnums = (q_l, q_s, q_r, q_ml, q_mr): the sector-grouped products when
    they apply (sector-contiguous virtual bonds; FP64, or FP32 for "fp32" and
    "bf16x6"), else the tile-mask products. Returns (out, grouped)."""
    a, w, l, r = as_dev(a), as_dev(w), as_dev(l), as_dev(r)
    Dl, d, Dr = a.shape
    Ml, Mr = w.shape[0], w.shape[3]
    _shape_check(w.shape == (Ml, d, d, Mr) and l.shape == (Dl, Ml, Dl) and r.shape == (Dr, Mr, Dr),
                 "apply_local_hamiltonian", a, w, l, r)
    op, keep = _local_op(_lib.PTN_OP_HEFF, precision, l, w, r, Dl, d, Dr, True, qnums)
    lib = _lib.load()
    out = torch.empty((Dl, d, Dr), dtype=torch.complex128, device=a.device)
    ws = workspace(lib.ptn_heff_workspace(op.precision, Dl, d, Dr, Ml, Mr))
    grouped = ctypes.c_int32(0)
    _check(lib.ptn_local_apply(ctypes.byref(op), ptr(a), ptr(out), ctypes.addressof(grouped),
                               ptr(ws), ws.numel(), stream()), "ptn_local_apply")
    del keep
    return out, bool(grouped.value)


def bond_apply(c, l, r, precision="fp64", block_sparse=False):
    """out[lb,rb] = sum L[l,m,lb] C[l,r] R[r,m,rb]."""
    c, l, r = as_dev(c), as_dev(l), as_dev(r)
    Dl, Dr = c.shape
    M = l.shape[1]
    _shape_check(l.shape == (Dl, M, Dl) and r.shape == (Dr, M, Dr),
                 "apply_local_bond_contraction", c, l, r)
    prec = precision_code(precision, block_sparse)
    lib = _lib.load()
    out = torch.empty((Dl, Dr), dtype=torch.complex128, device=c.device)
    ws = workspace(lib.ptn_bond_workspace(prec, Dl, Dr, M))
    _check(lib.ptn_bond_apply(prec, Dl, Dr, M, ptr(l), ptr(r), ptr(c), ptr(out), ptr(ws),
                              ws.numel(), stream()), "ptn_bond_apply")
    return out


def _env_sectors(side, a, b, w, blk, qnums, precision, shape):
    """Environment step through ptn_env_apply with U(1) labels (grouped
    products where they apply); returns (out, grouped)."""
    Dl, d, Dr = a.shape
    lib = _lib.load()
    left = side == _lib.PTN_ENV_LEFT
    op, keep = _local_op(_lib.PTN_OP_HEFF, precision, blk if left else None, w,
                         None if left else blk, Dl, d, Dr, True, qnums)
    out = torch.empty(shape, dtype=torch.complex128, device=a.device)
    ws = workspace(lib.ptn_env_workspace(op.precision, Dl, d, Dr, op.Ml, op.Mr))
    grouped = ctypes.c_int32(0)
    _check(lib.ptn_env_apply(ctypes.byref(op), side, ptr(a), ptr(b), ptr(out),
                             ctypes.addressof(grouped), ptr(ws), ws.numel(), stream()),
           "ptn_env_apply")
    del keep
    return out, bool(grouped.value)


def env_left(a, b, w, l, precision="fp64", block_sparse=False, qnums=None):
    """out[r,m',rb] = sum A[l,s,r] W[m,sb,s,m'] L[l,m,lb] conj(B[lb,sb,rb])."""
    a, b, w, l = as_dev(a), as_dev(b), as_dev(w), as_dev(l)
    Dl, d, Dr = a.shape
    Ml, Mr = w.shape[0], w.shape[3]
    _shape_check(b.shape == a.shape and w.shape == (Ml, d, d, Mr) and l.shape == (Dl, Ml, Dl),
                 "contraction_operator_step_left", a, b, w, l)
    if block_sparse and qnums is not None:
        return _env_sectors(_lib.PTN_ENV_LEFT, a, b, w, l, qnums, precision, (Dr, Mr, Dr))[0]
    prec = precision_code(precision, block_sparse)
    lib = _lib.load()
    out = torch.empty((Dr, Mr, Dr), dtype=torch.complex128, device=a.device)
    ws = workspace(lib.ptn_env_workspace(prec, Dl, d, Dr, Ml, Mr))
    _check(lib.ptn_env_left(prec, Dl, d, Dr, Ml, Mr, ptr(a), ptr(b), ptr(w), ptr(l), ptr(out),
                            ptr(ws), ws.numel(), stream()), "ptn_env_left")
    return out


def env_right(a, b, w, r, precision="fp64", block_sparse=False, qnums=None):
    """out[l,m,lb] = sum A[l,s,r] W[m,sb,s,m'] R[r,m',rb] conj(B[lb,sb,rb])."""
    a, b, w, r = as_dev(a), as_dev(b), as_dev(w), as_dev(r)
    Dl, d, Dr = a.shape
    Ml, Mr = w.shape[0], w.shape[3]
    _shape_check(b.shape == a.shape and w.shape == (Ml, d, d, Mr) and r.shape == (Dr, Mr, Dr),
                 "contraction_operator_step_right", a, b, w, r)
    if block_sparse and qnums is not None:
        return _env_sectors(_lib.PTN_ENV_RIGHT, a, b, w, r, qnums, precision, (Dl, Ml, Dl))[0]
    prec = precision_code(precision, block_sparse)
    lib = _lib.load()
    out = torch.empty((Dl, Ml, Dl), dtype=torch.complex128, device=a.device)
    ws = workspace(lib.ptn_env_workspace(prec, Dl, d, Dr, Ml, Mr))
    _check(lib.ptn_env_right(prec, Dl, d, Dr, Ml, Mr, ptr(a), ptr(b), ptr(w), ptr(r), ptr(out),
                             ptr(ws), ws.numel(), stream()), "ptn_env_right")
    return out


def _shape_check(ok, what, *ts):
    if not ok:
        raise AssertionError(f"{what}: incompatible shapes {[tuple(t.shape) for t in ts]}")


# ---------------------------------------------------------------------------
# strided complex products (K5, K7)

def matmul(A, B, conjA=False, conjB=False, out=None, accumulate=False):
    """C = op(A) @ op(B) for 2-D complex128 device tensors of any strides."""
    assert A.dim() == 2 and B.dim() == 2 and A.shape[1] == B.shape[0]
    assert A.dtype == torch.complex128 and B.dtype == torch.complex128
    M, K = A.shape
    N = B.shape[1]
    if out is None:
        out = torch.empty((M, N), dtype=torch.complex128, device=A.device)
        accumulate = False
    lib = _lib.load()
    _check(lib.ptn_zgemm(M, N, K, ptr(A), A.stride(0), A.stride(1), int(conjA),
                         ptr(B), B.stride(0), B.stride(1), int(conjB),
                         ptr(out), out.stride(0), out.stride(1), int(accumulate),
                         None, 0, stream()), "ptn_zgemm")
    return out


def merge_pair(a0, a1):
    """Theta[l, s0*d1+s1, r] = sum_k a0[l,s0,k] a1[k,s1,r]."""
    a0, a1 = as_dev(a0), as_dev(a1)
    Dl, d0, k = a0.shape
    k1, d1, Dr = a1.shape
    assert k == k1
    t = matmul(a0.reshape(Dl * d0, k), a1.reshape(k, d1 * Dr))
    return t.reshape(Dl, d0 * d1, Dr)


# ---------------------------------------------------------------------------
# Krylov solvers (K8-K10)

def _local_op(kind, precision, l, w, r, Dl, d, Dr, block_sparse=False, qnums=None):
    """ptn_local_op of a local operator, and the host label arrays it points to
    (keep them alive for the call). qnums = (q_l, q_s, q_r, q_ml, q_mr) or None."""
    op = _lib.LocalOp()
    op.kind = kind
    op.precision = precision_code(precision, block_sparse)
    op.Dl, op.d, op.Dr = Dl, d, Dr
    op.Ml = l.shape[1] if l is not None else w.shape[0]
    op.Mr = r.shape[1] if r is not None else w.shape[3]
    op.l = l.data_ptr() if l is not None else None
    op.w = w.data_ptr() if w is not None else None
    op.r = r.data_ptr() if r is not None else None
    keep = None
    if qnums is not None and block_sparse and kind == _lib.PTN_OP_HEFF:
        keep = [np.ascontiguousarray(np.asarray(q, dtype=np.int64).reshape(-1)) for q in qnums]
        want = (Dl, d, Dr, op.Ml, op.Mr)
        assert tuple(len(q) for q in keep) == want, "quantum numbers do not match the shapes"
        op.q_l, op.q_s, op.q_r, op.q_ml, op.q_mr = (q.ctypes.data for q in keep)
    return op, keep


def krylov_local(kind, v, l, w, r, numiter, mode, dt=0.0, precision="fp64", info=None,
                 block_sparse=False, qnums=None):
    """Lanczos on the local operator (PTN_OP_HEFF or PTN_OP_BOND):
    mode "expm" -> expm(dt*H) v; mode "eigh" -> (lowest Ritz vector, energy tensor).
    `info` (int32[4] device tensor) receives the status words; returns
    (out, energy or None, info). With block_sparse and the U(1) labels
    qnums = (q_l, q_s, q_r, q_ml, q_mr) of an H_eff operator the matvecs use
    the sector-grouped products where they apply."""
    v, l, r = as_dev(v), as_dev(l), as_dev(r)
    if w is not None:
        w = as_dev(w)
    if v.numel() == 0:
        # an empty start vector has norm zero (the reference asserts, krylov.py:30)
        raise AssertionError("Lanczos start vector has norm zero")
    if kind == _lib.PTN_OP_HEFF:
        Dl, d, Dr = v.shape
    else:
        (Dl, Dr), d = v.shape, 1
    op, keep = _local_op(kind, precision, l, w, r, Dl, d, Dr, block_sparse, qnums)
    lib = _lib.load()
    nb = lib.ptn_krylov_workspace(ctypes.byref(op), int(numiter))
    ws = workspace(nb)
    out = torch.empty_like(v)
    if info is None:
        info = torch.zeros(4, dtype=torch.int32, device=v.device)
    if mode == "expm":
        dt = complex(dt)
        _check(lib.ptn_krylov_expm(ctypes.byref(op), ptr(v), dt.real, dt.imag, int(numiter),
                                   ptr(out), ptr(info), None, ptr(ws), ws.numel(), stream()),
               "ptn_krylov_expm")
        return out, None, info
    energy = torch.zeros(1, dtype=torch.float64, device=v.device)
    _check(lib.ptn_krylov_eigh(ctypes.byref(op), ptr(v), int(numiter), ptr(out), ptr(energy),
                               ptr(info), None, ptr(ws), ws.numel(), stream()), "ptn_krylov_eigh")
    return out, energy, info


def report_krylov_status(info_rows):
    """Emit the reference's breakdown warnings (krylov.py:44-50) and raise on
    non-finite values or a zero start vector (krylov.py:30), or when the
    tridiagonal eigensolver did not converge (where the reference's
    np.linalg.eigh, krylov.py:150, raises LinAlgError), from a host copy of
    the status words (rows of 4 int32: vectors used, breakdown step + 1,
    bit 0 non-finite | bit 1 no convergence, zero start)."""
    for row in np.asarray(info_rows).reshape(-1, 4):
        if row[3]:
            raise AssertionError("Lanczos start vector has norm zero")
        if row[2] & 1:
            raise FloatingPointError("non-finite value encountered during Lanczos iteration")
        if row[2] & 2:
            raise np.linalg.LinAlgError("Eigenvalues did not converge (Lanczos tridiagonal "
                                        "matrix)")
        if row[1]:
            warnings.warn(f"beta[{row[1] - 1}] ~= 0 encountered during Lanczos iteration.",
                          RuntimeWarning)


# ---------------------------------------------------------------------------
# QR / SVD (K11, K12) and elementwise helpers

def qr(a):
    """Reduced Householder QR of a 2-D complex128 device tensor (LAPACK conventions)."""
    a = as_dev(a)
    m, n = a.shape
    k = min(m, n)
    q = torch.empty((m, k), dtype=torch.complex128, device=a.device)
    r = torch.empty((k, n), dtype=torch.complex128, device=a.device)
    if m == 0 or n == 0:
        return q, r
    lib = _lib.load()
    ws = workspace(lib.ptn_qr_workspace(m, n))
    _check(lib.ptn_qr(m, n, ptr(a), a.stride(0), ptr(q), k, ptr(r), n, ptr(ws), ws.numel(),
                      stream()), "ptn_qr")
    return q, r


def timing_begin():
    """Start the live launch timing of the complex products (HIP events on the
    stream of every launch; no synchronisation)."""
    _check(_lib.load().ptn_timing_begin(), "ptn_timing_begin")


def timing_end(max_rows=64):
    """Stop the launch timing; per kernel instance (largest total time first):
    dict(name, launches, total_ms, avg_ms, flop, bytes). Waits for the events."""
    lib = _lib.load()
    rows = (_lib.KernelTiming * max_rows)()
    n = ctypes.c_int(0)
    _check(lib.ptn_timing_end(rows, max_rows, ctypes.byref(n)), "ptn_timing_end")
    out = []
    for r in rows[:n.value]:
        out.append({"name": r.name.decode(), "launches": int(r.launches),
                    "total_ms": float(r.total_ms),
                    "avg_ms": float(r.total_ms) / max(int(r.launches), 1),
                    "flop": float(r.flop), "bytes": float(r.bytes)})
    return out


def set_graphs(on=True):
    """Switch the hipGraph capture of short local solves on or off (process-wide);
    returns the previous setting."""
    return bool(_lib.load().ptn_set_graphs(int(bool(on))))


def graph_launches():
    """Number of captured-solve graph launches so far in this process."""
    return int(_lib.load().ptn_graph_launches())


def set_autotune(on=True):
    """Switch the tile autotuner of the complex products on or off (process-wide;
    the first product of each shape then times the tile shapes once). Returns
    the previous setting."""
    return bool(_lib.load().ptn_set_autotune(int(bool(on))))


def svd(a, return_info=False, precond32=False):
    """Thin SVD of a 2-D complex128 device tensor: (u, s, vh), s descending (device float64).
    Synchronises the stream (the Jacobi sweeps are checked for convergence on the host).
    With return_info, also returns the status words (FP64 sweeps, not converged,
    non-finite, FP32 sweeps). With precond32, the Jacobi iteration starts with
    FP32 sweeps and finishes in FP64 (same output accuracy)."""
    a = as_dev(a)
    m, n = a.shape
    k = min(m, n)
    u = torch.empty((m, k), dtype=torch.complex128, device=a.device)
    vh = torch.empty((k, n), dtype=torch.complex128, device=a.device)
    s = torch.empty(k, dtype=torch.float64, device=a.device)
    if k == 0:
        return (u, s, vh, np.zeros(4, dtype=np.int32)) if return_info else (u, s, vh)
    lib = _lib.load()
    info = torch.zeros(4, dtype=torch.int32, device=a.device)
    if precond32:
        a = a.contiguous()
        ws = workspace(lib.ptn_svd_workspace_precond32(1, m, n))
        _check(lib.ptn_svd_precond32(1, m, n, ptr(a), ptr(u), ptr(s), ptr(vh), ptr(info), ptr(ws),
                                     ws.numel(), stream()), "ptn_svd_precond32")
    else:
        ws = workspace(lib.ptn_svd_workspace(m, n))
        _check(lib.ptn_svd(m, n, ptr(a), a.stride(0), ptr(u), k, ptr(s), ptr(vh), n, ptr(info),
                           ptr(ws), ws.numel(), stream()), "ptn_svd")
    inf = info.cpu().numpy()
    _svd_status(inf, f"a {m} x {n} matrix")
    if return_info:
        return u, s, vh, inf
    return u, s, vh


def svd_split(a, distr, tol=0.0, max_bond=None, precond32=False):
    """Truncating split of a 2-D complex128 device tensor (one sector):
    (a0, a1, s) with distr "right": a0 = U_k, a1 = Sigma_k V_k^H; "left":
    a0 = U_k Sigma_k, a1 = V_k^H; s = all singular values (host, descending).
    k follows the reference's discarded-weight rule with `tol` (bond_ops.py:23-38)
    and the cap `max_bond`. Synchronises the stream."""
    a = as_dev(a)
    m, n = a.shape
    code = {"left": _lib.PTN_SPLIT_LEFT, "right": _lib.PTN_SPLIT_RIGHT}[distr]
    q = min(m, n)
    kmax = q if max_bond is None else max(min(q, int(max_bond)), 0)
    a0 = torch.empty(max(m * kmax, 1), dtype=torch.complex128, device=a.device)
    a1 = torch.empty(max(kmax * n, 1), dtype=torch.complex128, device=a.device)
    s = np.zeros(q, dtype=np.float64)
    if q == 0 or kmax == 0:
        return a0[:0].reshape(m, 0), a1[:0].reshape(0, n), s
    lib = _lib.load()
    ws = workspace(lib.ptn_svd_split_workspace(m, n, code, int(bool(precond32))))
    k = ctypes.c_int64(0)
    info = np.zeros(4, dtype=np.int32)
    st = lib.ptn_svd_split(m, n, ptr(a), a.stride(0), code, float(tol),
                           0 if max_bond is None else int(max_bond), int(bool(precond32)),
                           ptr(a0), ptr(a1), s.ctypes.data_as(c_p), ctypes.byref(k),
                           info.ctypes.data_as(c_p), ptr(ws), ws.numel(), stream())
    if st == _lib.PTN_ERR_NONFINITE:
        raise FloatingPointError("non-finite singular value")
    _check(st, "ptn_svd_split")
    _svd_status(info, f"a {m} x {n} matrix")
    k = k.value
    return a0[:m * k].reshape(m, k), a1[:k * n].reshape(k, n), s


def scale_axis(t, s, axis):
    """In place: t (n0, n1, n2) *= s along axis 0 or 2 (s real float64 device tensor)."""
    assert t.is_contiguous() and t.dtype == torch.complex128
    s = s.to(device=t.device, dtype=torch.float64).contiguous()
    lib = _lib.load()
    n0, n1, n2 = t.shape
    _check(lib.ptn_scale_axis(n0, n1, n2, ptr(t), ptr(s), axis, stream()), "ptn_scale_axis")
    return t


def _qnums3(t, qnums):
    qs = [np.asarray(q, dtype=np.int64).reshape(-1) for q in qnums]
    shape = tuple(len(q) for q in qs)
    assert shape == tuple(t.shape), "quantum numbers must match the tensor shape"
    return [torch.as_tensor(q, device=t.device) for q in qs]


def _as3(t, qnums):
    """View a tensor with len(qnums) axes as 3 axes (merging trailing/leading axes)."""
    qs = [np.asarray(q, dtype=np.int64).reshape(-1) for q in qnums]
    if len(qs) == 3:
        return t, qs
    if len(qs) == 2:
        return t.reshape(t.shape[0], 1, t.shape[1]), [qs[0], np.zeros(1, np.int64), qs[1]]
    if len(qs) == 4:
        mid = np.add.outer(qs[1], qs[2]).reshape(-1)
        return t.reshape(t.shape[0], t.shape[1] * t.shape[2], t.shape[3]), [qs[0], mid, qs[3]]
    raise ValueError("quantum-number checks support 2, 3 or 4 axes")


def qsparse_violation(t, qnums):
    """max |entry| over entries whose quantum numbers do not sum to zero."""
    t = as_dev(t)
    t3, qs = _as3(t, qnums)
    q0, q1, q2 = _qnums3(t3, qs)
    out = torch.zeros(1, dtype=torch.float64, device=t.device)
    lib = _lib.load()
    _check(lib.ptn_qsparse3(*t3.shape, ptr(q0), ptr(q1), ptr(q2), ptr(t3), 0, ptr(out),
                            stream()), "ptn_qsparse3")
    return float(out.item())


def enforce_qsparsity(t, qnums):
    """Zero, in place, the entries of a complex128 device tensor whose quantum numbers do not sum to zero."""
    assert isinstance(t, torch.Tensor) and t.dtype == torch.complex128 and t.is_contiguous()
    t3, qs = _as3(t, qnums)
    q0, q1, q2 = _qnums3(t3, qs)
    out = torch.zeros(1, dtype=torch.float64, device=t.device)
    lib = _lib.load()
    _check(lib.ptn_qsparse3(*t3.shape, ptr(q0), ptr(q1), ptr(q2), ptr(t3), 1, ptr(out),
                            stream()), "ptn_qsparse3")


# ---------------------------------------------------------------------------
# block-sparse factorisations (block_sparse_util.py:106-180, 244-319)

def _sectors(q0, q1):
    """Host bookkeeping: stable sort permutations and the common sectors."""
    p0 = np.argsort(q0, kind="mergesort")
    p1 = np.argsort(q1, kind="mergesort")
    s0, s1 = q0[p0], q1[p1]
    common = np.intersect1d(q0, q1)
    out = []
    for qn in common:
        r = np.flatnonzero(s0 == qn)
        c = np.flatnonzero(s1 == qn)
        out.append((qn, p0[r[0]:r[-1] + 1], p1[c[0]:c[-1] + 1]))
    return out


def block_sparse_qr(a, q0, q1):
    """Per-sector QR; the sector blocks are gathered, factorised on the GPU and
    scattered back in sector order (the column order of the reference)."""
    a = as_dev(a)
    assert a.dim() == 2 and len(q0) == a.shape[0] and len(q1) == a.shape[1]
    q0 = np.asarray(q0)
    q1 = np.asarray(q1)
    if not np.any(q0) and not np.any(q1):
        # one sector holding the whole matrix (zero quantum numbers)
        q, r = qr(a)
        return q, r, np.zeros(q.shape[1], dtype=q0.dtype)
    secs = _sectors(q0, q1)
    if not secs:
        if qsparse_violation(a, [np.zeros(a.shape[0], np.int64), np.zeros(a.shape[1], np.int64)]) != 0:
            raise AssertionError("no common quantum numbers, but the matrix is not zero")
        q = torch.zeros((a.shape[0], 1), dtype=torch.complex128, device=a.device)
        q[0, 0] = 1
        r = torch.zeros((1, a.shape[1]), dtype=torch.complex128, device=a.device)
        return q, r, q0[:1]
    kmax = min(a.shape)
    q = torch.zeros((a.shape[0], kmax), dtype=torch.complex128, device=a.device)
    r = torch.zeros((kmax, a.shape[1]), dtype=torch.complex128, device=a.device)
    qi = np.zeros(kmax, dtype=q0.dtype)
    k = 0
    for qn, rows, cols in secs:
        ri = torch.as_tensor(rows, device=a.device)
        ci = torch.as_tensor(cols, device=a.device)
        blk = a.index_select(0, ri).index_select(1, ci).contiguous()
        qs, rs = qr(blk)
        kk = qs.shape[1]
        q[ri, k:k + kk] = qs
        r[k:k + kk, ci] = rs
        qi[k:k + kk] = qn
        k += kk
    return q[:, :k].contiguous(), r[:k, :].contiguous(), qi[:k]


def _svd_status(inf, what):
    if inf[2]:
        raise FloatingPointError("non-finite singular value")
    if inf[1]:
        warnings.warn(f"Jacobi SVD of {what} stopped after {inf[0]} sweeps before every column "
                      "pair met the orthogonality tolerance", RuntimeWarning)


def svd_batched(a, return_info=False, precond32=False):
    """Thin SVDs of a stack (B, m, n) of complex128 device matrices, all Jacobi
    sweeps run together: (u (B, m, k), s (B, k) device float64, vh (B, k, n))."""
    a = as_dev(a).contiguous()
    nb, m, n = a.shape
    k = min(m, n)
    u = torch.empty((nb, m, k), dtype=torch.complex128, device=a.device)
    vh = torch.empty((nb, k, n), dtype=torch.complex128, device=a.device)
    s = torch.empty((nb, k), dtype=torch.float64, device=a.device)
    if k == 0 or nb == 0:
        return (u, s, vh, np.zeros(4, dtype=np.int32)) if return_info else (u, s, vh)
    lib = _lib.load()
    info = torch.zeros(4, dtype=torch.int32, device=a.device)
    if precond32:
        ws = workspace(lib.ptn_svd_workspace_precond32(nb, m, n))
        fn, name = lib.ptn_svd_precond32, "ptn_svd_precond32"
    else:
        ws = workspace(lib.ptn_svd_workspace_batched(nb, m, n))
        fn, name = lib.ptn_svd_batched, "ptn_svd_batched"
    _check(fn(nb, m, n, ptr(a), ptr(u), ptr(s), ptr(vh), ptr(info), ptr(ws), ws.numel(),
              stream()), name)
    inf = info.cpu().numpy()
    _svd_status(inf, f"{nb} matrices of {m} x {n}")
    if return_info:
        return u, s, vh, inf
    return u, s, vh


def _svd_classes(shapes, merge_below=512):
    """Group sector shapes for batched SVDs: a wide block is factorised through
    its adjoint. Blocks whose smaller dimension is at most `merge_below` share
    one zero-padded batch (their Jacobi rounds and QR panels are bound by
    launch latency, so one batch at the largest shape costs about as much per
    round as the smallest would, and the number of dependent rounds is that of
    the largest block alone); larger blocks are grouped by half-octave of the
    smaller dimension."""
    groups = {}
    for idx, (ms, ns) in enumerate(shapes):
        small = min(ms, ns)
        key = -1 if small <= merge_below else int(np.ceil(2 * np.log2(small)))
        groups.setdefault(key, []).append(idx)
    return list(groups.values())


def block_sparse_svd(a, q0, q1, precond32=False):
    """Per-sector SVD (block_sparse_util.py:244-310): (u, s, v, q) with s a host
    float64 array (sector order, descending within a sector) and u, v device
    tensors. The sector blocks are factorised in batches: each block is placed,
    zero-padded, in a common (P, Q) shape (P >= Q, wide blocks as their
    adjoint). The padding stays exactly zero through the Householder and Jacobi
    steps (zero columns give identity reflectors and zero Gram entries, so no
    rotation touches them), hence the first min(m_s, n_s) singular triplets of
    a padded block are those of the sector block itself."""
    a = as_dev(a)
    assert a.dim() == 2 and len(q0) == a.shape[0] and len(q1) == a.shape[1]
    q0 = np.asarray(q0)
    q1 = np.asarray(q1)
    if not np.any(q0) and not np.any(q1):
        u, s, vh = svd(a, precond32=precond32)
        return u, s.cpu().numpy(), vh, np.zeros(len(s), dtype=q0.dtype)
    secs = _sectors(q0, q1)
    if not secs:
        u = torch.zeros((a.shape[0], 1), dtype=torch.complex128, device=a.device)
        if a.shape[0] > 0:
            u[0, 0] = 1
        v = torch.zeros((1, a.shape[1]), dtype=torch.complex128, device=a.device)
        return u, np.zeros(1), v, q0[:1]
    idx = [(torch.as_tensor(rows, device=a.device), torch.as_tensor(cols, device=a.device))
           for _, rows, cols in secs]
    shapes = [(len(rows), len(cols)) for _, rows, cols in secs]
    res = [None] * len(secs)
    for group in _svd_classes(shapes):
        P = max(max(shapes[j]) for j in group)
        Q = max(min(shapes[j]) for j in group)
        pad = torch.zeros((len(group), P, Q), dtype=torch.complex128, device=a.device)
        for b, j in enumerate(group):
            ri, ci = idx[j]
            blk = a.index_select(0, ri).index_select(1, ci)
            ms, ns = shapes[j]
            if ms >= ns:
                pad[b, :ms, :ns] = blk
            else:
                pad[b, :ns, :ms] = blk.mH
        ub, sb, vb = svd_batched(pad, precond32=precond32)
        for b, j in enumerate(group):
            ms, ns = shapes[j]
            kk = min(ms, ns)
            if ms >= ns:
                res[j] = (ub[b, :ms, :kk], sb[b, :kk], vb[b, :kk, :ns])
            else:
                res[j] = (vb[b, :kk, :ms].mH, sb[b, :kk], ub[b, :ns, :kk].mH)
    kmax = min(a.shape)
    u = torch.zeros((a.shape[0], kmax), dtype=torch.complex128, device=a.device)
    v = torch.zeros((kmax, a.shape[1]), dtype=torch.complex128, device=a.device)
    s_parts, qi = [], np.zeros(kmax, dtype=q0.dtype)
    k = 0
    for (qn, _, _), (ri, ci), (us, ss, vs) in zip(secs, idx, res):
        kk = len(ss)
        u[ri, k:k + kk] = us
        v[k:k + kk, ci] = vs
        s_parts.append(ss)
        qi[k:k + kk] = qn
        k += kk
    s = torch.cat(s_parts).cpu().numpy()
    return u[:, :k].contiguous(), s, v[:k, :].contiguous(), qi[:k]


# ---------------------------------------------------------------------------
# batched forms: many independent problems of identical shape, stacked along a
# leading batch axis (one launch sequence for all of them)

def matmul_batched(A, B, conjA=False, conjB=False):
    """C[b] = op(A[b]) @ op(B[b]) for 3-D complex128 device tensors of any strides."""
    assert A.dim() == 3 and B.dim() == 3 and A.shape[0] == B.shape[0] and A.shape[2] == B.shape[1]
    assert A.dtype == torch.complex128 and B.dtype == torch.complex128
    nb, M, K = A.shape
    N = B.shape[2]
    out = torch.empty((nb, M, N), dtype=torch.complex128, device=A.device)
    lib = _lib.load()
    _check(lib.ptn_zgemm_batched(nb, M, N, K, ptr(A), A.stride(1), A.stride(2), A.stride(0),
                                 int(conjA), ptr(B), B.stride(1), B.stride(2), B.stride(0),
                                 int(conjB), ptr(out), out.stride(1), out.stride(2),
                                 out.stride(0), 0, stream()), "ptn_zgemm_batched")
    return out


def env_left_batched(a, w, l, precision="fp64"):
    """Batched contraction_operator_step_left with the bra equal to the ket:
    a (B, Dl, d, Dr), w (B, Ml, d, d, Mr), l (B, Dl, Ml, Dl) -> (B, Dr, Mr, Dr)."""
    a, w, l = as_dev(a), as_dev(w), as_dev(l)
    nb, Dl, d, Dr = a.shape
    Ml, Mr = w.shape[1], w.shape[4]
    prec = precision_code(precision)
    lib = _lib.load()
    out = torch.empty((nb, Dr, Mr, Dr), dtype=torch.complex128, device=a.device)
    ws = workspace(lib.ptn_env_workspace_batched(prec, nb, Dl, d, Dr, Ml, Mr))
    _check(lib.ptn_env_left_batched(prec, nb, Dl, d, Dr, Ml, Mr, ptr(a), ptr(a), ptr(w), ptr(l),
                                    ptr(out), ptr(ws), ws.numel(), stream()),
           "ptn_env_left_batched")
    return out


def env_right_batched(a, w, r, precision="fp64"):
    """Batched contraction_operator_step_right with the bra equal to the ket:
    a (B, Dl, d, Dr), w (B, Ml, d, d, Mr), r (B, Dr, Mr, Dr) -> (B, Dl, Ml, Dl)."""
    a, w, r = as_dev(a), as_dev(w), as_dev(r)
    nb, Dl, d, Dr = a.shape
    Ml, Mr = w.shape[1], w.shape[4]
    prec = precision_code(precision)
    lib = _lib.load()
    out = torch.empty((nb, Dl, Ml, Dl), dtype=torch.complex128, device=a.device)
    ws = workspace(lib.ptn_env_workspace_batched(prec, nb, Dl, d, Dr, Ml, Mr))
    _check(lib.ptn_env_right_batched(prec, nb, Dl, d, Dr, Ml, Mr, ptr(a), ptr(a), ptr(w), ptr(r),
                                     ptr(out), ptr(ws), ws.numel(), stream()),
           "ptn_env_right_batched")
    return out


def krylov_local_batched(kind, v, l, w, r, numiter, mode, dt=0.0, precision="fp64"):
    """Batched krylov_local: v (B, ...), l (B, Dl, Ml, Dl), w (B, Ml, d, d, Mr) or None,
    r (B, Dr, Mr, Dr). Returns (out, energies (B,) or None, info (B, 4))."""
    v, l, r = as_dev(v), as_dev(l), as_dev(r)
    if w is not None:
        w = as_dev(w)
    nb = v.shape[0]
    if kind == _lib.PTN_OP_HEFF:
        Dl, d, Dr = v.shape[1:]
    else:
        (Dl, Dr), d = v.shape[1:], 1
    op = _lib.LocalOp()
    op.kind = kind
    op.precision = precision_code(precision)
    op.Dl, op.d, op.Dr = Dl, d, Dr
    op.Ml, op.Mr = l.shape[2], r.shape[2]
    op.l = l.data_ptr()
    op.w = w.data_ptr() if w is not None else None
    op.r = r.data_ptr()
    lib = _lib.load()
    ws = workspace(lib.ptn_krylov_workspace_batched(ctypes.byref(op), int(numiter), nb))
    out = torch.empty_like(v)
    info = torch.zeros((nb, 4), dtype=torch.int32, device=v.device)
    if mode == "expm":
        dt = complex(dt)
        _check(lib.ptn_krylov_expm_batched(ctypes.byref(op), nb, ptr(v), dt.real, dt.imag,
                                           int(numiter), ptr(out), ptr(info), ptr(ws), ws.numel(),
                                           stream()), "ptn_krylov_expm_batched")
        return out, None, info
    energy = torch.zeros(nb, dtype=torch.float64, device=v.device)
    _check(lib.ptn_krylov_eigh_batched(ctypes.byref(op), nb, ptr(v), int(numiter), ptr(out),
                                       ptr(energy), ptr(info), ptr(ws), ws.numel(), stream()),
           "ptn_krylov_eigh_batched")
    return out, energy, info


def qr_batched(a):
    """Reduced Householder QR of every matrix of a (B, m, n) complex128 device tensor."""
    a = as_dev(a)
    nb, m, n = a.shape
    k = min(m, n)
    q = torch.empty((nb, m, k), dtype=torch.complex128, device=a.device)
    r = torch.empty((nb, k, n), dtype=torch.complex128, device=a.device)
    if m == 0 or n == 0 or nb == 0:
        return q, r
    lib = _lib.load()
    ws = workspace(lib.ptn_qr_workspace_batched(nb, m, n))
    _check(lib.ptn_qr_batched(nb, m, n, ptr(a), ptr(q), ptr(r), ptr(ws), ws.numel(), stream()),
           "ptn_qr_batched")
    return q, r
