"""
The allocation-free local solve (ptn_krylov_expm_noalloc, called through the
C ABI) against ptn_krylov_expm with a caller workspace: two calls in a row on
the same GPU, the second on a smaller problem (the scratch array keeps the
first call's contents), each bit-identical to the workspace entry point, in
FP64 and FP32 contraction modes; a problem whose workspace exceeds the
scratch array is rejected.
"""

import ctypes

import numpy as np
import pytest
import torch

import oracle as orc
from pytenet_amd import backend, _lib

pytestmark = pytest.mark.gpu


def _problem(seed, Dl, d, Dr, M):
    rng = np.random.default_rng(seed)
    w = orc.crandn((M, d, d, M), rng)
    w = 0.5 * (w + w.transpose(0, 2, 1, 3).conj())

    def env(D):
        x = orc.crandn((D, M, D), rng) / D
        return 0.5 * (x + x.transpose(2, 1, 0).conj())
    return [backend.as_dev(x) for x in (orc.crandn((Dl, d, Dr), rng), w, env(Dl), env(Dr))]


def _op(prec, a, w, l, r):
    op = _lib.LocalOp()
    op.kind = _lib.PTN_OP_HEFF
    op.precision = backend.precision_code(prec)
    op.Dl, op.d, op.Dr = a.shape
    op.Ml, op.Mr = l.shape[1], r.shape[1]
    op.l, op.w, op.r = l.data_ptr(), w.data_ptr(), r.data_ptr()
    return op


def _expm_ws(lib, op, a, dt, numiter):
    ws = backend.workspace(lib.ptn_krylov_workspace(ctypes.byref(op), numiter))
    out = torch.empty_like(a)
    info = torch.zeros(4, dtype=torch.int32, device=a.device)
    assert lib.ptn_krylov_expm(ctypes.byref(op), backend.ptr(a), dt.real, dt.imag, numiter,
                               backend.ptr(out), backend.ptr(info), None, backend.ptr(ws),
                               ws.numel(), backend.stream()) == _lib.PTN_OK
    return out, info


def _expm_noalloc(lib, op, a, dt, numiter):
    out = torch.empty_like(a)
    info = torch.zeros(4, dtype=torch.int32, device=a.device)
    st = lib.ptn_krylov_expm_noalloc(ctypes.byref(op), backend.ptr(a), dt.real, dt.imag, numiter,
                                     backend.ptr(out), backend.ptr(info), None, backend.stream())
    return st, out, info


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_noalloc_expm_equals_workspace_expm(prec):
    lib = _lib.load()
    dt = complex(-0.05j)
    for seed, shape in ((1, (48, 4, 40, 5)), (2, (12, 2, 9, 3))):   # second call smaller
        a, w, l, r = _problem(seed, *shape)
        op = _op(prec, a, w, l, r)
        st, got, ginfo = _expm_noalloc(lib, op, a, dt, 8)
        assert st == _lib.PTN_OK
        ref, rinfo = _expm_ws(lib, op, a, dt, 8)
        torch.cuda.synchronize()
        assert torch.equal(got, ref) and torch.equal(ginfo, rinfo)
        assert int(ginfo[0]) == 8


def test_noalloc_rejects_problems_beyond_the_scratch_array():
    lib = _lib.load()
    a, w, l, r = _problem(3, 8, 4, 8, 5)
    op = _op("fp64", a, w, l, r)
    op.Dl = op.Dr = 2048          # workspace query only: nothing is launched
    assert lib.ptn_krylov_workspace(ctypes.byref(op), 8) > _lib.PTN_NOALLOC_SCRATCH_BYTES
    st, _, _ = _expm_noalloc(lib, op, a, complex(0.1j), 8)
    assert st == _lib.PTN_ERR_WORKSPACE
