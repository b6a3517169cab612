"""cdk_argmin_window_noalloc (device buffers in and out, working memory in a
fixed per-device array) against cdk_argmin_window on the same inputs: two
calls in a row, the second smaller, and a call past the fixed maximum refused
with nothing launched.  The test's own device buffers and stream come from
the HIP runtime libcdk.so links (the caller allocates; the entry point does
not)."""
import ctypes as C

import numpy as np
import pytest

import cdocker_amd as cdk
from cdocker_amd import _lib

pytestmark = pytest.mark.gpu

MAX_POSES = 16384
H2D, D2H = 1, 2


def say(*a):
    print("[gpu]", *a, flush=True)


class Hip:
    """hipMalloc / hipMemcpy / hipFree / streams from libamdhip64 (the caller's side)."""

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.rt.hipStreamSynchronize.argtypes = [C.c_void_p]
        self.rt.hipStreamDestroy.argtypes = [C.c_void_p]
        self.bufs = []

    def ok(self, rc, what):
        assert rc == 0, (what, rc)

    def alloc(self, nbytes):
        p = C.c_void_p()
        self.ok(self.rt.hipMalloc(C.byref(p), max(int(nbytes), 4)), "hipMalloc")
        self.bufs.append(p)
        return p

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        self.ok(self.rt.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D), "hipMemcpy H2D")
        return p

    def get(self, p, shape, dtype):
        out = np.empty(shape, dtype)
        self.ok(self.rt.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def close(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def grids_and_windows(rs, nq, n):
    E = rs.standard_normal((nq, n, n, n)).astype(np.float32)
    # planted equal minima: the first in x-major order must win
    E[0, 1, 2, 3] = E[0, 4, 0, 0] = np.float32(-50.0)
    win = rs.integers(1, n + 3, size=(nq, 3)).astype(np.int32)
    win[0] = (n, n, n)
    win[1] = (0, n, n)  # an empty window
    return E, win


def test_argmin_window_noalloc_matches_allocating_entry_point():
    ctx = cdk.Context(0)
    hip = Hip()
    stream = C.c_void_p()
    hip.ok(hip.rt.hipStreamCreate(C.byref(stream)), "hipStreamCreate")
    rs = np.random.default_rng(77)
    f = _lib.lib.cdk_argmin_window_noalloc
    try:
        for nq, n in ((40, 32), (7, 32)):
            say("no-alloc argmin: nq", nq, "n", n)
            E, win = grids_and_windows(rs, nq, n)
            e_ref, i_ref = ctx.argmin_window(E, win)
            dE, dW = hip.put(E), hip.put(win)
            de = hip.put(np.full(nq, 7.0, np.float32))
            di = hip.put(np.full((nq, 3), 9, np.int32))
            # on the caller's stream: enqueued, the caller waits
            rc = f(ctx.h, nq, n, dE, dW, de, di, stream)
            assert rc == 0, _lib.last_error()
            hip.ok(hip.rt.hipStreamSynchronize(stream), "hipStreamSynchronize")
            e = hip.get(de, (nq,), np.float32)
            i = hip.get(di, (nq, 3), np.int32)
            assert np.array_equal(e.view(np.uint32), e_ref.view(np.uint32)), (e, e_ref)
            assert np.array_equal(i, i_ref)
            assert not np.isfinite(e[1]) and list(i[1]) == [-1, -1, -1]
            assert list(i[0]) == [1, 2, 3]
            # on the context stream (NULL): the call returns with the outputs written
            de2 = hip.put(np.full(nq, 7.0, np.float32))
            di2 = hip.put(np.full((nq, 3), 9, np.int32))
            rc = f(ctx.h, nq, n, dE, dW, de2, di2, None)
            assert rc == 0, _lib.last_error()
            assert np.array_equal(hip.get(de2, (nq,), np.float32).view(np.uint32), e_ref.view(np.uint32))
            assert np.array_equal(hip.get(di2, (nq, 3), np.int32), i_ref)
        # past the fixed maximum: refused before anything is launched
        rc = f(ctx.h, MAX_POSES + 1, 32, dE, dW, de, di, None)
        assert rc == _lib.ERR_INVALID and "maximum" in _lib.last_error()
    finally:
        hip.rt.hipStreamDestroy(stream)
        hip.close()
        ctx.close()
