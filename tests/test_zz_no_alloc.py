"""The allocation-free CSR SpMV (lt_csr_spmv_device_noalloc) against the
allocating host entry point it varies (lt_csr_spmv_host): bit-identical y on
matrices with long (chunked) rows, on two calls in a row (the second with a
smaller matrix, so the fixed scratch keeps the first call's contents), and
against scipy; sizes beyond its fixed scratch are refused before any device
access."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

from helpers import rng

pytestmark = pytest.mark.gpu


def _matrix(nrows, ncols, long_rows, seed):
    g = rng(seed)
    lens = g.integers(0, 40, size=nrows)
    for r, n in long_rows.items():
        lens[r] = n
    rp = np.zeros(nrows + 1, dtype=np.int32)
    rp[1:] = np.cumsum(lens)
    ci = np.concatenate([np.sort(g.choice(ncols, size=n, replace=n > ncols)) for n in lens]).astype(np.int32)
    va = g.normal(size=rp[-1])
    return rp, ci, va, g.normal(size=ncols)


class _Dev:
    def __init__(self, B, a):
        self.B = B
        self.p = C.c_void_p()
        self.n = a.nbytes
        B.check(B.lib().lt_device_alloc(max(self.n, 8), C.byref(self.p)))
        B.check(B.lib().lt_device_upload(self.p, B.ptr(a), self.n))

    def get(self, like):
        out = np.empty_like(like)
        self.B.check(self.B.lib().lt_device_download(self.p, self.B.ptr(out), out.nbytes))
        return out

    def free(self):
        self.B.check(self.B.lib().lt_device_free(self.p))


def _noalloc(B, rp, ci, va, x):
    L = B.lib()
    L.lt_csr_spmv_device_noalloc.argtypes = [C.c_long, C.c_long, C.c_long] + [C.c_void_p] * 5
    L.lt_csr_spmv_device_noalloc.restype = C.c_int
    bufs = [_Dev(B, a) for a in (rp, ci, va, x)]
    y0 = np.full(rp.size - 1, np.nan)
    dy = _Dev(B, y0)
    B.check(L.lt_csr_spmv_device_noalloc(rp.size - 1, x.size, int(rp[-1]), *[b.p for b in bufs], dy.p))
    y = dy.get(y0)
    for b in bufs + [dy]:
        b.free()
    return y


def _host(B, rp, ci, va, x):
    y = np.zeros(rp.size - 1)
    B.check(B.lib().lt_csr_spmv_host(rp.size - 1, x.size, B.ptr(rp), B.ptr(ci), B.ptr(va), B.ptr(x), B.ptr(y)))
    return y


def test_noalloc_spmv_matches_host_entry(hip):
    B = hip
    cases = [(20000, 50000, {7: 9000, 1234: 30000, 19999: 2049, 500: 2048}, 61),
             (3000, 4000, {5: 5000, 2999: 4097}, 62)]
    for nrows, ncols, longs, seed in cases:
        rp, ci, va, x = _matrix(nrows, ncols, longs, seed)
        print(f'case {nrows} x {ncols}, nnz {rp[-1]}, long rows {sorted(longs.values())}', flush=True)
        y = _noalloc(B, rp, ci, va, x)
        ref = _host(B, rp, ci, va, x)
        assert np.array_equal(y, ref)
        A = sp.csr_matrix((va, ci, rp), shape=(nrows, ncols))
        yy = A @ x
        assert np.max(np.abs(y - yy)) <= 1e-13 * np.max(np.abs(yy))


def test_noalloc_spmv_refuses_oversize(hip):
    B = hip
    L = B.lib()
    L.lt_csr_spmv_device_noalloc.argtypes = [C.c_long, C.c_long, C.c_long] + [C.c_void_p] * 5
    L.lt_csr_spmv_device_noalloc.restype = C.c_int
    rp = np.zeros(2, dtype=np.int32)
    d = _Dev(B, rp)
    st = L.lt_csr_spmv_device_noalloc(1, 1, 1 << 30, d.p, d.p, d.p, d.p, d.p)
    d.free()
    assert st == B.LT_ERR_CAPACITY
    assert b'scratch' in L.lt_last_error_message()
