"""GPU SpMV kernels against scipy: every lane-group width, the streaming
(LDS-staged) kernel, the sliced copy and the product over the list of
nonempty rows, on matrices with the row patterns of the lens matrices and
their transposes (empty rows, trailing empty rows, rows that run across the
streaming kernel's 2048-nonzero ranges, rows of 65..512 entries for the
one-wave path, longer ones for the chunked path).  Every kernel sums a row in a fixed order, so repeated products are
bit-identical."""
import numpy as np
import pytest
import scipy.sparse as sp

from helpers import rng

pytestmark = pytest.mark.gpu

# 0 default, 1..64 lanes per row, 128 streaming, 256 sliced copy, -1..-64 nonempty-row list
KERNELS = [0, 1, 2, 4, 8, 16, 32, 64, 128, 256, -1, -4, -16, -64]


def _matrix(g, nrows, ncols, lengths):
    rp = np.zeros(nrows + 1, dtype=np.int32)
    rp[1:] = np.cumsum(lengths)
    nnz = int(rp[-1])
    ci = g.integers(0, ncols, size=nnz).astype(np.int32)
    va = g.normal(size=nnz)
    return rp, ci, va


def _cases():
    g = rng(101)
    out = []
    # short rows like L (0..4 entries), trailing empty rows
    n = 50000
    lens = g.choice([0, 1, 2, 3, 4], size=n, p=[0.2, 0.3, 0.3, 0.1, 0.1])
    lens[-300:] = 0
    out.append(('short', n, 4000, lens))
    # L's full mix: empty, 1..4 (the 4-wide table), 5..64 (lane groups), longer (chunks)
    n = 40000
    lens = g.choice([0, 1, 2, 3, 4, 5, 8, 17, 64, 65, 300], size=n,
                    p=[0.5, 0.01, 0.15, 0.15, 0.1, 0.04, 0.03, 0.01, 0.005, 0.003, 0.002])
    out.append(('lens mix', n, 16384, lens))
    # rows like L^T / M (10..60 entries) with a few long rows (chunked path)
    n = 6000
    lens = g.integers(10, 60, size=n)
    lens[[5, 777, 4000]] = [3000, 9000, 2049]
    lens[4001] = 2048
    lens[[10, 11, 3000, 5999]] = [65, 512, 513, 200]
    out.append(('medium+long', n, 20000, lens))
    # L^T's tail: many rows of 65..512 entries (one wave each) among 5..64
    n = 20000
    lens = g.integers(5, 64, size=n)
    lens[g.integers(0, n, size=1500)] = g.integers(65, 513, size=1500)
    out.append(('medium rows', n, 30000, lens))
    # rows of ~1000 entries: most cross a streaming range boundary
    n = 300
    lens = g.integers(900, 1100, size=n)
    out.append(('wide', n, 5000, lens))
    # a single long row; an all-empty matrix
    out.append(('one long row', 1, 100, np.array([5000])))
    out.append(('empty', 40, 10, np.zeros(40, dtype=np.int64)))
    return [(name, nr, nc, _matrix(g, nr, nc, l)) for name, nr, nc, l in out]


@pytest.mark.parametrize('case', _cases(), ids=lambda c: c[0])
def test_spmv_kernels_match_scipy(hip, case):
    B = hip
    name, nrows, ncols, (rp, ci, va) = case
    x = rng(102).normal(size=ncols)
    ref = sp.csr_matrix((va, ci, rp), shape=(nrows, ncols)) @ x
    scale = np.max(np.abs(ref)) if ref.size and np.max(np.abs(ref)) > 0 else 1.0
    got = {}
    for k in KERNELS:
        y = np.full(nrows, np.nan)
        B.check(B.lib().lt_csr_spmv_host_kernel(nrows, ncols, B.ptr(rp), B.ptr(ci), B.ptr(va), B.ptr(x),
                                                 B.ptr(y), k))
        assert np.all(np.isfinite(y)), (name, k)  # every row written
        assert np.max(np.abs(y - ref)) <= 1e-13 * scale, (name, k)
        got[k] = y
    for k in (1, 128, 256, -4):
        y = np.full(nrows, np.nan)
        B.check(B.lib().lt_csr_spmv_host_kernel(nrows, ncols, B.ptr(rp), B.ptr(ci), B.ptr(va), B.ptr(x),
                                                 B.ptr(y), k))
        assert np.array_equal(y, got[k]), (name, k)
