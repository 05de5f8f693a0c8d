// CSR sparse matrix kernels: SpMV and a deterministic transpose.
//
// SpMV assigns a power-of-two group of lanes to each row (chosen from the
// mean row length), strides the group over the row and reduces with
// wavefront shuffles, so every row is summed in a fixed order.  It is
// HBM-bound: (4 + 8) B per nonzero + 4 B per row + x gathers.
#include <algorithm>

#include "lt_internal.h"

namespace lt {

namespace {

template <int LANES>
__global__ void __launch_bounds__(256) k_spmv(long nrows, const int32_t *__restrict__ rp,
                                               const int32_t *__restrict__ ci, const double *__restrict__ va,
                                               const double *__restrict__ x, double *__restrict__ y, double alpha,
                                               double beta) {
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long row = gid / LANES;
    const int lane = (int)(gid % LANES);
    if (row >= nrows) return;  // whole lane groups leave together
    double s = 0.0;
    const int e = rp[row + 1];
    for (int k = rp[row] + lane; k < e; k += LANES) s += va[k] * x[ci[k]];
#pragma unroll
    for (int off = LANES / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, LANES);
    if (lane == 0) y[row] = (beta == 0.0) ? alpha * s : alpha * s + beta * y[row];
}

__global__ void k_col_count(long nnz, const int32_t *__restrict__ ci, int64_t *__restrict__ cnt) {
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nnz) atomicAdd((unsigned long long *)&cnt[ci[k]], 1ull);
}

__global__ void k_row_of(long nrows, const int32_t *__restrict__ rp, int32_t *__restrict__ rowof) {
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    for (int k = rp[r]; k < rp[r + 1]; ++k) rowof[k] = (int32_t)r;
}

__global__ void k_scatter_keys(long nnz, const int32_t *__restrict__ ci, const int64_t *__restrict__ off,
                               int64_t *__restrict__ cursor, uint64_t *__restrict__ keys) {
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    const int c = ci[k];
    const long pos = off[c] + (long)atomicAdd((unsigned long long *)&cursor[c], 1ull);
    keys[pos] = ((uint64_t)(uint32_t)c << 32) | (uint32_t)k;
}

__global__ void k_gather_t(long nnz, const uint64_t *__restrict__ keys, const int32_t *__restrict__ rowof,
                           const double *__restrict__ va, int32_t *__restrict__ tci, double *__restrict__ tva) {
    long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nnz) return;
    const uint32_t k = (uint32_t)keys[p];
    tci[p] = rowof[k];
    tva[p] = va[k];
}

__global__ void k_i64_to_i32(long n, const int64_t *__restrict__ a, int32_t *__restrict__ b) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (int32_t)a[i];
}

}  // namespace

int csr_pick_lanes(const DevCSR &A) {
    double mean = A.nrows ? (double)A.nnz / (double)A.nrows : 1.0;
    int L = 1;
    while (L < 64 && L < mean) L <<= 1;
    return L;
}

lt_status csr_spmv(const DevCSR &A, const double *x, double *y, double alpha, double beta_y, int lanes) {
    if (A.nrows == 0) return LT_OK;
    if (lanes <= 0) lanes = csr_pick_lanes(A);
    const long threads = A.nrows * lanes;
    const dim3 g(blocks_for(threads, 256)), b(256);
#define LT_SPMV_CASE(L)                                                                                        \
    case L:                                                                                                    \
        hipLaunchKernelGGL(k_spmv<L>, g, b, 0, stream(), A.nrows, A.rp.p, A.ci.p, A.va.p, x, y, alpha, beta_y); \
        break;
    switch (lanes) {
        LT_SPMV_CASE(1)
        LT_SPMV_CASE(2)
        LT_SPMV_CASE(4)
        LT_SPMV_CASE(8)
        LT_SPMV_CASE(16)
        LT_SPMV_CASE(32)
        LT_SPMV_CASE(64)
        default:
            set_error("spmv: lanes per row must be a power of two <= 64");
            return LT_ERR_ARG;
    }
#undef LT_SPMV_CASE
    LT_CHECK_LAUNCH();
    return LT_OK;
}

// Transpose: counting sort of the nonzeros by column, then a segmented sort
// inside each column on the original nonzero position, which orders every
// output row by ascending source row -- independent of atomic order.
lt_status csr_transpose(const DevCSR &A, DevCSR &B) {
    const long nnz = A.nnz, nc = A.ncols;
    B.nrows = nc;
    B.ncols = A.nrows;
    B.nnz = nnz;
    DevBuf<int64_t> cnt, off, cursor;
    DevBuf<int32_t> rowof;
    DevBuf<uint64_t> keys;
    LT_TRY(cnt.alloc(nc + 1));
    LT_TRY(cnt.zero());
    LT_TRY(off.alloc(nc + 1));
    LT_TRY(cursor.alloc(nc + 1));
    LT_TRY(cursor.zero());
    LT_TRY(rowof.alloc(std::max<long>(nnz, 1)));
    LT_TRY(keys.alloc(std::max<long>(nnz, 1)));
    LT_TRY(B.rp.alloc(nc + 1));
    LT_TRY(B.ci.alloc(std::max<long>(nnz, 1)));
    LT_TRY(B.va.alloc(std::max<long>(nnz, 1)));
    if (nnz > 0) {
        hipLaunchKernelGGL(k_col_count, dim3(blocks_for(nnz, 256)), dim3(256), 0, stream(), nnz, A.ci.p, cnt.p);
        LT_CHECK_LAUNCH();
    }
    LT_TRY(exclusive_scan_i64(cnt.p, off.p, nc));
    if (nnz > 0) {
        hipLaunchKernelGGL(k_row_of, dim3(blocks_for(A.nrows, 256)), dim3(256), 0, stream(), A.nrows, A.rp.p,
                           rowof.p);
        LT_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_scatter_keys, dim3(blocks_for(nnz, 256)), dim3(256), 0, stream(), nnz, A.ci.p, off.p,
                           cursor.p, keys.p);
        LT_CHECK_LAUNCH();
        std::vector<int64_t> seg(nc + 1);
        LT_TRY(off.download(seg.data(), nc + 1));
        LT_TRY(sync());
        LT_TRY(segmented_sort_u64(keys.p, off.p, seg));
        hipLaunchKernelGGL(k_gather_t, dim3(blocks_for(nnz, 256)), dim3(256), 0, stream(), nnz, keys.p, rowof.p,
                           A.va.p, B.ci.p, B.va.p);
        LT_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(k_i64_to_i32, dim3(blocks_for(nc + 1, 256)), dim3(256), 0, stream(), nc + 1, off.p, B.rp.p);
    LT_CHECK_LAUNCH();
    return sync();
}

}  // namespace lt
