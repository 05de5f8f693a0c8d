// CSR sparse matrix kernels: SpMV and a deterministic transpose.
//
// SpMV either assigns a power-of-two group of lanes to each row, strides the
// group over the row and reduces with wavefront shuffles, or streams fixed
// ranges of nonzeros through LDS (k_spmv_stream); the form is timed per
// matrix shape (csr_autotune).  Every row is summed in a fixed order.  Work:
// (4 + 8) B per nonzero + 4 B per row + x gathers.  The transpose is one radix
// sort of (column, position) keys.
#include <algorithm>

#include <cstring>

#include <rocprim/rocprim.hpp>

#include "lt_internal.h"

#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>

namespace lt {

namespace {

constexpr int kChunk = 2048;
// rows of long_row + 1 .. kMedRow entries (when long_row < kMedRow): one
// wave each (k_spmv_med); longer ones in kChunk chunks
constexpr int kMedRow = 512;

// Rows of a block-diagonal batch belong to samples: sample = row / rps, or
// rowseg[row] when given.  With act set, rows of inactive samples are
// skipped (their outputs are left as they are).
struct RowFilter {
    const int *act;
    long rps;
    const int32_t *rowseg;
    __device__ bool skip(long row) const {
        if (!act) return false;
        // rows and rows per sample fit 32 bits (int32 row pointers)
        const long s = rowseg ? (long)rowseg[row] : (long)((uint32_t)row / (uint32_t)rps);
        return act[s] == 0;
    }
};

// T: value and vector type (double, or float for the precision form).
// One row's product by a group of LANES lanes; lane 0 writes.  Rows above
// skip_long entries (> 0) are left to the chunked path.
template <int LANES, class T>
__device__ inline void row_spmv(long row, int lane, const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                const T *__restrict__ va, const T *__restrict__ x, T *__restrict__ y, T alpha,
                                T beta, int skip_long) {
    T s = (T)0;
    const int b = rp[row], e = rp[row + 1];
    if (skip_long > 0 && e - b > skip_long) return;  // summed by k_spmv_chunks / k_spmv_long
    for (int k = b + lane; k < e; k += LANES) s += va[k] * x[ci[k]];
#pragma unroll
    for (int off = LANES / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, LANES);
    if (lane == 0) y[row] = (beta == (T)0) ? alpha * s : alpha * s + beta * y[row];
}

// The same for the nonempty-row list (alpha 1, beta 0): a row of at most 64
// entries has every lane's loads (at most 64 / LANES) issued together, then
// the gathers, then the sum in the loop's order (the same bits); longer rows
// not left to the other paths take the loop.  `store` false computes
// without writing (a filtered row: its loads go out before the filter is
// known).
template <int LANES, class T>
__device__ inline void row_spmv_short(long row, int lane, const int32_t *__restrict__ rp,
                                      const int32_t *__restrict__ ci, const T *__restrict__ va,
                                      const T *__restrict__ x, T *__restrict__ y, int skip_long, bool store) {
    constexpr int PER = 64 / LANES;
    const int b = rp[row], e = rp[row + 1];
    if (skip_long > 0 && e - b > skip_long) return;  // summed by the medium-row / chunked paths
    if (e - b > 64) {  // whole lane groups
        if (store) row_spmv<LANES, T>(row, lane, rp, ci, va, x, y, (T)1, (T)0, skip_long);
        return;
    }
    int c[PER];
    T v[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int k = b + lane + LANES * u;
        c[u] = k < e ? ci[k] : -1;
        v[u] = k < e ? va[k] : (T)0;
    }
    T s = (T)0;
#pragma unroll
    for (int u = 0; u < PER; ++u)
        if (c[u] >= 0) s += v[u] * x[c[u]];
#pragma unroll
    for (int off = LANES / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, LANES);
    if (lane == 0 && store) y[row] = s;
}

template <int LANES, class T>
__global__ void __launch_bounds__(256) k_spmv(long row0, long nrows, const int32_t *__restrict__ rp,
                                               const int32_t *__restrict__ ci, const T *__restrict__ va,
                                               const T *__restrict__ x, T *__restrict__ y, T alpha, T beta,
                                               int skip_long, RowFilter F) {
    const long gid = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const long row = row0 + gid / LANES;
    if (row >= nrows || F.skip(row)) return;  // whole lane groups leave together
    row_spmv<LANES, T>(row, (int)(gid % LANES), rp, ci, va, x, y, alpha, beta, skip_long);
}

// Sample grid: workgroup row y takes sample alist[y] (skipped when its act
// flag is 0) and x runs over that sample's rows [srow[s], srow[s+1]), so a
// product over a few samples launches only their rows.
template <int LANES, class T>
__global__ void __launch_bounds__(256) k_spmv_sg(const int *__restrict__ alist, const int *__restrict__ act,
                                                  const int64_t *__restrict__ srow, const int32_t *__restrict__ rp,
                                                  const int32_t *__restrict__ ci, const T *__restrict__ va,
                                                  const T *__restrict__ x, T *__restrict__ y, T alpha, T beta,
                                                  int skip_long) {
    const int s = alist[blockIdx.y];
    if (act && act[s] == 0) return;
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long row = srow[s] + gid / LANES;
    if (row >= srow[s + 1]) return;
    row_spmv<LANES, T>(row, (int)(gid % LANES), rp, ci, va, x, y, alpha, beta, skip_long);
}

// The same product over a list of rows (the nonempty rows of a matrix whose
// other rows the caller keeps at zero): lane group g takes row nzr[g].
template <int LANES, class T>
__global__ void __launch_bounds__(256) k_spmv_rows(long i0, long n, const int32_t *__restrict__ nzr,
                                                    const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                    const T *__restrict__ va, const T *__restrict__ x,
                                                    T *__restrict__ y, int skip_long, RowFilter F) {
    const long gid = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const long i = i0 + gid / LANES;
    if (i >= n) return;
    const long row = nzr[i];
    row_spmv_short<LANES, T>(row, (int)(gid % LANES), rp, ci, va, x, y, skip_long, !F.skip(row));
}

// list entries of sample alist[y]: [snzr[s], snzr[s+1])
template <int LANES, class T>
__global__ void __launch_bounds__(256) k_spmv_rows_sg(const int *__restrict__ alist, const int *__restrict__ act,
                                                       const int64_t *__restrict__ snzr,
                                                       const int32_t *__restrict__ nzr, const int32_t *__restrict__ rp,
                                                       const int32_t *__restrict__ ci, const T *__restrict__ va,
                                                       const T *__restrict__ x, T *__restrict__ y, int skip_long) {
    const int s = alist[blockIdx.y];
    if (act && act[s] == 0) return;
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long i = snzr[s] + gid / LANES;
    if (i >= snzr[s + 1]) return;
    row_spmv_short<LANES, T>(nzr[i], (int)(gid % LANES), rp, ci, va, x, y, skip_long, true);
}

// Rows of 1..4 entries, one thread per row of the padded table: the row
// index, the four columns and the four values are independent aligned
// loads, then four gathers.  Terms are added in the row's order.
template <class T>
__device__ inline void ell4_row(long i, long row, const int4 *__restrict__ eci, const T *__restrict__ eva,
                                const T *__restrict__ x, T *__restrict__ y) {
    const int4 c = eci[i];
    const T v0 = eva[4 * i], v1 = eva[4 * i + 1], v2 = eva[4 * i + 2], v3 = eva[4 * i + 3];
    T s = v0 * x[c.x];
    s += v1 * x[c.y];
    s += v2 * x[c.z];
    s += v3 * x[c.w];
    y[row] = s;
}

template <class T>
__global__ void __launch_bounds__(256) k_spmv_ell4(long n, const int32_t *__restrict__ erow,
                                                    const int4 *__restrict__ eci, const T *__restrict__ eva,
                                                    const T *__restrict__ x, T *__restrict__ y, RowFilter F) {
    const long i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // the entry's loads go out with the row's (all in bounds whatever the
    // filter says) and the gathers with the filter's: two memory round trips
    // before the store instead of four
    const long row = erow[i];
    const int4 c = eci[i];
    const T v0 = eva[4 * i], v1 = eva[4 * i + 1], v2 = eva[4 * i + 2], v3 = eva[4 * i + 3];
    const bool sk = F.skip(row);
    T s = v0 * x[c.x];
    s += v1 * x[c.y];
    s += v2 * x[c.z];
    s += v3 * x[c.w];
    if (!sk) y[row] = s;
}

// table entries of sample alist[y]: [sell[s], sell[s+1])
template <class T>
__global__ void __launch_bounds__(256) k_spmv_ell4_sg(const int *__restrict__ alist, const int *__restrict__ act,
                                                       const int64_t *__restrict__ sell,
                                                       const int32_t *__restrict__ erow, const int4 *__restrict__ eci,
                                                       const T *__restrict__ eva, const T *__restrict__ x,
                                                       T *__restrict__ y) {
    const int s = alist[blockIdx.y];
    if (act && act[s] == 0) return;
    const long i = sell[s] + (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= sell[s + 1]) return;
    ell4_row<T>(i, erow[i], eci, eva, x, y);
}

// One thread per row over the sliced copy: entry j of the row's slice is
// read by the 64 rows of the slice together (coalesced); the row's terms
// are added in order.
template <class T>
__device__ inline void slice_row(long row, const int32_t *__restrict__ rp, const int64_t *__restrict__ sl_ptr,
                                 const int32_t *__restrict__ sci, const T *__restrict__ sva, const T *__restrict__ x,
                                 T *__restrict__ y, T alpha, T beta, int skip_long) {
    const int len = rp[row + 1] - rp[row];
    if (skip_long > 0 && len > skip_long) return;  // summed by k_spmv_chunks / k_spmv_long
    const long base = sl_ptr[row >> 6] + (row & 63);
    T s = (T)0;
    for (int j = 0; j < len; ++j) s += sva[base + 64L * j] * x[sci[base + 64L * j]];
    y[row] = (beta == (T)0) ? alpha * s : alpha * s + beta * y[row];
}

template <class T>
__global__ void __launch_bounds__(256) k_spmv_slice(long nrows, const int32_t *__restrict__ rp,
                                                     const int64_t *__restrict__ sl_ptr,
                                                     const int32_t *__restrict__ sci, const T *__restrict__ sva,
                                                     const T *__restrict__ x, T *__restrict__ y, T alpha, T beta,
                                                     int skip_long, RowFilter F) {
    const long row = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (row >= nrows || F.skip(row)) return;
    slice_row<T>(row, rp, sl_ptr, sci, sva, x, y, alpha, beta, skip_long);
}

template <class T>
__global__ void __launch_bounds__(256) k_spmv_slice_sg(const int *__restrict__ alist, const int *__restrict__ act,
                                                        const int64_t *__restrict__ srow,
                                                        const int32_t *__restrict__ rp,
                                                        const int64_t *__restrict__ sl_ptr,
                                                        const int32_t *__restrict__ sci, const T *__restrict__ sva,
                                                        const T *__restrict__ x, T *__restrict__ y, T alpha, T beta,
                                                        int skip_long) {
    const int s = alist[blockIdx.y];
    if (act && act[s] == 0) return;
    const long row = srow[s] + (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= srow[s + 1]) return;
    slice_row<T>(row, rp, sl_ptr, sci, sva, x, y, alpha, beta, skip_long);
}

// slice widths: 64 x the longest row of each slice (rows above long_row
// excluded when the chunked path sums them)
__global__ void k_slice_width(long nslice, long nrows, const int32_t *__restrict__ rp, int skip_long,
                              int64_t *__restrict__ width) {
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nslice) return;
    int w = 0;
    for (long r = q * 64; r < min(nrows, q * 64 + 64); ++r) {
        const int len = rp[r + 1] - rp[r];
        if (!(skip_long > 0 && len > skip_long)) w = max(w, len);
    }
    width[q] = 64L * w;
}

__global__ void k_slice_fill(long nrows, const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                             const double *__restrict__ va, const int64_t *__restrict__ sl_ptr, int skip_long,
                             int32_t *__restrict__ sci, double *__restrict__ sva) {
    const long row = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= nrows) return;
    const int b = rp[row], len = rp[row + 1] - b;
    if (skip_long > 0 && len > skip_long) return;
    const long base = sl_ptr[row >> 6] + (row & 63);
    for (int j = 0; j < len; ++j) {
        sci[base + 64L * j] = ci[b + j];
        sva[base + 64L * j] = va[b + j];
    }
}

// first position of a sorted list at or above each key
__global__ void k_lower_bound(int nkey, const int64_t *__restrict__ key, const int32_t *__restrict__ list, long n,
                              int64_t *__restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nkey) return;
    long lo = 0, hi = n;
    while (lo < hi) {
        const long mid = (lo + hi) / 2;
        if (list[mid] < key[t])
            lo = mid + 1;
        else
            hi = mid;
    }
    out[t] = lo;
}

__global__ void k_row_class(long n, const int32_t *__restrict__ rp, int64_t *__restrict__ fe,
                            int64_t *__restrict__ fo) {
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int len = rp[r + 1] - rp[r];
    fe[r] = (len >= 1 && len <= 4) ? 1 : 0;
    fo[r] = len > 4 ? 1 : 0;
}

__global__ void k_row_split(long n, const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                            const double *__restrict__ va, const int64_t *__restrict__ pe,
                            const int64_t *__restrict__ po, int32_t *__restrict__ erow, int32_t *__restrict__ eci,
                            double *__restrict__ eva, int32_t *__restrict__ nzr) {
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int b = rp[r], e = rp[r + 1];
    if (e - b > 4) {
        nzr[po[r]] = (int32_t)r;
    } else if (e > b) {
        const long i = pe[r];
        erow[i] = (int32_t)r;
        for (int j = 0; j < 4; ++j) {
            eci[4 * i + j] = b + j < e ? ci[b + j] : ci[b];
            eva[4 * i + j] = b + j < e ? va[b + j] : 0.0;
        }
    }
}

// Streaming SpMV (CSR-stream): workgroup b owns the nonzeros
// [b*kStreamNnz, (b+1)*kStreamNnz) and the rows that start among them
// (rblk[b] .. rblk[b+1]).  It first forms every product va[k] * x[ci[k]] of
// its range and of the kStreamTail nonzeros after it with coalesced loads
// into LDS, then one thread per row adds the row's products in order (the
// row running past the range finds its tail staged when it has at most
// kStreamTail + 1 entries, else reads the rest from memory).  Long rows are
// left to the medium-row and chunked paths.  With a row filter, workgroups
// whose rows all belong to inactive samples return at once.
constexpr int kStreamNnz = 2048;
constexpr int kStreamTail = 256;

__device__ inline long row_sample(const RowFilter &F, long row) {
    return F.rowseg ? (long)F.rowseg[row] : row / F.rps;
}

// TV: the stored value type (T, or FP32 values against FP64 vectors for a
// preconditioner's matrix, csr_spmv_fv); products and sums in T
template <class T, class TV = T>
__global__ void __launch_bounds__(256) k_spmv_stream(long nnz, const int32_t *__restrict__ rblk,
                                                      const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                      const TV *__restrict__ va, const T *__restrict__ x,
                                                      T *__restrict__ y, T alpha, T beta, int skip_long, RowFilter F) {
    __shared__ T prod[kStreamNnz + kStreamTail];
    __shared__ int32_t rps[kStreamNnz + 1];
    const long b = xcd_block(blockIdx.x, gridDim.x);
    const long r0 = rblk[b], r1 = rblk[b + 1];
    if (r0 >= r1) return;  // inside a long row: no row starts here
    if (F.act) {
        bool any = false;
        for (long sm = row_sample(F, r0), s1 = row_sample(F, r1 - 1); sm <= s1 && !any; ++sm) any = F.act[sm] != 0;
        if (!any) return;
    }
    const long k0 = b * kStreamNnz, k1 = min(k0 + kStreamNnz + kStreamTail, nnz);  // staged range
    constexpr int PER = kStreamNnz / 256;
    static_assert(kStreamTail <= 256, "the tail is one partial load per thread");
    // the row pointers of the block's rows go to LDS alongside the products
    const long nr = r1 - r0;
    const bool rp_lds = nr < kStreamNnz;
    if (rp_lds)
        for (long r = threadIdx.x; r <= nr; r += 256) rps[r] = rp[r0 + r];
    if (k1 - k0 == kStreamNnz + kStreamTail) {  // full range: all loads of a thread in flight together
        int c[PER + 1];
        T v[PER + 1], xv[PER + 1];  // values widened to T as loaded
        const bool tl = threadIdx.x < kStreamTail;
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            c[u] = ci[k0 + threadIdx.x + 256 * u];
            v[u] = (T)va[k0 + threadIdx.x + 256 * u];
        }
        if (tl) {
            c[PER] = ci[k0 + kStreamNnz + threadIdx.x];
            v[PER] = (T)va[k0 + kStreamNnz + threadIdx.x];
        }
#pragma unroll
        for (int u = 0; u < PER; ++u) xv[u] = x[c[u]];
        if (tl) xv[PER] = x[c[PER]];
#pragma unroll
        for (int u = 0; u < PER; ++u) prod[threadIdx.x + 256 * u] = v[u] * xv[u];
        if (tl) prod[kStreamNnz + threadIdx.x] = v[PER] * xv[PER];
    } else {
        for (long k = k0 + threadIdx.x; k < k1; k += 256) prod[k - k0] = (T)va[k] * x[ci[k]];
    }
    __syncthreads();
    for (long r = r0 + threadIdx.x; r < r1; r += 256) {
        if (F.skip(r)) continue;
        const int rb = rp_lds ? rps[r - r0] : rp[r], re = rp_lds ? rps[r - r0 + 1] : rp[r + 1];
        if (skip_long > 0 && re - rb > skip_long) continue;
        const int ein = (int)min((long)re, k1);
        T s = (T)0;
        for (int k = rb; k < ein; ++k) s += prod[k - k0];
        for (int k = ein; k < re; ++k) s += (T)va[k] * x[ci[k]];
        y[r] = (beta == (T)0) ? alpha * s : alpha * s + beta * y[r];
    }
}

// first row starting at or after nonzero b * kStreamNnz, per workgroup
__global__ void k_stream_blocks(long nblk, long nrows, const int32_t *__restrict__ rp, int32_t *__restrict__ rblk) {
    const long b = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (b > nblk) return;
    if (b == nblk) {
        rblk[b] = (int32_t)nrows;
        return;
    }
    const long k0 = b * kStreamNnz;
    long lo = 0, hi = nrows;
    while (lo < hi) {
        const long mid = (lo + hi) >> 1;
        if (rp[mid] < k0)
            lo = mid + 1;
        else
            hi = mid;
    }
    rblk[b] = (int32_t)lo;
}

// sum of one chunk [b, min(b + kChunk, row end)) of a long row by a whole
// 256-thread workgroup, in a fixed order; thread 0 writes *out
template <class T>
__device__ inline void chunk_sum(int b, int row, const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                 const T *__restrict__ va, const T *__restrict__ x, T *__restrict__ out) {
    __shared__ T sh[4];
    const int e = min(b + kChunk, rp[row + 1]);
    // every load of a thread issued together, then the products added in
    // increasing position
    constexpr int PER = kChunk / 256;
    int c[PER];
    T v[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int k = b + threadIdx.x + 256 * i;
        c[i] = k < e ? ci[k] : -1;
        v[i] = k < e ? va[k] : (T)0;
    }
    T s = (T)0;
#pragma unroll
    for (int i = 0; i < PER; ++i)
        if (c[i] >= 0) s += v[i] * x[c[i]];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) *out = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// partial sums of fixed-size chunks of the long rows, one workgroup each
template <class T>
__global__ void __launch_bounds__(256) k_spmv_chunks(long nchunk, const int32_t *__restrict__ chunk_beg,
                                                      const int32_t *__restrict__ chunk_end_row,
                                                      const int32_t *__restrict__ rp,
                                                      const int32_t *__restrict__ ci, const T *__restrict__ va,
                                                      const T *__restrict__ x, T *__restrict__ out, RowFilter F) {
    const long c = blockIdx.x;
    if (F.skip(chunk_end_row[c])) return;
    chunk_sum<T>(chunk_beg[c], chunk_end_row[c], rp, ci, va, x, out + c);
}

// medium rows (long_row < length <= kMedRow): one wave per row, four rows
// per workgroup; each lane's entries b + lane + 64 i loaded together, summed
// in increasing i, then a butterfly sum across the wave read at lane 0 (a
// fixed order)
template <class T>
__global__ void __launch_bounds__(256) k_spmv_med(long nmed, const int32_t *__restrict__ rows,
                                                   const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                   const T *__restrict__ va, const T *__restrict__ x,
                                                   T *__restrict__ y, T alpha, T beta, RowFilter F) {
    const long w = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int l = threadIdx.x & 63;
    if (w >= nmed) return;  // whole waves
    const int row = rows[w];
    if (F.skip(row)) return;
    const int b = rp[row], e = rp[row + 1];
    constexpr int PER = kMedRow / 64;
    int c[PER];
    T v[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int k = b + l + 64 * i;
        c[i] = k < e ? ci[k] : -1;
        v[i] = k < e ? va[k] : (T)0;
    }
    T s = (T)0;
#pragma unroll
    for (int i = 0; i < PER; ++i)
        if (c[i] >= 0) s += v[i] * x[c[i]];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (l == 0) y[row] = (beta == (T)0) ? alpha * s : alpha * s + beta * y[row];
}

// long rows: one wave per row (four per workgroup): lane i adds the row's chunk
// sums i, i + 64, ... in order, then a butterfly sum read at lane 0 (a fixed
// order; the allocation-free path sums the same way)
template <class T>
__device__ inline T chunks_total(const T *__restrict__ part, int c0, int c1, int lane) {
    T s = (T)0;
    for (int c = c0 + lane; c < c1; c += 64) s += part[c];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    return s;
}

template <class T>
__global__ void __launch_bounds__(256) k_spmv_long(long nlong, const int32_t *__restrict__ rows,
                                                    const int32_t *__restrict__ chunk0, const T *__restrict__ part,
                                                    T *__restrict__ y, T alpha, T beta, RowFilter F) {
    const long l = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (l >= nlong) return;  // whole waves
    const int row = rows[l];
    if (F.skip(row)) return;
    const T s = chunks_total(part, chunk0[l], chunk0[l + 1], lane);
    if (lane == 0) y[row] = (beta == (T)0) ? alpha * s : alpha * s + beta * y[row];
}

__global__ void k_d2f_sp(long n, const double *__restrict__ a, float *__restrict__ b) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

__global__ void k_long_rows_flag(long n, const int32_t *__restrict__ rp, int long_row, int64_t *__restrict__ flag,
                                 int64_t *__restrict__ nch, int64_t *__restrict__ mflag) {
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int len = rp[r + 1] - rp[r];
    const bool md = len > long_row && len <= kMedRow;
    const bool lg = len > long_row && !md;
    flag[r] = lg ? 1 : 0;
    nch[r] = lg ? (len + kChunk - 1) / kChunk : 0;
    mflag[r] = md ? 1 : 0;
}

__global__ void k_flag_list(long n, const int64_t *__restrict__ pos, int32_t *__restrict__ out) {
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n && pos[r + 1] != pos[r]) out[pos[r]] = (int32_t)r;
}

__global__ void k_long_rows_fill(long n, const int32_t *__restrict__ rp, const int64_t *__restrict__ pos,
                                 const int64_t *__restrict__ choff, int32_t *__restrict__ rows,
                                 int32_t *__restrict__ chunk0, int32_t *__restrict__ chunk_beg,
                                 int32_t *__restrict__ chunk_row) {
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n || pos[r + 1] == pos[r]) return;
    const long l = pos[r];
    rows[l] = (int32_t)r;
    chunk0[l] = (int32_t)choff[r];
    const int b = rp[r], e = rp[r + 1];
    long c = choff[r];
    for (int k = b; k < e; k += kChunk, ++c) {
        chunk_beg[c] = k;
        chunk_row[c] = (int32_t)r;
    }
}


__global__ void k_row_of(long nrows, const int32_t *__restrict__ rp, int32_t *__restrict__ rowof) {
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    for (int k = rp[r]; k < rp[r + 1]; ++k) rowof[k] = (int32_t)r;
}


__global__ void k_nonzero_flags(long nnz, const double *__restrict__ va, int64_t *__restrict__ flag) {
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nnz) flag[k] = va[k] != 0.0 ? 1 : 0;
}

// kept entries of row r = pos[rp[r+1]] - pos[rp[r]] (pos: exclusive scan of the flags)
__global__ void k_row_kept(long n, const int32_t *__restrict__ rp, const int64_t *__restrict__ pos,
                           int64_t *__restrict__ cnt) {
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) cnt[r] = pos[rp[r + 1]] - pos[rp[r]];
}

__global__ void k_i64_rowptr(long n, const int64_t *__restrict__ off, int32_t *__restrict__ rp) {
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r <= n) rp[r] = (int32_t)off[r];
}

__global__ void k_keep_entries(long nnz, const int32_t *__restrict__ ci, const double *__restrict__ va,
                               const int64_t *__restrict__ pos, int32_t *__restrict__ oci, double *__restrict__ ova) {
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nnz && pos[k + 1] != pos[k]) {
        oci[pos[k]] = ci[k];
        ova[pos[k]] = va[k];
    }
}

__global__ void k_col_keys(long nnz, const int32_t *__restrict__ ci, int kb, uint64_t *__restrict__ keys) {
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nnz) keys[k] = ((uint64_t)(uint32_t)ci[k] << kb) | (uint64_t)k;
}

__global__ void k_col_starts(long nc, long nnz, const uint64_t *__restrict__ keys, int kb, int32_t *__restrict__ rp) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c > nc) return;
    const uint64_t key = (uint64_t)c << kb;
    long lo = 0, hi = nnz;
    while (lo < hi) {
        const long mid = (lo + hi) >> 1;
        if (keys[mid] < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    rp[c] = (int32_t)lo;
}

__global__ void k_gather_t(long nnz, const uint64_t *__restrict__ keys, const int32_t *__restrict__ rowof,
                           const double *__restrict__ va, int32_t *__restrict__ tci, double *__restrict__ tva,
                           uint64_t kmask) {
    long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nnz) return;
    const long k = (long)(keys[p] & kmask);
    tci[p] = rowof[k];
    tva[p] = va[k];
}


}  // namespace

int csr_pick_lanes(const DevCSR &A) {
    double mean = A.nrows ? (double)A.nnz / (double)A.nrows : 1.0;
    int L = 1;
    while (L < 64 && L < mean) L <<= 1;
    return L;
}

lt_status csr_prepare(DevCSR &A) {
    int32_t last = 0;  // host source of an async copy: lives until the final sync
    A.lanes = csr_pick_lanes(A);
    A.nlong = A.nchunk = A.nmed = 0;
    if (A.nrows > 0) {
        const long n = A.nrows;
        DevBuf<int64_t> flag, pos, nch, choff, mflag, mpos;
        LT_TRY(flag.alloc(n + 1));
        LT_TRY(pos.alloc(n + 1));
        LT_TRY(nch.alloc(n + 1));
        LT_TRY(choff.alloc(n + 1));
        LT_TRY(mflag.alloc(n + 1));
        LT_TRY(mpos.alloc(n + 1));
        hipLaunchKernelGGL(k_long_rows_flag, dim3(blocks_for(n, 256)), dim3(256), 0, stream(), n, A.rp.p, A.long_row,
                           flag.p, nch.p, mflag.p);
        LT_CHECK_LAUNCH();
        LT_TRY(exclusive_scan_i64(flag.p, pos.p, n));
        LT_TRY(exclusive_scan_i64(nch.p, choff.p, n));
        LT_TRY(exclusive_scan_i64(mflag.p, mpos.p, n));
        int64_t nl = 0, nc = 0, nm = 0;
        LT_TRY(pos.download(&nl, 1, n));
        LT_TRY(choff.download(&nc, 1, n));
        LT_TRY(mpos.download(&nm, 1, n));
        LT_TRY(sync());
        A.nlong = nl;
        A.nchunk = nc;
        A.nmed = nm;
        if (nm > 0) {
            LT_TRY(A.med_rows.alloc(nm));
            hipLaunchKernelGGL(k_flag_list, dim3(blocks_for(n, 256)), dim3(256), 0, stream(), n, mpos.p,
                               A.med_rows.p);
            LT_CHECK_LAUNCH();
        }
        if (nl > 0) {
            DevBuf<int32_t> chunk_row;
            LT_TRY(A.long_rows.alloc(nl));
            LT_TRY(A.long_chunk0.alloc(nl + 1));
            LT_TRY(A.chunk_beg.alloc(nc));
            LT_TRY(chunk_row.alloc(nc));
            LT_TRY(A.chunk_sum.alloc(nc));
            hipLaunchKernelGGL(k_long_rows_fill, dim3(blocks_for(n, 256)), dim3(256), 0, stream(), n, A.rp.p, pos.p,
                               choff.p, A.long_rows.p, A.long_chunk0.p, A.chunk_beg.p, chunk_row.p);
            LT_CHECK_LAUNCH();
            last = (int32_t)nc;
            LT_HIP(hipMemcpyAsync(A.long_chunk0.p + nl, &last, sizeof(int32_t), hipMemcpyHostToDevice, stream()));
            // chunk -> row mapping is needed by the chunk kernel for the row end
            A.long_rows_of_chunk = std::move(chunk_row);
        }
    }
    A.nblk = 0;
    if (A.nrows > 0 && A.nnz > 0) {
        A.nblk = (A.nnz + kStreamNnz - 1) / kStreamNnz;
        LT_TRY(A.rblk.alloc(A.nblk + 1));
        hipLaunchKernelGGL(k_stream_blocks, dim3(blocks_for(A.nblk + 1, 256)), dim3(256), 0, stream(), A.nblk,
                           A.nrows, A.rp.p, A.rblk.p);
        LT_CHECK_LAUNCH();
    }
    A.prepared = true;
    return sync();
}

// calls f(std::integral_constant<int, L>) for a lane-group width L
template <class Fn>
static lt_status with_lanes(int lanes, Fn &&f) {
    switch (lanes) {
        case 1: f(std::integral_constant<int, 1>{}); break;
        case 2: f(std::integral_constant<int, 2>{}); break;
        case 4: f(std::integral_constant<int, 4>{}); break;
        case 8: f(std::integral_constant<int, 8>{}); break;
        case 16: f(std::integral_constant<int, 16>{}); break;
        case 32: f(std::integral_constant<int, 32>{}); break;
        case 64: f(std::integral_constant<int, 64>{}); break;
        default:
            set_error("spmv: lanes per row must be a power of two <= 64");
            return LT_ERR_ARG;
    }
    LT_CHECK_LAUNCH();
    return LT_OK;
}

// alist / nact (device list, its length): when the matrix has a sample
// partition and the list holds fewer than all samples, the row kernels
// launch a sample grid over the listed samples only (same row products,
// same bits); the streaming and long-row kernels keep their filters.
template <class T>
static lt_status spmv_t(const DevCSR &A, const T *va, T *chunk_sum, const T *x, T *y, T alpha, T beta_y, int lanes,
                        const int *act, long rows_per_sample, const int32_t *rowseg, bool nz_only = false,
                        const int *alist = nullptr, int nact = 0) {
    if (A.nrows == 0) return LT_OK;
    const RowFilter F{act, rows_per_sample > 0 ? rows_per_sample : 1, rowseg};
    const bool sg = alist && A.ns > 0 && nact < A.ns;
    if (sg && nact <= 0) return LT_OK;
    if (lanes <= 0) lanes = A.prepared ? A.lanes : csr_pick_lanes(A);
    // row length above which the medium-row and chunked paths sum
    const int skip = (A.prepared && (A.nlong > 0 || A.nmed > 0)) ? A.long_row : 0;
    if (nz_only) {
        if (A.nnzr < 0 || !A.prepared || alpha != (T)1 || beta_y != (T)0 || lanes > 64) {
            set_error("spmv: nonempty-row product needs csr_nonempty_rows, alpha 1, beta 0, lanes <= 64");
            return LT_ERR_ARG;
        }
        const T *eva = std::is_same<T, float>::value ? (const T *)A.ell_va32.p : (const T *)A.ell_va.p;
        if (sg) {
            if (A.sell_max > 0) {
                hipLaunchKernelGGL((k_spmv_ell4_sg<T>), dim3(blocks_for(A.sell_max, 256), nact), dim3(256), 0,
                                   stream(), alist, act, A.sell.p, A.ell_row.p, (const int4 *)A.ell_ci.p, eva, x, y);
                LT_CHECK_LAUNCH();
            }
            if (A.snzr_max > 0)
                LT_TRY(with_lanes(lanes, [&](auto LC) {
                    constexpr int L = decltype(LC)::value;
                    hipLaunchKernelGGL((k_spmv_rows_sg<L, T>), dim3(blocks_for(A.snzr_max * L, 256), nact),
                                       dim3(256), 0, stream(), alist, act, A.snzr.p, A.nzr.p, A.rp.p, A.ci.p, va, x,
                                       y, skip);
                }));
        } else {
            if (A.nell > 0) {
                hipLaunchKernelGGL((k_spmv_ell4<T>), dim3(blocks_for(A.nell, 256)), dim3(256), 0, stream(), A.nell,
                                   A.ell_row.p, (const int4 *)A.ell_ci.p, eva, x, y, F);
                LT_CHECK_LAUNCH();
            }
            // a dispatch covers at most 2^30 work-items (the grid size is 32-bit)
            const long per = (1L << 30) / lanes;
            for (long i0 = 0; i0 < A.nnzr; i0 += per) {
                const long ni = std::min(per, A.nnzr - i0);
                LT_TRY(with_lanes(lanes, [&](auto LC) {
                    constexpr int L = decltype(LC)::value;
                    hipLaunchKernelGGL((k_spmv_rows<L, T>), dim3(blocks_for(ni * L, 256)), dim3(256), 0, stream(),
                                       i0, i0 + ni, A.nzr.p, A.rp.p, A.ci.p, va, x, y, skip, F);
                }));
            }
        }
    } else if (lanes == kLanesSlice && A.nslice > 0) {
        const T *sva = std::is_same<T, float>::value ? (const T *)A.sl_va32.p : (const T *)A.sl_va.p;
        if (sg) {
            if (A.srow_max > 0) {
                hipLaunchKernelGGL((k_spmv_slice_sg<T>), dim3(blocks_for(A.srow_max, 256), nact), dim3(256), 0,
                                   stream(), alist, act, A.srow.p, A.rp.p, A.sl_ptr.p, A.sl_ci.p, sva, x, y, alpha,
                                   beta_y, skip);
                LT_CHECK_LAUNCH();
            }
        } else {
            hipLaunchKernelGGL((k_spmv_slice<T>), dim3(blocks_for(A.nrows, 256)), dim3(256), 0, stream(), A.nrows,
                               A.rp.p, A.sl_ptr.p, A.sl_ci.p, sva, x, y, alpha, beta_y, skip, F);
            LT_CHECK_LAUNCH();
        }
    } else if (lanes == kLanesStream && A.prepared && A.nblk > 0) {
        hipLaunchKernelGGL((k_spmv_stream<T>), dim3((unsigned)A.nblk), dim3(256), 0, stream(), A.nnz, A.rblk.p,
                           A.rp.p, A.ci.p, va, x, y, alpha, beta_y, skip, F);
        LT_CHECK_LAUNCH();
    } else {
        if (lanes == kLanesStream || lanes == kLanesSlice) lanes = csr_pick_lanes(A);
        if (sg) {
            if (A.srow_max > 0)
                LT_TRY(with_lanes(lanes, [&](auto LC) {
                    constexpr int L = decltype(LC)::value;
                    hipLaunchKernelGGL((k_spmv_sg<L, T>), dim3(blocks_for(A.srow_max * L, 256), nact), dim3(256), 0,
                                       stream(), alist, act, A.srow.p, A.rp.p, A.ci.p, va, x, y, alpha, beta_y, skip);
                }));
        } else {
            const long rows_per_launch = (1L << 30) / std::min(lanes, 64);
            for (long r0 = 0; r0 < A.nrows; r0 += rows_per_launch) {
                const long nr = std::min(rows_per_launch, A.nrows - r0);
                LT_TRY(with_lanes(lanes, [&](auto LC) {
                    constexpr int L = decltype(LC)::value;
                    hipLaunchKernelGGL((k_spmv<L, T>), dim3(blocks_for(nr * L, 256)), dim3(256), 0, stream(), r0,
                                       r0 + nr, A.rp.p, A.ci.p, va, x, y, alpha, beta_y, skip, F);
                }));
            }
        }
    }
    if (skip && A.nmed > 0) {
        hipLaunchKernelGGL((k_spmv_med<T>), dim3((unsigned)((A.nmed + 3) / 4)), dim3(256), 0, stream(), A.nmed,
                           A.med_rows.p, A.rp.p, A.ci.p, va, x, y, alpha, beta_y, F);
        LT_CHECK_LAUNCH();
    }
    if (skip && A.nlong > 0) {
        hipLaunchKernelGGL((k_spmv_chunks<T>), dim3((unsigned)A.nchunk), dim3(256), 0, stream(), A.nchunk,
                           A.chunk_beg.p, A.long_rows_of_chunk.p, A.rp.p, A.ci.p, va, x, chunk_sum, F);
        LT_CHECK_LAUNCH();
        hipLaunchKernelGGL((k_spmv_long<T>), dim3((unsigned)((A.nlong + 3) / 4)), dim3(256), 0, stream(), A.nlong,
                           A.long_rows.p, A.long_chunk0.p, chunk_sum, y, alpha, beta_y, F);
        LT_CHECK_LAUNCH();
    }
    return LT_OK;
}

lt_status csr_make_slices(DevCSR &A) {
    if (A.nslice > 0 || A.nrows == 0 || A.nnz == 0) return LT_OK;
    if (!A.prepared) LT_TRY(csr_prepare(A));
    const long nsl = (A.nrows + 63) / 64;
    const int skip = (A.nlong > 0 || A.nmed > 0) ? A.long_row : 0;
    DevBuf<int64_t> width;
    LT_TRY(width.alloc(nsl + 1));
    LT_TRY(A.sl_ptr.alloc(nsl + 1));
    hipLaunchKernelGGL(k_slice_width, dim3(blocks_for(nsl, 256)), dim3(256), 0, stream(), nsl, A.nrows, A.rp.p, skip,
                       width.p);
    LT_CHECK_LAUNCH();
    LT_TRY(exclusive_scan_i64(width.p, A.sl_ptr.p, nsl));
    int64_t tot = 0;
    LT_TRY(A.sl_ptr.download(&tot, 1, nsl));
    LT_TRY(A.sl_ci.alloc(std::max<long>(tot, 1)));
    LT_TRY(A.sl_va.alloc(std::max<long>(tot, 1)));
    hipLaunchKernelGGL(k_slice_fill, dim3(blocks_for(A.nrows, 256)), dim3(256), 0, stream(), A.nrows, A.rp.p, A.ci.p,
                       A.va.p, A.sl_ptr.p, skip, A.sl_ci.p, A.sl_va.p);
    LT_CHECK_LAUNCH();
    if (A.va32.p) {
        LT_TRY(A.sl_va32.alloc(std::max<long>(tot, 1)));
        if (tot > 0) {  // no entries when every row is long (chunked) or empty
            hipLaunchKernelGGL(k_d2f_sp, dim3(blocks_for(tot, 256)), dim3(256), 0, stream(), tot, A.sl_va.p,
                               A.sl_va32.p);
            LT_CHECK_LAUNCH();
        }
    }
    LT_TRY(sync());
    A.nslice = nsl;
    return LT_OK;
}

lt_status csr_set_samples(DevCSR &A, const int64_t *row_off, int S) {
    if (S <= 0 || row_off[S] != A.nrows) {
        set_error("spmv: sample partition does not cover the matrix rows");
        return LT_ERR_ARG;
    }
    LT_TRY(A.srow.upload(row_off, S + 1));
    A.srow_max = 0;
    for (int s = 0; s < S; ++s) A.srow_max = std::max<long>(A.srow_max, row_off[s + 1] - row_off[s]);
    if (A.nnzr >= 0) {
        std::vector<int64_t> he(S + 1), hn(S + 1);
        LT_TRY(A.sell.alloc(S + 1));
        LT_TRY(A.snzr.alloc(S + 1));
        hipLaunchKernelGGL(k_lower_bound, dim3(blocks_for(S + 1, 64)), dim3(64), 0, stream(), S + 1, A.srow.p,
                           A.ell_row.p, A.nell, A.sell.p);
        hipLaunchKernelGGL(k_lower_bound, dim3(blocks_for(S + 1, 64)), dim3(64), 0, stream(), S + 1, A.srow.p,
                           A.nzr.p, A.nnzr, A.snzr.p);
        LT_CHECK_LAUNCH();
        LT_TRY(A.sell.download(he.data(), S + 1));
        LT_TRY(A.snzr.download(hn.data(), S + 1));
        LT_TRY(sync());
        A.sell_max = A.snzr_max = 0;
        for (int s = 0; s < S; ++s) {
            A.sell_max = std::max<long>(A.sell_max, he[s + 1] - he[s]);
            A.snzr_max = std::max<long>(A.snzr_max, hn[s + 1] - hn[s]);
        }
    }
    A.ns = S;
    return sync();
}

lt_status csr_spmv(const DevCSR &A, const double *x, double *y, double alpha, double beta_y, int lanes,
                   const int *act, long rows_per_sample, const int32_t *rowseg, const int *alist, int nact) {
    return spmv_t<double>(A, A.va.p, A.chunk_sum.p, x, y, alpha, beta_y, lanes, act, rows_per_sample, rowseg, false,
                          alist, nact);
}

lt_status csr_spmv_fv(DevCSR &A, const double *x, double *y, const int *act, const int32_t *rowseg) {
    if (A.nrows == 0) return LT_OK;
    if (!A.prepared || A.nblk <= 0) return csr_spmv(A, x, y, 1.0, 0.0, 0, act, 0, rowseg);
    if (!A.va32.p) LT_TRY(csr_make_f32(A));
    const RowFilter F{act, 1, rowseg};
    hipLaunchKernelGGL((k_spmv_stream<double, float>), dim3((unsigned)A.nblk), dim3(256), 0, stream(), A.nnz,
                       A.rblk.p, A.rp.p, A.ci.p, A.va32.p, x, y, 1.0, 0.0, 0, F);
    LT_CHECK_LAUNCH();
    return LT_OK;
}

lt_status csr_spmv_ff(DevCSR &A, const float *x, float *y, const int *act, const int32_t *rowseg) {
    if (A.nrows == 0) return LT_OK;
    if (!A.prepared || A.nblk <= 0) {
        set_error("spmv: the FP32 streaming product needs a prepared matrix");
        return LT_ERR_ARG;
    }
    if (!A.va32.p) LT_TRY(csr_make_f32(A));
    const RowFilter F{act, 1, rowseg};
    hipLaunchKernelGGL((k_spmv_stream<float, float>), dim3((unsigned)A.nblk), dim3(256), 0, stream(), A.nnz, A.rblk.p,
                       A.rp.p, A.ci.p, A.va32.p, x, y, 1.f, 0.f, 0, F);
    LT_CHECK_LAUNCH();
    return LT_OK;
}

lt_status csr_spmv_nz(const DevCSR &A, const double *x, double *y, const int *act, long rows_per_sample,
                      const int *alist, int nact) {
    return spmv_t<double>(A, A.va.p, A.chunk_sum.p, x, y, 1.0, 0.0, A.lanes_nz > 0 ? A.lanes_nz : 4, act,
                          rows_per_sample, nullptr, true, alist, nact);
}

lt_status csr_spmv32_nz(const DevCSR &A, const float *x, float *y, const int *act, long rows_per_sample,
                        const int *alist, int nact) {
    if ((A.nnz > 0 && !A.va32.p) || (A.nell > 0 && !A.ell_va32.p)) {
        set_error("spmv: FP32 values not prepared");
        return LT_ERR_ARG;
    }
    const int lanes = A.lanes_nz32 > 0 ? A.lanes_nz32 : (A.lanes_nz > 0 ? A.lanes_nz : 4);
    return spmv_t<float>(A, A.va32.p, A.chunk_sum32.p, x, y, 1.f, 0.f, lanes, act, rows_per_sample, nullptr, true,
                         alist, nact);
}

lt_status csr_nonempty_rows(DevCSR &A) {
    if (A.nnzr >= 0) return LT_OK;
    const long n = A.nrows;
    DevBuf<int64_t> fe, fo, pe, po;
    LT_TRY(fe.alloc(n + 1));
    LT_TRY(fo.alloc(n + 1));
    LT_TRY(pe.alloc(n + 1));
    LT_TRY(po.alloc(n + 1));
    if (n > 0) {
        hipLaunchKernelGGL(k_row_class, dim3(blocks_for(n, 256)), dim3(256), 0, stream(), n, A.rp.p, fe.p, fo.p);
        LT_CHECK_LAUNCH();
    }
    LT_TRY(exclusive_scan_i64(fe.p, pe.p, n));
    LT_TRY(exclusive_scan_i64(fo.p, po.p, n));
    int64_t ne = 0, no = 0;
    LT_TRY(pe.download(&ne, 1, n));
    LT_TRY(po.download(&no, 1, n));
    LT_TRY(A.nzr.alloc(std::max<long>(no, 1)));
    LT_TRY(A.ell_row.alloc(std::max<long>(ne, 1)));
    LT_TRY(A.ell_ci.alloc(4 * std::max<long>(ne, 1)));
    LT_TRY(A.ell_va.alloc(4 * std::max<long>(ne, 1)));
    if (n > 0) {
        hipLaunchKernelGGL(k_row_split, dim3(blocks_for(n, 256)), dim3(256), 0, stream(), n, A.rp.p, A.ci.p, A.va.p,
                           pe.p, po.p, A.ell_row.p, A.ell_ci.p, A.ell_va.p, A.nzr.p);
        LT_CHECK_LAUNCH();
    }
    LT_TRY(sync());
    A.nnzr = no;
    A.nell = ne;
    if (A.va32.p) LT_TRY(csr_make_f32(A));  // the table's FP32 values
    return LT_OK;
}

lt_status csr_make_f32(DevCSR &A) {
    if (A.nslice > 0 && !A.sl_va32.p) {
        int64_t tot = 0;
        LT_TRY(A.sl_ptr.download(&tot, 1, A.nslice));
        LT_TRY(sync());
        LT_TRY(A.sl_va32.alloc(std::max<long>(tot, 1)));
        if (tot > 0) {  // no entries when every row is long (chunked) or empty
            hipLaunchKernelGGL(k_d2f_sp, dim3(blocks_for(tot, 256)), dim3(256), 0, stream(), tot, A.sl_va.p,
                               A.sl_va32.p);
            LT_CHECK_LAUNCH();
        }
    }
    if (A.nell > 0 && !A.ell_va32.p) {
        LT_TRY(A.ell_va32.alloc(4 * A.nell));
        hipLaunchKernelGGL(k_d2f_sp, dim3(blocks_for(4 * A.nell, 256)), dim3(256), 0, stream(), 4 * A.nell,
                           A.ell_va.p, A.ell_va32.p);
        LT_CHECK_LAUNCH();
    }
    if (A.va32.p || A.nnz == 0) return LT_OK;
    LT_TRY(A.va32.alloc(A.nnz));
    hipLaunchKernelGGL(k_d2f_sp, dim3(blocks_for(A.nnz, 256)), dim3(256), 0, stream(), A.nnz, A.va.p, A.va32.p);
    LT_CHECK_LAUNCH();
    if (A.nchunk > 0) LT_TRY(A.chunk_sum32.alloc(A.nchunk));
    return LT_OK;
}

lt_status csr_spmv32(const DevCSR &A, const float *x, float *y, float alpha, float beta_y, const int *act,
                     long rows_per_sample, const int32_t *rowseg, const int *alist, int nact) {
    if (A.nnz > 0 && !A.va32.p) {
        set_error("spmv: FP32 values not prepared");
        return LT_ERR_ARG;
    }
    return spmv_t<float>(A, A.va32.p, A.chunk_sum32.p, x, y, alpha, beta_y, A.lanes32, act, rows_per_sample,
                         rowseg, false, alist, nact);
}

// Autotuner: times the product with every lane-group width (1..64 lanes per
// row) on this matrix and keeps the fastest (A.lanes, or A.lanes32 for the
// FP32 values).  A few milliseconds per matrix.
template <class T>
static lt_status tune_lanes(DevCSR &A, const T *va, T *chunk_sum, bool nz, int &best_out) {
    DevBuf<T> x, y;
    LT_TRY(x.alloc(std::max<long>(A.ncols, 1)));
    LT_TRY(y.alloc(std::max<long>(A.nrows, 1)));
    LT_TRY(x.zero());
    hipEvent_t e0, e1;
    LT_HIP(hipEventCreate(&e0));
    LT_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    int best_l = A.lanes;
    for (int L = 1; L <= kLanesSlice; L <<= 1) {
        if (L == kLanesStream && (A.nblk == 0 || nz)) continue;
        if (L == kLanesSlice && (A.nslice == 0 || nz)) continue;
        LT_TRY(spmv_t<T>(A, va, chunk_sum, x.p, y.p, (T)1, (T)0, L, nullptr, 0, nullptr, nz));  // warm-up
        LT_HIP(hipEventRecord(e0, stream()));
        for (int r = 0; r < 3; ++r)
            LT_TRY(spmv_t<T>(A, va, chunk_sum, x.p, y.p, (T)1, (T)0, L, nullptr, 0, nullptr, nz));
        LT_HIP(hipEventRecord(e1, stream()));
        LT_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        LT_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) {
            best = ms;
            best_l = L;
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    best_out = best_l;
    return LT_OK;
}

// Process-wide memo of tuned lane counts: matrices of the same shape class
// (device, precision, row count, mean row length to a quarter octave,
// presence of long rows) reuse the first measurement, so rebuilding lens
// matrices of a new batch of the same geometry does not re-run the sweep.
namespace {
// shape class of a product: device, precision, mean row length (quarter
// octaves), long rows present, product over the nonempty rows, row count
// (half octaves: successive lens batches of one workload share a class)
struct TuneKey {
    int dev, fp32, lenq, has_long, nz;
    long nrows;
    bool operator<(const TuneKey &o) const {
        return std::tie(dev, fp32, lenq, has_long, nz, nrows) <
               std::tie(o.dev, o.fp32, o.lenq, o.has_long, o.nz, o.nrows);
    }
};
std::mutex g_tune_mu;
std::map<TuneKey, int> g_tune;
}  // namespace

lt_status csr_autotune(DevCSR &A, bool fp32) {
    if (A.nrows == 0 || A.nnz == 0) return LT_OK;
    if (!A.prepared) LT_TRY(csr_prepare(A));
    // the sliced copy is a candidate for matrices of several entries per row
    // (not for those multiplied over their nonempty rows); built only when
    // tuning or when a memoized choice needs it
    // (slices as wide as their longest row: only for rows of up to 64 entries)
    const bool slice_cand = A.nnzr < 0 && (double)A.nnz >= 4.0 * (double)A.nrows && A.long_row <= 64;
    if (fp32) LT_TRY(csr_make_f32(A));
    int dev = 0;
    LT_HIP(hipGetDevice(&dev));
    // the full product, then (when the nonempty-row list exists) the
    // product over the listed rows, each with its own memo entry
    for (int nz = 0; nz <= (A.nnzr >= 0 ? 1 : 0); ++nz) {
        const double mean = (double)A.nnz / (double)(nz ? std::max(1L, A.nnzr) : A.nrows);
        const long rows = std::max(1L, nz ? A.nnzr : A.nrows);
        const TuneKey key{dev, fp32 ? 1 : 0, (int)std::lround(4.0 * std::log2(mean)),
                          (A.nlong > 0 || A.nmed > 0) ? 1 : 0, nz, std::lround(2.0 * std::log2((double)rows))};
        int &slot = nz ? (fp32 ? A.lanes_nz32 : A.lanes_nz) : (fp32 ? A.lanes32 : A.lanes);
        bool hit = false;
        {
            std::lock_guard<std::mutex> g(g_tune_mu);
            auto it = g_tune.find(key);
            if (it != g_tune.end()) {
                slot = it->second;
                hit = true;
            }
        }
        if (hit) {
            if (slot == kLanesSlice) {
                LT_TRY(csr_make_slices(A));
                if (fp32) LT_TRY(csr_make_f32(A));  // the slices' FP32 values
            }
            continue;
        }
        if (slice_cand) {
            LT_TRY(csr_make_slices(A));
            if (fp32) LT_TRY(csr_make_f32(A));
        }
        int best = 0;
        if (fp32)
            LT_TRY(tune_lanes<float>(A, A.va32.p, A.chunk_sum32.p, nz != 0, best));
        else
            LT_TRY(tune_lanes<double>(A, A.va.p, A.chunk_sum.p, nz != 0, best));
        slot = best;
        std::lock_guard<std::mutex> g(g_tune_mu);
        g_tune[key] = best;
    }
    return LT_OK;
}

// Transpose: counting sort of the nonzeros by column, then a segmented sort
// inside each column on the original nonzero position, which orders every
// output row by ascending source row -- independent of atomic order.
lt_status csr_drop_zeros(const DevCSR &A, DevCSR &B) {
    const long nnz = A.nnz, n = A.nrows;
    B = DevCSR();
    B.nrows = n;
    B.ncols = A.ncols;
    LT_TRY(B.rp.alloc(n + 1));
    DevBuf<int64_t> flag, pos, cnt, off;
    LT_TRY(flag.alloc(nnz + 1));
    LT_TRY(pos.alloc(nnz + 1));
    LT_TRY(cnt.alloc(n + 1));
    LT_TRY(off.alloc(n + 1));
    if (nnz > 0) {
        hipLaunchKernelGGL(k_nonzero_flags, dim3(blocks_for(nnz, 256)), dim3(256), 0, stream(), nnz, A.va.p, flag.p);
        LT_CHECK_LAUNCH();
    }
    LT_TRY(exclusive_scan_i64(flag.p, pos.p, nnz));
    int64_t kept = 0;
    LT_TRY(pos.download(&kept, 1, nnz));
    hipLaunchKernelGGL(k_row_kept, dim3(blocks_for(n, 256)), dim3(256), 0, stream(), n, A.rp.p, pos.p, cnt.p);
    LT_CHECK_LAUNCH();
    LT_TRY(exclusive_scan_i64(cnt.p, off.p, n));
    hipLaunchKernelGGL(k_i64_rowptr, dim3(blocks_for(n + 1, 256)), dim3(256), 0, stream(), n, off.p, B.rp.p);
    LT_CHECK_LAUNCH();
    LT_TRY(sync());
    B.nnz = kept;
    LT_TRY(B.ci.alloc(std::max<long>(kept, 1)));
    LT_TRY(B.va.alloc(std::max<long>(kept, 1)));
    if (nnz > 0) {
        hipLaunchKernelGGL(k_keep_entries, dim3(blocks_for(nnz, 256)), dim3(256), 0, stream(), nnz, A.ci.p, A.va.p,
                           pos.p, B.ci.p, B.va.p);
        LT_CHECK_LAUNCH();
    }
    return sync();
}

lt_status csr_transpose(const DevCSR &A, DevCSR &B) {
    const long nnz = A.nnz, nc = A.ncols;
    B.nrows = nc;
    B.ncols = A.nrows;
    B.nnz = nnz;
    LT_TRY(B.rp.alloc(nc + 1));
    LT_TRY(B.ci.alloc(std::max<long>(nnz, 1)));
    LT_TRY(B.va.alloc(std::max<long>(nnz, 1)));
    if (nnz == 0) {
        LT_TRY(B.rp.zero());
        return sync();
    }
    // one radix sort of (column, position) keys: ascending columns, and
    // inside a column ascending source rows, independent of any atomics
    int kb = 1, cb = 1;
    while ((1L << kb) < nnz) ++kb;
    while ((1L << cb) < nc) ++cb;
    DevBuf<int32_t> rowof;
    DevBuf<uint64_t> keys, keys2;
    LT_TRY(rowof.alloc(nnz));
    LT_TRY(keys.alloc(nnz));
    LT_TRY(keys2.alloc(nnz));
    hipLaunchKernelGGL(k_row_of, dim3(blocks_for(A.nrows, 256)), dim3(256), 0, stream(), A.nrows, A.rp.p, rowof.p);
    LT_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_col_keys, dim3(blocks_for(nnz, 256)), dim3(256), 0, stream(), nnz, A.ci.p, kb, keys.p);
    LT_CHECK_LAUNCH();
    size_t tb = 0;
    LT_HIP(rocprim::radix_sort_keys(nullptr, tb, keys.p, keys2.p, (size_t)nnz, 0, kb + cb, stream()));
    DevBuf<uint8_t> tmp;
    LT_TRY(tmp.alloc(std::max<size_t>(tb, 1)));
    LT_HIP(rocprim::radix_sort_keys(tmp.p, tb, keys.p, keys2.p, (size_t)nnz, 0, kb + cb, stream()));
    hipLaunchKernelGGL(k_gather_t, dim3(blocks_for(nnz, 256)), dim3(256), 0, stream(), nnz, keys2.p, rowof.p, A.va.p,
                       B.ci.p, B.va.p, (1ull << kb) - 1);
    LT_CHECK_LAUNCH();
    // row pointers of the transpose: first sorted key of each column
    hipLaunchKernelGGL(k_col_starts, dim3(blocks_for(nc + 1, 256)), dim3(256), 0, stream(), nc, nnz, keys2.p, kb,
                       B.rp.p);
    LT_CHECK_LAUNCH();
    return sync();
}

// ---------------------------------------------------------------------------
// SpMV that allocates nothing (lt_csr_spmv_device_noalloc): the long-row plan
// and its chunk sums live in `scratch`, one fixed array per GPU, instead of
// pool memory.  Same kernels and summation order as lt_csr_spmv_host.
// ---------------------------------------------------------------------------

namespace {

constexpr long kNaMaxChunks = 65536;  // chunk sums of rows above kNaLongRow entries
constexpr long kNaMaxLong = 32768;    // rows above kNaLongRow entries
constexpr int kNaLongRow = 2048;      // DevCSR::long_row default
// byte offsets into scratch: two counters, then the plan and the sums
constexpr size_t kNaRows = 256;
constexpr size_t kNaChunk0 = kNaRows + 4 * kNaMaxLong;
constexpr size_t kNaNch = kNaChunk0 + 4 * kNaMaxLong;
constexpr size_t kNaBeg = kNaNch + 4 * kNaMaxLong;
constexpr size_t kNaRow = kNaBeg + 4 * kNaMaxChunks;
constexpr size_t kNaSum = kNaRow + 4 * kNaMaxChunks;
constexpr size_t kNaBytes = kNaSum + 8 * kNaMaxChunks;  // 1.38 MiB

// per-GPU device array; host code uses its address from hipGetSymbolAddress
static __device__ __align__(256) unsigned char scratch[kNaBytes];

struct NaPlan {
    int *cnt;  // [0] long rows, [1] chunks
    int32_t *rows, *chunk0, *nch, *beg, *row;
    double *sum;
};

NaPlan na_plan(void *base) {
    unsigned char *b = (unsigned char *)base;
    return NaPlan{(int *)b,
                  (int32_t *)(b + kNaRows),
                  (int32_t *)(b + kNaChunk0),
                  (int32_t *)(b + kNaNch),
                  (int32_t *)(b + kNaBeg),
                  (int32_t *)(b + kNaRow),
                  (double *)(b + kNaSum)};
}

// counters to zero by a kernel (no memset node if a caller captures the work)
__global__ void k_na_reset(int *cnt) {
    if (threadIdx.x < 2) cnt[threadIdx.x] = 0;
}

// every row above long_row entries takes a slot and a contiguous range of
// chunk slots; the slots' order varies from run to run, a row's sum does not
// (its chunks are summed in a fixed order by k_na_long)
__global__ void k_na_plan(long nrows, const int32_t *__restrict__ rp, NaPlan P) {
    const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    const int b = rp[r], e = rp[r + 1];
    if (e - b <= kNaLongRow) return;
    const int nc = (e - b + kChunk - 1) / kChunk;
    const int l = atomicAdd(&P.cnt[0], 1);
    const int c0 = atomicAdd(&P.cnt[1], nc);
    if (l >= kNaMaxLong || c0 + nc > kNaMaxChunks) return;  // excluded by the host's size check
    P.rows[l] = (int32_t)r;
    P.chunk0[l] = c0;
    P.nch[l] = nc;
    for (int k = 0; k < nc; ++k) {
        P.beg[c0 + k] = b + k * kChunk;
        P.row[c0 + k] = (int32_t)r;
    }
}

__global__ void __launch_bounds__(256) k_na_chunks(const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                    const double *__restrict__ va, const double *__restrict__ x,
                                                    NaPlan P) {
    const int c = blockIdx.x;
    if (c >= P.cnt[1]) return;  // whole workgroups leave together
    chunk_sum<double>(P.beg[c], P.row[c], rp, ci, va, x, P.sum + c);
}

__global__ void __launch_bounds__(256) k_na_long(NaPlan P, double *__restrict__ y) {
    const int l = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (l >= P.cnt[0]) return;  // whole waves
    const double s = chunks_total(P.sum, P.chunk0[l], P.chunk0[l] + P.nch[l], lane);
    if (lane == 0) y[P.rows[l]] = 1.0 * s;
}

std::mutex g_na_mu;
void *g_na_ptr[64] = {};

}  // namespace

// Called by the library's device set-up with `device` current: the device
// address of this GPU's copy of scratch.
lt_status noalloc_setup(int device) {
    if (device < 0 || device >= 64) {
        set_error("noalloc: device %d out of range", device);
        return LT_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (g_na_ptr[device]) return LT_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        set_error("noalloc: hipGetSymbolAddress(scratch) failed on device %d: %s", device, hipGetErrorString(e));
        return LT_ERR_HIP;
    }
    g_na_ptr[device] = p;
    return LT_OK;
}

}  // namespace lt

using namespace lt;

extern "C" {

lt_status lt_csr_spmv_device_noalloc(long nrows, long ncols, long nnz, const int *d_row_ptr, const int *d_cols,
                                     const double *d_vals, const double *d_x, double *d_y) {
    LT_TRY(ensure_init());
    if (nrows < 0 || ncols < 0 || nnz < 0 || !d_row_ptr || (nrows > 0 && !d_y) || (nnz > 0 && (!d_cols || !d_vals)) ||
        (ncols > 0 && nnz > 0 && !d_x) || nrows > (1L << 31) - 2 || nnz > (1L << 31) - 1) {
        set_error("lt_csr_spmv_device_noalloc: bad arguments");
        return LT_ERR_ARG;
    }
    // every row above kNaLongRow entries has at least kNaLongRow + 1, so at
    // most nnz / (kNaLongRow + 1) such rows and nnz / kChunk + that many chunks
    const long max_long = nnz / (kNaLongRow + 1), max_chunks = nnz / kChunk + max_long;
    if (max_long > kNaMaxLong || max_chunks > kNaMaxChunks) {
        set_error("lt_csr_spmv_device_noalloc: %ld nonzeros exceed the fixed scratch (at most %ld chunks of %d "
                  "entries for rows longer than %d)", nnz, kNaMaxChunks, kChunk, kNaLongRow);
        return LT_ERR_CAPACITY;
    }
    int dev = 0;
    LT_HIP(hipGetDevice(&dev));
    void *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_na_mu);
        base = (dev >= 0 && dev < 64) ? g_na_ptr[dev] : nullptr;
    }
    if (!base) {
        set_error("lt_csr_spmv_device_noalloc: scratch not set up on device %d", dev);
        return LT_ERR_HIP;
    }
    if (nrows == 0) return LT_OK;
    int32_t last = 0;
    LT_HIP(hipMemcpyAsync(&last, d_row_ptr + nrows, sizeof(int32_t), hipMemcpyDeviceToHost, stream()));
    LT_TRY(sync());
    if ((long)last != nnz) {
        set_error("lt_csr_spmv_device_noalloc: row_ptr[nrows] = %d but nnz = %ld", last, nnz);
        return LT_ERR_ARG;
    }
    const NaPlan P = na_plan(base);
    hipLaunchKernelGGL(k_na_reset, dim3(1), dim3(64), 0, stream(), P.cnt);
    hipLaunchKernelGGL(k_na_plan, dim3(blocks_for(nrows, 256)), dim3(256), 0, stream(), nrows, d_row_ptr, P);
    LT_CHECK_LAUNCH();
    // rows up to kNaLongRow entries: the lane-group product lt_csr_spmv_host
    // picks from the mean row length (csr_pick_lanes)
    double mean = (double)nnz / (double)nrows;
    int lanes = 1;
    while (lanes < 64 && lanes < mean) lanes <<= 1;
    const RowFilter F{nullptr, 1, nullptr};
    const long rows_per_launch = (1L << 30) / lanes;
    for (long r0 = 0; r0 < nrows; r0 += rows_per_launch) {
        const long nr = std::min(rows_per_launch, nrows - r0);
        LT_TRY(with_lanes(lanes, [&](auto LC) {
            constexpr int L = decltype(LC)::value;
            hipLaunchKernelGGL((k_spmv<L, double>), dim3(blocks_for(nr * L, 256)), dim3(256), 0, stream(), r0,
                               r0 + nr, d_row_ptr, d_cols, d_vals, d_x, d_y, 1.0, 0.0, kNaLongRow, F);
        }));
    }
    if (max_chunks > 0) {
        hipLaunchKernelGGL(k_na_chunks, dim3((unsigned)max_chunks), dim3(256), 0, stream(), d_row_ptr, d_cols, d_vals,
                           d_x, P);
        hipLaunchKernelGGL(k_na_long, dim3((unsigned)((std::max(max_long, 1L) + 3) / 4)), dim3(256), 0, stream(), P,
                           d_y);
        LT_CHECK_LAUNCH();
    }
    return sync();
}

}  // extern "C"
