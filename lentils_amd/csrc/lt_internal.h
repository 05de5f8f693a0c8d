// Internal helpers shared by the HIP translation units of liblentils_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/lentils_hip.h"

namespace lt {

// ---------------------------------------------------------------------------
// error reporting: every lt_* entry point returns a code and leaves a
// message retrievable through lt_last_error_message()
// ---------------------------------------------------------------------------
void set_error(const char *fmt, ...);
const char *last_error();

#define LT_HIP(expr)                                                                            \
    do {                                                                                        \
        hipError_t _e = (expr);                                                                 \
        if (_e != hipSuccess) {                                                                 \
            ::lt::set_error("%s:%d %s: %s", __FILE__, __LINE__, #expr, hipGetErrorString(_e)); \
            return LT_ERR_HIP;                                                                  \
        }                                                                                       \
    } while (0)

#define LT_TRY(expr)                  \
    do {                              \
        lt_status _s = (expr);        \
        if (_s != LT_OK) return _s;   \
    } while (0)

#define LT_CHECK_LAUNCH() LT_HIP(hipGetLastError())

// library-wide device and stream (default priority; one process per GPU)
lt_status ensure_init();
// device set-up of the allocation-free SpMV's fixed scratch (sparse.hip);
// called with `device` current
lt_status noalloc_setup(int device);
hipStream_t stream();
int device_cu_count();

// Device memory pool behind DevBuf (capi.hip): freed blocks are kept per
// device and size class and handed to later allocations of the same class.
hipError_t pool_get(void **p, size_t bytes);
void pool_put(void *p, size_t bytes);

// owning device buffer
template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void release() {
        if (p) pool_put(p, n * sizeof(T));
        p = nullptr;
        n = 0;
    }
    lt_status alloc(size_t count) {
        if (count == n && p) return LT_OK;
        release();
        if (count == 0) return LT_OK;
        LT_HIP(pool_get((void **)&p, count * sizeof(T)));
        n = count;
        return LT_OK;
    }
    lt_status zero() {
        if (n) LT_HIP(hipMemsetAsync(p, 0, n * sizeof(T), stream()));
        return LT_OK;
    }
    lt_status upload(const T *h, size_t count) {
        LT_TRY(alloc(count));
        if (count) LT_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, stream()));
        return LT_OK;
    }
    lt_status download(T *h, size_t count, size_t offset = 0) const {
        if (count) LT_HIP(hipMemcpyAsync(h, p + offset, count * sizeof(T), hipMemcpyDeviceToHost, stream()));
        return LT_OK;
    }
};

inline lt_status sync() {
    LT_HIP(hipStreamSynchronize(stream()));
    return LT_OK;
}

inline unsigned blocks_for(long n, int threads) { return (unsigned)((n + threads - 1) / threads); }

// ---------------------------------------------------------------------------
// device-wide exclusive scan of int64 counts (scan.hip)
// ---------------------------------------------------------------------------
// out[i] = sum_{j<i} in[j], out[n] = total.  `in` and `out` may alias.
lt_status exclusive_scan_i64(const int64_t *in, int64_t *out, long n);
// inclusive running max of int32 values (used for "last non-empty row")
lt_status inclusive_max_i32(const int32_t *in, int32_t *out, long n);

// ---------------------------------------------------------------------------
// per-lens constants, precomputed on the host with the host libm so that the
// per-lens pow() of deflect.c:284 is bit-identical to the reference
// ---------------------------------------------------------------------------
struct DevLens {
    int type;  // 0 PEMD, 1 shear
    int pad;
    double x, y;
    // PEMD
    double q, qq, qt, t, tm2, b, bb, cr, ci;
    // shear and Keeton gradient terms
    double ss, c2, s2, gks, c_gka, s_gka, gss, c_gsa, s_gsa;
};

lt_status pack_lenses(const generic_mass_model *in, int n, std::vector<DevLens> &out);
// deflected points and per (point, PEMD component) the 12 parameter
// derivatives at dv[(p * npemd + c) * 12 + k] (deflect.hip)
lt_status deflect_points_deriv_device(const std::vector<DevLens> &lens, const double *pin_dev, double *pout_dev,
                                      double *dv_dev, long n);

// CSR on device (block-diagonal batch of matrices in one array)
struct DevCSR {
    long nrows = 0, ncols = 0, nnz = 0;
    DevBuf<int32_t> rp;  // nrows + 1 (global offsets)
    DevBuf<int32_t> ci;  // global column indices
    DevBuf<double> va;
    // long-row plan (csr_prepare): rows above long_row nonzeros are summed
    // by one wave each up to 512 entries (med_rows), longer ones in
    // fixed-size chunks by whole workgroups, then combined in order
    int long_row = 2048;
    long nmed = 0;
    DevBuf<int32_t> med_rows;
    bool prepared = false;
    int lanes = 1;  // lane-group width of k_spmv, kLanesStream (streaming kernel) or kLanesSlice (sliced copy)
    long nblk = 0;        // streaming-kernel workgroups (ceil(nnz / its nonzero range))
    DevBuf<int32_t> rblk;  // first row of each streaming workgroup, nblk + 1
    long nlong = 0, nchunk = 0;
    DevBuf<int32_t> long_rows, long_chunk0, chunk_beg, long_rows_of_chunk;
    DevBuf<double> chunk_sum;
    DevBuf<float> va32, chunk_sum32;  // FP32 values (csr_make_f32)
    int lanes32 = 0;                  // lane-group width of the FP32 product (0: as lanes)
    // nonempty rows (csr_nonempty_rows), for products that write those rows
    // only: rows of 1..4 entries as a padded 4-wide table (ell_*: row, four
    // columns, four values; padding value 0 at the row's first column), the
    // longer ones listed in nzr
    long nnzr = -1, nell = 0;
    DevBuf<int32_t> nzr, ell_row, ell_ci;
    DevBuf<double> ell_va;
    DevBuf<float> ell_va32;
    int lanes_nz = 0, lanes_nz32 = 0;
    // sample partition (csr_set_samples): sample s owns rows [srow[s],
    // srow[s+1]) and the table / list entries [sell[s], sell[s+1]) and
    // [snzr[s], snzr[s+1]); *_max: the largest per-sample extent
    int ns = 0;
    DevBuf<int64_t> srow, sell, snzr;
    long srow_max = 0, sell_max = 0, snzr_max = 0;
    // sliced copy (csr_make_slices): slice q = rows 64q .. 64q+63, entry j of
    // its row t at sl_ptr[q] + 64 j + t (slice width = its longest row that
    // is not left to the chunked path), for one thread per row with
    // coalesced loads
    long nslice = 0;
    DevBuf<int64_t> sl_ptr;
    DevBuf<int32_t> sl_ci;
    DevBuf<double> sl_va;
    DevBuf<float> sl_va32;
};

constexpr int kLanesStream = 128;  // pseudo lane count: the streaming (LDS-staged) SpMV kernel
constexpr int kLanesSlice = 256;   // pseudo lane count: one thread per row over the sliced copy

// analyse row lengths once (lane group size, long-row chunk plan)
lt_status csr_prepare(DevCSR &A);
// y = alpha A x + beta y (FP64); lanes_per_row <= 0 uses the prepared choice
// act (device, per sample): skip the rows of samples with act == 0; a row's
// sample is row / rows_per_sample, or rowseg[row] when rowseg is given
// alist / nact: the device list of samples the caller launches for (with a
// sample partition and nact below the sample count, only their rows run)
lt_status csr_spmv(const DevCSR &A, const double *x, double *y, double alpha, double beta_y, int lanes_per_row,
                   const int *act = nullptr, long rows_per_sample = 0, const int32_t *rowseg = nullptr,
                   const int *alist = nullptr, int nact = 0);
int csr_pick_lanes(const DevCSR &A);
// FP32 copy of the values (after csr_prepare) and y = alpha A x + beta y in FP32
lt_status csr_make_f32(DevCSR &A);
// y = A x with FP32 copies of the values (made on first use) and FP64 x, y,
// products and sums, by the streaming kernel (every row summed whole by the
// workgroup its first entry falls in): for a preconditioner's matrix, where
// the values' rounding only changes the preconditioner.  Rows of samples with
// act == 0 (rowseg: row -> sample) are skipped.
lt_status csr_spmv_fv(DevCSR &A, const double *x, double *y, const int *act, const int32_t *rowseg);
// y = A x in FP32 (FP32 values and vectors), the streaming kernel, no long-row path
lt_status csr_spmv_ff(DevCSR &A, const float *x, float *y, const int *act, const int32_t *rowseg);
// pick the fastest lane-group width for this matrix by timing (FP64 or FP32)
lt_status csr_autotune(DevCSR &A, bool fp32);
lt_status csr_spmv32(const DevCSR &A, const float *x, float *y, float alpha, float beta_y, const int *act = nullptr,
                     long rows_per_sample = 0, const int32_t *rowseg = nullptr, const int *alist = nullptr,
                     int nact = 0);
// B = A^T with column order inside each row ascending (deterministic)
lt_status csr_transpose(const DevCSR &A, DevCSR &B);
// B = A without its entries whose value is exactly zero (order kept)
lt_status csr_drop_zeros(const DevCSR &A, DevCSR &B);
// the nonempty rows of A, after csr_prepare: rows of 1..4 entries as the
// 4-wide table (A.ell_*, A.nell), longer ones as a list (A.nzr, A.nnzr)
lt_status csr_nonempty_rows(DevCSR &A);
// y = A x on the nonempty rows only (csr_nonempty_rows): the other rows of y
// are not written (the caller keeps them zero).  Same filter as csr_spmv.
lt_status csr_spmv_nz(const DevCSR &A, const double *x, double *y, const int *act = nullptr,
                      long rows_per_sample = 0, const int *alist = nullptr, int nact = 0);
lt_status csr_spmv32_nz(const DevCSR &A, const float *x, float *y, const int *act = nullptr,
                        long rows_per_sample = 0, const int *alist = nullptr, int nact = 0);
// the sliced copy of A (after csr_prepare; FP32 values too when present)
lt_status csr_make_slices(DevCSR &A);
// sample partition of A's rows (row_off: S + 1 host offsets, ascending);
// after csr_nonempty_rows when that is used
lt_status csr_set_samples(DevCSR &A, const int64_t *row_off, int S);

// Workgroup order for the 8 XCDs (each with its own L2): the dispatcher
// hands workgroup b to XCD b % 8, so renumbering the i-th workgroup of XCD x
// to start_x + i gives every XCD one contiguous range of the logical
// workgroups, processed in increasing order.  A bijection for any grid size.
__device__ inline long xcd_block(long b, long nb) {
    constexpr long X = 8;
    const long x = b % X, i = b / X, q = nb / X, r = nb % X;
    return x * q + (x < r ? x : r) + i;
}

}  // namespace lt
