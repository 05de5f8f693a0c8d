// Host-pointer entry points for the stand-alone operators of the drop-in
// surface: 2-D real FFTs with the FFTOperator conventions
// (fourier_operators.py:12-24: rfft2 unnormalised, irfft2 with norm='forward',
// i.e. also unnormalised) and CSR SpMV (Operator.apply on an explicit
// matrix, operator_base.py:96-98).
#include <rocfft/rocfft.h>

#include <algorithm>

#include "lt_internal.h"

namespace lt {
namespace {

__global__ void k_herm_col(int nu, int half_nv, int col, double *__restrict__ X) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > nu / 2) return;
    const int m = (nu - k) % nu;
    double ar = X[2 * ((long)k * half_nv + col)], ai = X[2 * ((long)k * half_nv + col) + 1];
    double br = X[2 * ((long)m * half_nv + col)], bi = X[2 * ((long)m * half_nv + col) + 1];
    X[2 * ((long)k * half_nv + col)] = 0.5 * (ar + br);
    X[2 * ((long)k * half_nv + col) + 1] = 0.5 * (ai - bi);
    if (m != k) {
        X[2 * ((long)m * half_nv + col)] = 0.5 * (br + ar);
        X[2 * ((long)m * half_nv + col) + 1] = 0.5 * (bi - ai);
    }
}

// plain streaming copy (16 B per thread per iteration, grid-stride)
__global__ void k_stream_copy(long n4, const double4 *__restrict__ a, double4 *__restrict__ b) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (long)gridDim.x * blockDim.x) b[i] = a[i];
}

lt_status run_fft(int nx, int ny, int batch, bool forward, const double *in, double *out) {
    LT_TRY(ensure_init());
    if (nx <= 0 || ny <= 0 || batch <= 0) {
        set_error("fft: bad shape");
        return LT_ERR_ARG;
    }
    const int hn = ny / 2 + 1;
    const long nr = (long)nx * ny * batch, nc = 2L * nx * hn * batch;
    rocfft_setup();
    rocfft_plan plan = nullptr;
    rocfft_execution_info info = nullptr;
    const size_t len[2] = {(size_t)ny, (size_t)nx};
    if (rocfft_plan_create(&plan, rocfft_placement_notinplace,
                           forward ? rocfft_transform_type_real_forward : rocfft_transform_type_real_inverse,
                           rocfft_precision_double, 2, len, (size_t)batch, nullptr) != rocfft_status_success) {
        set_error("rocFFT plan creation failed");
        return LT_ERR_FFT;
    }
    size_t ws = 0;
    rocfft_plan_get_work_buffer_size(plan, &ws);
    DevBuf<uint8_t> work;
    DevBuf<double> din, dout;
    lt_status st = work.alloc(ws ? ws : 1);
    if (st == LT_OK) st = din.upload(in, forward ? nr : nc);
    if (st == LT_OK) st = dout.alloc(forward ? nc : nr);
    if (st == LT_OK && rocfft_execution_info_create(&info) != rocfft_status_success) st = LT_ERR_FFT;
    if (st == LT_OK && ws) rocfft_execution_info_set_work_buffer(info, work.p, ws);
    if (st == LT_OK) rocfft_execution_info_set_stream(info, stream());
    if (st == LT_OK && !forward) {
        // c2r: symmetrise columns 0 and ny/2 as numpy's irfft2 does implicitly
        for (int b = 0; b < batch && st == LT_OK; ++b) {
            double *X = din.p + 2L * nx * hn * b;
            hipLaunchKernelGGL(k_herm_col, dim3(blocks_for(nx / 2 + 1, 256)), dim3(256), 0, stream(), nx, hn, 0, X);
            if (ny % 2 == 0)
                hipLaunchKernelGGL(k_herm_col, dim3(blocks_for(nx / 2 + 1, 256)), dim3(256), 0, stream(), nx, hn,
                                   ny / 2, X);
            if (hipGetLastError() != hipSuccess) st = LT_ERR_HIP;
        }
    }
    if (st == LT_OK) {
        void *i[1] = {din.p};
        void *o[1] = {dout.p};
        if (rocfft_execute(plan, i, o, info) != rocfft_status_success) st = LT_ERR_FFT;
    }
    if (st == LT_OK) st = dout.download(out, forward ? nc : nr);
    if (st == LT_OK) st = sync();
    if (info) rocfft_execution_info_destroy(info);
    rocfft_plan_destroy(plan);
    if (st == LT_ERR_FFT) set_error("rocFFT execution failed");
    return st;
}

}  // namespace
}  // namespace lt

using namespace lt;

extern "C" {

lt_status lt_fft_r2c_host(int nx, int ny, int batch, const double *in, double *out) {
    return run_fft(nx, ny, batch, true, in, out);
}

lt_status lt_fft_c2r_host(int nx, int ny, int batch, const double *in, double *out) {
    return run_fft(nx, ny, batch, false, in, out);
}

lt_status lt_stream_copy(long nbytes, int reps, double *ms_per_launch) {
    LT_TRY(ensure_init());
    if (nbytes <= 0 || nbytes % 32 || reps <= 0 || !ms_per_launch) {
        set_error("lt_stream_copy: bad arguments");
        return LT_ERR_ARG;
    }
    const long n4 = nbytes / 32;
    DevBuf<double4> a, b;
    LT_TRY(a.alloc(n4));
    LT_TRY(b.alloc(n4));
    LT_TRY(a.zero());
    const unsigned grid = (unsigned)std::min<long>(blocks_for(n4, 256), 8192);
    hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, stream(), n4, a.p, b.p);
    LT_CHECK_LAUNCH();
    hipEvent_t e0, e1;
    LT_HIP(hipEventCreate(&e0));
    LT_HIP(hipEventCreate(&e1));
    LT_HIP(hipEventRecord(e0, stream()));
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, stream(), n4, a.p, b.p);
    LT_HIP(hipEventRecord(e1, stream()));
    LT_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    LT_HIP(hipEventElapsedTime(&ms, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    LT_CHECK_LAUNCH();
    *ms_per_launch = ms / reps;
    return LT_OK;
}

lt_status lt_csr_spmv_host(long nrows, long ncols, const int *rp, const int *ci, const double *va, const double *x,
                           double *y) {
    return lt_csr_spmv_host_kernel(nrows, ncols, rp, ci, va, x, y, 0);
}

lt_status lt_csr_spmv_host_kernel(long nrows, long ncols, const int *rp, const int *ci, const double *va,
                                  const double *x, double *y, int kernel) {
    LT_TRY(ensure_init());
    const int kl = kernel < 0 ? -kernel : kernel;
    const bool pow2 = kl > 0 && kl <= 64 && (kl & (kl - 1)) == 0;
    if (nrows < 0 || ncols < 0 || !rp || !y || (ncols > 0 && !x) ||
        !(kernel == 0 || pow2 || kernel == kLanesStream || kernel == kLanesSlice)) {
        set_error("lt_csr_spmv_host: bad arguments");
        return LT_ERR_ARG;
    }
    DevCSR A;
    A.nrows = nrows;
    A.ncols = ncols;
    A.nnz = rp[nrows];
    LT_TRY(A.rp.upload(rp, nrows + 1));
    LT_TRY(A.ci.upload(ci, A.nnz));
    LT_TRY(A.va.upload(va, A.nnz));
    // a chosen kernel as in the solver's L and L^T: rows of 65..512 entries
    // one wave each, longer ones chunked; the default (lt_csr_spmv_host):
    // rows above 2048 entries chunked
    if (kernel != 0) A.long_row = 64;
    LT_TRY(csr_prepare(A));
    DevBuf<double> dx, dy;
    LT_TRY(dx.upload(x, ncols));
    LT_TRY(dy.alloc(nrows > 0 ? nrows : 1));
    if (kernel < 0) {  // the product over the nonempty rows, the rest of y zero
        LT_TRY(csr_nonempty_rows(A));
        A.lanes_nz = kl;
        LT_TRY(dy.zero());
        LT_TRY(csr_spmv_nz(A, dx.p, dy.p));
    } else {
        if (kernel == kLanesSlice) LT_TRY(csr_make_slices(A));
        LT_TRY(csr_spmv(A, dx.p, dy.p, 1.0, 0.0, kernel));
    }
    LT_TRY(dy.download(y, nrows));
    return sync();
}

}  // extern "C"
