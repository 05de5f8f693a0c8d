// Small CSR builders: curvature and gradient finite-difference regularizers
// (lentils/backend/nufft.c:163-229) and the dense DFT matrix
// (nufft.c:128-161).  Count pass -> scan -> fill pass, all on the GPU.
#include <cmath>

#include "lt_internal.h"

namespace lt {

namespace {

// entries of curvature row (i, j): up, left, centre, right, down
__device__ inline int curv_count(int i, int j, int nx, int ny) {
    return 1 + (i > 0) + (j > 0) + (j < ny - 1) + (i < nx - 1);
}

__global__ void k_curv_count(int nx, int ny, int64_t *__restrict__ cnt) {
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= (long)nx * ny) return;
    int i = (int)(r / ny), j = (int)(r % ny);
    cnt[r] = curv_count(i, j, nx, ny);
}

__global__ void k_curv_fill(int nx, int ny, double h, int32_t coloff, const int64_t *__restrict__ off,
                            int32_t *__restrict__ cols, double *__restrict__ vals) {
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= (long)nx * ny) return;
    int i = (int)(r / ny), j = (int)(r % ny);
    long k = off[r];
    if (i > 0) { vals[k] = h; cols[k++] = coloff + ny * (i - 1) + j; }
    if (j > 0) { vals[k] = h; cols[k++] = coloff + ny * i + (j - 1); }
    vals[k] = -4.0 * h;
    cols[k++] = coloff + ny * i + j;
    if (j < ny - 1) { vals[k] = h; cols[k++] = coloff + ny * i + (j + 1); }
    if (i < nx - 1) { vals[k] = h; cols[k++] = coloff + ny * (i + 1) + j; }
}

__global__ void k_grad_count(int nx, int ny, int64_t *__restrict__ cnt) {
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long n = (long)nx * ny;
    if (r >= 2 * n) return;
    long p = r < n ? r : r - n;
    int i = (int)(p / ny), j = (int)(p % ny);
    cnt[r] = (r < n) ? (i < nx - 1 ? 2 : 0) : (j < ny - 1 ? 2 : 0);
}

__global__ void k_grad_fill(int nx, int ny, double ix, double iy, const int64_t *__restrict__ off,
                            int32_t *__restrict__ cols, double *__restrict__ vals) {
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long n = (long)nx * ny;
    if (r >= 2 * n) return;
    long p = r < n ? r : r - n;
    int i = (int)(p / ny), j = (int)(p % ny);
    long k = off[r];
    if (r < n) {
        if (i < nx - 1) {
            vals[k] = -ix; cols[k] = ny * i + j;
            vals[k + 1] = ix; cols[k + 1] = ny * (i + 1) + j;
        }
    } else if (j < ny - 1) {
        vals[k] = -iy; cols[k] = ny * i + j;
        vals[k + 1] = iy; cols[k + 1] = ny * i + (j + 1);
    }
}

__global__ void k_i64_i32(long n, const int64_t *__restrict__ a, int32_t *__restrict__ b) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (int32_t)a[i];
}

// DFT row 2v (cos) / 2v+1 (sin) of -2 pi ARCSEC (u x + v y) over the masked
// pixels; mpix lists the masked pixel indices in x-major order
__global__ void k_dft_fill(long nrows2, const double *__restrict__ uv, double ch, int ny, double xmin, double dx,
                           double ymin, double dy, const int32_t *__restrict__ mpix, int nm,
                           int32_t *__restrict__ cols, double *__restrict__ vals) {
    long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nrows2 * nm) return;
    long row = g / nm;
    int c = (int)(g - row * nm);
    long vi = row / 2;
    int odd = (int)(row % 2);
    double u = uv[3 * vi] * ch;
    double v = -uv[3 * vi + 1] * ch;
    int col = mpix[c];
    int i = col / ny, j = col % ny;
    double x = i * dx + xmin;
    double y = j * dy + ymin;
    double arg = -2.0 * M_PI * 4.8481368111e-6 * (u * x + v * y);
    vals[g] = odd ? sin(arg) : cos(arg);
    cols[g] = col;
}

}  // namespace

lt_status build_curvature(int nx, int ny, double dx, double dy, int32_t coloff, DevCSR &C) {
    const long n = (long)nx * ny;
    DevBuf<int64_t> off;
    LT_TRY(off.alloc(n + 1));
    hipLaunchKernelGGL(k_curv_count, dim3(blocks_for(n, 256)), dim3(256), 0, stream(), nx, ny, off.p);
    LT_CHECK_LAUNCH();
    LT_TRY(exclusive_scan_i64(off.p, off.p, n));
    int64_t nnz = 0;
    LT_TRY(off.download(&nnz, 1, n));
    LT_TRY(sync());
    C.nrows = n;
    C.ncols = n;
    C.nnz = nnz;
    LT_TRY(C.rp.alloc(n + 1));
    LT_TRY(C.ci.alloc(nnz));
    LT_TRY(C.va.alloc(nnz));
    const double h = 1.0 / (dx * dy);
    hipLaunchKernelGGL(k_curv_fill, dim3(blocks_for(n, 256)), dim3(256), 0, stream(), nx, ny, h, coloff, off.p,
                       C.ci.p, C.va.p);
    LT_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_i64_i32, dim3(blocks_for(n + 1, 256)), dim3(256), 0, stream(), n + 1, off.p, C.rp.p);
    LT_CHECK_LAUNCH();
    return sync();
}

lt_status build_gradient(int nx, int ny, double dx, double dy, DevCSR &G) {
    const long n = 2L * nx * ny;
    DevBuf<int64_t> off;
    LT_TRY(off.alloc(n + 1));
    hipLaunchKernelGGL(k_grad_count, dim3(blocks_for(n, 256)), dim3(256), 0, stream(), nx, ny, off.p);
    LT_CHECK_LAUNCH();
    LT_TRY(exclusive_scan_i64(off.p, off.p, n));
    int64_t nnz = 0;
    LT_TRY(off.download(&nnz, 1, n));
    LT_TRY(sync());
    G.nrows = n;
    G.ncols = (long)nx * ny;
    G.nnz = nnz;
    LT_TRY(G.rp.alloc(n + 1));
    LT_TRY(G.ci.alloc(nnz > 0 ? nnz : 1));
    LT_TRY(G.va.alloc(nnz > 0 ? nnz : 1));
    hipLaunchKernelGGL(k_grad_fill, dim3(blocks_for(n, 256)), dim3(256), 0, stream(), nx, ny, 1.0 / dx, 1.0 / dy,
                       off.p, G.ci.p, G.va.p);
    LT_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_i64_i32, dim3(blocks_for(n + 1, 256)), dim3(256), 0, stream(), n + 1, off.p, G.rp.p);
    LT_CHECK_LAUNCH();
    return sync();
}

lt_status build_dft(const visibility_space &vs, const image_space &im, int *row_inds, int *cols, double *vals) {
    std::vector<int32_t> mpix;
    const long npix = (long)im.nx * im.ny;
    for (long p = 0; p < npix; ++p)
        if (!im.mask || im.mask[p]) mpix.push_back((int32_t)p);
    const int nm = (int)mpix.size();
    const long nr2 = 2L * vs.nrows;
    const long nnz = nr2 * nm;
    if (nnz >= INT32_MAX) {
        set_error("dft matrix: %ld entries exceed the int32 range", nnz);
        return LT_ERR_CAPACITY;
    }
    for (long r = 0; r <= nr2; ++r) row_inds[r] = (int)(r * nm);
    if (nnz == 0) return LT_OK;
    DevBuf<double> duv;
    DevBuf<int32_t> dm, dc;
    DevBuf<double> dv;
    LT_TRY(duv.upload(vs.uv, 3L * vs.nrows));
    LT_TRY(dm.upload(mpix.data(), nm));
    LT_TRY(dc.alloc(nnz));
    LT_TRY(dv.alloc(nnz));
    hipLaunchKernelGGL(k_dft_fill, dim3(blocks_for(nnz, 256)), dim3(256), 0, stream(), nr2, duv.p, vs.channels[0],
                       im.ny, im.xmin, im.dx, im.ymin, im.dy, dm.p, nm, dc.p, dv.p);
    LT_CHECK_LAUNCH();
    LT_TRY(dc.download(cols, nnz));
    LT_TRY(dv.download(vals, nnz));
    return sync();
}

}  // namespace lt
