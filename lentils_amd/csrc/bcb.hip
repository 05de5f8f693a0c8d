// Dirty beam and blurred covariance operator of a radio data set
// (RadioDataset, lentils/common/data_base.py:115-129,148-172):
//
//   beam space  = 2nx x 2ny pixels, bounds +-dx*n (the x pixel size on both
//                 axes, data_base.py:159-160)
//   dirty beam  = N_b^T C 1, N_b the NUFFT of the beam space with the
//                 reference's defaults (pad 2, kernel support 4, Beatty's beta)
//   BCB x       = zpad^T irfft2( rfft2(roll(beam, (nx, ny))) / (4 nx ny)
//                                * rfft2(zpad x) )
//
// BCB is a circular convolution of the zero-padded image with the dirty beam;
// because the padded grid is twice the image, it equals the linear
// convolution over every pair of image pixels, i.e. D^T C D up to the beam's
// own NUFFT error.  The transforms are rocFFT calls; the rest is elementwise.
#include <rocfft/rocfft.h>

#include <algorithm>
#include <memory>

#include "lt_internal.h"
#include "nufft_plan.h"

namespace lt {

fourier_space padded_fourier_space(const image_space &im, int pad);
double default_kb_beta(int wsup, int pad);
__global__ void k_gather_weights(long n, const int32_t *__restrict__ perm, const double *__restrict__ w_in,
                                 double *__restrict__ w_out);

namespace {

__global__ void k_fill_one(long n, double *__restrict__ v) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        v[2 * i] = 1.0;
        v[2 * i + 1] = 0.0;
    }
}

// np.roll(in, (sx, sy), axis=(0, 1)) of an nx x ny array
__global__ void k_roll2(int nx, int ny, int sx, int sy, const double *__restrict__ in, double *__restrict__ out) {
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)nx * ny) return;
    const int i = (int)(g / ny), j = (int)(g % ny);
    out[(long)((i + sx) % nx) * ny + (j + sy) % ny] = in[g];
}

__global__ void k_scale2(long n, double a, double *__restrict__ v) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] *= a;
}

// image b of `batch` (nx x ny) into the centre of its zeroed (nxp x nyp) slot
__global__ void k_zpad(int nx, int ny, int padx, int pady, int nxp, int nyp, const double *__restrict__ img,
                       double *__restrict__ out) {
    const int b = blockIdx.y;
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)nxp * nyp) return;
    const int i = (int)(g / nyp) - padx, j = (int)(g % nyp) - pady;
    double v = 0.0;
    if (i >= 0 && i < nx && j >= 0 && j < ny) v = img[(long)b * nx * ny + (long)i * ny + j];
    out[(long)b * nxp * nyp + g] = v;
}

__global__ void k_crop(int nx, int ny, int padx, int pady, int nxp, int nyp, const double *__restrict__ in,
                       double *__restrict__ img) {
    const int b = blockIdx.y;
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)nx * ny) return;
    const int i = (int)(g / ny), j = (int)(g % ny);
    img[(long)b * nx * ny + g] = in[(long)b * nxp * nyp + (long)(i + padx) * nyp + (j + pady)];
}

// X_b *= mul (complex, elementwise) for every image b
__global__ void k_cmul(long n, const double *__restrict__ mul, double *__restrict__ X) {
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    double *Y = X + 2 * ((long)blockIdx.y * n + g);
    const double mr = mul[2 * g], mi = mul[2 * g + 1], xr = Y[0], xi = Y[1];
    Y[0] = xr * mr - xi * mi;
    Y[1] = xr * mi + xi * mr;
}

// Hermitian part of column `col` along u (numpy's irfft2 reads only that part
// of columns 0 and nv/2)
__global__ void k_herm_col_b(int nu, int half_nv, int col, long stride, double *__restrict__ X) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > nu / 2) return;
    double *Y = X + 2 * (long)blockIdx.y * stride;
    const int m = (nu - k) % nu;
    const long a = 2 * ((long)k * half_nv + col), c = 2 * ((long)m * half_nv + col);
    const double ar = Y[a], ai = Y[a + 1], br = Y[c], bi = Y[c + 1];
    Y[a] = 0.5 * (ar + br);
    Y[a + 1] = 0.5 * (ai - bi);
    if (m != k) {
        Y[c] = 0.5 * (br + ar);
        Y[c + 1] = 0.5 * (bi - ai);
    }
}

struct FftPair {
    rocfft_plan fwd = nullptr, inv = nullptr;
    rocfft_execution_info fi = nullptr, ii = nullptr;
    DevBuf<uint8_t> fw, iw;
    ~FftPair() {
        if (fwd) rocfft_plan_destroy(fwd);
        if (inv) rocfft_plan_destroy(inv);
        if (fi) rocfft_execution_info_destroy(fi);
        if (ii) rocfft_execution_info_destroy(ii);
    }
};

lt_status fft_fail(const char *what) {
    set_error("rocFFT %s failed", what);
    return LT_ERR_FFT;
}

// batched (nxp, nyp) real <-> (nxp, nyp/2+1) complex FP64 plans on the library stream
lt_status make_pair(int nxp, int nyp, int batch, FftPair &P) {
    rocfft_setup();
    const size_t len[2] = {(size_t)nyp, (size_t)nxp};
    if (rocfft_plan_create(&P.fwd, rocfft_placement_notinplace, rocfft_transform_type_real_forward,
                           rocfft_precision_double, 2, len, (size_t)batch, nullptr) != rocfft_status_success)
        return fft_fail("plan (r2c)");
    if (rocfft_plan_create(&P.inv, rocfft_placement_notinplace, rocfft_transform_type_real_inverse,
                           rocfft_precision_double, 2, len, (size_t)batch, nullptr) != rocfft_status_success)
        return fft_fail("plan (c2r)");
    if (rocfft_execution_info_create(&P.fi) != rocfft_status_success ||
        rocfft_execution_info_create(&P.ii) != rocfft_status_success)
        return fft_fail("execution info");
    size_t wf = 0, wi = 0;
    rocfft_plan_get_work_buffer_size(P.fwd, &wf);
    rocfft_plan_get_work_buffer_size(P.inv, &wi);
    if (wf) {
        LT_TRY(P.fw.alloc(wf));
        rocfft_execution_info_set_work_buffer(P.fi, P.fw.p, wf);
    }
    if (wi) {
        LT_TRY(P.iw.alloc(wi));
        rocfft_execution_info_set_work_buffer(P.ii, P.iw.p, wi);
    }
    rocfft_execution_info_set_stream(P.fi, stream());
    rocfft_execution_info_set_stream(P.ii, stream());
    return LT_OK;
}

lt_status check_even(const image_space &im) {
    if (im.nx < 2 || im.ny < 2 || (im.nx & 1) || (im.ny & 1)) {
        set_error("BCB needs even image sizes (the padded grid must be twice the image, as the beam is): %d x %d",
                  im.nx, im.ny);
        return LT_ERR_ARG;
    }
    return LT_OK;
}

}  // namespace

// RadioDataset._make_beam_space, data_base.py:148-160
image_space beam_space_of(const image_space &im) {
    image_space bs{};
    bs.nx = 2 * im.nx;
    bs.ny = 2 * im.ny;
    bs.nchannels = 1;
    bs.nstokes = 1;
    const double rx = im.dx * im.nx, ry = im.dx * im.ny;
    bs.xmin = -rx;
    bs.xmax = rx;
    bs.ymin = -ry;
    bs.ymax = ry;
    bs.dx = (bs.xmax - bs.xmin) / bs.nx;
    bs.dy = (bs.ymax - bs.ymin) / bs.ny;
    return bs;
}

// Dirty beam N_b^T C 1 (data_base.py:163-172) into beam (device, 2nx x 2ny,
// x-major); weights: nchannels*nrows, caller order, host.  wsup <= 0 / beta
// <= 0: the reference's NUFFTOperator defaults (4; Beatty's formula).
lt_status dirty_beam_dev(const visibility_space &vs, const image_space &im, const double *weights, int wsup,
                         double beta, double *beam, double *peak) {
    if (wsup <= 0) wsup = 4;
    if (beta <= 0.0) beta = default_kb_beta(wsup, 2);
    const image_space bs = beam_space_of(im);
    NufftPlan B;
    LT_TRY(B.create(vs, padded_fourier_space(bs, 2), wsup, beta, LT_F64));
    LT_TRY(B.attach_image(bs, 2));
    const long nv = B.nvis, nb = (long)bs.nx * bs.ny;
    if (nv == 0) {
        LT_HIP(hipMemsetAsync(beam, 0, nb * sizeof(double), stream()));
        if (peak) *peak = 0.0;
        return sync();
    }
    DevBuf<double> raw, w;
    LT_TRY(raw.upload(weights, nv));
    LT_TRY(w.alloc(nv));
    hipLaunchKernelGGL(k_gather_weights, dim3(blocks_for(nv, 256)), dim3(256), 0, stream(), nv, B.perm.p, raw.p, w.p);
    hipLaunchKernelGGL(k_fill_one, dim3(blocks_for(nv, 256)), dim3(256), 0, stream(), nv, (double *)B.visbuf.p);
    LT_CHECK_LAUNCH();
    LT_TRY(B.adjoint(B.visbuf.p, w.p, beam, 1.0, false));
    if (peak) {
        std::vector<double> hb(nb);
        LT_HIP(hipMemcpyAsync(hb.data(), beam, nb * sizeof(double), hipMemcpyDeviceToHost, stream()));
        LT_TRY(sync());
        double m = hb[0];
        for (double v : hb) m = std::max(m, v);
        *peak = m;
    }
    return sync();
}

// rfft2(roll(beam, (nxp/2, nyp/2))) / (nxp nyp) (data_base.py:124-126): the
// complex nxp x (nyp/2+1) multiplier of BCB, into spec (device)
lt_status beam_multiplier(int nxp, int nyp, const double *beam, DevBuf<double> &spec) {
    const long nb = (long)nxp * nyp, ns = 2L * nxp * (nyp / 2 + 1);
    DevBuf<double> rolled;
    LT_TRY(rolled.alloc(nb));
    LT_TRY(spec.alloc(ns));
    hipLaunchKernelGGL(k_roll2, dim3(blocks_for(nb, 256)), dim3(256), 0, stream(), nxp, nyp, nxp / 2, nyp / 2, beam,
                       rolled.p);
    LT_CHECK_LAUNCH();
    FftPair F;
    LT_TRY(make_pair(nxp, nyp, 1, F));
    void *in[1] = {rolled.p};
    void *out[1] = {spec.p};
    if (rocfft_execute(F.fwd, in, out, F.fi) != rocfft_status_success) return fft_fail("execute (beam)");
    hipLaunchKernelGGL(k_scale2, dim3(blocks_for(ns, 256)), dim3(256), 0, stream(), ns, 1.0 / (double)nb, spec.p);
    LT_CHECK_LAUNCH();
    return sync();
}

}  // namespace lt

using namespace lt;

extern "C" {

lt_status lt_dirty_beam_host(visibility_space vs, image_space im, const double *weights, int wsup, double kb_beta,
                             double *beam, double *peak) {
    LT_TRY(ensure_init());
    if (!weights || !beam || im.nx <= 0 || im.ny <= 0 || vs.nrows < 0 || vs.nchannels <= 0 || !vs.channels ||
        (vs.nrows > 0 && !vs.uv)) {
        set_error("lt_dirty_beam_host: bad arguments");
        return LT_ERR_ARG;
    }
    const long nb = 4L * im.nx * im.ny;
    DevBuf<double> d;
    LT_TRY(d.alloc(nb));
    LT_TRY(dirty_beam_dev(vs, im, weights, wsup, kb_beta, d.p, peak));
    LT_TRY(d.download(beam, nb));
    return sync();
}

lt_status lt_bcb_apply_host(image_space im, const double *beam, int nimg, const double *img, double *out) {
    LT_TRY(ensure_init());
    if (!beam || !img || !out || nimg <= 0) {
        set_error("lt_bcb_apply_host: bad arguments");
        return LT_ERR_ARG;
    }
    LT_TRY(check_even(im));
    const int nx = im.nx, ny = im.ny, nxp = 2 * nx, nyp = 2 * ny, hv = nyp / 2 + 1;
    const int padx = nx / 2, pady = ny / 2;  // fourier_operators.py:34-35 with pad 2, even sizes
    const long npix = (long)nx * ny, nr = (long)nxp * nyp, nc = (long)nxp * hv;
    DevBuf<double> dbeam, mul, dimg, real, spec;
    LT_TRY(dbeam.upload(beam, nr));
    LT_TRY(beam_multiplier(nxp, nyp, dbeam.p, mul));
    LT_TRY(dimg.upload(img, npix * nimg));
    LT_TRY(real.alloc(nr * nimg));
    LT_TRY(spec.alloc(2 * nc * nimg));
    FftPair F;
    LT_TRY(make_pair(nxp, nyp, nimg, F));
    hipLaunchKernelGGL(k_zpad, dim3(blocks_for(nr, 256), nimg), dim3(256), 0, stream(), nx, ny, padx, pady, nxp, nyp,
                       dimg.p, real.p);
    LT_CHECK_LAUNCH();
    void *r[1] = {real.p};
    void *c[1] = {spec.p};
    if (rocfft_execute(F.fwd, r, c, F.fi) != rocfft_status_success) return fft_fail("execute (r2c)");
    hipLaunchKernelGGL(k_cmul, dim3(blocks_for(nc, 256), nimg), dim3(256), 0, stream(), nc, mul.p, spec.p);
    const int cols[2] = {0, nyp / 2};
    for (int col : cols)
        hipLaunchKernelGGL(k_herm_col_b, dim3(blocks_for(nxp / 2 + 1, 256), nimg), dim3(256), 0, stream(), nxp, hv,
                           col, nc, spec.p);
    LT_CHECK_LAUNCH();
    if (rocfft_execute(F.inv, c, r, F.ii) != rocfft_status_success) return fft_fail("execute (c2r)");
    hipLaunchKernelGGL(k_crop, dim3(blocks_for(npix, 256), nimg), dim3(256), 0, stream(), nx, ny, padx, pady, nxp,
                       nyp, real.p, dimg.p);
    LT_CHECK_LAUNCH();
    LT_TRY(dimg.download(out, npix * nimg));
    return sync();
}

}  // extern "C"
