// Batched conjugate-gradient source inversion on the GPU.
//
// For every lens sample s (all advanced together, one launch per step):
//   A_s x = L_s^T D^T C D L_s x + lambda_c Cc_s^T Cc_s x + lambda_t x
//   b_s   = L_s^T D^T C d
// as the reference's tests assemble it (lentils/tests/test_all.py:57-72 with
// PriorCovarianceOperator, operator_base.py:200-251) and solve it with
// scipy's cg.  D^T C D is applied either explicitly (fused degrid-weight-grid
// per uv-tile chunk, LT_OP_NUFFT) or through the dirty-beam convolution
// (LT_OP_BCB, data_base.py:115-129).  Inner products are FP64 per sample,
// reduced in a fixed order inside one workgroup per sample; the solve stops
// when the FP64 true residual ||b - A x|| / ||b|| <= rtol.
#include <algorithm>
#include <cmath>
#include <memory>

#include "nufft_plan.h"
#include "raster_batch.h"

namespace lt {

lt_status build_curvature(int nx, int ny, double dx, double dy, int32_t coloff, DevCSR &C);
LensBatch *lensbatch_of(lt_lensbatch h);
NufftPlan *nufft_of(lt_nufft h);
lt_status grid_c128_to_plan(NufftPlan &P, const double *src, void *dst);

namespace {

constexpr int kDotThreads = 256;

// out[s] = sum over segment s of a*b (a fixed-order block reduction)
__global__ void __launch_bounds__(kDotThreads) k_seg_dot(const int64_t *__restrict__ off, const double *__restrict__ a,
                                                          const double *__restrict__ b, double *__restrict__ out) {
    __shared__ double sh[kDotThreads / 64];
    const int s = blockIdx.x;
    const long b0 = off[s], b1 = off[s + 1];
    double acc = 0.0;
    for (long i = b0 + threadIdx.x; i < b1; i += kDotThreads) acc += a[i] * b[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < kDotThreads / 64; ++k) t += sh[k];
        out[s] = t;
    }
}

// segment id of every element
__global__ void k_seg_ids(const int64_t *__restrict__ off, int ns, int32_t *__restrict__ seg) {
    const int s = blockIdx.y;
    if (s >= ns) return;
    for (long i = off[s] + (long)blockIdx.x * blockDim.x + threadIdx.x; i < off[s + 1];
         i += (long)gridDim.x * blockDim.x)
        seg[i] = s;
}

// y = a*x + b*y (elementwise, whole vector)
__global__ void k_axpby(long n, double a, const double *__restrict__ x, double b, double *__restrict__ y) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = a * x[i] + b * y[i];
}

__global__ void k_copy(long n, const double *__restrict__ x, double *__restrict__ y) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = x[i];
}

__global__ void k_mul(long n, const double *__restrict__ m, double *__restrict__ x) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] *= m[i];
}

// z = M^{-1} r (Jacobi) or z = r
__global__ void k_precond(long n, const double *__restrict__ dinv, const double *__restrict__ r,
                          double *__restrict__ z) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) z[i] = dinv ? dinv[i] * r[i] : r[i];
}

// x += alpha_s p; r -= alpha_s Ap  on active samples
__global__ void k_cg_xr(long n, const int32_t *__restrict__ seg, const int *__restrict__ act,
                        const double *__restrict__ alpha, const double *__restrict__ p,
                        const double *__restrict__ Ap, double *__restrict__ x, double *__restrict__ r) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = seg[i];
    if (!act[s]) return;
    const double a = alpha[s];
    x[i] += a * p[i];
    r[i] -= a * Ap[i];
}

// p = z + beta_s p on active samples
__global__ void k_cg_p(long n, const int32_t *__restrict__ seg, const int *__restrict__ act,
                       const double *__restrict__ beta, const double *__restrict__ z, double *__restrict__ p) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = seg[i];
    if (!act[s]) return;
    p[i] = z[i] + beta[s] * p[i];
}

// per-sample scalars: alpha = rz / pAp
__global__ void k_cg_alpha(int ns, const int *__restrict__ act, const double *__restrict__ rz,
                           const double *__restrict__ pAp, double *__restrict__ alpha) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    alpha[s] = (act[s] && pAp[s] != 0.0) ? rz[s] / pAp[s] : 0.0;
}

// beta = rz_new / rz; convergence test on the recurrence residual
__global__ void k_cg_beta(int ns, int *__restrict__ act, double *__restrict__ rz, const double *__restrict__ rz_new,
                          const double *__restrict__ rr, const double *__restrict__ bb, double rtol2,
                          double *__restrict__ beta, int *__restrict__ iters) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    if (!act[s]) return;
    iters[s] += 1;
    beta[s] = rz[s] != 0.0 ? rz_new[s] / rz[s] : 0.0;
    rz[s] = rz_new[s];
    if (!(rr[s] > rtol2 * bb[s])) act[s] = 0;  // also stops on NaN
}

// Jacobi diagonal of L^T B L with the data operator B approximated by its
// (translation-invariant) response beta to a unit impulse at the image
// centre: diag_j = sum_{a,b in column j of L} L_a L_b beta(p_a - p_b).
// One wavefront per source pixel (row of L^T).
__global__ void __launch_bounds__(256) k_jacobi_diag(long nrow, const int32_t *__restrict__ rp,
                                                      const int32_t *__restrict__ ci, const double *__restrict__ va,
                                                      const int32_t *__restrict__ seg, long npix, int nx, int ny,
                                                      const double *__restrict__ beta, double *__restrict__ diag) {
    const long row = ((long)blockIdx.x * blockDim.x + threadIdx.x) / 64;
    const int lane = threadIdx.x & 63;
    if (row >= nrow) return;
    const long base = (long)seg[row] * npix;
    const int b0 = rp[row], b1 = rp[row + 1];
    const int ic = nx / 2, jc = ny / 2;
    double acc = 0.0;
    for (int a = b0 + lane; a < b1; a += 64) {
        const long pa = ci[a] - base;
        const int ia = (int)(pa / ny), ja = (int)(pa % ny);
        const double va_ = va[a];
        for (int b = b0; b < b1; ++b) {
            const long pb = ci[b] - base;
            const int di = (int)(pb / ny) - ia + ic, dj = (int)(pb % ny) - ja + jc;
            if (di >= 0 && di < nx && dj >= 0 && dj < ny) acc += va_ * va[b] * beta[(long)di * ny + dj];
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (lane == 0) diag[row] = acc;
}

// sum of squares of each CSR row (diag of C^T C for a symmetric C)
__global__ void k_row_sumsq(long nrow, const int32_t *__restrict__ rp, const double *__restrict__ va,
                            double *__restrict__ out) {
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrow) return;
    double s = 0.0;
    for (int k = rp[r]; k < rp[r + 1]; ++k) s += va[k] * va[k];
    out[r] = s;
}

__global__ void k_jacobi_inv(long n, const double *__restrict__ dd, const double *__restrict__ dc, double lc, double lt,
                             double *__restrict__ dinv) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double d = dd[i] + lc * dc[i] + lt;
    dinv[i] = (d > 0.0 && isfinite(d)) ? 1.0 / d : 1.0;
}

__global__ void k_delta(long n, long at, double *__restrict__ v) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = (i == at) ? 1.0 : 0.0;
}

}  // namespace

struct Solver {
    LensBatch *lb = nullptr;
    NufftPlan *nf = nullptr;
    int S = 0;
    long npix = 0, ntot = 0;
    int mode = LT_OP_NUFFT;
    double lam_c = 0, lam_t = 0;
    std::vector<int64_t> soff_host;
    DevBuf<int64_t> soff;
    DevBuf<int32_t> seg;
    DevCSR C;                 // block-diagonal curvature of every source grid
    DevBuf<double> w;         // visibility weights in plan order
    DevBuf<double> dirty;     // D^T C d_s (ndirty x nx*ny)
    int ndirty = 1;           // 1: one data set shared by every sample; S: one per sample
    DevBuf<double> beamspec;  // BCB multiplier, nu x half_nv complex
    DevBuf<double> b, x, r, z, p, Ap, img1, img2, tmp, dinv;
    DevBuf<double> rz, rz_new, pAp, rr, bb, alpha, beta;
    DevBuf<int> act, iters, all_active;
    double beam_peak = 0.0;
    NormalStencil stencil;
    bool use_stencil = false;

    lt_status create(LensBatch *lb_, NufftPlan *nf_, const double *weights, const double *data,
                     const double *dirty_in, int ndirty_in, double lc, double lt, int op_mode);
    lt_status apply(const double *xin, double *yout, const int *active_list, int nactive);
    lt_status solve(double rtol, int maxiter, int precond, double *xout, int *it_out, double *res_out);
    lt_status jacobi_setup();
};

fourier_space padded_fourier_space(const image_space &im, int pad);

namespace {

lt_status beam_spectrum(Solver &S, const double *weights_caller_order);

}  // namespace

lt_status Solver::create(LensBatch *lb_, NufftPlan *nf_, const double *weights, const double *data,
                         const double *dirty_in, int ndirty_in, double lc, double lt_, int op_mode) {
    lb = lb_;
    nf = nf_;
    S = lb->nsamples;
    npix = lb->npix;
    mode = op_mode;
    lam_c = lc;
    lam_t = lt_;
    if (!nf->has_image || nf->nx != lb->im.nx || nf->ny != lb->im.ny) {
        set_error("solver: NUFFT image space does not match the lens image space");
        return LT_ERR_ARG;
    }
    if (dirty_in && ndirty_in != 1 && ndirty_in != S) {
        set_error("solver: %d dirty images for %d samples (1 or one per sample)", ndirty_in, S);
        return LT_ERR_ARG;
    }
    ndirty = dirty_in ? ndirty_in : 1;
    if (op_mode != LT_OP_NUFFT && op_mode != LT_OP_BCB && op_mode != LT_OP_NUFFT_EXPLICIT) {
        set_error("solver: unknown operator mode %d", op_mode);
        return LT_ERR_ARG;
    }
    soff_host = lb->srcoff_host;
    ntot = soff_host[S];
    LT_TRY(soff.upload(soff_host.data(), S + 1));
    LT_TRY(seg.alloc(ntot));
    hipLaunchKernelGGL(k_seg_ids, dim3(64, S), dim3(256), 0, stream(), soff.p, S, seg.p);
    LT_CHECK_LAUNCH();
    LT_TRY(lb->ensure_transpose());
    // block-diagonal curvature matrix
    {
        std::vector<DevCSR> parts(S);
        long nnz = 0;
        for (int s = 0; s < S; ++s) {
            const SrcGrid &g = lb->src_host[s];
            LT_TRY(build_curvature(g.nx, g.ny, g.dx, g.dy, (int32_t)soff_host[s], parts[s]));
            nnz += parts[s].nnz;
        }
        C.nrows = C.ncols = ntot;
        C.nnz = nnz;
        LT_TRY(C.rp.alloc(ntot + 1));
        LT_TRY(C.ci.alloc(nnz));
        LT_TRY(C.va.alloc(nnz));
        long ro = 0, no = 0;
        for (int s = 0; s < S; ++s) {
            std::vector<int32_t> rp(parts[s].nrows + 1);
            LT_TRY(parts[s].rp.download(rp.data(), rp.size()));
            LT_TRY(sync());
            for (auto &v : rp) v += (int32_t)no;
            LT_HIP(hipMemcpy(C.rp.p + ro, rp.data(), rp.size() * sizeof(int32_t), hipMemcpyHostToDevice));
            LT_HIP(hipMemcpyAsync(C.ci.p + no, parts[s].ci.p, parts[s].nnz * sizeof(int32_t),
                                  hipMemcpyDeviceToDevice, stream()));
            LT_HIP(hipMemcpyAsync(C.va.p + no, parts[s].va.p, parts[s].nnz * sizeof(double), hipMemcpyDeviceToDevice,
                                  stream()));
            LT_TRY(sync());
            ro += parts[s].nrows;
            no += parts[s].nnz;
        }
        LT_TRY(csr_prepare(C));
    }
    // weights and data into plan order; dirty image D^T C d
    const long nv = nf->nvis;
    {
        DevBuf<double> dc;
        LT_TRY(w.alloc(std::max<long>(nv, 1)));
        if (nv) {
            std::vector<int32_t> hp(nv);
            LT_TRY(nf->perm.download(hp.data(), nv));
            LT_TRY(sync());
            std::vector<double> hs(nv);
            for (long q = 0; q < nv; ++q) hs[q] = weights[hp[q]];
            LT_TRY(w.upload(hs.data(), nv));
        }
        LT_TRY(dirty.alloc(npix * ndirty));
        if (dirty_in) {
            LT_TRY(dirty.upload(dirty_in, npix * ndirty));
        } else if (nv) {
            LT_TRY(dc.upload(data, 2 * nv));
            LT_TRY(nf->to_sorted(dc.p, nf->visbuf.p));
            LT_TRY(nf->adjoint(nf->visbuf.p, w.p, dirty.p, 1.0, false));
        } else {
            LT_TRY(dirty.zero());
        }
    }
    LT_TRY(nf->ensure_batch(S));
    if (mode == LT_OP_BCB) LT_TRY(beam_spectrum(*this, weights));
    if (mode == LT_OP_NUFFT && nf->prec == LT_F64 && nf->W <= 4 && nv > 0) {
        LT_TRY(stencil.build(*nf, w.p));
        use_stencil = true;
    }
    // work vectors
    LT_TRY(b.alloc(ntot));
    LT_TRY(x.alloc(ntot));
    LT_TRY(r.alloc(ntot));
    LT_TRY(z.alloc(ntot));
    LT_TRY(p.alloc(ntot));
    LT_TRY(Ap.alloc(ntot));
    LT_TRY(tmp.alloc(ntot));
    LT_TRY(img1.alloc(npix * S));
    LT_TRY(img2.alloc(npix * S));
    for (auto *v : {&rz, &rz_new, &pAp, &rr, &bb, &alpha, &beta}) LT_TRY(v->alloc(S));
    LT_TRY(act.alloc(S));
    LT_TRY(iters.alloc(S));
    LT_TRY(all_active.alloc(S));
    std::vector<int> ids(S);
    for (int s = 0; s < S; ++s) ids[s] = s;
    LT_TRY(all_active.upload(ids.data(), S));
    // rhs b_s = L_s^T dirty_s (a shared dirty image is replicated)
    for (int s = 0; s < S; ++s)
        LT_HIP(hipMemcpyAsync(img2.p + (long)s * npix, dirty.p + (ndirty == 1 ? 0L : (long)s * npix),
                              npix * sizeof(double), hipMemcpyDeviceToDevice, stream()));
    LT_TRY(csr_spmv(lb->LT, img2.p, b.p, 1.0, 0.0, 0));
    return sync();
}

lt_status Solver::apply(const double *xin, double *yout, const int *alist, int nact) {
    LT_TRY(csr_spmv(lb->L, xin, img1.p, 1.0, 0.0, 0));
    if (mode != LT_OP_BCB)
        LT_TRY(nf->normal_batched(img1.p, img2.p, w.p, alist, nact, 1.0, false, use_stencil ? &stencil : nullptr));
    else
        LT_TRY(nf->convolve_batched(img1.p, img2.p, beamspec.p, alist, nact, 1.0, false));
    LT_TRY(csr_spmv(lb->LT, img2.p, yout, 1.0, 0.0, 0));
    if (lam_c != 0.0) {
        LT_TRY(csr_spmv(C, xin, tmp.p, 1.0, 0.0, 0));
        LT_TRY(csr_spmv(C, tmp.p, yout, lam_c, 1.0, 0));
    }
    if (lam_t != 0.0) {
        hipLaunchKernelGGL(k_axpby, dim3(blocks_for(ntot, 256)), dim3(256), 0, stream(), ntot, lam_t, xin, 1.0, yout);
        LT_CHECK_LAUNCH();
    }
    return LT_OK;
}

lt_status Solver::jacobi_setup() {
    // impulse response of the data operator at the image centre (sample 0)
    const long c = (long)(lb->im.nx / 2) * lb->im.ny + lb->im.ny / 2;
    hipLaunchKernelGGL(k_delta, dim3(blocks_for(npix * S, 256)), dim3(256), 0, stream(), npix * S, c, img1.p);
    LT_CHECK_LAUNCH();
    if (mode != LT_OP_BCB)
        LT_TRY(nf->normal_batched(img1.p, img2.p, w.p, all_active.p, 1, 1.0, false, use_stencil ? &stencil : nullptr));
    else
        LT_TRY(nf->convolve_batched(img1.p, img2.p, beamspec.p, all_active.p, 1, 1.0, false));
    DevBuf<double> dd, dc;
    LT_TRY(dd.alloc(ntot));
    LT_TRY(dc.alloc(ntot));
    LT_TRY(dinv.alloc(ntot));
    const DevCSR &T = lb->LT;
    hipLaunchKernelGGL(k_jacobi_diag, dim3(blocks_for(T.nrows * 64, 256)), dim3(256), 0, stream(), T.nrows, T.rp.p,
                       T.ci.p, T.va.p, seg.p, npix, lb->im.nx, lb->im.ny, img2.p, dd.p);
    LT_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_row_sumsq, dim3(blocks_for(ntot, 256)), dim3(256), 0, stream(), ntot, C.rp.p, C.va.p, dc.p);
    LT_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_jacobi_inv, dim3(blocks_for(ntot, 256)), dim3(256), 0, stream(), ntot, dd.p, dc.p, lam_c,
                       lam_t, dinv.p);
    LT_CHECK_LAUNCH();
    return sync();
}

lt_status Solver::solve(double rtol, int maxiter, int precond, double *xout, int *it_out, double *res_out) {
    const unsigned gb = blocks_for(ntot, 256), gs = blocks_for(S, 64);
    // Jacobi: diag(A) ~ beam peak * diag(L^T L) + lambda_c diag(Cc^T Cc) + lambda_t
    const double *dinv_p = nullptr;
    if (precond == LT_PC_JACOBI) {
        if (!dinv.p) LT_TRY(jacobi_setup());
        dinv_p = dinv.p;
    }
    std::vector<int> ones(S, 1), zeros(S, 0);
    LT_TRY(act.upload(ones.data(), S));
    LT_TRY(iters.upload(zeros.data(), S));
    LT_TRY(x.zero());
    hipLaunchKernelGGL(k_copy, dim3(gb), dim3(256), 0, stream(), ntot, b.p, r.p);
    hipLaunchKernelGGL(k_precond, dim3(gb), dim3(256), 0, stream(), ntot, dinv_p, r.p, z.p);
    hipLaunchKernelGGL(k_copy, dim3(gb), dim3(256), 0, stream(), ntot, z.p, p.p);
    LT_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_seg_dot, dim3(S), dim3(kDotThreads), 0, stream(), soff.p, b.p, b.p, bb.p);
    hipLaunchKernelGGL(k_seg_dot, dim3(S), dim3(kDotThreads), 0, stream(), soff.p, r.p, z.p, rz.p);
    LT_CHECK_LAUNCH();
    const double rtol2 = rtol * rtol;
    std::vector<int> hact(S, 1), alist(S);
    int nact = S;
    for (int s = 0; s < S; ++s) alist[s] = s;
    DevBuf<int> dlist;
    LT_TRY(dlist.upload(alist.data(), S));
    const int check_every = 5;
    int outer = 0;
    for (int it = 0; it < maxiter && nact > 0; ++it) {
        LT_TRY(apply(p.p, Ap.p, dlist.p, nact));
        hipLaunchKernelGGL(k_seg_dot, dim3(S), dim3(kDotThreads), 0, stream(), soff.p, p.p, Ap.p, pAp.p);
        hipLaunchKernelGGL(k_cg_alpha, dim3(gs), dim3(64), 0, stream(), S, act.p, rz.p, pAp.p, alpha.p);
        hipLaunchKernelGGL(k_cg_xr, dim3(gb), dim3(256), 0, stream(), ntot, seg.p, act.p, alpha.p, p.p, Ap.p, x.p, r.p);
        hipLaunchKernelGGL(k_precond, dim3(gb), dim3(256), 0, stream(), ntot, dinv_p, r.p, z.p);
        hipLaunchKernelGGL(k_seg_dot, dim3(S), dim3(kDotThreads), 0, stream(), soff.p, r.p, z.p, rz_new.p);
        hipLaunchKernelGGL(k_seg_dot, dim3(S), dim3(kDotThreads), 0, stream(), soff.p, r.p, r.p, rr.p);
        hipLaunchKernelGGL(k_cg_beta, dim3(gs), dim3(64), 0, stream(), S, act.p, rz.p, rz_new.p, rr.p, bb.p, rtol2,
                           beta.p, iters.p);
        hipLaunchKernelGGL(k_cg_p, dim3(gb), dim3(256), 0, stream(), ntot, seg.p, act.p, beta.p, z.p, p.p);
        LT_CHECK_LAUNCH();
        ++outer;
        if (outer % check_every == 0 || it == maxiter - 1) {
            LT_TRY(act.download(hact.data(), S));
            LT_TRY(sync());
            nact = 0;
            for (int s = 0; s < S; ++s)
                if (hact[s]) alist[nact++] = s;
            if (nact > 0) LT_TRY(dlist.upload(alist.data(), nact));
        }
    }
    // FP64 true residual ||b - A x|| / ||b||
    LT_TRY(apply(x.p, Ap.p, all_active.p, S));
    hipLaunchKernelGGL(k_copy, dim3(gb), dim3(256), 0, stream(), ntot, b.p, r.p);
    hipLaunchKernelGGL(k_axpby, dim3(gb), dim3(256), 0, stream(), ntot, -1.0, Ap.p, 1.0, r.p);
    hipLaunchKernelGGL(k_seg_dot, dim3(S), dim3(kDotThreads), 0, stream(), soff.p, r.p, r.p, rr.p);
    LT_CHECK_LAUNCH();
    std::vector<double> hrr(S), hbb(S);
    std::vector<int> hit(S);
    LT_TRY(rr.download(hrr.data(), S));
    LT_TRY(bb.download(hbb.data(), S));
    LT_TRY(iters.download(hit.data(), S));
    if (xout) LT_TRY(x.download(xout, ntot));
    LT_TRY(sync());
    for (int s = 0; s < S; ++s) {
        if (it_out) it_out[s] = hit[s];
        if (res_out) res_out[s] = hbb[s] > 0 ? std::sqrt(hrr[s] / hbb[s]) : std::sqrt(hrr[s]);
    }
    return LT_OK;
}

namespace {

__global__ void k_roll(int nx, int ny, int sx, int sy, const double *__restrict__ in, double *__restrict__ out) {
    long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)nx * ny) return;
    const int i = (int)(g / ny), j = (int)(g % ny);
    const int ti = (i + sx) % nx, tj = (j + sy) % ny;
    out[(long)ti * ny + tj] = in[g];
}

__global__ void k_fill_c128(long n, double re, double im, double *__restrict__ v) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        v[2 * i] = re;
        v[2 * i + 1] = im;
    }
}

__global__ void k_scale(long n, double a, double *__restrict__ v) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] *= a;
}

// Dirty beam D_b^T C 1 on the (2nx, 2ny) beam space (data_base.py:149-172),
// rolled by half the padded grid and transformed:
// beamspec = rfft2(roll(beam)) / (nxp nyp)   (data_base.py:115-129).
// The beam plan uses the same kernel support and beta as the main plan.
lt_status beam_spectrum(Solver &S, const double *weights) {
    NufftPlan &M = *S.nf;
    if (M.prec != LT_F64) {
        set_error("BCB mode needs an FP64 NUFFT plan");
        return LT_ERR_UNSUPPORTED;
    }
    if (M.nxp != 2 * M.nx || M.nyp != 2 * M.ny) {
        set_error("BCB mode needs the padded grid to be exactly twice the image (pad 2, even sizes)");
        return LT_ERR_UNSUPPORTED;
    }
    const image_space &im = S.lb->im;
    image_space bs{};
    bs.nx = 2 * im.nx;
    bs.ny = 2 * im.ny;
    const double rx = im.dx * im.nx, ry = im.dx * im.ny;  // the reference uses dx for both axes
    bs.xmin = -rx;
    bs.xmax = rx;
    bs.ymin = -ry;
    bs.ymax = ry;
    bs.dx = (bs.xmax - bs.xmin) / bs.nx;
    bs.dy = (bs.ymax - bs.ymin) / bs.ny;
    visibility_space vs{};
    vs.nrows = M.nrows;
    vs.nchannels = M.nch;
    vs.nstokes = 1;
    vs.uv = M.uvw_host.data();
    vs.channels = M.chan_host.data();
    NufftPlan B;
    LT_TRY(B.create(vs, padded_fourier_space(bs, 2), M.W, M.beta, LT_F64));
    LT_TRY(B.attach_image(bs, 2));
    const long nv = B.nvis;
    DevBuf<double> wB, beam, rolled;
    {
        std::vector<int32_t> hp(std::max<long>(nv, 1));
        if (nv) LT_TRY(B.perm.download(hp.data(), nv));
        LT_TRY(sync());
        std::vector<double> hs(std::max<long>(nv, 1));
        for (long q = 0; q < nv; ++q) hs[q] = weights[hp[q]];
        LT_TRY(wB.upload(hs.data(), std::max<long>(nv, 1)));
    }
    const long nb = (long)bs.nx * bs.ny;
    LT_TRY(beam.alloc(nb));
    LT_TRY(rolled.alloc(nb));
    if (nv) {
        hipLaunchKernelGGL(k_fill_c128, dim3(blocks_for(nv, 256)), dim3(256), 0, stream(), nv, 1.0, 0.0,
                           (double *)B.visbuf.p);
        LT_CHECK_LAUNCH();
        LT_TRY(B.adjoint(B.visbuf.p, wB.p, beam.p, 1.0, false));
    } else {
        LT_TRY(beam.zero());
    }
    hipLaunchKernelGGL(k_roll, dim3(blocks_for(nb, 256)), dim3(256), 0, stream(), bs.nx, bs.ny, M.nxp / 2, M.nyp / 2,
                       beam.p, rolled.p);
    LT_CHECK_LAUNCH();
    // rfft2 on the main plan's transform (same nxp x nyp shape)
    LT_HIP(hipMemcpyAsync(M.realbuf.p, rolled.p, nb * sizeof(double), hipMemcpyDeviceToDevice, stream()));
    void *in[1] = {M.realbuf.p};
    void *out[1] = {M.spec.p};
    if (rocfft_execute(M.fwd, in, out, M.fwd_info) != rocfft_status_success) {
        set_error("rocFFT execute (beam) failed");
        return LT_ERR_FFT;
    }
    const long ns = 2L * M.nu * M.half_nv;
    LT_TRY(S.beamspec.alloc(ns));
    LT_HIP(hipMemcpyAsync(S.beamspec.p, M.spec.p, ns * sizeof(double), hipMemcpyDeviceToDevice, stream()));
    hipLaunchKernelGGL(k_scale, dim3(blocks_for(ns, 256)), dim3(256), 0, stream(), ns, 1.0 / (double)nb,
                       S.beamspec.p);
    LT_CHECK_LAUNCH();
    S.beam_peak = 0.0;
    {
        std::vector<double> hb(nb);
        LT_TRY(beam.download(hb.data(), nb));
        LT_TRY(sync());
        for (double v : hb) S.beam_peak = std::max(S.beam_peak, v);
    }
    return sync();
}

}  // namespace

}  // namespace lt

using namespace lt;

struct lt_solver_s {
    Solver s;
};

extern "C" {

lt_status lt_solver_create(lt_lensbatch lbh, lt_nufft nfh, const double *weights, const double *data, double lambda_curv,
                           double lambda_tikh, int op_mode, lt_solver *out) {
    LensBatch *lb = lensbatch_of(lbh);
    NufftPlan *nf = nufft_of(nfh);
    if (!lb || !nf || !out || (nf->nvis > 0 && (!weights || !data))) {
        set_error("lt_solver_create: bad arguments");
        return LT_ERR_ARG;
    }
    auto h = std::make_unique<lt_solver_s>();
    LT_TRY(h->s.create(lb, nf, weights, data, nullptr, 1, lambda_curv, lambda_tikh, op_mode));
    *out = h.release();
    return LT_OK;
}

lt_status lt_solver_create_dirty(lt_lensbatch lbh, lt_nufft nfh, const double *weights, const double *dirty,
                                 int ndirty, double lambda_curv, double lambda_tikh, int op_mode, lt_solver *out) {
    LensBatch *lb = lensbatch_of(lbh);
    NufftPlan *nf = nufft_of(nfh);
    if (!lb || !nf || !out || !dirty || (nf->nvis > 0 && !weights)) {
        set_error("lt_solver_create_dirty: bad arguments");
        return LT_ERR_ARG;
    }
    auto h = std::make_unique<lt_solver_s>();
    LT_TRY(h->s.create(lb, nf, weights, nullptr, dirty, ndirty, lambda_curv, lambda_tikh, op_mode));
    *out = h.release();
    return LT_OK;
}

lt_status lt_solver_solve(lt_solver h, double rtol, int maxiter, int precond, double *x, int *iters, double *resid) {
    if (!h || rtol <= 0 || maxiter < 0) {
        set_error("lt_solver_solve: bad arguments");
        return LT_ERR_ARG;
    }
    return h->s.solve(rtol, maxiter, precond, x, iters, resid);
}

lt_status lt_solver_rhs(lt_solver h, double *rhs) {
    if (!h || !rhs) return LT_ERR_ARG;
    LT_TRY(h->s.b.download(rhs, h->s.ntot));
    return sync();
}

lt_status lt_solver_apply(lt_solver h, const double *x, double *y) {
    if (!h || !x || !y) return LT_ERR_ARG;
    Solver &S = h->s;
    DevBuf<double> dx, dy;
    LT_TRY(dx.upload(x, S.ntot));
    LT_TRY(dy.alloc(S.ntot));
    LT_TRY(S.apply(dx.p, dy.p, S.all_active.p, S.S));
    LT_TRY(dy.download(y, S.ntot));
    return sync();
}

lt_status lt_solver_destroy(lt_solver h) {
    delete h;
    return LT_OK;
}

}  // extern "C"
