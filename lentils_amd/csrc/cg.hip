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
#include <cstring>

#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <map>
#include <tuple>
#include <cmath>
#include <memory>
#include <random>

#include "nufft_plan.h"
#include "pcnet.h"
#include "raster_batch.h"

namespace lt {

lt_status build_curvature(int nx, int ny, double dx, double dy, int32_t coloff, DevCSR &C);
LensBatch *lensbatch_of(lt_lensbatch h);
NufftPlan *nufft_of(lt_nufft h);
lt_status grid_c128_to_plan(NufftPlan &P, const double *src, void *dst);
__global__ void k_gather_weights(long n, const int32_t *__restrict__ perm, const double *__restrict__ w_in,
                                 double *__restrict__ w_out);

namespace {

constexpr int kDotThreads = 256;

// Segmented dot products out[s] = sum over segment s of a*b in two passes:
// gridDim.x parts per segment each reduce a contiguous slice to part[s][p]
// (fixed-order block reduction), then one thread per segment adds its parts
// in order.  The order depends only on the segment bounds and the part
// count, so results are run-to-run identical.
template <class TV = double>
__global__ void __launch_bounds__(kDotThreads) k_seg_dot_part(const int64_t *__restrict__ off,
                                                               const TV *__restrict__ a,
                                                               const TV *__restrict__ b,
                                                               double *__restrict__ part) {
    __shared__ double sh[kDotThreads / 64];
    const int s = blockIdx.y, P = gridDim.x, q = blockIdx.x;
    const long b0 = off[s], len = off[s + 1] - b0;
    const long per = (len + P - 1) / P;
    const long lo = b0 + min(len, per * q), hi = b0 + min(len, per * (q + 1));
    double acc = 0.0;
    for (long i = lo + threadIdx.x; i < hi; i += kDotThreads) acc += (double)a[i] * (double)b[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < kDotThreads / 64; ++k) t += sh[k];
        part[(long)s * P + q] = t;
    }
}

// image rows (per sample, x index) holding a nonempty row of L: the others
// are zero in L x and not read by L^T
__global__ void k_row_mask(long n, const int32_t *__restrict__ rp, int ny, uint8_t *__restrict__ mask) {
    const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n && rp[r + 1] > rp[r]) mask[r / ny] = 1;
}

// one workgroup per sample (many samples): the whole dot in one launch
constexpr int kDotOne = 1024;
__global__ void __launch_bounds__(kDotOne) k_seg_dot_one(const int64_t *__restrict__ off,
                                                          const double *__restrict__ a,
                                                          const double *__restrict__ b, double *__restrict__ out) {
    __shared__ double sh[kDotOne / 64];
    const int s = blockIdx.x;
    const long lo = off[s], hi = off[s + 1];
    double acc = 0.0;
    for (long i = lo + threadIdx.x; i < hi; i += kDotOne) acc += a[i] * b[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < kDotOne / 64; ++k) t += sh[k];
        out[s] = t;
    }
}

// sum of sample s's P partials in increasing q; the loads of each block of
// 16 issued together (one latency per block instead of one per partial)
__device__ inline double parts_sum(const double *__restrict__ part, int s, int P) {
    const double *ps = part + (long)s * P;
    double t = 0.0;
    for (int q0 = 0; q0 < P; q0 += 16) {
        double v[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) v[k] = q0 + k < P ? ps[q0 + k] : 0.0;
#pragma unroll
        for (int k = 0; k < 16; ++k)
            if (q0 + k < P) t += v[k];
    }
    return t;
}

__global__ void k_seg_dot_fin(int S, int P, const double *__restrict__ part, double *__restrict__ out) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    out[s] = parts_sum(part, s, P);
}

// copy a CSR block into a block-diagonal matrix at (row0, col0), its
// entries starting at nz0
__global__ void k_csr_place(long nrows, long nnz, const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                            const double *__restrict__ va, long row0, long nz0, int32_t col0,
                            int32_t *__restrict__ orp, int32_t *__restrict__ oci, double *__restrict__ ova) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nrows) orp[row0 + i] = (int32_t)(rp[i] + nz0);
    if (i < nnz) {
        oci[nz0 + i] = ci[i] + col0;
        ova[nz0 + i] = va[i];
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

// the same two updates over the listed samples only (grid y: the list)
__global__ void k_cg_xr_sg(const int *__restrict__ list, const int64_t *__restrict__ off, const int *__restrict__ act,
                           const double *__restrict__ alpha, const double *__restrict__ p,
                           const double *__restrict__ Ap, double *__restrict__ x, double *__restrict__ r) {
    const int s = list[blockIdx.y];
    if (!act[s]) return;
    const long i = off[s] + (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= off[s + 1]) return;
    const double a = alpha[s];
    x[i] += a * p[i];
    r[i] -= a * Ap[i];
}

__global__ void k_cg_p_sg(const int *__restrict__ list, const int64_t *__restrict__ off, const int *__restrict__ act,
                          const double *__restrict__ beta, const double *__restrict__ z, double *__restrict__ p) {
    const int s = list[blockIdx.y];
    if (!act[s]) return;
    const long i = off[s] + (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= off[s + 1]) return;
    p[i] = z[i] + beta[s] * p[i];
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

// probing vectors for the diagonal: 1 on source pixels (i, j) with
// i mod m == cx and j mod m == cy of every sample, else 0
__global__ void k_probe_set(long n, const int32_t *__restrict__ seg, const int64_t *__restrict__ off,
                            const int *__restrict__ ny, int m, int cx, int cy, double *__restrict__ v) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = seg[i];
    const long q = i - off[s];
    const int a = (int)(q / ny[s]), b = (int)(q - (long)a * ny[s]);
    v[i] = (a % m == cx && b % m == cy) ? 1.0 : 0.0;
}

__global__ void k_probe_take(long n, const int32_t *__restrict__ seg, const int64_t *__restrict__ off,
                             const int *__restrict__ ny, int m, int cx, int cy, const double *__restrict__ w,
                             double *__restrict__ d) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = seg[i];
    const long q = i - off[s];
    const int a = (int)(q / ny[s]), b = (int)(q - (long)a * ny[s]);
    if (a % m == cx && b % m == cy) d[i] = w[i];
}

__global__ void k_jacobi_inv(long n, const double *__restrict__ dd, const double *__restrict__ dc, double lc, double lt,
                             double *__restrict__ dinv) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double d = dd[i] + lc * dc[i] + lt;
    dinv[i] = (d > 0.0 && isfinite(d)) ? 1.0 / d : 1.0;
}

// flexible (Polak-Ribiere) beta for a preconditioner that changes between
// iterations: beta = z_new . (r_new - r_old) / (z_old . r_old)
//                  = -alpha (z_new . A p) / (z_old . r_old)
__global__ void k_cg_beta_flex(int ns, int *__restrict__ act, double *__restrict__ rz,
                               const double *__restrict__ rz_new, const double *__restrict__ zap,
                               const double *__restrict__ alpha, const double *__restrict__ rr,
                               const double *__restrict__ bb, double rtol2, double *__restrict__ beta,
                               int *__restrict__ iters) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    if (!act[s]) return;
    iters[s] += 1;
    beta[s] = rz[s] != 0.0 ? -alpha[s] * zap[s] / rz[s] : 0.0;
    rz[s] = rz_new[s];
    if (!(rr[s] > rtol2 * bb[s])) act[s] = 0;
}

// truncated flexible CG: p = z - sum_j (z . A p_j / p_j . A p_j) p_j over
// the nj stored directions (ring slots, n values each) of every active
// sample, making p A-orthogonal to all of them and not only to the last
__global__ void k_fcg_p(long n, const int32_t *__restrict__ seg, const int *__restrict__ act, int nj, int S,
                        const double *__restrict__ d, const double *__restrict__ pap, const double *__restrict__ P,
                        const double *__restrict__ z, double *__restrict__ p) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = seg[i];
    if (!act[s]) return;
    double v = z[i];
    for (int j = 0; j < nj; ++j) {
        const double q = pap[(long)j * S + s];
        if (q != 0.0) v -= d[(long)j * S + s] / q * P[(long)j * n + i];
    }
    p[i] = v;
}

// y += lam C (C x) on the listed samples (list == NULL: all; inactive ones
// skipped), C the curvature matrix of each sample's source grid: the 5-point
// Laplacian times 1 / (dx dy) with the terms beyond the grid's edge dropped
// (the matrix build_curvature makes).  Both products in one pass over a
// 16 x 16 tile of the output with its halo staged in LDS; each row's terms
// are summed in the matrix's column order (up, left, centre, right, down).
constexpr int kCvT = 16;
__device__ inline double curv_row(int i, int j, int nx, int ny, double h, double up, double lf, double c, double rt,
                                  double dn) {
    double s = 0.0;
    if (i > 0) s += h * up;
    if (j > 0) s += h * lf;
    s += (-4.0 * h) * c;
    if (j < ny - 1) s += h * rt;
    if (i < nx - 1) s += h * dn;
    return s;
}

__global__ void __launch_bounds__(256) k_curv2(const int *__restrict__ list, const int *__restrict__ act,
                                               const int64_t *__restrict__ off, const int *__restrict__ nxs,
                                               const int *__restrict__ nys, const double *__restrict__ hs, double lam,
                                               const double *__restrict__ x, double *__restrict__ y) {
    __shared__ double xs[kCvT + 4][kCvT + 4];
    __shared__ double ts[kCvT + 2][kCvT + 2];
    const int s = list ? list[blockIdx.y] : (int)blockIdx.y;
    if (act && !act[s]) return;
    const int nx = nxs[s], ny = nys[s];
    const int ntj = (ny + kCvT - 1) / kCvT;
    const int ti = (int)blockIdx.x / ntj, tj = (int)blockIdx.x - ti * ntj;
    if (ti * kCvT >= nx) return;
    const int i0 = ti * kCvT - 2, j0 = tj * kCvT - 2;
    const double *xv = x + off[s];
    for (int q = threadIdx.x; q < (kCvT + 4) * (kCvT + 4); q += blockDim.x) {
        const int a = q / (kCvT + 4), b = q - a * (kCvT + 4);
        const int i = i0 + a, j = j0 + b;
        xs[a][b] = (i >= 0 && i < nx && j >= 0 && j < ny) ? xv[(long)i * ny + j] : 0.0;
    }
    __syncthreads();
    const double h = hs[s];
    for (int q = threadIdx.x; q < (kCvT + 2) * (kCvT + 2); q += blockDim.x) {
        const int a = q / (kCvT + 2), b = q - a * (kCvT + 2);
        const int i = i0 + 1 + a, j = j0 + 1 + b;
        double t = 0.0;
        if (i >= 0 && i < nx && j >= 0 && j < ny)
            t = curv_row(i, j, nx, ny, h, xs[a][b + 1], xs[a + 1][b], xs[a + 1][b + 1], xs[a + 1][b + 2],
                         xs[a + 2][b + 1]);
        ts[a][b] = t;
    }
    __syncthreads();
    const int a = threadIdx.x / kCvT, b = threadIdx.x - a * kCvT;
    const int i = ti * kCvT + a, j = tj * kCvT + b;
    if (i >= nx || j >= ny) return;
    const double v = curv_row(i, j, nx, ny, h, ts[a][b + 1], ts[a + 1][b], ts[a + 1][b + 1], ts[a + 1][b + 2],
                              ts[a + 2][b + 1]);
    double *yo = y + off[s] + (long)i * ny + j;
    *yo = lam * v + *yo;
}

__global__ void k_recip(long n, const double *__restrict__ a, double *__restrict__ b) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = 1.0 / a[i];
}

__global__ void k_fill(long n, double v, double *__restrict__ a) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// Learned preconditioner: a correction by the network on top of the
// lens-structured preconditioner, z = m + y rms(m) with m ~ M^-1 r.  Input
// per sample (FP32 [s][2][HW]): r / rms(r) and m / rms(m); rr = r.r and
// mm = m.m per sample.
__global__ void k_pc_in(long n, const int32_t *__restrict__ seg, const int64_t *__restrict__ off,
                        const double *__restrict__ r, const double *__restrict__ rr, const double *__restrict__ m,
                        const double *__restrict__ mm, int HW, float *__restrict__ x0) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = seg[i];
    const long pix = i - off[s];
    const double N = (double)(off[s + 1] - off[s]);
    const double rs = sqrt(rr[s] / N), ms = sqrt(mm[s] / N);
    x0[((long)s * 2) * HW + pix] = rs > 0.0 ? (float)(r[i] / rs) : 0.f;
    x0[((long)s * 2 + 1) * HW + pix] = ms > 0.0 ? (float)(m[i] / ms) : 0.f;
}

// z = m + y rms(m)
__global__ void k_pc_out(long n, const int32_t *__restrict__ seg, const int64_t *__restrict__ off,
                         const float *__restrict__ y, const double *__restrict__ m, const double *__restrict__ mm,
                         int HW, double *__restrict__ z) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = seg[i];
    const long pix = i - off[s];
    const double N = (double)(off[s + 1] - off[s]);
    z[i] = m[i] + (double)y[(long)s * HW + pix] * sqrt(mm[s] / N);
}

// active samples (act == NULL: all) whose corrected direction is not a
// descent direction (r.z <= 0 or not finite, e.g. a network that produced
// NaN) take the lens-structured direction m this iteration
__global__ void k_pc_guard(long n, const int32_t *__restrict__ seg, const int *__restrict__ act,
                           const double *__restrict__ rz, const double *__restrict__ m, double *__restrict__ z,
                           int *__restrict__ nfall) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = seg[i];
    if (act && !act[s]) return;
    if (!(rz[s] > 0.0) || !isfinite(rz[s])) {
        z[i] = m[i];
        if (i == 0 || seg[i - 1] != s) atomicAdd(&nfall[s], 1);
    }
}

// Gradient of the training loss of the learned preconditioner,
//   L = (1/S) sum_s -(z_s^T r_s)^2 / (z_s^T A z_s  c_s),  c_s = r_s^T m_s,
// with respect to the network output y (z = m + y rms(m)).  -(z^T r)^2 /
// z^T A z is the energy z'^T A z' - 2 z'^T r of the best multiple z' of z
// (the A-norm error ||z' - A^-1 r||_A^2 up to a constant): the loss grades
// the direction of z, which is all that CG uses, and not its scale.  Its
// floor, at z ~ A^-1 r, is -r^T A^-1 r / c.
__global__ void k_energy_grad(long n, int S, const int32_t *__restrict__ seg, const int64_t *__restrict__ off,
                              const double *__restrict__ Az, const double *__restrict__ r,
                              const double *__restrict__ mm, const double *__restrict__ rm,
                              const double *__restrict__ zr, const double *__restrict__ zAz, int HW,
                              float *__restrict__ gy) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = seg[i];
    const long pix = i - off[s];
    const double N = (double)(off[s + 1] - off[s]);
    const double c = rm[s], q = zAz[s], t = zr[s];
    double g = 0.0;
    if (c > 0.0 && q > 0.0) g = -2.0 * t / (q * c) * (r[i] - (t / q) * Az[i]) * sqrt(mm[s] / N) / S;
    gy[(long)s * HW + pix] = (float)g;
}

// x32 = x / rms_s per sample (the FP32 operator works on unit-scale input)
__global__ void k_to_f32(long n, const int32_t *__restrict__ seg, const int64_t *__restrict__ off,
                         const double *__restrict__ xx, const double *__restrict__ a, float *__restrict__ b) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = seg[i];
    const double rms = sqrt(xx[s] / (double)(off[s + 1] - off[s]));
    b[i] = rms > 0.0 ? (float)(a[i] / rms) : 0.f;
}

// y = rms_s (cscale * yd + lc * yc) + lt * x: the FP32 data and curvature
// terms back into the FP64 CG
__global__ void k_from_f32(long n, const int32_t *__restrict__ seg, const int64_t *__restrict__ off,
                           const double *__restrict__ xx, const float *__restrict__ yd, double cscale,
                           const float *__restrict__ yc, double lc, double lt, const double *__restrict__ x,
                           double *__restrict__ y) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = seg[i];
    const double rms = sqrt(xx[s] / (double)(off[s + 1] - off[s]));
    double v = cscale * (double)yd[i];
    if (yc) v += lc * (double)yc[i];
    y[i] = rms * v + lt * x[i];
}

// r = b - A x on active samples (reliable update of the recurrence residual)
__global__ void k_resid(long n, const int32_t *__restrict__ seg, const int *__restrict__ act,
                        const double *__restrict__ b, const double *__restrict__ Ax, double *__restrict__ r) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && act[seg[i]]) r[i] = b[i] - Ax[i];
}

__global__ void k_take_flagged(int ns, const int *__restrict__ flag, const double *__restrict__ a,
                               double *__restrict__ b) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < ns && flag[s]) b[s] = a[s];
}

// restart samples flagged in act: p = z
__global__ void k_restart_p(long n, const int32_t *__restrict__ seg, const int *__restrict__ act,
                            const double *__restrict__ z, double *__restrict__ p) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && act[seg[i]]) p[i] = z[i];
}

// --- lens-structured preconditioner ---------------------------------------
// M = lambda_c Cc^T Cc + c L_s^T L_s, L_s = L without its long rows (the
// lens-centre pixels that touch most of the source plane): the reference's
// preconditioner R + w L^T L (test_all.py:63-65, there factorised with
// splu), applied here approximately by a few Jacobi-preconditioned CG steps
// on the explicitly built M per outer iteration (flexible CG outside).

// Explicit M = c L_s^T L_s + lambda_c Cc^T Cc as the sum over the rows r of
// the stacked matrix [sqrt(c) L_s; sqrt(lambda_c) Cc] of their outer
// products: row r with m entries gives m^2 (row, col, w v_a v_b) pairs
__global__ void k_pair_count(long nl, const int32_t *__restrict__ lrp, int maxrow, long nc,
                             const int32_t *__restrict__ crp, int64_t *__restrict__ cnt) {
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nl + nc) return;
    long m;
    if (r < nl) {
        m = lrp[r + 1] - lrp[r];
        if (m > maxrow) m = 0;
    } else {
        m = crp[r - nl + 1] - crp[r - nl];
    }
    cnt[r] = m * m;
}

// key = row << nb | col: both are source-pixel indices of the batch, < 2^nb
__global__ void k_pair_emit(long nl, const int32_t *__restrict__ lrp, const int32_t *__restrict__ lci,
                            const double *__restrict__ lva, int maxrow, double c, long nc, int nb,
                            const int32_t *__restrict__ crp, const int32_t *__restrict__ cci,
                            const double *__restrict__ cva, double lam, const int64_t *__restrict__ off,
                            uint64_t *__restrict__ key, double *__restrict__ val) {
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nl + nc) return;
    const int32_t *ci;
    const double *va;
    int b0, b1;
    double w;
    if (r < nl) {
        b0 = lrp[r];
        b1 = lrp[r + 1];
        if (b1 - b0 > maxrow) return;
        ci = lci;
        va = lva;
        w = c;
    } else {
        b0 = crp[r - nl];
        b1 = crp[r - nl + 1];
        ci = cci;
        va = cva;
        w = lam;
    }
    long o = off[r];
    for (int a = b0; a < b1; ++a)
        for (int b = b0; b < b1; ++b, ++o) {
            key[o] = ((uint64_t)(uint32_t)ci[a] << nb) | (uint64_t)(uint32_t)ci[b];
            val[o] = w * va[a] * va[b];
        }
}

__global__ void k_key_flags(long n, const uint64_t *__restrict__ key, int64_t *__restrict__ flag) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = (i == 0 || key[i] != key[i - 1]) ? 1 : 0;
}

// position of the first pair of every distinct (row, col)
__global__ void k_key_heads(long n, const int64_t *__restrict__ flag, const int64_t *__restrict__ gpos,
                            int64_t *__restrict__ hpos) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flag[i]) hpos[gpos[i]] = i;
}

// one thread per distinct (row, col): the in-order sum of its pairs
__global__ void k_group_pairs(long nz, int nb, const uint64_t *__restrict__ key, const double *__restrict__ val,
                              const int64_t *__restrict__ hpos, int32_t *__restrict__ col, double *__restrict__ out,
                              int32_t *__restrict__ urow) {
    long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nz) return;
    const long h0 = hpos[g], h1 = hpos[g + 1];
    double s = 0.0;
    for (long j = h0; j < h1; ++j) s += val[j];
    const uint64_t k = key[h0];
    col[g] = (int32_t)(k & ((1ull << nb) - 1));
    urow[g] = (int32_t)(k >> nb);
    out[g] = s;
}

// row pointers of the sorted distinct entries: rp[r] = first entry of row >= r
__global__ void k_rowptr_search(long nrows, long nz, const int32_t *__restrict__ urow, int32_t *__restrict__ rp) {
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r > nrows) return;
    long lo = 0, hi = nz;
    while (lo < hi) {
        const long mid = (lo + hi) >> 1;
        if (urow[mid] < r)
            lo = mid + 1;
        else
            hi = mid;
    }
    rp[r] = (int32_t)lo;
}

__global__ void k_csr_diag(long n, const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                           const double *__restrict__ va, double *__restrict__ d) {
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double v = 0.0;
    for (int k = rp[r]; k < rp[r + 1]; ++k)
        if (ci[k] == r) v = va[k];
    d[r] = v > 0.0 ? v : 1.0;
}

// The inner Jacobi-PCG steps on M spread every sample over P workgroups
// (gridDim.x = P, blockIdx.y = sample), each taking the contiguous slice
// k_seg_dot_part takes.  A per-sample dot is left as its P block partials;
// the kernel that needs the scalar sums the P partials of its sample in
// order 0..P-1 in every workgroup (the same value in all of them), so the
// results do not depend on scheduling and no workgroup waits for another.
__device__ inline void part_slice(const int64_t *__restrict__ off, int s, int q, int P, long &lo, long &hi) {
    const long b0 = off[s], len = off[s + 1] - b0;
    const long per = (len + P - 1) / P;
    lo = b0 + min(len, per * q);
    hi = b0 + min(len, per * (q + 1));
}

// fixed-order sum over the block (kDotThreads threads); valid in thread 0
__device__ inline double block_sum_part(double acc) {
    __shared__ double sh[kDotThreads / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < kDotThreads / 64; ++k) t += sh[k];
    return t;
}


// Plain CG's scalar steps over all CUs: the two dots of an iteration as P
// workgroup partials per sample, summed in a fixed order by their consumers.
// Workgroup (q, listed sample): p . Ap over its slice.
__global__ void __launch_bounds__(kDotThreads) k_cg_dot_part(const int64_t *__restrict__ off,
                                                              const int *__restrict__ list,
                                                              const double *__restrict__ a,
                                                              const double *__restrict__ b, double *__restrict__ part,
                                                              int P) {
    const int s = list ? list[blockIdx.y] : (int)blockIdx.y, q = blockIdx.x;
    long lo, hi;
    part_slice(off, s, q, P, lo, hi);
    double acc = 0.0;
    for (long i = lo + threadIdx.x; i < hi; i += kDotThreads) acc += a[i] * b[i];
    const double t = block_sum_part(acc);
    if (threadIdx.x == 0) part[(long)s * P + q] = t;
}

// alpha = rz / p.Ap from the partials (each workgroup of the sample forms
// it the same way), x += alpha p and r -= alpha Ap over the slice (active
// samples), and the partials of the new r . r (of every listed sample)
__global__ void __launch_bounds__(kDotThreads) k_cg_xr_part(const int64_t *__restrict__ off,
                                                             const int *__restrict__ list,
                                                             const int *__restrict__ act,
                                                             const double *__restrict__ rz,
                                                             const double *__restrict__ part_pap, int P,
                                                             const double *__restrict__ p,
                                                             const double *__restrict__ Ap, double *__restrict__ x,
                                                             double *__restrict__ r, double *__restrict__ part_rr,
                                                             double *__restrict__ alpha_out) {
    const int s = list ? list[blockIdx.y] : (int)blockIdx.y, q = blockIdx.x;
    long lo, hi;
    part_slice(off, s, q, P, lo, hi);
    double acc = 0.0;
    if (alpha_out && q == 0 && threadIdx.x == 0 && !act[s]) alpha_out[s] = 0.0;
    if (act[s]) {
        const double pap = parts_sum(part_pap, s, P);
        const double a = pap != 0.0 ? rz[s] / pap : 0.0;
        if (alpha_out && q == 0 && threadIdx.x == 0) alpha_out[s] = a;
        for (long i = lo + threadIdx.x; i < hi; i += kDotThreads) {
            x[i] += a * p[i];
            const double ri = r[i] - a * Ap[i];
            r[i] = ri;
            acc += ri * ri;
        }
    } else {
        for (long i = lo + threadIdx.x; i < hi; i += kDotThreads) acc += r[i] * r[i];
    }
    const double t = block_sum_part(acc);
    if (threadIdx.x == 0) part_rr[(long)s * P + q] = t;
}

// r . z and z . Ap of flexible CG in one pass: partials of a . b and c . d
__global__ void __launch_bounds__(kDotThreads) k_cg_dot2_part(const int64_t *__restrict__ off,
                                                               const int *__restrict__ list,
                                                               const double *__restrict__ a,
                                                               const double *__restrict__ b,
                                                               const double *__restrict__ c,
                                                               const double *__restrict__ d,
                                                               double *__restrict__ part_ab,
                                                               double *__restrict__ part_cd, int P) {
    const int s = list ? list[blockIdx.y] : (int)blockIdx.y, q = blockIdx.x;
    long lo, hi;
    part_slice(off, s, q, P, lo, hi);
    double acc1 = 0.0, acc2 = 0.0;
    for (long i = lo + threadIdx.x; i < hi; i += kDotThreads) {
        acc1 += a[i] * b[i];
        acc2 += c[i] * d[i];
    }
    const double t1 = block_sum_part(acc1);
    __syncthreads();  // block_sum_part's shared slots are reused
    const double t2 = block_sum_part(acc2);
    if (threadIdx.x == 0) {
        part_ab[(long)s * P + q] = t1;
        part_cd[(long)s * P + q] = t2;
    }
}

// flexible CG's beta step from the partials (one thread per listed
// sample): rr; for active samples the iteration count, r . z and z . Ap,
// beta = -alpha z.Ap / rz, rz = r . z and the convergence test on rr (as
// k_cg_beta_flex)
__global__ void k_cg_beta_flex_part(const int *__restrict__ list, int n, const double *__restrict__ part_rr,
                                    const double *__restrict__ part_rz, const double *__restrict__ part_zap, int P,
                                    int *__restrict__ act, double *__restrict__ rz, double *__restrict__ rz_new,
                                    const double *__restrict__ alpha, double *__restrict__ beta,
                                    double *__restrict__ rr, const double *__restrict__ bb, double rtol2,
                                    int *__restrict__ iters) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int s = list ? list[k] : k;
    const double t = parts_sum(part_rr, s, P);
    rr[s] = t;
    if (!act[s]) return;
    iters[s] += 1;
    const double rzn = parts_sum(part_rz, s, P), zap = parts_sum(part_zap, s, P);
    rz_new[s] = rzn;
    beta[s] = rz[s] != 0.0 ? -alpha[s] * zap / rz[s] : 0.0;
    rz[s] = rzn;
    if (!(t > rtol2 * bb[s])) act[s] = 0;
}

// one thread per listed sample: rr from its partials; for active samples
// the iteration count, beta = rr / rz, rz = rr and the convergence test on
// the recurrence residual (as k_cg_beta with rz_new = rr)
__global__ void k_cg_beta_part(const int *__restrict__ list, int n, const double *__restrict__ part_rr, int P,
                               int *__restrict__ act, double *__restrict__ rz, double *__restrict__ beta,
                               double *__restrict__ rr, const double *__restrict__ bb, double rtol2,
                               int *__restrict__ iters) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int s = list ? list[k] : k;
    const double t = parts_sum(part_rr, s, P);
    rr[s] = t;
    if (!act[s]) return;
    iters[s] += 1;
    beta[s] = rz[s] != 0.0 ? t / rz[s] : 0.0;
    rz[s] = t;
    if (!(t > rtol2 * bb[s])) act[s] = 0;  // also stops on NaN
}

// The inner solve's vectors r, z, p, Mp in TV (FP32: the preconditioner
// only approximates M^-1 r), its iterate x (the outer z) and its dots in
// FP64.  start: x = 0, r = r0, z = p = r0 / dM; partials of r . z
template <class TV>
__global__ void __launch_bounds__(kDotThreads) k_inner_start_part(const int64_t *__restrict__ off,
                                                                   const double *__restrict__ r0,
                                                                   const double *__restrict__ dM,
                                                                   double *__restrict__ x, TV *__restrict__ r,
                                                                   TV *__restrict__ z, TV *__restrict__ p,
                                                                   double *__restrict__ part_rz) {
    const int s = blockIdx.y, q = blockIdx.x, P = gridDim.x;
    long lo, hi;
    part_slice(off, s, q, P, lo, hi);
    double acc = 0.0;
    for (long i = lo + threadIdx.x; i < hi; i += kDotThreads) {
        const TV ri = (TV)r0[i], zi = (TV)((double)ri / dM[i]);
        x[i] = 0.0;
        r[i] = ri;
        z[i] = zi;
        p[i] = zi;
        acc += (double)ri * (double)zi;
    }
    const double t = block_sum_part(acc);
    if (threadIdx.x == 0) part_rz[(long)s * P + q] = t;
}

// step: a = rz / pMp; x += a p; r -= a Mp; z = r / dM; partials of r . z
template <class TV>
__global__ void __launch_bounds__(kDotThreads) k_inner_step_part(const int64_t *__restrict__ off,
                                                                  const double *__restrict__ part_rz,
                                                                  const double *__restrict__ part_pMp,
                                                                  const TV *__restrict__ p,
                                                                  const TV *__restrict__ Mp,
                                                                  const double *__restrict__ dM,
                                                                  double *__restrict__ x, TV *__restrict__ r,
                                                                  TV *__restrict__ z,
                                                                  double *__restrict__ part_rz_new) {
    __shared__ double a_sh;
    const int s = blockIdx.y, q = blockIdx.x, P = gridDim.x;
    if (threadIdx.x == 0) {
        const double rz = parts_sum(part_rz, s, P), pMp = parts_sum(part_pMp, s, P);
        a_sh = pMp != 0.0 ? rz / pMp : 0.0;
    }
    __syncthreads();
    const double a = a_sh;
    long lo, hi;
    part_slice(off, s, q, P, lo, hi);
    double acc = 0.0;
    for (long i = lo + threadIdx.x; i < hi; i += kDotThreads) {
        x[i] += a * (double)p[i];
        const TV ri = (TV)((double)r[i] - a * (double)Mp[i]);
        r[i] = ri;
        const TV zi = (TV)((double)ri / dM[i]);
        z[i] = zi;
        acc += (double)ri * (double)zi;
    }
    const double t = block_sum_part(acc);
    if (threadIdx.x == 0) part_rz_new[(long)s * P + q] = t;
}

// p = z + (rz_new / rz) p
template <class TV>
__global__ void __launch_bounds__(kDotThreads) k_inner_p_part(const int64_t *__restrict__ off,
                                                               const double *__restrict__ part_rz,
                                                               const double *__restrict__ part_rz_new,
                                                               const TV *__restrict__ z, TV *__restrict__ p) {
    __shared__ double b_sh;
    const int s = blockIdx.y, q = blockIdx.x, P = gridDim.x;
    if (threadIdx.x == 0) {
        const double rz = parts_sum(part_rz, s, P), rzn = parts_sum(part_rz_new, s, P);
        b_sh = rz != 0.0 ? rzn / rz : 0.0;
    }
    __syncthreads();
    const double b = b_sh;
    long lo, hi;
    part_slice(off, s, q, P, lo, hi);
    for (long i = lo + threadIdx.x; i < hi; i += kDotThreads) p[i] = (TV)((double)z[i] + b * (double)p[i]);
}

// samples whose network was switched off take the lens-structured direction
__global__ void k_pc_lensonly(long n, const int32_t *__restrict__ seg, const int *__restrict__ usenet,
                              const double *__restrict__ m, double *__restrict__ z) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !usenet[seg[i]]) z[i] = m[i];
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
    DevBuf<int> cnx, cny;     // per-sample source grid shape and 1 / (dx dy) (k_curv2)
    DevBuf<double> ch;
    int cv_tiles = 0;         // 16 x 16 tiles of the largest source grid
    DevBuf<double> w;         // visibility weights in plan order (own copy)
    const double *wp = nullptr;  // the weights used: w.p, or the plan's resident ones
    DevBuf<double> dirty;     // D^T C d_s (ndirty x nx*ny)
    int ndirty = 1;           // 1: one data set shared by every sample; S: one per sample
    DevBuf<double> beamspec;  // BCB multiplier, nu x half_nv complex
    DevBuf<double> b, x, r, z, p, Ap, img1, img2, tmp, dinv;
    DevBuf<uint8_t> rowmask;  // S x nx: image rows L reaches (k_row_mask)
    DevBuf<double> rz, rz_new, pAp, rr, bb, alpha, beta;
    DevBuf<int> act, iters, all_active;
    double beam_peak = 0.0;
    NormalStencil own_stencil;
    NormalStencil *stp = nullptr;  // &own_stencil, or the plan's shared stencil
    bool use_stencil = false;
    // precision form: FP32 operator (SpMV, FFT, stencil) inside FP64 CG,
    // true FP64 residual refreshed every refine_every iterations
    int prec_mode = LT_F64;
    int refine_every = 20;
    DevBuf<float> x32, i1_32, i2_32, y32, t32, c32;
    DevBuf<double> Ax, xx;
    // lens-structured preconditioner (LT_PC_LENS)
    int lens_inner = 3;            // inner Jacobi-PCG steps on M per application
    int lens_maxrow = 64;          // image rows with more entries are left out of L_s
    double lens_c = 0.0;           // weight of L_s^T L_s (0: the data operator's impulse-response peak)
    double lens_c_used = 0.0;
    const int *cur_act = nullptr;  // active flags of the running solve (sparse products skip the rest)
    DevBuf<double> dM, ix, ir, iz, ip, iMp;
    // inner dots as partials: inner_parts workgroups per sample (lens_setup)
    int inner_parts = 1;
    DevBuf<double> ipart_rz, ipart_rz2, ipart_pMp;
    bool lens_ready = false;
    DevCSR Mpc;                    // explicit M (block-diagonal over samples)
    lt_status lens_setup();
    lt_status lens_build_matrix();
    lt_status lens_apply(const int *af, const int *alist, int nact);  // z = approx M^-1 r
    PcNet *net = nullptr;          // learned preconditioner (LT_PC_LEARNED)
    DevBuf<double> pm, pmm, prm;   // m ~ M^-1 r, m.m and r.m per sample (network input and scale)
    DevBuf<double> zap;            // z . A p
    DevBuf<int> nfall;             // guard fallbacks to m taken per sample
    DevBuf<int> usenet;            // 1: network correction on, 0: switched off (M alone)
    // truncated flexible CG (fcg_depth > 1): each new direction is made
    // A-orthogonal to the last fcg_depth directions, and alpha = p.r / p.A p
    int fcg_depth = 1;
    DevBuf<double> fP, fAP, fpAp, fd, fpr;  // ring of p, A p, p.A p; z.A p_j; p.r
    std::vector<int> nfall_host, usenet_host;
    int net_on = 0;                // samples with the network on (0: M alone, no forward pass)
    int net_iters = 0;             // network on for the first net_iters iterations (0: no limit)

    // CG blocks of check_every iterations captured as HIP graphs, one per
    // launch width (active samples rounded up to a multiple of 8)
    bool use_graphs = true;
    DevBuf<int> dlist;  // samples launched by the running solve (captured graphs read it)
    std::map<int, hipGraphExec_t> graphs;
    std::map<int, int> graph_warm;
    void drop_graphs() {
        for (auto &g : graphs) (void)hipGraphExecDestroy(g.second);
        graphs.clear();
        graph_warm.clear();
    }
    ~Solver() { drop_graphs(); }
    DevBuf<double> dotpart;  // S x dot_parts partial sums of seg_dot
    int dot_parts = 1;
    // plain CG's dot partials: p . Ap, then r . r (2 x S x cg_parts; a
    // member, so that its address stays fixed while captured graphs refer
    // to it)
    DevBuf<double> cgpart;
    int cg_parts = 16;
    // out[s] = a_s . b_s for every sample s (deterministic order): one
    // workgroup per sample when there are many samples, else partial sums
    // over dot_parts workgroups per sample and a second pass
    lt_status seg_dot(const double *a, const double *b, double *out) {
        if (dot_parts == 1) {
            hipLaunchKernelGGL(k_seg_dot_one, dim3(S), dim3(kDotOne), 0, stream(), soff.p, a, b, out);
            LT_CHECK_LAUNCH();
            return LT_OK;
        }
        hipLaunchKernelGGL(k_seg_dot_part<double>, dim3(dot_parts, S), dim3(kDotThreads), 0, stream(), soff.p, a, b,
                           dotpart.p);
        LT_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_seg_dot_fin, dim3(blocks_for(S, 64)), dim3(64), 0, stream(), S, dot_parts, dotpart.p,
                           out);
        LT_CHECK_LAUNCH();
        return LT_OK;
    }
    lt_status create(LensBatch *lb_, NufftPlan *nf_, const double *weights, const double *data,
                     const double *dirty_in, int ndirty_in, double lc, double lt, int op_mode,
                     bool dirty_on_device = false);
    // y = A x for the samples in active_list; with actflags (per-sample 0/1)
    // the sparse products skip the rows of inactive samples as well
    lt_status apply(const double *xin, double *yout, const int *active_list, int nactive,
                    const int *actflags = nullptr);
    lt_status solve(double rtol, int maxiter, int precond, double *xout, int *it_out, double *res_out);
    lt_status apply32(const double *xin, double *yout, const int *active_list, int nactive, const int *actflags);
    lt_status prepare32();
    lt_status jacobi_setup();
    lt_status learned_setup();
    // train (epochs > 0) or evaluate the network on residual vectors R
    // (nres x ntot, host) with the energy loss of this solver's operators
    lt_status train_pcnet(const double *R, int nres, int epochs, float lr, float lr_final, unsigned seed,
                          double *loss_hist);
    // z from r; alist / nact: the samples launched for (all when null)
    lt_status precondition(int precond, const double *dinv_p, const int *alist = nullptr, int nact = 0);
};

fourier_space padded_fourier_space(const image_space &im, int pad);
lt_status dirty_beam_dev(const visibility_space &vs, const image_space &im, const double *weights, int wsup,
                         double beta, double *beam, double *peak);
lt_status beam_multiplier(int nxp, int nyp, const double *beam, DevBuf<double> &spec);

namespace {

lt_status beam_spectrum(Solver &S, const double *weights_caller_order);

}  // namespace

lt_status Solver::create(LensBatch *lb_, NufftPlan *nf_, const double *weights, const double *data,
                         const double *dirty_in, int ndirty_in, double lc, double lt_, int op_mode,
                         bool dirty_on_device) {
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
    dot_parts = S >= 32 ? 1 : std::max(1, std::min(64, (1024 + S - 1) / S));
    cg_parts = std::max(1, std::min(64, (1024 + S - 1) / S));
    LT_TRY(dotpart.alloc((long)S * dot_parts));
    LT_TRY(seg.alloc(ntot));
    hipLaunchKernelGGL(k_seg_ids, dim3(64, S), dim3(256), 0, stream(), soff.p, S, seg.p);
    LT_CHECK_LAUNCH();
    LT_TRY(lb->ensure_transpose());
    // block-diagonal curvature matrix: one block per distinct source grid,
    // placed at every sample's row / column offset
    {
        std::map<std::tuple<int, int, double, double>, int> uniq;
        std::vector<DevCSR> blocks;
        std::vector<int> bsel(S);
        std::vector<long> nzoff(S + 1, 0);
        for (int s = 0; s < S; ++s) {
            const SrcGrid &g = lb->src_host[s];
            const auto key = std::make_tuple(g.nx, g.ny, g.dx, g.dy);
            auto it = uniq.find(key);
            if (it == uniq.end()) {
                blocks.emplace_back();
                LT_TRY(build_curvature(g.nx, g.ny, g.dx, g.dy, 0, blocks.back()));
                it = uniq.emplace(key, (int)blocks.size() - 1).first;
            }
            bsel[s] = it->second;
            nzoff[s + 1] = nzoff[s] + blocks[bsel[s]].nnz;
        }
        const long nnz = nzoff[S];
        C.nrows = C.ncols = ntot;
        C.nnz = nnz;
        LT_TRY(C.rp.alloc(ntot + 1));
        LT_TRY(C.ci.alloc(std::max<long>(nnz, 1)));
        LT_TRY(C.va.alloc(std::max<long>(nnz, 1)));
        for (int s = 0; s < S; ++s) {
            const DevCSR &B = blocks[bsel[s]];
            const long n = std::max<long>(B.nrows, B.nnz);
            if (n == 0) continue;
            hipLaunchKernelGGL(k_csr_place, dim3(blocks_for(n, 256)), dim3(256), 0, stream(), B.nrows, B.nnz, B.rp.p,
                               B.ci.p, B.va.p, soff_host[s], nzoff[s], (int32_t)soff_host[s], C.rp.p, C.ci.p, C.va.p);
            LT_CHECK_LAUNCH();
        }
        const int32_t last = (int32_t)nnz;
        LT_HIP(hipMemcpyAsync(C.rp.p + ntot, &last, sizeof(int32_t), hipMemcpyHostToDevice, stream()));
        LT_TRY(sync());
        LT_TRY(csr_prepare(C));
        std::vector<int> hnx(S), hny(S);
        std::vector<double> hh(S);
        cv_tiles = 0;
        for (int s = 0; s < S; ++s) {
            const SrcGrid &g = lb->src_host[s];
            hnx[s] = g.nx;
            hny[s] = g.ny;
            hh[s] = 1.0 / (g.dx * g.dy);
            cv_tiles = std::max(cv_tiles, ((g.nx + kCvT - 1) / kCvT) * ((g.ny + kCvT - 1) / kCvT));
        }
        LT_TRY(cnx.upload(hnx.data(), S));
        LT_TRY(cny.upload(hny.data(), S));
        LT_TRY(ch.upload(hh.data(), S));
    }
    // weights and data into plan order; dirty image D^T C d
    const long nv = nf->nvis;
    const bool shared_w = (weights == nullptr);
    if (shared_w && nv > 0 && nf->wplan.n < (size_t)nv) {
        set_error("solver: no weights given and none resident in the NUFFT plan (lt_nufft_set_weights)");
        return LT_ERR_ARG;
    }
    if (shared_w && mode == LT_OP_BCB) {
        set_error("solver: the BCB operator needs the weights argument");
        return LT_ERR_ARG;
    }
    {
        DevBuf<double> dc;
        if (shared_w) {
            wp = nf->wplan.p;
        } else {
            LT_TRY(w.alloc(std::max<long>(nv, 1)));
            if (nv) {
                DevBuf<double> raw;
                LT_TRY(raw.upload(weights, nv));
                hipLaunchKernelGGL(k_gather_weights, dim3(blocks_for(nv, 256)), dim3(256), 0, stream(), nv,
                                   nf->perm.p, raw.p, w.p);
                LT_CHECK_LAUNCH();
                LT_TRY(sync());
            }
            wp = w.p;
        }
        LT_TRY(dirty.alloc(npix * ndirty));
        if (dirty_in && dirty_on_device) {
            LT_HIP(hipMemcpyAsync(dirty.p, dirty_in, npix * ndirty * sizeof(double), hipMemcpyDeviceToDevice,
                                  stream()));
        } else if (dirty_in) {
            LT_TRY(dirty.upload(dirty_in, npix * ndirty));
        } else if (nv) {
            LT_TRY(dc.upload(data, 2 * nv));
            LT_TRY(nf->to_sorted(dc.p, nf->visbuf.p));
            LT_TRY(nf->adjoint(nf->visbuf.p, wp, dirty.p, 1.0, false));
        } else {
            LT_TRY(dirty.zero());
        }
    }
    LT_TRY(nf->ensure_batch(S));
    if (mode == LT_OP_BCB) LT_TRY(beam_spectrum(*this, weights));
    if (mode == LT_OP_NUFFT && nf->prec == LT_F64 && nf->W <= 4 && nv > 0) {
        if (shared_w) {
            stp = nf->wstencil.get();
        } else {
            LT_TRY(own_stencil.build(*nf, w.p));
            stp = &own_stencil;
        }
        use_stencil = stp != nullptr;
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
    // L x writes only the nonempty rows of L: the others stay zero here
    LT_TRY(img1.zero());
    LT_TRY(rowmask.alloc((long)S * lb->im.nx));
    LT_TRY(rowmask.zero());
    hipLaunchKernelGGL(k_row_mask, dim3(blocks_for(lb->Lp.nrows, 256)), dim3(256), 0, stream(), lb->Lp.nrows,
                       lb->Lp.rp.p, lb->im.ny, rowmask.p);
    LT_CHECK_LAUNCH();
    for (auto *v : {&rz, &rz_new, &pAp, &rr, &bb, &alpha, &beta}) LT_TRY(v->alloc(S));
    LT_TRY(act.alloc(S));
    LT_TRY(iters.alloc(S));
    LT_TRY(all_active.alloc(S));
    std::vector<int> ids(S);
    for (int s = 0; s < S; ++s) ids[s] = s;
    LT_TRY(all_active.upload(ids.data(), S));
    // autotuned launch shapes: SpMV lane groups of L, L^T, Cc; stencil batching
    LT_TRY(csr_autotune(lb->Lp, false));
    LT_TRY(csr_autotune(lb->LT, false));
    LT_TRY(csr_autotune(C, false));
    // sample partitions: products over some of the samples launch only their rows
    {
        std::vector<int64_t> poff(S + 1);
        for (int s = 0; s <= S; ++s) poff[s] = (int64_t)s * npix;
        LT_TRY(csr_set_samples(lb->Lp, poff.data(), S));
        LT_TRY(csr_set_samples(lb->LT, soff_host.data(), S));
        LT_TRY(csr_set_samples(C, soff_host.data(), S));
    }
    if (use_stencil && !(shared_w && nf->wstencil_tuned_batch == S)) {
        LT_TRY(stencil_autotune(*nf, *stp, img1.p, img2.p, all_active.p, S));
        if (shared_w) nf->wstencil_tuned_batch = S;
    }
    // rhs b_s = L_s^T dirty_s (a shared dirty image is replicated)
    for (int s = 0; s < S; ++s)
        LT_HIP(hipMemcpyAsync(img2.p + (long)s * npix, dirty.p + (ndirty == 1 ? 0L : (long)s * npix),
                              npix * sizeof(double), hipMemcpyDeviceToDevice, stream()));
    LT_TRY(csr_spmv(lb->LT, img2.p, b.p, 1.0, 0.0, 0));
    return sync();
}

lt_status Solver::apply(const double *xin, double *yout, const int *alist, int nact, const int *af) {
    LT_TRY(csr_spmv_nz(lb->Lp, xin, img1.p, af, npix, alist, nact));
    if (mode != LT_OP_BCB)
        LT_TRY(nf->normal_batched(img1.p, img2.p, wp, alist, nact, 1.0, false, use_stencil ? stp : nullptr,
                                  rowmask.p));
    else
        LT_TRY(nf->convolve_batched(img1.p, img2.p, beamspec.p, alist, nact, 1.0, false));
    LT_TRY(csr_spmv(lb->LT, img2.p, yout, 1.0, 0.0, 0, af, 0, seg.p, alist, nact));
    if (lam_c != 0.0 && cv_tiles > 0) {
        // alist with nact == 0 means all samples
        const bool listed = alist && nact > 0;
        hipLaunchKernelGGL(k_curv2, dim3(cv_tiles, listed ? nact : S), dim3(kCvT * kCvT), 0, stream(),
                           listed ? alist : nullptr, af, soff.p, cnx.p, cny.p, ch.p, lam_c, xin, yout);
        LT_CHECK_LAUNCH();
    }
    if (lam_t != 0.0) {
        hipLaunchKernelGGL(k_axpby, dim3(blocks_for(ntot, 256)), dim3(256), 0, stream(), ntot, lam_t, xin, 1.0, yout);
        LT_CHECK_LAUNCH();
    }
    return LT_OK;
}

// diag(A) by probing: for each of the m x m residue classes c of source
// pixels (i mod m, j mod m), v_c = indicator of the class and
// (A v_c)_j = A_jj + sum of A_jk over the other pixels k of j's class, so the
// estimate is exact where pixels m apart are not coupled (the dirty beam
// and the lens mapping keep couplings local; multiple images add small
// distant terms).  m^2 applications of A for every sample at once.
lt_status Solver::jacobi_setup() {
    constexpr int m = 8;
    std::vector<int> hny(S);
    for (int s = 0; s < S; ++s) hny[s] = lb->src_host[s].ny;
    DevBuf<int> dny;
    DevBuf<double> v, wv, dd;
    LT_TRY(dny.upload(hny.data(), S));
    LT_TRY(v.alloc(ntot));
    LT_TRY(wv.alloc(ntot));
    LT_TRY(dd.alloc(ntot));
    LT_TRY(dinv.alloc(ntot));
    const unsigned gb = blocks_for(ntot, 256);
    for (int cx = 0; cx < m; ++cx)
        for (int cy = 0; cy < m; ++cy) {
            hipLaunchKernelGGL(k_probe_set, dim3(gb), dim3(256), 0, stream(), ntot, seg.p, soff.p, dny.p, m, cx, cy,
                               v.p);
            LT_CHECK_LAUNCH();
            LT_TRY(apply(v.p, wv.p, all_active.p, S));
            hipLaunchKernelGGL(k_probe_take, dim3(gb), dim3(256), 0, stream(), ntot, seg.p, soff.p, dny.p, m, cx, cy,
                               wv.p, dd.p);
            LT_CHECK_LAUNCH();
        }
    hipLaunchKernelGGL(k_jacobi_inv, dim3(gb), dim3(256), 0, stream(), ntot, dd.p, dd.p, 0.0, 0.0, dinv.p);
    LT_CHECK_LAUNCH();
    return sync();
}

lt_status Solver::lens_setup() {
    if (lens_ready) return LT_OK;
    const unsigned gb = blocks_for(ntot, 256);
    for (auto *v : {&dM, &ix, &ir, &iz, &ip, &iMp}) LT_TRY(v->alloc(ntot));
    // enough workgroups to spread the inner vectors over the chip: about
    // 1024 in all, at most 64 per sample
    inner_parts = std::max(1, std::min(64, (1024 + S - 1) / S));
    for (auto *v : {&ipart_rz, &ipart_rz2, &ipart_pMp}) LT_TRY(v->alloc((long)S * inner_parts));
    double c = lens_c;
    if (!(c > 0.0)) {
        // impulse response of the data operator at the image centre (sample
        // 0's slot): its peak is the data weight of one image pixel
        const long ctr = (long)(lb->im.nx / 2) * lb->im.ny + lb->im.ny / 2;
        LT_TRY(img1.zero());
        const double one = 1.0;
        LT_HIP(hipMemcpyAsync(img1.p + ctr, &one, sizeof(double), hipMemcpyHostToDevice, stream()));
        if (mode != LT_OP_BCB)
            LT_TRY(nf->normal_batched(img1.p, img2.p, wp, all_active.p, 1, 1.0, false, use_stencil ? stp : nullptr));
        else
            LT_TRY(nf->convolve_batched(img1.p, img2.p, beamspec.p, all_active.p, 1, 1.0, false));
        LT_TRY(img2.download(&c, 1, ctr));
        LT_TRY(img1.zero());  // L x relies on zero rows outside L's nonempty rows
        LT_TRY(sync());
        if (!(c > 0.0)) c = 1.0;
    }
    lens_c_used = c;
    LT_TRY(lens_build_matrix());
    hipLaunchKernelGGL(k_csr_diag, dim3(gb), dim3(256), 0, stream(), ntot, Mpc.rp.p, Mpc.ci.p, Mpc.va.p, dM.p);
    LT_CHECK_LAUNCH();
    lens_ready = true;
    return sync();
}

lt_status Solver::lens_build_matrix() {
    int nb = 1;
    while ((1L << nb) < ntot) ++nb;  // ntot < 2^31, so 2 nb <= 62 key bits
    const DevCSR &L = lb->Lp;
    const long nl = L.nrows, nc = C.nrows, nr = nl + nc;
    DevBuf<int64_t> cnt, off;
    LT_TRY(cnt.alloc(nr));
    LT_TRY(off.alloc(nr + 1));
    hipLaunchKernelGGL(k_pair_count, dim3(blocks_for(nr, 256)), dim3(256), 0, stream(), nl, L.rp.p, lens_maxrow, nc,
                       C.rp.p, cnt.p);
    LT_CHECK_LAUNCH();
    LT_TRY(exclusive_scan_i64(cnt.p, off.p, nr));
    int64_t np = 0;
    LT_TRY(off.download(&np, 1, nr));
    LT_TRY(sync());
    if (np == 0) {
        set_error("lens preconditioner: empty matrix");
        return LT_ERR_ARG;
    }
    DevBuf<uint64_t> k1, k2;
    DevBuf<double> v1, v2;
    LT_TRY(k1.alloc(std::max<int64_t>(np, 1)));
    LT_TRY(k2.alloc(std::max<int64_t>(np, 1)));
    LT_TRY(v1.alloc(std::max<int64_t>(np, 1)));
    LT_TRY(v2.alloc(std::max<int64_t>(np, 1)));
    hipLaunchKernelGGL(k_pair_emit, dim3(blocks_for(nr, 256)), dim3(256), 0, stream(), nl, L.rp.p, L.ci.p, L.va.p,
                       lens_maxrow, lens_c_used, nc, nb, C.rp.p, C.ci.p, C.va.p, lam_c, off.p, k1.p, v1.p);
    LT_CHECK_LAUNCH();
    {
        size_t tb = 0;
        LT_HIP(rocprim::radix_sort_pairs(nullptr, tb, k1.p, k2.p, v1.p, v2.p, (size_t)np, 0, 2 * nb, stream()));
        DevBuf<uint8_t> tmpb;
        LT_TRY(tmpb.alloc(std::max<size_t>(tb, 1)));
        LT_HIP(rocprim::radix_sort_pairs(tmpb.p, tb, k1.p, k2.p, v1.p, v2.p, (size_t)np, 0, 2 * nb, stream()));
    }
    k1.release();
    v1.release();
    DevBuf<int64_t> flag, gpos;
    LT_TRY(flag.alloc(std::max<int64_t>(np, 1)));
    LT_TRY(gpos.alloc(np + 1));
    hipLaunchKernelGGL(k_key_flags, dim3(blocks_for(np, 256)), dim3(256), 0, stream(), np, k2.p, flag.p);
    LT_CHECK_LAUNCH();
    LT_TRY(exclusive_scan_i64(flag.p, gpos.p, np));
    int64_t nz = 0;
    LT_TRY(gpos.download(&nz, 1, np));
    LT_TRY(sync());
    if (nz >= INT32_MAX) {
        set_error("lens preconditioner: matrix too large");
        return LT_ERR_CAPACITY;
    }
    Mpc = DevCSR();
    Mpc.nrows = Mpc.ncols = ntot;
    Mpc.nnz = nz;
    LT_TRY(Mpc.ci.alloc(std::max<int64_t>(nz, 1)));
    LT_TRY(Mpc.va.alloc(std::max<int64_t>(nz, 1)));
    LT_TRY(Mpc.rp.alloc(ntot + 1));
    DevBuf<int64_t> hpos;
    DevBuf<int32_t> urow;
    LT_TRY(hpos.alloc(nz + 1));
    LT_TRY(urow.alloc(std::max<int64_t>(nz, 1)));
    hipLaunchKernelGGL(k_key_heads, dim3(blocks_for(np, 256)), dim3(256), 0, stream(), np, flag.p, gpos.p, hpos.p);
    LT_CHECK_LAUNCH();
    LT_HIP(hipMemcpyAsync(hpos.p + nz, &np, sizeof(int64_t), hipMemcpyHostToDevice, stream()));
    hipLaunchKernelGGL(k_group_pairs, dim3(blocks_for(nz, 256)), dim3(256), 0, stream(), nz, nb, k2.p, v2.p, hpos.p,
                       Mpc.ci.p, Mpc.va.p, urow.p);
    LT_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_rowptr_search, dim3(blocks_for(ntot + 1, 256)), dim3(256), 0, stream(), ntot, nz, urow.p,
                       Mpc.rp.p);
    LT_CHECK_LAUNCH();
    LT_TRY(csr_prepare(Mpc));
    LT_TRY(csr_make_f32(Mpc));  // the inner solve's products (csr_spmv_fv)
    LT_TRY(csr_autotune(Mpc, false));
    LT_TRY(csr_set_samples(Mpc, soff_host.data(), S));
    return sync();
}

// z ~ M^-1 r: lens_inner steps of Jacobi-preconditioned CG on
// M = lam_c Cc^T Cc + c L_s^T L_s from z = 0, every sample at once
lt_status Solver::lens_apply(const int *af, const int *alist, int nact) {
    const unsigned gb = blocks_for(ntot, 256);
    // the inner iterate is z itself (no final copy); with no inner steps z is
    // the Jacobi step r / dM
    double *xi = lens_inner > 0 ? z.p : ix.p;
    // per-sample dots as partials of inner_parts workgroups per sample: r . z
    // of the current and the next inner step (swapped, not copied), p . Mp
    double *rzc = ipart_rz.p, *rzn = ipart_rz2.p;
    const int P = inner_parts;
    const dim3 grid((unsigned)P, (unsigned)S);
    if (lens_inner == 0) {
        hipLaunchKernelGGL(k_inner_start_part<double>, grid, dim3(kDotThreads), 0, stream(), soff.p, r.p, dM.p, xi,
                           ir.p, iz.p, ip.p, rzc);
        LT_CHECK_LAUNCH();
    } else {
        // FP32 inner vectors; M's products with its FP32 values
        float *r32 = (float *)ir.p, *z32 = (float *)iz.p, *p32 = (float *)ip.p, *Mp32 = (float *)iMp.p;
        hipLaunchKernelGGL(k_inner_start_part<float>, grid, dim3(kDotThreads), 0, stream(), soff.p, r.p, dM.p, xi,
                           r32, z32, p32, rzc);
        LT_CHECK_LAUNCH();
        for (int k = 0; k < lens_inner; ++k) {
            // Mp with the explicit M = c L_s^T L_s + lam_c Cc^T Cc
            LT_TRY(csr_spmv_ff(Mpc, p32, Mp32, af, seg.p));
            hipLaunchKernelGGL(k_seg_dot_part<float>, grid, dim3(kDotThreads), 0, stream(), soff.p, p32, Mp32,
                               ipart_pMp.p);
            hipLaunchKernelGGL(k_inner_step_part<float>, grid, dim3(kDotThreads), 0, stream(), soff.p, rzc,
                               ipart_pMp.p, p32, Mp32, dM.p, xi, r32, z32, rzn);
            LT_CHECK_LAUNCH();
            if (k + 1 < lens_inner) {
                hipLaunchKernelGGL(k_inner_p_part<float>, grid, dim3(kDotThreads), 0, stream(), soff.p, rzc, rzn, z32,
                                   p32);
                LT_CHECK_LAUNCH();
                std::swap(rzc, rzn);
            }
        }
    }
    if (lens_inner == 0) {
        hipLaunchKernelGGL(k_copy, dim3(gb), dim3(256), 0, stream(), ntot, iz.p, z.p);
        LT_CHECK_LAUNCH();
    }
    return LT_OK;
}

lt_status Solver::precondition(int precond, const double *dinv_p, const int *alist, int nact) {
    const unsigned gb = blocks_for(ntot, 256);
    if (precond == LT_PC_LENS) return lens_apply(cur_act, alist, nact);
    if (precond != LT_PC_LEARNED) {
        hipLaunchKernelGGL(k_precond, dim3(gb), dim3(256), 0, stream(), ntot, dinv_p, r.p, z.p);
        LT_CHECK_LAUNCH();
        return LT_OK;
    }
    // z = m + y rms(m), m ~ M^-1 r (the lens-structured preconditioner);
    // rr = r . r must be current (network input scale)
    LT_TRY(lens_apply(cur_act, alist, nact));
    if (net_on == 0) return LT_OK;  // every sample's network is off: z = m
    hipLaunchKernelGGL(k_copy, dim3(gb), dim3(256), 0, stream(), ntot, z.p, pm.p);
    LT_CHECK_LAUNCH();
    LT_TRY(seg_dot(pm.p, pm.p, pmm.p));
    const int HW = net->H * net->W;
    hipLaunchKernelGGL(k_pc_in, dim3(gb), dim3(256), 0, stream(), ntot, seg.p, soff.p, r.p, rr.p, pm.p, pmm.p, HW,
                       net->act[PcNet::X0].p);
    LT_CHECK_LAUNCH();
    LT_TRY(net->forward(S));
    hipLaunchKernelGGL(k_pc_out, dim3(gb), dim3(256), 0, stream(), ntot, seg.p, soff.p, net->act[PcNet::Y].p, pm.p,
                       pmm.p, HW, z.p);
    LT_TRY(seg_dot(r.p, z.p, rz_new.p));
    hipLaunchKernelGGL(k_pc_guard, dim3(gb), dim3(256), 0, stream(), ntot, seg.p, cur_act, rz_new.p, pm.p, z.p,
                       nfall.p);
    hipLaunchKernelGGL(k_pc_lensonly, dim3(gb), dim3(256), 0, stream(), ntot, seg.p, usenet.p, pm.p, z.p);
    LT_CHECK_LAUNCH();
    return LT_OK;
}

lt_status Solver::learned_setup() {
    if (!net) {
        set_error("solver: LT_PC_LEARNED needs a network (lt_solver_set_pcnet)");
        return LT_ERR_ARG;
    }
    if (S > net->nbcap) {
        set_error("solver: %d samples exceed the network's batch capacity %d", S, net->nbcap);
        return LT_ERR_ARG;
    }
    for (int s = 0; s < S; ++s)
        if (lb->src_host[s].nx != net->H || lb->src_host[s].ny != net->W) {
            set_error("solver: sample %d source grid %dx%d differs from the network's %dx%d", s, lb->src_host[s].nx,
                      lb->src_host[s].ny, net->H, net->W);
            return LT_ERR_ARG;
        }
    LT_TRY(lens_setup());
    if (pm.n != (size_t)ntot) LT_TRY(pm.alloc(ntot));
    for (auto *v : {&pmm, &prm, &zap}) LT_TRY(v->alloc(S));
    LT_TRY(nfall.alloc(S));
    LT_TRY(usenet.alloc(S));
    return LT_OK;
}

lt_status Solver::train_pcnet(const double *R, int nres, int epochs, float lr, float lr_final, unsigned seed,
                              double *loss_hist) {
    LT_TRY(learned_setup());
    cur_act = nullptr;  // every sample's rows
    const unsigned gb = blocks_for(ntot, 256);
    const int HW = net->H * net->W;
    DevBuf<double> pool, zAz, zr;
    LT_TRY(pool.upload(R, (long)nres * ntot));
    LT_TRY(zAz.alloc(S));
    LT_TRY(zr.alloc(S));
    std::vector<int> order(nres);
    for (int k = 0; k < nres; ++k) order[k] = k;
    std::mt19937 gen(seed);
    const int nep = epochs > 0 ? epochs : 1;
    const long total = (long)nep * nres;
    long it = 0;
    std::vector<double> hzz(S), hzr(S), hrm(S);
    for (int ep = 0; ep < nep; ++ep) {
        if (epochs > 0) std::shuffle(order.begin(), order.end(), gen);
        double acc = 0.0;
        for (int q = 0; q < nres; ++q, ++it) {
            hipLaunchKernelGGL(k_copy, dim3(gb), dim3(256), 0, stream(), ntot, pool.p + (long)order[q] * ntot, r.p);
            LT_CHECK_LAUNCH();
            LT_TRY(seg_dot(r.p, r.p, rr.p));
            LT_TRY(lens_apply(nullptr, nullptr, 0));
            hipLaunchKernelGGL(k_copy, dim3(gb), dim3(256), 0, stream(), ntot, z.p, pm.p);
            LT_CHECK_LAUNCH();
            LT_TRY(seg_dot(pm.p, pm.p, pmm.p));
            LT_TRY(seg_dot(r.p, pm.p, prm.p));
            hipLaunchKernelGGL(k_pc_in, dim3(gb), dim3(256), 0, stream(), ntot, seg.p, soff.p, r.p, rr.p, pm.p, pmm.p,
                               HW, net->act[PcNet::X0].p);
            LT_CHECK_LAUNCH();
            LT_TRY(net->forward(S));
            hipLaunchKernelGGL(k_pc_out, dim3(gb), dim3(256), 0, stream(), ntot, seg.p, soff.p,
                               net->act[PcNet::Y].p, pm.p, pmm.p, HW, z.p);
            LT_CHECK_LAUNCH();
            LT_TRY(apply(z.p, Ap.p, all_active.p, S));
            LT_TRY(seg_dot(z.p, Ap.p, zAz.p));
            LT_TRY(seg_dot(z.p, r.p, zr.p));
            LT_CHECK_LAUNCH();
            if (epochs > 0) {
                hipLaunchKernelGGL(k_energy_grad, dim3(gb), dim3(256), 0, stream(), ntot, S, seg.p, soff.p, Ap.p,
                                   r.p, pmm.p, prm.p, zr.p, zAz.p, HW, net->gact[PcNet::Y].p);
                LT_CHECK_LAUNCH();
                LT_TRY(net->backward(S));
                const float fr = total > 1 ? (float)it / (float)(total - 1) : 1.f;
                const float lrt = lr_final + 0.5f * (lr - lr_final) * (1.f + std::cos((float)M_PI * fr));
                LT_TRY(net->adam_step(lrt, 1.0f));
            }
            LT_TRY(zAz.download(hzz.data(), S));
            LT_TRY(zr.download(hzr.data(), S));
            LT_TRY(prm.download(hrm.data(), S));
            LT_TRY(sync());
            double l = 0.0;
            for (int s = 0; s < S; ++s)
                l += (hrm[s] > 0 && hzz[s] > 0) ? -hzr[s] * hzr[s] / (hzz[s] * hrm[s]) : 0.0;
            acc += l / S;
        }
        if (loss_hist) loss_hist[ep] = acc / nres;
    }
    return sync();
}

lt_status Solver::prepare32() {
    if (!use_stencil) {
        set_error("solver: the FP32 precision form needs the precomputed stencil operator (LT_OP_NUFFT, FP64 plan, "
                  "kernel_support <= 4)");
        return LT_ERR_UNSUPPORTED;
    }
    if (x32.p) return LT_OK;
    LT_TRY(csr_autotune(lb->Lp, true));
    LT_TRY(csr_autotune(lb->LT, true));
    LT_TRY(csr_autotune(C, true));
    LT_TRY(x32.alloc(ntot));
    LT_TRY(y32.alloc(ntot));
    LT_TRY(t32.alloc(ntot));
    LT_TRY(c32.alloc(ntot));
    LT_TRY(xx.alloc(S));
    LT_TRY(i1_32.alloc(npix * S));
    LT_TRY(i1_32.zero());  // rows of L without entries are never written
    LT_TRY(i2_32.alloc(npix * S));
    LT_TRY(Ax.alloc(ntot));
    if (stp->nb_sel32 == 0) LT_TRY(stencil_autotune32(*nf, *stp, i1_32.p, i2_32.p, all_active.p, S));
    return LT_OK;
}

lt_status Solver::apply32(const double *xin, double *yout, const int *alist, int nact, const int *af) {
    const unsigned gb = blocks_for(ntot, 256);
    LT_TRY(seg_dot(xin, xin, xx.p));
    hipLaunchKernelGGL(k_to_f32, dim3(gb), dim3(256), 0, stream(), ntot, seg.p, soff.p, xx.p, xin, x32.p);
    LT_CHECK_LAUNCH();
    LT_TRY(csr_spmv32_nz(lb->Lp, x32.p, i1_32.p, af, npix, alist, nact));
    LT_TRY(nf->normal_batched32(i1_32.p, i2_32.p, alist, nact, 1.f, false, *stp, rowmask.p));
    LT_TRY(csr_spmv32(lb->LT, i2_32.p, y32.p, 1.f, 0.f, af, 0, seg.p, alist, nact));
    // the curvature prior: the fused FP64 tile kernel on x itself where the
    // source grids allow it, else two FP32 products with the CSR matrix
    const bool curv32 = lam_c != 0.0 && cv_tiles == 0;
    if (curv32) {
        LT_TRY(csr_spmv32(C, x32.p, t32.p, 1.f, 0.f, af, 0, seg.p, alist, nact));
        LT_TRY(csr_spmv32(C, t32.p, c32.p, 1.f, 0.f, af, 0, seg.p, alist, nact));
    }
    hipLaunchKernelGGL(k_from_f32, dim3(gb), dim3(256), 0, stream(), ntot, seg.p, soff.p, xx.p, y32.p,
                       stp->coef32_scale, curv32 ? c32.p : nullptr, lam_c, lam_t, xin, yout);
    LT_CHECK_LAUNCH();
    if (lam_c != 0.0 && !curv32) {
        const bool listed = alist && nact > 0;
        hipLaunchKernelGGL(k_curv2, dim3(cv_tiles, listed ? nact : S), dim3(kCvT * kCvT), 0, stream(),
                           listed ? alist : nullptr, af, soff.p, cnx.p, cny.p, ch.p, lam_c, xin, yout);
        LT_CHECK_LAUNCH();
    }
    return LT_OK;
}

lt_status Solver::solve(double rtol, int maxiter, int precond, double *xout, int *it_out, double *res_out) {
    const unsigned gb = blocks_for(ntot, 256), gs = blocks_for(S, 64);
    const bool mixed = (prec_mode == LT_F32);
    if (mixed) LT_TRY(prepare32());
    if (precond != LT_PC_NONE && precond != LT_PC_JACOBI && precond != LT_PC_LEARNED && precond != LT_PC_LENS) {
        set_error("solver: unknown preconditioner %d", precond);
        return LT_ERR_ARG;
    }
    // Jacobi: diag(A) ~ beam peak * diag(L^T L) + lambda_c diag(Cc^T Cc) + lambda_t
    const double *dinv_p = nullptr;
    if (precond == LT_PC_JACOBI) {
        if (!dinv.p) LT_TRY(jacobi_setup());
        dinv_p = dinv.p;
    }
    const bool learned = (precond == LT_PC_LEARNED);
    const bool flexible = learned || precond == LT_PC_LENS;
    // plain CG: z = r, so r . z is the r . r just formed; the iterations use
    // r and rr in place of z and rz_new (the same values, two launches fewer)
    const bool plain = (precond == LT_PC_NONE);
    const double *zz = plain ? r.p : z.p;
    const double *rzn = plain ? rr.p : rz_new.p;
    // plain CG: p . Ap as partials over all CUs, then x, r and the r . r
    // partials in one pass (alpha formed there), then the beta step (one
    // thread per sample)
    const bool fused = plain;

    long smax = 0;  // the largest source, for launches over listed samples
    for (int q = 0; q < S; ++q) smax = std::max<long>(smax, soff_host[q + 1] - soff_host[q]);
    const unsigned gsmax = blocks_for(std::max<long>(smax, 1), 256);
    if (precond == LT_PC_LENS || precond == LT_PC_LEARNED) LT_TRY(lens_setup());
    if (flexible && zap.n < (size_t)S) LT_TRY(zap.alloc(S));  // z . A p of the flexible beta
    const bool deep = flexible && fcg_depth > 1 && !mixed;
    const bool flexp = flexible && !deep;  // flexible CG with the partial-sum steps
    if ((fused || flexp) && cgpart.n < (size_t)4 * S * cg_parts) LT_TRY(cgpart.alloc((size_t)4 * S * cg_parts));
    int nstored = 0;  // directions in the ring (truncated flexible CG)
    if (deep) {
        const long dn = (long)fcg_depth * ntot, ds = (long)fcg_depth * S;
        if (fP.n != (size_t)dn) {
            LT_TRY(fP.alloc(dn));
            LT_TRY(fAP.alloc(dn));
        }
        if (fpAp.n != (size_t)ds) {
            LT_TRY(fpAp.alloc(ds));
            LT_TRY(fd.alloc(ds));
        }
        LT_TRY(fpr.alloc(S));
    }
    cur_act = nullptr;
    if (learned) {
        LT_TRY(learned_setup());
        LT_TRY(nfall.zero());
        usenet_host.assign(S, 1);
        net_on = S;
        LT_TRY(usenet.upload(usenet_host.data(), S));
    }
    // fallback of the learned preconditioner: a sample whose residual norm
    // has not fallen 10x over the last kWindow iterations switches its
    // network off and goes on with z = M^-1 r
    constexpr int kWindow = 100;
    std::vector<std::vector<double>> rr_hist(S);
    std::vector<double> hrr_chk(S);
    std::vector<int> hflag(S), reflist(S);
    std::vector<double> rr_ref(S);
    DevBuf<int> dflag, dref;
    if (mixed) LT_TRY(dflag.alloc(S));
    if (mixed) LT_TRY(dref.alloc(S));
    std::vector<int> ones(S, 1), zeros(S, 0);
    LT_TRY(act.upload(ones.data(), S));
    cur_act = act.p;
    LT_TRY(iters.upload(zeros.data(), S));
    LT_TRY(x.zero());
    hipLaunchKernelGGL(k_copy, dim3(gb), dim3(256), 0, stream(), ntot, b.p, r.p);
    LT_TRY(seg_dot(r.p, r.p, rr.p));
    LT_CHECK_LAUNCH();
    LT_TRY(precondition(precond, dinv_p));
    hipLaunchKernelGGL(k_copy, dim3(gb), dim3(256), 0, stream(), ntot, z.p, p.p);
    LT_CHECK_LAUNCH();
    LT_TRY(seg_dot(b.p, b.p, bb.p));
    LT_TRY(seg_dot(r.p, z.p, rz.p));
    LT_CHECK_LAUNCH();
    if (mixed) {
        LT_TRY(bb.download(rr_ref.data(), S));
        LT_TRY(sync());
    }
    const double rtol2 = rtol * rtol;
    auto plain_steps = [&](const int *list, int ny) -> lt_status {
        double *p1 = cgpart.p, *p2 = cgpart.p + (long)S * cg_parts;
        hipLaunchKernelGGL(k_cg_dot_part, dim3(cg_parts, ny), dim3(kDotThreads), 0, stream(), soff.p, list, p.p, Ap.p,
                           p1, cg_parts);
        hipLaunchKernelGGL(k_cg_xr_part, dim3(cg_parts, ny), dim3(kDotThreads), 0, stream(), soff.p, list, act.p,
                           rz.p, p1, cg_parts, p.p, Ap.p, x.p, r.p, p2, (double *)nullptr);
        hipLaunchKernelGGL(k_cg_beta_part, dim3(blocks_for(ny, 64)), dim3(64), 0, stream(), list, ny, p2, cg_parts,
                           act.p, rz.p, beta.p, rr.p, bb.p, rtol2, iters.p);
        LT_CHECK_LAUNCH();
        return LT_OK;
    };
    // flexible CG (one direction): the same first half, alpha kept for the
    // beta step; after the preconditioner r . z and z . Ap in one pass, then
    // the beta step from the partials
    auto flex_steps_a = [&](const int *list, int ny) -> lt_status {
        double *p1 = cgpart.p, *p2 = cgpart.p + (long)S * cg_parts;
        hipLaunchKernelGGL(k_cg_dot_part, dim3(cg_parts, ny), dim3(kDotThreads), 0, stream(), soff.p, list, p.p, Ap.p,
                           p1, cg_parts);
        hipLaunchKernelGGL(k_cg_xr_part, dim3(cg_parts, ny), dim3(kDotThreads), 0, stream(), soff.p, list, act.p,
                           rz.p, p1, cg_parts, p.p, Ap.p, x.p, r.p, p2, alpha.p);
        LT_CHECK_LAUNCH();
        return LT_OK;
    };
    auto flex_steps_b = [&](const int *list, int ny) -> lt_status {
        double *p2 = cgpart.p + (long)S * cg_parts, *p3 = cgpart.p + 2L * S * cg_parts,
               *p4 = cgpart.p + 3L * S * cg_parts;
        hipLaunchKernelGGL(k_cg_dot2_part, dim3(cg_parts, ny), dim3(kDotThreads), 0, stream(), soff.p, list, r.p, z.p,
                           z.p, Ap.p, p3, p4, cg_parts);
        hipLaunchKernelGGL(k_cg_beta_flex_part, dim3(blocks_for(ny, 64)), dim3(64), 0, stream(), list, ny, p2, p3, p4,
                           cg_parts, act.p, rz.p, rz_new.p, alpha.p, beta.p, rr.p, bb.p, rtol2, iters.p);
        LT_CHECK_LAUNCH();
        return LT_OK;
    };
    std::vector<int> hact(S, 1), alist(S);
    int nact = S;
    for (int s = 0; s < S; ++s) alist[s] = s;
    // the device list of samples to launch: a member of fixed capacity S,
    // allocated once, so that its address never changes while a captured
    // graph refers to it
    if (dlist.n != (size_t)S) LT_TRY(dlist.alloc(S));
    LT_HIP(hipMemcpyAsync(dlist.p, alist.data(), (size_t)S * sizeof(int), hipMemcpyHostToDevice, stream()));
    auto set_list = [&](const int *h, int n) -> lt_status {
        if (n > 0) LT_HIP(hipMemcpyAsync(dlist.p, h, (size_t)n * sizeof(int), hipMemcpyHostToDevice, stream()));
        return LT_OK;
    };
    const int check_every = 5;
    int outer = 0;
    int it = 0;
    std::vector<double> hrr(S), hbb(S);
    // one FP64 CG iteration over the samples in `list` (nl entries; samples
    // that are not active ride along with zero steps)
    auto iter_body = [&](const int *list, int nl) -> lt_status {
        LT_TRY(apply(p.p, Ap.p, list, nl, act.p));
        // with fewer listed samples than samples, the updates and dots run
        // over the list only (same values)
        const bool part = nl < S;
        if (fused) {
            LT_TRY(plain_steps(part ? list : nullptr, part ? nl : S));
        } else if (flexp) {
            LT_TRY(flex_steps_a(part ? list : nullptr, part ? nl : S));
        } else {
            LT_TRY(seg_dot(p.p, Ap.p, pAp.p));
            hipLaunchKernelGGL(k_cg_alpha, dim3(gs), dim3(64), 0, stream(), S, act.p, rz.p, pAp.p, alpha.p);
            if (part)
                hipLaunchKernelGGL(k_cg_xr_sg, dim3(gsmax, nl), dim3(256), 0, stream(), list, soff.p, act.p, alpha.p,
                                   p.p, Ap.p, x.p, r.p);
            else
                hipLaunchKernelGGL(k_cg_xr, dim3(gb), dim3(256), 0, stream(), ntot, seg.p, act.p, alpha.p, p.p, Ap.p,
                                   x.p, r.p);
            LT_TRY(seg_dot(r.p, r.p, rr.p));
        }
        LT_CHECK_LAUNCH();
        if (!plain) {
            LT_TRY(precondition(precond, dinv_p, list, nl));
            if (!flexp) LT_TRY(seg_dot(r.p, z.p, rz_new.p));
        }
        if (flexp) {
            LT_TRY(flex_steps_b(part ? list : nullptr, part ? nl : S));
        } else if (flexible) {
            LT_TRY(seg_dot(z.p, Ap.p, zap.p));
            hipLaunchKernelGGL(k_cg_beta_flex, dim3(gs), dim3(64), 0, stream(), S, act.p, rz.p, rz_new.p, zap.p,
                               alpha.p, rr.p, bb.p, rtol2, beta.p, iters.p);
        } else if (!fused) {
            hipLaunchKernelGGL(k_cg_beta, dim3(gs), dim3(64), 0, stream(), S, act.p, rz.p, rzn, rr.p, bb.p, rtol2,
                               beta.p, iters.p);
        }
        if (part)
            hipLaunchKernelGGL(k_cg_p_sg, dim3(gsmax, nl), dim3(256), 0, stream(), list, soff.p, act.p, beta.p, zz,
                               p.p);
        else
            hipLaunchKernelGGL(k_cg_p, dim3(gb), dim3(256), 0, stream(), ntot, seg.p, act.p, beta.p, zz, p.p);
        LT_CHECK_LAUNCH();
        return LT_OK;
    };
    // Graph mode: blocks of check_every iterations between the host's
    // convergence checks are captured once per launch width (active samples
    // rounded up to 8, padded with inactive ones) and replayed; the first
    // block at a new width runs uncaptured so that every lazily created
    // buffer and FFT plan exists before capture.  Same arithmetic as the
    // per-iteration loop.
    drop_graphs();
    const bool graph_mode = use_graphs && !mixed && !learned && !deep;
    std::vector<int> padded(S);
    for (int round = 0;; ++round) {
    while (graph_mode && it < maxiter && nact > 0) {
        const int nb = std::min(check_every, maxiter - it);
        const int nl = std::min(S, (nact + 7) / 8 * 8);
        int q = 0;
        for (int s = 0; s < S; ++s)
            if (hact[s]) padded[q++] = s;
        for (int s = 0; s < S && q < nl; ++s)
            if (!hact[s]) padded[q++] = s;
        LT_TRY(set_list(padded.data(), nl));
        bool done = false;
        if (nb == check_every && use_graphs) {
            auto g = graphs.find(nl);
            if (g == graphs.end() && graph_warm[nl] > 0) {
                hipGraph_t gr = nullptr;
                LT_HIP(hipStreamBeginCapture(stream(), hipStreamCaptureModeThreadLocal));
                lt_status st = LT_OK;
                for (int k = 0; k < nb && st == LT_OK; ++k) st = iter_body(dlist.p, nl);
                const hipError_t e = hipStreamEndCapture(stream(), &gr);
                hipGraphExec_t ge = nullptr;
                if (st == LT_OK && e == hipSuccess && gr &&
                    hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0) == hipSuccess) {
                    graphs[nl] = ge;
                    g = graphs.find(nl);
                } else {
                    use_graphs = false;  // capture not possible here: plain launches from now on
                    (void)hipGetLastError();
                }
                if (gr) (void)hipGraphDestroy(gr);
            }
            if (g != graphs.end()) {
                LT_HIP(hipGraphLaunch(g->second, stream()));
                done = true;
            }
        }
        if (!done) {
            for (int k = 0; k < nb; ++k) LT_TRY(iter_body(dlist.p, nl));
            ++graph_warm[nl];
        }
        it += nb;
        outer += nb;
        LT_TRY(act.download(hact.data(), S));
        LT_TRY(sync());
        nact = 0;
        for (int s = 0; s < S; ++s)
            if (hact[s]) alist[nact++] = s;
    }
    if (graph_mode && nact > 0) LT_TRY(set_list(alist.data(), nact));
    for (; !graph_mode && it < maxiter && nact > 0; ++it) {
        if (mixed)
            LT_TRY(apply32(p.p, Ap.p, dlist.p, nact, act.p));
        else
            LT_TRY(apply(p.p, Ap.p, dlist.p, nact, act.p));
        // with fewer listed samples than samples, the updates and dots run
        // over the list only (same values)
        const bool part = nact < S;
        if (fused) {
            LT_TRY(plain_steps(part ? dlist.p : nullptr, part ? nact : S));
        } else if (flexp) {
            LT_TRY(flex_steps_a(part ? dlist.p : nullptr, part ? nact : S));
        } else if (deep) {
            // alpha = p.r / p.A p; the direction joins the ring
            LT_TRY(seg_dot(p.p, Ap.p, pAp.p));
            LT_TRY(seg_dot(p.p, r.p, fpr.p));
            hipLaunchKernelGGL(k_cg_alpha, dim3(gs), dim3(64), 0, stream(), S, act.p, fpr.p, pAp.p, alpha.p);
            LT_CHECK_LAUNCH();
            const int slot = nstored % fcg_depth;
            LT_HIP(hipMemcpyAsync(fP.p + (long)slot * ntot, p.p, ntot * sizeof(double), hipMemcpyDeviceToDevice,
                                  stream()));
            LT_HIP(hipMemcpyAsync(fAP.p + (long)slot * ntot, Ap.p, ntot * sizeof(double), hipMemcpyDeviceToDevice,
                                  stream()));
            LT_HIP(hipMemcpyAsync(fpAp.p + (long)slot * S, pAp.p, S * sizeof(double), hipMemcpyDeviceToDevice,
                                  stream()));
            ++nstored;
        } else {
            LT_TRY(seg_dot(p.p, Ap.p, pAp.p));
            hipLaunchKernelGGL(k_cg_alpha, dim3(gs), dim3(64), 0, stream(), S, act.p, rz.p, pAp.p, alpha.p);
        }
        if (!fused && !flexp) {
            if (part)
                hipLaunchKernelGGL(k_cg_xr_sg, dim3(gsmax, nact), dim3(256), 0, stream(), dlist.p, soff.p, act.p,
                                   alpha.p, p.p, Ap.p, x.p, r.p);
            else
                hipLaunchKernelGGL(k_cg_xr, dim3(gb), dim3(256), 0, stream(), ntot, seg.p, act.p, alpha.p, p.p, Ap.p,
                                   x.p, r.p);
            LT_TRY(seg_dot(r.p, r.p, rr.p));
        }
        LT_CHECK_LAUNCH();
        if (!plain) {
            LT_TRY(precondition(precond, dinv_p, dlist.p, nact));
            if (!flexp) LT_TRY(seg_dot(r.p, z.p, rz_new.p));
        }
        if (flexp) {
            LT_TRY(flex_steps_b(part ? dlist.p : nullptr, part ? nact : S));
        } else if (flexible) {
            LT_TRY(seg_dot(z.p, Ap.p, zap.p));
            hipLaunchKernelGGL(k_cg_beta_flex, dim3(gs), dim3(64), 0, stream(), S, act.p, rz.p, rz_new.p, zap.p,
                               alpha.p, rr.p, bb.p, rtol2, beta.p, iters.p);
        } else if (!fused) {
            hipLaunchKernelGGL(k_cg_beta, dim3(gs), dim3(64), 0, stream(), S, act.p, rz.p, rzn, rr.p, bb.p, rtol2,
                               beta.p, iters.p);
        }
        if (deep) {
            const int nj = std::min(nstored, fcg_depth);
            for (int j = 0; j < nj; ++j) LT_TRY(seg_dot(z.p, fAP.p + (long)j * ntot, fd.p + (long)j * S));
            hipLaunchKernelGGL(k_fcg_p, dim3(gb), dim3(256), 0, stream(), ntot, seg.p, act.p, nj, S, fd.p, fpAp.p,
                               fP.p, z.p, p.p);
        } else {
            hipLaunchKernelGGL(k_cg_p, dim3(gb), dim3(256), 0, stream(), ntot, seg.p, act.p, beta.p, zz, p.p);
        }
        LT_CHECK_LAUNCH();
        ++outer;
        if (learned && net_iters > 0 && outer == net_iters && net_on > 0) {
            usenet_host.assign(S, 0);
            net_on = 0;
            LT_TRY(usenet.upload(usenet_host.data(), S));
        }
        if (mixed && refine_every > 0 && outer % refine_every == 0) {
            // reliable update: a sample whose recurrence residual fell 10x
            // since its last update gets the FP64 true residual b - A x in
            // place of the FP32 operator's recurrence one (direction kept)
            LT_TRY(rr.download(hrr_chk.data(), S));
            LT_TRY(act.download(hact.data(), S));
            LT_TRY(sync());
            int nref = 0;
            for (int s = 0; s < S; ++s) {
                hflag[s] = (hact[s] && hrr_chk[s] < 1e-2 * rr_ref[s]) ? 1 : 0;
                if (hflag[s]) {
                    reflist[nref++] = s;
                    rr_ref[s] = hrr_chk[s];
                }
            }
            if (nref) {
                LT_TRY(dflag.upload(hflag.data(), S));
                LT_TRY(dref.upload(reflist.data(), nref));
                LT_TRY(apply(x.p, Ax.p, dref.p, nref, dflag.p));
                hipLaunchKernelGGL(k_resid, dim3(gb), dim3(256), 0, stream(), ntot, seg.p, dflag.p, b.p, Ax.p, r.p);
                LT_TRY(seg_dot(r.p, r.p, rr.p));
                LT_CHECK_LAUNCH();
                LT_TRY(precondition(precond, dinv_p));
                LT_TRY(seg_dot(r.p, z.p, rz_new.p));
                hipLaunchKernelGGL(k_take_flagged, dim3(gs), dim3(64), 0, stream(), S, dflag.p, rz_new.p, rz.p);
                LT_CHECK_LAUNCH();
            }
        }
        if (outer % check_every == 0 || it == maxiter - 1) {
            LT_TRY(act.download(hact.data(), S));
            if (learned) LT_TRY(rr.download(hrr_chk.data(), S));
            LT_TRY(sync());
            nact = 0;
            for (int s = 0; s < S; ++s)
                if (hact[s]) alist[nact++] = s;
            if (nact > 0) LT_TRY(set_list(alist.data(), nact));
            if (learned) {
                const size_t back = kWindow / check_every;
                int nsw = 0;
                for (int s = 0; s < S; ++s) {
                    rr_hist[s].push_back(hrr_chk[s]);
                    const size_t h = rr_hist[s].size();
                    if (hact[s] && usenet_host[s] && h > back && !(hrr_chk[s] < 1e-2 * rr_hist[s][h - 1 - back])) {
                        usenet_host[s] = 0;
                        --net_on;
                        ++nsw;
                    }
                }
                // flexible CG takes the change of preconditioner in its
                // beta (z . A p); no restart
                if (nsw) LT_TRY(usenet.upload(usenet_host.data(), S));
            }
        }
    }
    // FP64 true residual ||b - A x|| / ||b||
    LT_TRY(apply(x.p, Ap.p, all_active.p, S));
    hipLaunchKernelGGL(k_copy, dim3(gb), dim3(256), 0, stream(), ntot, b.p, r.p);
    hipLaunchKernelGGL(k_axpby, dim3(gb), dim3(256), 0, stream(), ntot, -1.0, Ap.p, 1.0, r.p);
    LT_TRY(seg_dot(r.p, r.p, rr.p));
    LT_CHECK_LAUNCH();
    LT_TRY(rr.download(hrr.data(), S));
    LT_TRY(bb.download(hbb.data(), S));
    LT_TRY(sync());
    // samples whose true residual is above the target (the recurrence drifted,
    // or the FP32 operator's error showed) restart from their current x
    nact = 0;
    for (int s = 0; s < S; ++s) {
        hact[s] = (hrr[s] > rtol2 * hbb[s] && std::isfinite(hrr[s])) ? 1 : 0;
        if (hact[s]) alist[nact++] = s;
    }
    if (nact == 0 || it >= maxiter || round >= 3) break;
    LT_TRY(act.upload(hact.data(), S));
    LT_TRY(set_list(alist.data(), nact));
    LT_TRY(precondition(precond, dinv_p));
    hipLaunchKernelGGL(k_restart_p, dim3(gb), dim3(256), 0, stream(), ntot, seg.p, act.p, z.p, p.p);
    nstored = 0;  // a restart forgets the stored directions
    LT_TRY(seg_dot(r.p, z.p, rz.p));
    LT_CHECK_LAUNCH();
    }
    // the captured graphs are tied to this solve's launch widths and
    // buffers: none outlives it
    drop_graphs();
    std::vector<int> hit(S);
    LT_TRY(iters.download(hit.data(), S));
    if (xout) LT_TRY(x.download(xout, ntot));
    nfall_host.assign(S, 0);
    if (learned) LT_TRY(nfall.download(nfall_host.data(), S));
    LT_TRY(sync());
    for (int s = 0; s < S; ++s) {
        if (it_out) it_out[s] = hit[s];
        if (res_out) res_out[s] = hbb[s] > 0 ? std::sqrt(hrr[s] / hbb[s]) : std::sqrt(hrr[s]);
    }
    return LT_OK;
}

namespace {

// BCB multiplier of the reference's RadioDataset (data_base.py:115-129,
// 163-172): the dirty beam N_b^T C 1 on the (2nx, 2ny) beam space with the
// reference's default NUFFT (kernel support 4, Beatty's beta), rolled by half
// the padded grid and transformed, rfft2(roll(beam)) / (nxp nyp).
lt_status beam_spectrum(Solver &S, const double *weights) {
    NufftPlan &M = *S.nf;
    if (M.nxp != 2 * M.nx || M.nyp != 2 * M.ny) {
        set_error("BCB mode needs the padded grid to be exactly twice the image (pad 2, even sizes)");
        return LT_ERR_UNSUPPORTED;
    }
    visibility_space vs{};
    vs.nrows = M.nrows;
    vs.nchannels = M.nch;
    vs.nstokes = 1;
    vs.uv = M.uvw_host.data();
    vs.channels = M.chan_host.data();
    DevBuf<double> beam;
    LT_TRY(beam.alloc(4L * M.nx * M.ny));
    LT_TRY(dirty_beam_dev(vs, S.lb->im, weights, 0, 0.0, beam.p, &S.beam_peak));
    LT_TRY(beam_multiplier(M.nxp, M.nyp, beam.p, S.beamspec));
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
    if (!lb || !nf || !out || !dirty) {
        set_error("lt_solver_create_dirty: bad arguments");
        return LT_ERR_ARG;
    }
    auto h = std::make_unique<lt_solver_s>();
    LT_TRY(h->s.create(lb, nf, weights, nullptr, dirty, ndirty, lambda_curv, lambda_tikh, op_mode));
    *out = h.release();
    return LT_OK;
}

lt_status lt_solver_create_resident(lt_lensbatch lbh, lt_nufft nfh, const double *d_dirty, int ndirty,
                                   double lambda_curv, double lambda_tikh, int op_mode, lt_solver *out) {
    LensBatch *lb = lensbatch_of(lbh);
    NufftPlan *nf = nufft_of(nfh);
    if (!lb || !nf || !out || !d_dirty) {
        set_error("lt_solver_create_resident: bad arguments");
        return LT_ERR_ARG;
    }
    hipPointerAttribute_t attr{};
    if (hipPointerGetAttributes(&attr, d_dirty) != hipSuccess || attr.type != hipMemoryTypeDevice) {
        (void)hipGetLastError();
        set_error("lt_solver_create_resident: d_dirty is not device memory");
        return LT_ERR_ARG;
    }
    auto h = std::make_unique<lt_solver_s>();
    LT_TRY(h->s.create(lb, nf, nullptr, nullptr, d_dirty, ndirty, lambda_curv, lambda_tikh, op_mode, true));
    *out = h.release();
    return LT_OK;
}

lt_status lt_nufft_set_weights(lt_nufft nfh, const double *weights) {
    NufftPlan *nf = nufft_of(nfh);
    if (!nf) {
        set_error("lt_nufft_set_weights: bad arguments");
        return LT_ERR_ARG;
    }
    return nf->set_weights(weights);
}

lt_status lt_nufft_set_stencil_shape(lt_nufft nfh, int samples_per_group) {
    NufftPlan *nf = nufft_of(nfh);
    const int m = samples_per_group < 0 ? -samples_per_group : samples_per_group;
    if (!nf || m > 8 || (m & (m - 1))) {
        set_error("lt_nufft_set_stencil_shape: 0, 1, 2, 4, 8 (vector form) or -1, -2, -4, -8 (matrix-core form)");
        return LT_ERR_ARG;
    }
    nf->force_nb = samples_per_group;
    return LT_OK;
}

lt_status lt_nufft_set_degrid_form(lt_nufft nfh, int form) {
    NufftPlan *nf = nufft_of(nfh);
    if (!nf || (form != 0 && form != 1)) {
        set_error("lt_nufft_set_degrid_form: 0 (per-offset taps) or 1 (window taps)");
        return LT_ERR_ARG;
    }
    nf->degrid_half = form == 1;
    return LT_OK;
}

lt_status lt_nufft_set_col_form(lt_nufft nfh, int form) {
    NufftPlan *nf = nufft_of(nfh);
    if (!nf || (form != 0 && form != 1)) {
        set_error("lt_nufft_set_col_form: 0 (rocFFT) or 1 (own radix-4 column transforms)");
        return LT_ERR_ARG;
    }
    nf->own_cols = form == 1;
    return LT_OK;
}

lt_status lt_nufft_set_row_form(lt_nufft nfh, int form) {
    NufftPlan *nf = nufft_of(nfh);
    if (!nf || (form != 0 && form != 1)) {
        set_error("lt_nufft_set_row_form: 0 (256-thread LDS radix-4 form) or 1 (register-fed radix-4 form)");
        return LT_ERR_ARG;
    }
    nf->wave_rows = form == 1;
    return LT_OK;
}

lt_status lt_device_alloc(long nbytes, void **dptr) {
    if (!dptr || nbytes < 0) return LT_ERR_ARG;
    *dptr = nullptr;
    if (nbytes == 0) return LT_OK;
    LT_HIP(hipMalloc(dptr, (size_t)nbytes));
    return LT_OK;
}

lt_status lt_device_upload(void *dptr, const void *host, long nbytes) {
    if ((!dptr || !host) && nbytes > 0) return LT_ERR_ARG;
    if (nbytes > 0) LT_HIP(hipMemcpy(dptr, host, (size_t)nbytes, hipMemcpyHostToDevice));
    return LT_OK;
}

lt_status lt_device_download(const void *dptr, void *host, long nbytes) {
    if ((!dptr || !host) && nbytes > 0) return LT_ERR_ARG;
    if (nbytes > 0) LT_HIP(hipMemcpy(host, dptr, (size_t)nbytes, hipMemcpyDeviceToHost));
    return LT_OK;
}

lt_status lt_device_free(void *dptr) {
    if (dptr) LT_HIP(hipFree(dptr));
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

lt_status lt_solver_set_precision(lt_solver h, int precision, int refine_every) {
    if (!h || (precision != LT_F64 && precision != LT_F32) || refine_every < 0) {
        set_error("lt_solver_set_precision: bad arguments");
        return LT_ERR_ARG;
    }
    h->s.prec_mode = precision;
    h->s.refine_every = refine_every;
    if (precision == LT_F32) return h->s.prepare32();
    return LT_OK;
}

lt_status lt_solver_set_lens_pc(lt_solver h, int inner_steps, int max_row, double c) {
    if (!h || inner_steps < 0 || inner_steps > 64 || max_row < 1) {
        set_error("lt_solver_set_lens_pc: bad arguments");
        return LT_ERR_ARG;
    }
    Solver &S = h->s;
    S.lens_inner = inner_steps;
    if (S.lens_maxrow != max_row || S.lens_c != c) S.lens_ready = false;
    S.lens_maxrow = max_row;
    S.lens_c = c;
    return LT_OK;
}

lt_status lt_solver_lens_pc_matrix(lt_solver h, long *nnz, double *c_used, int *row_ptr, int *cols, double *vals) {
    if (!h || !nnz) {
        set_error("lt_solver_lens_pc_matrix: bad arguments");
        return LT_ERR_ARG;
    }
    Solver &S = h->s;
    LT_TRY(S.lens_setup());
    *nnz = (long)S.Mpc.nnz;
    if (c_used) *c_used = S.lens_c_used;
    if (row_ptr) LT_TRY(S.Mpc.rp.download(row_ptr, S.ntot + 1));
    if (cols && S.Mpc.nnz > 0) LT_TRY(S.Mpc.ci.download(cols, S.Mpc.nnz));
    if (vals && S.Mpc.nnz > 0) LT_TRY(S.Mpc.va.download(vals, S.Mpc.nnz));
    return sync();
}

lt_status lt_solver_set_graphs(lt_solver h, int on) {
    if (!h) return LT_ERR_ARG;
    h->s.use_graphs = on != 0;
    return LT_OK;
}

lt_status lt_solver_set_learned_iterations(lt_solver h, int n) {
    if (!h || n < 0) return LT_ERR_ARG;
    h->s.net_iters = n;
    return LT_OK;
}

lt_status lt_solver_set_flexible_depth(lt_solver h, int depth) {
    if (!h) return LT_ERR_ARG;
    if (depth < 1 || depth > 32) {
        set_error("solver: flexible CG depth %d outside 1..32", depth);
        return LT_ERR_ARG;
    }
    h->s.fcg_depth = depth;
    return LT_OK;
}

lt_status lt_solver_set_rhs(lt_solver h, const double *rhs) {
    if (!h || !rhs) return LT_ERR_ARG;
    LT_TRY(h->s.b.upload(rhs, h->s.ntot));
    return sync();
}

lt_status lt_solver_apply(lt_solver h, const double *x, double *y) {
    if (!h || !x || !y) return LT_ERR_ARG;
    Solver &S = h->s;
    DevBuf<double> dx, dy;
    LT_TRY(dx.upload(x, S.ntot));
    LT_TRY(dy.alloc(S.ntot));
    if (S.prec_mode == LT_F32)
        LT_TRY(S.apply32(dx.p, dy.p, S.all_active.p, S.S, nullptr));
    else
        LT_TRY(S.apply(dx.p, dy.p, S.all_active.p, S.S));
    LT_TRY(dy.download(y, S.ntot));
    return sync();
}

lt_status lt_solver_time_products(lt_solver h, int reps, double *ms_L, double *ms_LT, long *shape_L,
                                  long *shape_LT) {
    if (!h || reps <= 0) return LT_ERR_ARG;
    Solver &S = h->s;
    LensBatch *lb = S.lb;
    DevBuf<double> x, y;
    LT_TRY(x.alloc(S.ntot));
    LT_TRY(y.alloc(S.ntot));
    hipLaunchKernelGGL(k_fill, dim3(blocks_for(S.ntot, 256)), dim3(256), 0, stream(), S.ntot, 1.0, x.p);
    LT_CHECK_LAUNCH();
    hipEvent_t e0, e1;
    LT_HIP(hipEventCreate(&e0));
    LT_HIP(hipEventCreate(&e1));
    float t = 0.f;
    // L x over the nonempty rows (as the solver forms it), then L^T of that image
    LT_TRY(csr_spmv_nz(lb->Lp, x.p, S.img1.p, nullptr, S.npix, nullptr, 0));
    LT_HIP(hipEventRecord(e0, stream()));
    for (int r = 0; r < reps; ++r) LT_TRY(csr_spmv_nz(lb->Lp, x.p, S.img1.p, nullptr, S.npix, nullptr, 0));
    LT_HIP(hipEventRecord(e1, stream()));
    LT_HIP(hipEventSynchronize(e1));
    LT_HIP(hipEventElapsedTime(&t, e0, e1));
    if (ms_L) *ms_L = t / reps;
    LT_TRY(csr_spmv(lb->LT, S.img1.p, y.p, 1.0, 0.0, 0));
    LT_HIP(hipEventRecord(e0, stream()));
    for (int r = 0; r < reps; ++r) LT_TRY(csr_spmv(lb->LT, S.img1.p, y.p, 1.0, 0.0, 0));
    LT_HIP(hipEventRecord(e1, stream()));
    LT_HIP(hipEventSynchronize(e1));
    LT_HIP(hipEventElapsedTime(&t, e0, e1));
    if (ms_LT) *ms_LT = t / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (shape_L) {
        shape_L[0] = lb->Lp.nrows;
        shape_L[1] = lb->Lp.ncols;
        shape_L[2] = lb->Lp.nnz;
    }
    if (shape_LT) {
        shape_LT[0] = lb->LT.nrows;
        shape_LT[1] = lb->LT.ncols;
        shape_LT[2] = lb->LT.nnz;
    }
    return sync();
}

lt_status lt_solver_set_pcnet(lt_solver h, lt_pcnet net) {
    if (!h) return LT_ERR_ARG;
    h->s.net = pcnet_of(net);
    return LT_OK;
}

lt_status lt_solver_train_pcnet(lt_solver h, const double *residuals, int nres, int epochs, float lr, float lr_final,
                                unsigned seed, double *loss_hist) {
    if (!h || !residuals || nres <= 0 || epochs < 0) {
        set_error("lt_solver_train_pcnet: bad arguments");
        return LT_ERR_ARG;
    }
    return h->s.train_pcnet(residuals, nres, epochs, lr, lr_final, seed, loss_hist);
}

lt_status lt_solver_diag(lt_solver h, double *diag) {
    if (!h || !diag) return LT_ERR_ARG;
    Solver &S = h->s;
    if (!S.dinv.p) LT_TRY(S.jacobi_setup());
    std::vector<double> d(S.ntot);
    LT_TRY(S.dinv.download(d.data(), S.ntot));
    LT_TRY(sync());
    for (long i = 0; i < S.ntot; ++i) diag[i] = 1.0 / d[i];
    return LT_OK;
}

lt_status lt_solver_pc_fallbacks(lt_solver h, int *counts, int *plain) {
    if (!h || !counts) return LT_ERR_ARG;
    Solver &S = h->s;
    for (int s = 0; s < S.S; ++s) {
        counts[s] = s < (int)S.nfall_host.size() ? S.nfall_host[s] : 0;
        if (plain) plain[s] = s < (int)S.usenet_host.size() ? !S.usenet_host[s] : 0;
    }
    return LT_OK;
}

lt_status lt_solver_destroy(lt_solver h) {
    delete h;
    return LT_OK;
}

}  // extern "C"
