// Delaunay source planes (lentils/backend/triangles.c): the weighted
// gradient rows of a triangulated source (triangle_gradient_csr) and the lens
// matrix that maps each image pixel either to a casted triangulation vertex
// or, through barycentric weights, to the triangle its deflected position
// falls in (delaunay_lens_matrix_csr).  The triangulation and the point
// location stay with the caller (the reference uses scipy.spatial.Delaunay
// and find_simplex, lens_operators.py:66-70).
//
// Arithmetic per entry is the reference's, operation for operation (compiled
// without FMA contraction), so the CSR is bit-identical; row pointers and
// the casted / uncasted pixel numbering come from device-wide scans instead
// of the reference's sequential counters.
#include <algorithm>
#include <cmath>
#include <vector>

#include "lt_internal.h"

namespace lt {
namespace {

// barycentric weights of s in the triangle (v0, v1, v2), triangles.c:7-31
__device__ inline void tri_weights(double sx, double sy, const double *__restrict__ pts, int p0, int p1, int p2,
                                   double w[3]) {
    const double a0 = pts[2 * p0] - sx, a1 = pts[2 * p0 + 1] - sy;
    const double b0 = pts[2 * p1] - sx, b1 = pts[2 * p1 + 1] - sy;
    const double c0 = pts[2 * p2] - sx, c1 = pts[2 * p2 + 1] - sy;
    double w0 = b0 * c1 - c0 * b1;
    double w1 = c0 * a1 - a0 * c1;
    double w2 = a0 * b1 - b0 * a1;
    const double area = w0 + w1 + w2;
    const double inv = 1.0 / area;
    w[0] = w0 * inv;
    w[1] = w1 * inv;
    w[2] = w2 * inv;
}

// two rows (x, y) of three entries per triangle, triangles.c:66-103
__global__ void k_tri_gradient(long ntri, const double *__restrict__ pts, const int *__restrict__ tris,
                               int32_t *__restrict__ rp, int32_t *__restrict__ ci, double *__restrict__ va) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t > ntri) return;
    if (t == ntri) {
        rp[2 * ntri] = (int32_t)(6 * ntri);
        return;
    }
    const int v0 = tris[3 * t], v1 = tris[3 * t + 1], v2 = tris[3 * t + 2];
    const double p0x = pts[2 * v0], p0y = pts[2 * v0 + 1];
    const double p1x = pts[2 * v1], p1y = pts[2 * v1 + 1];
    const double p2x = pts[2 * v2], p2y = pts[2 * v2 + 1];
    const double area = 0.5 * (p0y * (p1x - p2x) + p1y * p2x - p1x * p2y + p0x * (-p1y + p2y));
    double weight = 1.0 / (2 * area);
    weight *= sqrt(fabs(area));
    const long i = 6 * t;
    rp[2 * t] = (int32_t)i;
    rp[2 * t + 1] = (int32_t)(i + 3);
    va[i] = (-p1y + p2y) * weight;
    va[i + 1] = (p0y - p2y) * weight;
    va[i + 2] = (-p0y + p1y) * weight;
    va[i + 3] = (p1x - p2x) * weight;
    va[i + 4] = (-p0x + p2x) * weight;
    va[i + 5] = (p0x - p1x) * weight;
    ci[i] = ci[i + 3] = v0;
    ci[i + 1] = ci[i + 4] = v1;
    ci[i + 2] = ci[i + 5] = v2;
}

// per pixel (x-major): 1 if it is a casted vertex, 1 in `unc` if it is an
// uncasted (interpolated) pixel; masked-out pixels are neither
__global__ void k_dl_flags(int nx, int ny, int ncast, const char *__restrict__ mask, int64_t *__restrict__ cast,
                           int64_t *__restrict__ unc) {
    const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= (long)nx * ny) return;
    const int i = (int)(r / ny), j = (int)(r % ny);
    const bool on = !mask || mask[r];
    const bool c = (i % ncast == 0) && (j % ncast == 0);
    cast[r] = (on && c) ? 1 : 0;
    unc[r] = (on && !c) ? 1 : 0;
}

// entries per row: 1 (casted), 3 (uncasted inside the triangulation) or 0
__global__ void k_dl_count(long n, const int64_t *__restrict__ cast_pos, const int64_t *__restrict__ unc_pos,
                           const int *__restrict__ unc_tri, int64_t *__restrict__ cnt) {
    const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    int64_t c = 0;
    if (cast_pos[r + 1] != cast_pos[r])
        c = 1;
    else if (unc_pos[r + 1] != unc_pos[r] && unc_tri[unc_pos[r]] >= 0)
        c = 3;
    cnt[r] = c;
}

__global__ void k_dl_fill(long n, const int64_t *__restrict__ cast_pos, const int64_t *__restrict__ unc_pos,
                          const int64_t *__restrict__ off, const double *__restrict__ pts, const int *__restrict__ tris,
                          const double *__restrict__ unc_pts, const int *__restrict__ unc_tri,
                          int32_t *__restrict__ rp, int32_t *__restrict__ ci, double *__restrict__ va) {
    const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r > n) return;
    rp[r] = (int32_t)off[r];
    if (r == n) return;
    const long k = off[r];
    if (cast_pos[r + 1] != cast_pos[r]) {
        va[k] = 1.0;
        ci[k] = (int32_t)cast_pos[r];
    } else if (off[r + 1] != k) {
        const long u = unc_pos[r];
        const int t = unc_tri[u];
        const int p0 = tris[3 * t], p1 = tris[3 * t + 1], p2 = tris[3 * t + 2];
        double w[3];
        tri_weights(unc_pts[2 * u], unc_pts[2 * u + 1], pts, p0, p1, p2, w);
        va[k] = w[0];
        va[k + 1] = w[1];
        va[k + 2] = w[2];
        ci[k] = p0;
        ci[k + 1] = p1;
        ci[k + 2] = p2;
    }
}

__global__ void k_tri_geometry(long n, const double *__restrict__ s, const int *__restrict__ tri_inds,
                               const double *__restrict__ pos, double *__restrict__ w_out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double w[3];
    tri_weights(s[2 * i], s[2 * i + 1], pos, tri_inds[3 * i], tri_inds[3 * i + 1], tri_inds[3 * i + 2], w);
    w_out[3 * i] = w[0];
    w_out[3 * i + 1] = w[1];
    w_out[3 * i + 2] = w[2];
}

lt_status check_triangles(const delaunay_space &ds) {
    if (ds.num_points < 0 || ds.num_tris < 0 || (ds.num_points > 0 && !ds.points) ||
        (ds.num_tris > 0 && !ds.triangles)) {
        set_error("delaunay space: bad sizes or NULL arrays");
        return LT_ERR_ARG;
    }
    for (long k = 0; k < 3L * ds.num_tris; ++k)
        if (ds.triangles[k] < 0 || ds.triangles[k] >= ds.num_points) {
            set_error("delaunay space: triangle vertex %d outside 0..%d", ds.triangles[k], ds.num_points - 1);
            return LT_ERR_ARG;
        }
    return LT_OK;
}

}  // namespace
}  // namespace lt

using namespace lt;

extern "C" {

lt_status lt_triangle_gradient_csr(delaunay_space ds, int *row_inds, int *cols, double *vals) {
    LT_TRY(ensure_init());
    LT_TRY(check_triangles(ds));
    if (!row_inds || (ds.num_tris > 0 && (!cols || !vals))) {
        set_error("lt_triangle_gradient_csr: NULL output");
        return LT_ERR_ARG;
    }
    const long nt = ds.num_tris, nnz = 6 * nt;
    DevBuf<double> pts, va;
    DevBuf<int> tris;
    DevBuf<int32_t> rp, ci;
    const std::vector<double> zd(2, 0.0);
    const std::vector<int> zi(3, 0);
    LT_TRY(pts.upload(ds.num_points > 0 ? ds.points : zd.data(), std::max(2L * ds.num_points, 2L)));
    LT_TRY(tris.upload(nt > 0 ? ds.triangles : zi.data(), std::max(3 * nt, 3L)));
    LT_TRY(rp.alloc(2 * nt + 1));
    LT_TRY(ci.alloc(std::max(nnz, 1L)));
    LT_TRY(va.alloc(std::max(nnz, 1L)));
    hipLaunchKernelGGL(k_tri_gradient, dim3(blocks_for(nt + 1, 256)), dim3(256), 0, stream(), nt, pts.p, tris.p, rp.p,
                       ci.p, va.p);
    LT_CHECK_LAUNCH();
    LT_TRY(rp.download(row_inds, 2 * nt + 1));
    if (nnz) {
        LT_TRY(ci.download(cols, nnz));
        LT_TRY(va.download(vals, nnz));
    }
    return sync();
}

lt_status lt_delaunay_lens_matrix_csr(image_space im, delaunay_space ds, int ncast, const double *unc_pts,
                                      const int *unc_tri, int *row_inds, int *cols, double *vals, long capacity,
                                      long *nnz_out) {
    LT_TRY(ensure_init());
    LT_TRY(check_triangles(ds));
    if (im.nx <= 0 || im.ny <= 0 || ncast < 1 || !row_inds) {
        set_error("lt_delaunay_lens_matrix_csr: bad arguments");
        return LT_ERR_ARG;
    }
    const long n = (long)im.nx * im.ny;
    // the casted and uncasted pixels in x-major order (the order the caller
    // deflected and located them in)
    long ncst = 0, nunc = 0;
    for (int i = 0; i < im.nx; ++i)
        for (int j = 0; j < im.ny; ++j) {
            if (im.mask && !im.mask[(long)i * im.ny + j]) continue;
            if (i % ncast == 0 && j % ncast == 0)
                ++ncst;
            else
                ++nunc;
        }
    if (ncst > ds.num_points) {
        set_error("lt_delaunay_lens_matrix_csr: %ld casted pixels but %d triangulation points", ncst, ds.num_points);
        return LT_ERR_ARG;
    }
    if (nunc > 0 && (!unc_pts || !unc_tri)) {
        set_error("lt_delaunay_lens_matrix_csr: NULL uncasted points or triangle indices");
        return LT_ERR_ARG;
    }
    for (long u = 0; u < nunc; ++u)
        if (unc_tri[u] >= ds.num_tris) {
            set_error("lt_delaunay_lens_matrix_csr: triangle index %d of point %ld outside 0..%d", unc_tri[u], u,
                      ds.num_tris - 1);
            return LT_ERR_ARG;
        }
    DevBuf<char> mask;
    if (im.mask) LT_TRY(mask.upload(im.mask, n));
    DevBuf<double> pts, up, va;
    DevBuf<int> tris, ut;
    DevBuf<int64_t> cf, uf, cpos, upos, cnt, off;
    const std::vector<double> zd(2, 0.0);
    const std::vector<int> zi(3, 0), none(1, -1);
    LT_TRY(pts.upload(ds.num_points > 0 ? ds.points : zd.data(), std::max(2L * ds.num_points, 2L)));
    LT_TRY(tris.upload(ds.num_tris > 0 ? ds.triangles : zi.data(), std::max(3L * ds.num_tris, 3L)));
    LT_TRY(up.upload(nunc > 0 ? unc_pts : zd.data(), std::max(2 * nunc, 2L)));
    LT_TRY(ut.upload(nunc > 0 ? unc_tri : none.data(), std::max(nunc, 1L)));
    for (auto *b : {&cf, &uf, &cpos, &upos, &cnt, &off}) LT_TRY(b->alloc(n + 1));
    hipLaunchKernelGGL(k_dl_flags, dim3(blocks_for(n, 256)), dim3(256), 0, stream(), im.nx, im.ny, ncast,
                       im.mask ? mask.p : nullptr, cf.p, uf.p);
    LT_CHECK_LAUNCH();
    LT_TRY(exclusive_scan_i64(cf.p, cpos.p, n));
    LT_TRY(exclusive_scan_i64(uf.p, upos.p, n));
    hipLaunchKernelGGL(k_dl_count, dim3(blocks_for(n, 256)), dim3(256), 0, stream(), n, cpos.p, upos.p, ut.p, cnt.p);
    LT_CHECK_LAUNCH();
    LT_TRY(exclusive_scan_i64(cnt.p, off.p, n));
    int64_t nnz = 0;
    LT_TRY(off.download(&nnz, 1, n));
    LT_TRY(sync());
    if (nnz_out) *nnz_out = nnz;
    if (capacity >= 0 && nnz > capacity) {
        set_error("Delaunay lens matrix has %ld entries, caller capacity is %ld", (long)nnz, capacity);
        return LT_ERR_CAPACITY;
    }
    if (nnz > 0 && (!cols || !vals)) {
        set_error("lt_delaunay_lens_matrix_csr: NULL output");
        return LT_ERR_ARG;
    }
    DevBuf<int32_t> rp, ci;
    LT_TRY(rp.alloc(n + 1));
    LT_TRY(ci.alloc(std::max<long>(nnz, 1)));
    LT_TRY(va.alloc(std::max<long>(nnz, 1)));
    hipLaunchKernelGGL(k_dl_fill, dim3(blocks_for(n + 1, 256)), dim3(256), 0, stream(), n, cpos.p, upos.p, off.p,
                       pts.p, tris.p, up.p, ut.p, rp.p, ci.p, va.p);
    LT_CHECK_LAUNCH();
    LT_TRY(rp.download(row_inds, n + 1));
    if (nnz) {
        LT_TRY(ci.download(cols, nnz));
        LT_TRY(va.download(vals, nnz));
    }
    return sync();
}

void triangle_gradient_csr(delaunay_space ds, int *row_inds, int *cols, double *vals) {
    if (lt_triangle_gradient_csr(ds, row_inds, cols, vals) != LT_OK) fprintf(stderr, "lentils_hip: %s\n", last_error());
}

void delaunay_lens_matrix_csr(image_space im, delaunay_space ds, int ncast, double *uncasted_points,
                              int *uncasted_tri_inds, int *row_inds, int *cols, double *vals) {
    if (lt_delaunay_lens_matrix_csr(im, ds, ncast, uncasted_points, uncasted_tri_inds, row_inds, cols, vals, -1,
                                    nullptr) != LT_OK)
        fprintf(stderr, "lentils_hip: %s\n", last_error());
}

void triangle_geometry_from_inds(double *s, int *tri_inds, double *weights_out, int num_points, double *pos) {
    lt_status st = ensure_init();
    if (st == LT_OK && num_points > 0 && (!s || !tri_inds || !weights_out || !pos)) {
        set_error("triangle_geometry_from_inds: NULL argument");
        st = LT_ERR_ARG;
    }
    if (st != LT_OK || num_points <= 0) {
        if (st != LT_OK) fprintf(stderr, "lentils_hip: %s\n", last_error());
        return;
    }
    int vmax = 0;
    for (long k = 0; k < 3L * num_points; ++k) {
        if (tri_inds[k] < 0) {
            set_error("triangle_geometry_from_inds: negative vertex index");
            fprintf(stderr, "lentils_hip: %s\n", last_error());
            return;
        }
        vmax = std::max(vmax, tri_inds[k]);
    }
    DevBuf<double> ds, dp, dw;
    DevBuf<int> di;
    st = ds.upload(s, 2L * num_points);
    if (st == LT_OK) st = di.upload(tri_inds, 3L * num_points);
    if (st == LT_OK) st = dp.upload(pos, 2L * (vmax + 1));
    if (st == LT_OK) st = dw.alloc(3L * num_points);
    if (st == LT_OK) {
        hipLaunchKernelGGL(k_tri_geometry, dim3(blocks_for(num_points, 256)), dim3(256), 0, stream(),
                           (long)num_points, ds.p, di.p, dp.p, dw.p);
        st = hipGetLastError() == hipSuccess ? LT_OK : LT_ERR_HIP;
    }
    if (st == LT_OK) st = dw.download(weights_out, 3L * num_points);
    if (st == LT_OK) st = sync();
    if (st != LT_OK) fprintf(stderr, "lentils_hip: %s\n", last_error());
}

}  // extern "C"
