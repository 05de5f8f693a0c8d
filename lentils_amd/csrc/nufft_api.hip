// C ABI of the NUFFT plan and the grid_cpu drop-in symbol.
#include <cmath>
#include <memory>

#include "nufft_plan.h"

namespace lt {
lt_status grid_c128_to_plan(NufftPlan &P, const double *src, void *dst);
lt_status grid_plan_add_c128(NufftPlan &P, const void *src, double *dst);

// fourier_space of the zero-padded image, space_base.py:62-84 and
// fourier_operators.py:27-41
fourier_space padded_fourier_space(const image_space &im, int pad) {
    const double arcsec = 4.8481368111e-6;
    const int padx = ((pad - 1) * im.nx + 1) / 2, pady = ((pad - 1) * im.ny + 1) / 2;
    const double xmin = im.xmin - padx * im.dx, xmax = im.xmax + padx * im.dx;
    const double ymin = im.ymin - pady * im.dy, ymax = im.ymax + pady * im.dy;
    fourier_space fs{};
    fs.nu = im.nx + 2 * padx;
    fs.nv = im.ny + 2 * pady;
    fs.half_nv = fs.nv / 2 + 1;
    fs.nchannels = 1;
    fs.nstokes = 1;
    fs.du = 1.0 / (arcsec * (xmax - xmin));
    fs.dv = 1.0 / (arcsec * (ymax - ymin));
    fs.gcx = arcsec * 0.5 * (xmax + xmin);
    fs.gcy = arcsec * 0.5 * (ymax + ymin);
    fs.channels = nullptr;
    return fs;
}

double default_kb_beta(int wsup, int pad) {  // fourier_operators.py:108
    const double a = 2.0 * wsup / pad;
    return M_PI * std::sqrt(a * a * (pad - 0.5) * (pad - 0.5) - 0.8);
}

}  // namespace lt

using namespace lt;

struct lt_nufft_s {
    NufftPlan plan;
};

namespace lt {
namespace {

// I0 by its power series and the separable KB kernel, nufft.c:11-21 (the
// reference evaluates I0 with GSL)
__device__ double wrap_i0(double x) {
    const double q = 0.25 * x * x;
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < 500; ++k) {
        term *= q / ((double)k * (double)k);
        sum += term;
        if (term < 1e-17 * sum) break;
    }
    return sum;
}

__device__ double wrap_kb(double ak, double aq, double beta, double w) {
    if (!(ak >= -w && ak <= w && aq >= -w && aq <= w)) return 0.0;
    double sk = 1.0 - (ak / w) * (ak / w);
    double sq = 1.0 - (aq / w) * (aq / w);
    if (sk < 0.0) sk = 0.0;
    if (sq < 0.0) sq = 0.0;
    return 1.0 / (4.0 * w * w) * wrap_i0(beta * sqrt(sk)) * wrap_i0(beta * sqrt(sq));
}

struct WrapGeo {
    int nu, nv, half_nv, nrows, nch, wsup;
    double du, dv, gcx, gcy, beta;
};

// grid_cpu's loop for one visibility (nufft.c:54-117), taps wrapped
// periodically: direction 0 adds into this visibility (its taps in order),
// direction 1 adds into the grid with FP64 atomics (the order of additions
// from different visibilities varies, as in the reference's OpenMP loop)
__global__ void k_grid_wrap(WrapGeo G, const double *__restrict__ uv, const double *__restrict__ chan,
                            int direction, double *__restrict__ vis, double *__restrict__ grid) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)G.nrows * G.nch) return;
    const int ch = (int)(e / G.nrows);
    const long i = e - (long)ch * G.nrows;
    const double u = uv[3 * i] * chan[ch];
    const double v = -uv[3 * i + 1] * chan[ch];
    const double ph = 6.28318530718 * (u * G.gcx + v * G.gcy);
    const double cosp = cos(ph), sinp = sin(ph);
    const double tu = u / G.du, tv = v / G.dv;
    if (!(fabs(tu) < 1e9 && fabs(tv) < 1e9)) return;
    const int ngu = (int)round(tu), ngv = (int)round(tv);
    double vr = direction == 0 ? vis[2 * e] : 0.0, vi = direction == 0 ? vis[2 * e + 1] : 0.0;
    const double xr = direction == 1 ? vis[2 * e] : 0.0, xi = direction == 1 ? vis[2 * e + 1] : 0.0;
    for (int a = ngu - G.wsup; a <= ngu + G.wsup; ++a)
        for (int b = ngv - G.wsup; b <= ngv + G.wsup; ++b) {
            const double kern = wrap_kb(tu - a, tv - b, G.beta, (double)G.wsup);
            const double pc = (double)(1 - 2 * ((a + b) & 1));
            double wr = 0.5 * kern * pc, wi = 0.5 * kern * pc;
            int iu = ((a % G.nu) + G.nu) % G.nu, iv = ((b % G.nv) + G.nv) % G.nv;
            if (iv >= G.half_nv) {
                iu = (G.nu - iu) % G.nu;
                iv = G.nv - iv;
                wi *= -1;
            }
            const long f = (long)G.half_nv * iu + iv;
            if (direction == 0) {
                const double tr = wr * grid[2 * f], ti = wi * grid[2 * f + 1];
                vr += 2 * (cosp * tr + sinp * ti);
                vi += 2 * (-sinp * tr + cosp * ti);
            } else {
                if (iv == 0) {
                    wr *= 2;
                    wi *= 2;
                }
                atomicAdd(&grid[2 * f], wr * (cosp * xr - sinp * xi));
                atomicAdd(&grid[2 * f + 1], wi * (sinp * xr + cosp * xi));
            }
        }
    if (direction == 0) {
        vis[2 * e] = vr;
        vis[2 * e + 1] = vi;
    }
}

}  // namespace

lt_status grid_wrap(const visibility_space &vs, double *vis, const fourier_space &gs, double *grid, int wsup,
                    double kb_beta, int direction) {
    LT_TRY(ensure_init());
    if (vs.nrows < 0 || vs.nchannels <= 0 || gs.nu <= 0 || gs.nv <= 0 || gs.half_nv != gs.nv / 2 + 1 || wsup < 1 ||
        (direction != 0 && direction != 1)) {
        set_error("grid_cpu: bad arguments");
        return LT_ERR_ARG;
    }
    const long nvis = (long)vs.nrows * vs.nchannels, ng = 2L * gs.nu * gs.half_nv;
    DevBuf<double> duv, dch, dv, dg;
    LT_TRY(duv.upload(vs.uv, std::max(3L * vs.nrows, 1L)));
    LT_TRY(dch.upload(vs.channels, vs.nchannels));
    LT_TRY(dv.upload(vis, std::max(2 * nvis, 1L)));
    LT_TRY(dg.upload(grid, ng));
    const WrapGeo G{gs.nu, gs.nv, gs.half_nv, vs.nrows, vs.nchannels, wsup, gs.du, gs.dv, gs.gcx, gs.gcy, kb_beta};
    if (nvis > 0) {
        hipLaunchKernelGGL(k_grid_wrap, dim3(blocks_for(nvis, 128)), dim3(128), 0, stream(), G, duv.p, dch.p,
                           direction, dv.p, dg.p);
        LT_CHECK_LAUNCH();
    }
    if (direction == 0)
        LT_TRY(dv.download(vis, 2 * nvis));
    else
        LT_TRY(dg.download(grid, ng));
    return sync();
}

NufftPlan *nufft_of(lt_nufft h) { return h ? &h->plan : nullptr; }
lt_status stencil_time(NufftPlan &P, int nb, int reps, double *ms, double *bytes, double *flops, long *rows,
                       long *win_cells);
}  // namespace lt

extern "C" {

lt_status lt_nufft_create(visibility_space vs, image_space im, int pad, int wsup, double beta, int precision,
                          lt_nufft *out) {
    if (!out || im.nx <= 0 || im.ny <= 0 || vs.nrows < 0 || vs.nchannels <= 0 || !vs.channels ||
        (vs.nrows > 0 && !vs.uv)) {
        set_error("lt_nufft_create: bad arguments");
        return LT_ERR_ARG;
    }
    if (beta <= 0) beta = default_kb_beta(wsup, pad);
    auto h = std::make_unique<lt_nufft_s>();
    fourier_space fs = padded_fourier_space(im, pad);
    LT_TRY(h->plan.create(vs, fs, wsup, beta, precision));
    LT_TRY(h->plan.attach_image(im, pad));
    *out = h.release();
    return LT_OK;
}

lt_status lt_nufft_info(lt_nufft h, int *nu, int *nv, int *half_nv, double *beta, long *nvis) {
    if (!h) return LT_ERR_ARG;
    if (nu) *nu = h->plan.nu;
    if (nv) *nv = h->plan.nv;
    if (half_nv) *half_nv = h->plan.half_nv;
    if (beta) *beta = h->plan.beta;
    if (nvis) *nvis = h->plan.nvis;
    return LT_OK;
}

lt_status lt_nufft_forward_host(lt_nufft h, const double *img, double *vis) {
    if (!h) return LT_ERR_ARG;
    NufftPlan &P = h->plan;
    DevBuf<double> out;
    LT_TRY(P.imgbuf.upload(img, (long)P.nx * P.ny));
    LT_TRY(out.alloc(2 * std::max<long>(P.nvis, 1)));
    LT_TRY(P.forward(P.imgbuf.p, P.visbuf.p, nullptr));
    LT_TRY(P.from_sorted(P.visbuf.p, out.p, false));
    LT_TRY(out.download(vis, 2 * P.nvis));
    return sync();
}

lt_status lt_nufft_adjoint_host(lt_nufft h, const double *vis, double *img) {
    if (!h) return LT_ERR_ARG;
    NufftPlan &P = h->plan;
    DevBuf<double> in;
    LT_TRY(in.upload(vis, 2 * P.nvis));
    LT_TRY(P.to_sorted(in.p, P.visbuf.p));
    LT_TRY(P.adjoint(P.visbuf.p, nullptr, P.imgbuf.p, 1.0, false));
    LT_TRY(P.imgbuf.download(img, (long)P.nx * P.ny));
    return sync();
}

lt_status lt_nufft_degrid_host(lt_nufft h, const double *grid, double *vis) {
    if (!h) return LT_ERR_ARG;
    NufftPlan &P = h->plan;
    const long ng = 2L * P.nu * P.half_nv;
    DevBuf<double> g, v;
    LT_TRY(g.upload(grid, ng));
    LT_TRY(v.upload(vis, 2 * P.nvis));
    LT_TRY(grid_c128_to_plan(P, g.p, P.spec.p));
    LT_TRY(P.degrid(P.spec.p, P.visbuf.p, nullptr, false));
    LT_TRY(P.from_sorted(P.visbuf.p, v.p, true));
    LT_TRY(v.download(vis, 2 * P.nvis));
    return sync();
}

lt_status lt_nufft_grid_host(lt_nufft h, const double *vis, double *grid) {
    if (!h) return LT_ERR_ARG;
    NufftPlan &P = h->plan;
    const long ng = 2L * P.nu * P.half_nv;
    DevBuf<double> g, v;
    LT_TRY(g.upload(grid, ng));
    LT_TRY(v.upload(vis, 2 * P.nvis));
    LT_TRY(P.to_sorted(v.p, P.visbuf.p));
    LT_TRY(P.grid(P.visbuf.p, nullptr, P.spec.p));
    LT_TRY(grid_plan_add_c128(P, P.spec.p, g.p));
    LT_TRY(g.download(grid, ng));
    return sync();
}

lt_status lt_kb_taps(double kb_beta, int wsup, int precision, double *coeffs, double *max_err) {
    if (wsup < 1 || wsup > 8 || !(kb_beta > 0.0) || (precision != LT_F64 && precision != LT_F32)) {
        set_error("lt_kb_taps: bad arguments");
        return LT_ERR_ARG;
    }
    std::vector<double> mono;
    const double err = kb_taps_fit(kb_beta, wsup, tap_degree(precision == LT_F64, wsup), mono);
    if (coeffs) std::copy(mono.begin(), mono.end(), coeffs);
    if (max_err) *max_err = err;
    return LT_OK;
}

lt_status lt_kb_taps_window(double kb_beta, int wsup, int precision, double *coeffs, double *max_err) {
    if (wsup < 1 || wsup > 8 || !(kb_beta > 0.0) || (precision != LT_F64 && precision != LT_F32)) {
        set_error("lt_kb_taps_window: bad arguments");
        return LT_ERR_ARG;
    }
    std::vector<double> mono;
    const double err = kb_taps_fit_half(kb_beta, wsup, tap_degree(precision == LT_F64, wsup), mono);
    if (coeffs) std::copy(mono.begin(), mono.end(), coeffs);
    if (max_err) *max_err = err;
    return LT_OK;
}

lt_status lt_nufft_time(lt_nufft h, int direction, int reps, double *ms) {
    if (!h || reps <= 0 || !ms) return LT_ERR_ARG;
    NufftPlan &P = h->plan;
    hipEvent_t e0, e1;
    LT_HIP(hipEventCreate(&e0));
    LT_HIP(hipEventCreate(&e1));
    // warm-up launch, then reps launches bracketed by events on the plan's stream
    if (direction == 0)
        LT_TRY(P.degrid(P.spec.p, P.visbuf.p, nullptr, false));
    else
        LT_TRY(P.grid(P.visbuf.p, nullptr, P.spec.p));
    LT_HIP(hipEventRecord(e0, stream()));
    for (int r = 0; r < reps; ++r) {
        if (direction == 0)
            LT_TRY(P.degrid(P.spec.p, P.visbuf.p, nullptr, false));
        else
            LT_TRY(P.grid(P.visbuf.p, nullptr, P.spec.p));
    }
    LT_HIP(hipEventRecord(e1, stream()));
    LT_HIP(hipEventSynchronize(e1));
    float t = 0;
    LT_HIP(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = (double)t / reps;
    return LT_OK;
}

lt_status lt_nufft_stencil_time(lt_nufft h, int nb, int reps, double *ms, double *bytes, double *flops, long *rows,
                                long *win_cells) {
    if (!h) return LT_ERR_ARG;
    LT_TRY(ensure_init());
    return stencil_time(h->plan, nb, reps, ms, bytes, flops, rows, win_cells);
}

lt_status lt_nufft_stencil_form(lt_nufft h, int *form) {
    if (!h || !form) return LT_ERR_ARG;
    const NormalStencil *St = h->plan.wstencil.get();
    *form = h->plan.force_nb != 0 ? h->plan.force_nb : (St ? St->nb_sel : 0);
    return LT_OK;
}

lt_status lt_nufft_destroy(lt_nufft h) {
    delete h;
    return LT_OK;
}

void grid_cpu(visibility_space vs, double *vis, fourier_space gs, double *grid, int wsup, double kb_beta,
              int direction) {
    NufftPlan P;
    lt_status s = P.create(vs, gs, wsup, kb_beta, LT_F64);
    if (s == LT_ERR_NYQUIST) {
        // taps beyond the band: the reference wraps them (nufft.c:83-86)
        s = grid_wrap(vs, vis, gs, grid, wsup, kb_beta, direction);
        if (s != LT_OK) fprintf(stderr, "lentils_hip: grid_cpu: %s\n", last_error());
        return;
    }
    if (s == LT_OK) {
        const long ng = 2L * P.nu * P.half_nv;
        DevBuf<double> g, v;
        s = g.upload(grid, ng);
        if (s == LT_OK) s = v.upload(vis, 2 * P.nvis);
        if (s == LT_OK) {
            if (direction == 0) {
                s = P.degrid(g.p, P.visbuf.p, nullptr, false);
                if (s == LT_OK) s = P.from_sorted(P.visbuf.p, v.p, true);
                if (s == LT_OK) s = v.download(vis, 2 * P.nvis);
            } else {
                s = P.to_sorted(v.p, P.visbuf.p);
                if (s == LT_OK) s = P.grid(P.visbuf.p, nullptr, P.spec.p);
                if (s == LT_OK) s = grid_plan_add_c128(P, P.spec.p, g.p);
                if (s == LT_OK) s = g.download(grid, ng);
            }
        }
        if (s == LT_OK) s = sync();
    }
    if (s != LT_OK) fprintf(stderr, "lentils_hip: grid_cpu: %s\n", last_error());
}

}  // extern "C"
