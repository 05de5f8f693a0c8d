/*
 * lentils_oracle.c -- test infrastructure, not product code.
 *
 * An FP64 CPU restatement of the Lentils forward-model C backend:
 *   - PEMD deflection by the truncated 2F1 series, external shear
 *       (lentils/backend/deflect.c:26-123, 179-339)
 *   - pixel -> source triangle rasterization into the CSR lens matrix
 *       (lentils/backend/raster.c:28-395)
 *   - Kaiser-Bessel gridding / degridding (lentils/backend/nufft.c:11-122)
 *   - DFT, curvature and gradient finite-difference CSR matrices
 *       (lentils/backend/nufft.c:128-229)
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library, and only as the checker or the CPU timing baseline.
 *
 * Arithmetic is written out component by component in the same operation
 * order as the reference so that results are bit-identical to a gcc build of
 * the reference (this is checked against oracle/_ref by the tests).  Build
 * with -O2 -fwrapv -ffp-contract=off: the reference relies on x86 wrapping
 * int arithmetic and on uncontracted double arithmetic.
 *
 * Known divergence: I0 is computed here by its power series instead of GSL's
 * gsl_sf_bessel_I0 (absent from this image).  Agreement with
 * scipy.special.i0 is ~1e-16 relative; the exact GSL bits are unpinned.
 */
#include <math.h>
#include <complex.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "oracle.h"

#define OR_ARCSEC 4.8481368111e-6     /* lentils/backend/common.h:7 */
#define OR_TWO_PI 6.28318530718       /* lentils/backend/common.h:9 (truncated) */

/* ------------------------------------------------------------------------- */
/* complex helpers with the exact lowering gcc uses for C99 complex           */
/* ------------------------------------------------------------------------- */

/* product of two general complex numbers: gcc expands it inline as
 * (ac - bd, ad + bc) and only calls __muldc3 when both parts are NaN. */
static inline void or_cmul(double a, double b, double c, double d, double *re, double *im) {
    double x = a * c - b * d;
    double y = a * d + b * c;
    if (isnan(x) && isnan(y)) {
        double complex r = (a + I * b) * (c + I * d); /* defers to libgcc recovery */
        x = creal(r);
        y = cimag(r);
    }
    *re = x;
    *im = y;
}

/* real / complex: gcc lowers it to __divdc3(a, 0, c, d).  Smith's method as
 * implemented by libgcc up to gcc 11, including its NaN recovery. */
static void or_cdiv(double a, double b, double c, double d, double *re, double *im) {
    double ratio, denom, x, y;
    if (fabs(c) < fabs(d)) {
        ratio = c / d;
        denom = (c * ratio) + d;
        x = ((a * ratio) + b) / denom;
        y = ((b * ratio) - a) / denom;
    } else {
        ratio = d / c;
        denom = (d * ratio) + c;
        x = ((b * ratio) + a) / denom;
        y = (b - (a * ratio)) / denom;
    }
    if (isnan(x) && isnan(y)) {
        if (c == 0.0 && d == 0.0 && (!isnan(a) || !isnan(b))) {
            x = copysign(INFINITY, c) * a;
            y = copysign(INFINITY, c) * b;
        } else if ((isinf(a) || isinf(b)) && isfinite(c) && isfinite(d)) {
            a = copysign(isinf(a) ? 1 : 0, a);
            b = copysign(isinf(b) ? 1 : 0, b);
            x = INFINITY * (a * c + b * d);
            y = INFINITY * (b * c - a * d);
        } else if ((isinf(c) || isinf(d)) && isfinite(a) && isfinite(b)) {
            c = copysign(isinf(c) ? 1 : 0, c);
            d = copysign(isinf(d) ? 1 : 0, d);
            x = 0.0 * (a * c + b * d);
            y = 0.0 * (b * c - a * d);
        }
    }
    *re = x;
    *im = y;
}

/* ------------------------------------------------------------------------- */
/* deflection                                                                 */
/* ------------------------------------------------------------------------- */

/* 2F1(1, 2-t; 2-t/2; u) by its 20-term series, u = (1 - sqrt(1 - w2))/2,
 * w2 = ((1-q^2)/q^2) (q^2 x^2 + y^2) / z^2.   deflect.c:179-239 */
static void or_hyp2f1(double t, double q, double zr, double zi, double *fr, double *fi) {
    double x = zr, y = zi;
    double r2 = q * q * x * x + y * y;
    double qt = (1 - q * q) / (q * q);
    double zzr, zzi, w2r, w2i;
    or_cmul(x, y, x, y, &zzr, &zzi);
    or_cdiv(qt * r2, 0.0, zzr, zzi, &w2r, &w2i);
    double complex s = csqrt((1.0 - w2r) + I * (-w2i));
    double ur = 0.5 * (1.0 - creal(s));
    double ui = 0.5 * (-cimag(s));
    double sr = 0.0, si = 0.0, pr = 1.0, pi = 0.0, an = 1.0;
    for (int n = 0; n < 20; n++) {
        sr += an * pr;
        si += an * pi;
        double den = 2 * n + 4 - t;
        double nr, ni;
        or_cmul(pr, pi, ur, ui, &nr, &ni);
        pr = nr;
        pi = ni;
        an *= (den - t) / den;
    }
    *fr = sr;
    *fi = si;
}

/* PEMD deflection added into a[2].  deflect.c:243-301 (want_deriv = 0) */
static void or_pemd(const or_mass_model *m, const double *p, double *a) {
    const double *f = m->fpars;
    double lx = f[0], ly = f[1], q = f[2], sn = f[4], cs = f[5], bb = f[6], qh = f[7];
    double t = 2 * qh;
    double b = pow(0.5 * bb * pow(q, t - 0.5) * (3.0 - t) / (2.0 - t), 1.0 / t);
    double cr = cs, ci = -sn;
    double zr, zi;
    or_cmul(p[0] - lx, p[1] - ly, cr, ci, &zr, &zi);
    double x = zr, y = zi;
    double rs2 = q * q * x * x + y * y;
    double rs = sqrt(rs2);
    double dr, di;
    or_cdiv(b * b, 0.0, q * zr, q * zi, &dr, &di);
    double pw = pow(b / rs, t - 2);
    double Ar = dr * pw, Ai = di * pw;
    double Fr, Fi;
    or_hyp2f1(t, q, zr, zi, &Fr, &Fi);
    double gr, gi, hr, hi;
    or_cmul(Ar, Ai, Fr, Fi, &gr, &gi);
    or_cmul(gr, gi, cr, ci, &hr, &hi);
    a[0] += hr;
    a[1] += -hi;
}

/* The same series with its derivatives (the want_deriv branch of
 * deflect.c:179-239): F, and d[0..7] = dF/dt, dF/dq, dF/dx, dF/dy (complex,
 * interleaved), where x, y are the rotated, centred coordinates.  The
 * recurrences follow deflect.c:213-224 term by term; the chain rule
 * deflect.c:227-236, with the real/complex mixed products the way gcc
 * evaluates them (componentwise) and every complex quotient by Smith's
 * method. */
static void or_hyp2f1_d(double t, double q, double zr, double zi, double *fr, double *fi, double *d) {
    double x = zr, y = zi;
    double r2 = q * q * x * x + y * y;
    double qt = (1 - q * q) / (q * q);
    double zzr, zzi, w2r, w2i;
    or_cmul(x, y, x, y, &zzr, &zzi);
    or_cdiv(qt * r2, 0.0, zzr, zzi, &w2r, &w2i);
    double complex s = csqrt((1.0 - w2r) + I * (-w2i));
    double ur = 0.5 * (1.0 - creal(s));
    double ui = 0.5 * (-cimag(s));
    double sr = 0.0, si = 0.0, pr = 1.0, pi = 0.0, an = 1.0;
    double dtr = 0.0, dti = 0.0, dur = 0.0, dui = 0.0, da = 0.0, p1r = 0.0, p1i = 0.0;
    for (int n = 0; n < 20; n++) {
        sr += an * pr;
        si += an * pi;
        double den = 2 * n + 4 - t;
        dtr += da * pr; /* df_dt += da_dt * u_n */
        dti += da * pi;
        double ann = an * n; /* df_du += a_n * n * u_{n-1} */
        dur += ann * p1r;
        dui += ann * p1i;
        p1r = pr;
        p1i = pi;
        da = -2 * (2 + n) / (den * den) * an + 2 * (2 + n - t) / den * da;
        double nr, ni;
        or_cmul(pr, pi, ur, ui, &nr, &ni);
        pr = nr;
        pi = ni;
        an *= (den - t) / den;
    }
    *fr = sr;
    *fi = si;
    /* du/dw2 = 1 / (4 sqrt(1 - w2)); dF/dw2 = dF/du du/dw2 */
    double qr, qi, gr, gi;
    or_cdiv(1.0, 0.0, 4 * creal(s), 4 * cimag(s), &qr, &qi);
    or_cmul(dur, dui, qr, qi, &gr, &gi);
    /* dw2/dq = -2 (q^4 x^2 + y^2) / (q^3 z^2) */
    double q3 = q * q * q, ar, ai, br, bi, cr, ci;
    or_cmul(q3 * zr, q3 * zi, zr, zi, &ar, &ai);
    or_cdiv(-2 * (q * q * q * q * x * x + y * y), 0.0, ar, ai, &br, &bi);
    /* z^3 */
    double z3r, z3i;
    or_cmul(zzr, zzi, zr, zi, &z3r, &z3i);
    /* dw2/dx = -2 q_ y (y - i q^2 x) / z^3 */
    double c1 = -2 * qt * y;
    double xr_, xi_;
    or_cdiv(c1 * y, c1 * (-q * q * x), z3r, z3i, &xr_, &xi_);
    /* dw2/dy = -2 q_ i x (q^2 x + i y) / z^3: (i c2) (q^2 x + i y) = (-c2 y, c2 q^2 x) */
    double c2 = -2 * qt * x;
    double yr_, yi_;
    or_cdiv(-(y * c2), (q * q * x) * c2, z3r, z3i, &yr_, &yi_);
    d[0] = dtr;
    d[1] = dti;
    or_cmul(gr, gi, br, bi, &cr, &ci);
    d[2] = cr;
    d[3] = ci;
    or_cmul(gr, gi, xr_, xi_, &cr, &ci);
    d[4] = cr;
    d[5] = ci;
    or_cmul(gr, gi, yr_, yi_, &cr, &ci);
    d[6] = cr;
    d[7] = ci;
}

/* PEMD deflection added into a[2] and its parameter derivatives added into
 * dv[12] = d alpha / d(b, qh, f, x, y, th) as (x, y) pairs, the want_deriv
 * branch of deflect.c:243-338 (b the record's Einstein-radius parameter, qh
 * the slope (t = 2 qh), f the axis ratio, x/y the lens centre, th in
 * degrees). */
static void or_pemd_d(const or_mass_model *m, const double *p, double *a, double *dv) {
    const double *f = m->fpars;
    double lx = f[0], ly = f[1], q = f[2], sn = f[4], cs = f[5], bb = f[6], qh = f[7];
    double t = 2 * qh;
    double b = pow(0.5 * bb * pow(q, t - 0.5) * (3.0 - t) / (2.0 - t), 1.0 / t);
    double cr = cs, ci = -sn;
    double zr, zi;
    or_cmul(p[0] - lx, p[1] - ly, cr, ci, &zr, &zi);
    double x = zr, y = zi;
    double rs2 = q * q * x * x + y * y;
    double rs = sqrt(rs2);
    double dr, di;
    or_cdiv(b * b, 0.0, q * zr, q * zi, &dr, &di);
    double pw = pow(b / rs, t - 2);
    double Ar = dr * pw, Ai = di * pw;
    double Fr, Fi, fd[8];
    or_hyp2f1_d(t, q, zr, zi, &Fr, &Fi, fd);
    double gr, gi, hr, hi;
    or_cmul(Ar, Ai, Fr, Fi, &gr, &gi);
    or_cmul(gr, gi, cr, ci, &hr, &hi);
    double alr = hr, ali = -hi; /* alpha = conj(A F crot) */
    a[0] += alr;
    a[1] += ali;
    /* dA/db = A t / b, dA/dt = A log(b / rs) */
    double Abr = Ar * t / b, Abi = Ai * t / b;
    double lg = log(b / rs);
    double Atr = Ar * lg, Ati = Ai * lg;
    /* dA/dq = -A ((t-1) q^2 x^2 + y^2) / (q rs2) */
    double kq = (t - 1) * q * q * x * x + y * y, qrs = q * rs2;
    double Aqr = -Ar * kq / qrs, Aqi = -Ai * kq / qrs;
    /* dA/dx = -A (y^2 + q^2 x (x (t-1) + i (t-2) y)) / (z rs2) */
    double zrsr = zr * rs2, zrsi = zi * rs2, nr_, ni_, Axr, Axi, Ayr, Ayi;
    double qqx = q * q * x;
    or_cmul(-Ar, -Ai, y * y + qqx * (x * (t - 1)), qqx * ((t - 2) * y), &nr_, &ni_);
    or_cdiv(nr_, ni_, zrsr, zrsi, &Axr, &Axi);
    /* dA/dy = -A (i q^2 x^2 + y ((t-2) x + i (t-1) y)) / (z rs2) */
    or_cmul(-Ar, -Ai, y * ((t - 2) * x), q * q * x * x + y * ((t - 1) * y), &nr_, &ni_);
    or_cdiv(nr_, ni_, zrsr, zrsi, &Ayr, &Ayi);
    /* chain rule, conjugate, rotation back */
    double ur, ui, vr, vi, wr, wi;
    or_cmul(Fr, Fi, Abr, Abi, &ur, &ui);
    or_cmul(ur, ui, cr, ci, &wr, &wi);
    double dbr = wr, dbi = -wi;
    or_cmul(Fr, Fi, Atr, Ati, &ur, &ui);
    or_cmul(Ar, Ai, fd[0], fd[1], &vr, &vi);
    or_cmul(ur + vr, ui + vi, cr, ci, &wr, &wi);
    double dtr = 2 * wr, dti = 2 * (-wi);
    or_cmul(Fr, Fi, Aqr, Aqi, &ur, &ui);
    or_cmul(Ar, Ai, fd[2], fd[3], &vr, &vi);
    or_cmul(ur + vr, ui + vi, cr, ci, &wr, &wi);
    double dqr = wr, dqi = -wi;
    or_cmul(Fr, Fi, Axr, Axi, &ur, &ui);
    or_cmul(Ar, Ai, fd[4], fd[5], &vr, &vi);
    or_cmul(ur + vr, ui + vi, cr, ci, &wr, &wi);
    double dxsr = wr, dxsi = -wi;
    or_cmul(Fr, Fi, Ayr, Ayi, &ur, &ui);
    or_cmul(Ar, Ai, fd[6], fd[7], &vr, &vi);
    or_cmul(ur + vr, ui + vi, cr, ci, &wr, &wi);
    double dysr = wr, dysi = -wi;
    double dxr = -(dxsr * cs - dysr * sn), dxi = -(dxsi * cs - dysi * sn);
    double dyr = -(dxsr * sn + dysr * cs), dyi = -(dxsi * sn + dysi * cs);
    /* d/dth = pi/180 (dalpha/dxs y - dalpha/dys x + i alpha) */
    double thr = M_PI / 180 * (dxsr * y - dysr * x + -ali);
    double thi = M_PI / 180 * (dxsi * y - dysi * x + alr);
    dv[0] += dbr;
    dv[1] += dbi;
    dv[2] += dtr;
    dv[3] += dti;
    dv[4] += dqr;
    dv[5] += dqi;
    dv[6] += dxr;
    dv[7] += dxi;
    dv[8] += dyr;
    dv[9] += dyi;
    dv[10] += thr;
    dv[11] += thi;
}

/* external shear plus Keeton-2001 gradient terms.  deflect.c:70-108 */
static void or_shear(const or_mass_model *m, const double *p, double *a) {
    const double *f = m->fpars;
    double sx = p[0] - f[0], sy = p[1] - f[1];
    double ss = f[2], s_sa = f[4], c_sa = f[5];
    double gks = f[6], s_gka = f[8], c_gka = f[9];
    double gss = f[10], s_gsa = f[12], c_gsa = f[13];
    double c2 = -c_sa * c_sa + s_sa * s_sa;
    double s2 = -2 * s_sa * c_sa;
    a[0] += ss * (c2 * sx + s2 * sy);
    a[1] += ss * (s2 * sx - c2 * sy);
    double xr = sx * c_gka + sy * s_gka;
    double yr = -sx * s_gka + sy * c_gka;
    double ex = -0.25 * gks * (3 * xr * xr + yr * yr);
    double ey = -0.5 * gks * xr * yr;
    a[0] += ex * c_gka - ey * s_gka;
    a[1] += ex * s_gka + ey * c_gka;
    xr = sx * c_gsa + sy * s_gsa;
    yr = -sx * s_gsa + sy * c_gsa;
    ex = -0.5 * gss * (xr * xr - yr * yr);
    ey = 1.0 * gss * xr * yr;
    a[0] += ex * c_gsa - ey * s_gsa;
    a[1] += ex * s_gsa + ey * c_gsa;
}

/* single-plane lens equation x_out = x_in - sum_i alpha_i(x_in).
 * deflect.c:26-57.  Returns nonzero on an unsupported lens type. */
static int or_deflect_one(or_lens_model glm, const double *xin, double *xout) {
    xout[0] = xin[0];
    xout[1] = xin[1];
    for (int i = 0; i < glm.num_lenses; ++i) {
        double a[2] = {0.0, 0.0};
        const or_mass_model *m = &glm.lenses[i];
        if (m->type == 0)
            or_pemd(m, xin, a);
        else if (m->type == 1)
            or_shear(m, xin, a);
        else
            return 1;
        xout[0] -= a[0];
        xout[1] -= a[1];
    }
    return 0;
}

int oracle_deflect_points(or_lens_model glm, const double *pin, int n, double *pout) {
    int bad = 0;
    for (int p = 0; p < n; ++p) bad |= or_deflect_one(glm, &pin[2 * p], &pout[2 * p]);
    return bad;
}

/* deflect_points with want_deriv = 1 (deflect.c:26-66, 304-338): every
 * point's PEMD derivatives are added into the same deriv[0..11] (points in
 * order, components in order), as the reference does; per_point (may be
 * NULL) receives each point's own sums over the PEMD components,
 * [point][12].  The shear's derivative block is disabled in the reference
 * (deflect.c:111-122), so shears add nothing. */
int oracle_deflect_points_deriv(or_lens_model glm, const double *pin, int n, double *pout, double *deriv,
                                double *per_point) {
    int bad = 0;
    for (int p = 0; p < n; ++p) {
        const double *xin = &pin[2 * p];
        double *xout = &pout[2 * p];
        double own[12] = {0};
        xout[0] = xin[0];
        xout[1] = xin[1];
        for (int i = 0; i < glm.num_lenses; ++i) {
            double a[2] = {0.0, 0.0};
            const or_mass_model *m = &glm.lenses[i];
            if (m->type == 0) {
                double dv[12] = {0};
                or_pemd_d(m, xin, a, dv);
                for (int k = 0; k < 12; ++k) {
                    deriv[k] += dv[k];
                    own[k] += dv[k];
                }
            } else if (m->type == 1) {
                or_shear(m, xin, a);
            } else {
                bad = 1;
            }
            xout[0] -= a[0];
            xout[1] -= a[1];
        }
        if (per_point) memcpy(&per_point[12 * p], own, sizeof(own));
    }
    return bad;
}

/* ------------------------------------------------------------------------- */
/* Delaunay source planes (lentils/backend/triangles.c)                      */
/* ------------------------------------------------------------------------- */

/* barycentric weights of point s in triangle tri (3 x 2) and its signed
 * doubled area, triangles.c:7-31 */
static void or_tri_geometry(const double *s, const double *tri, double *w_out, double *area) {
    double a0 = tri[0] - s[0], a1 = tri[1] - s[1];
    double b0 = tri[2] - s[0], b1 = tri[3] - s[1];
    double c0 = tri[4] - s[0], c1 = tri[5] - s[1];
    double w0 = b0 * c1 - c0 * b1;
    double w1 = c0 * a1 - a0 * c1;
    double w2 = a0 * b1 - b0 * a1;
    *area = w0 + w1 + w2;
    double inv = 1.0 / (*area);
    w_out[0] = w0 * inv;
    w_out[1] = w1 * inv;
    w_out[2] = w2 * inv;
}

/* triangles.c:35-62 */
void oracle_triangle_geometry_from_inds(const double *s, const int *tri_inds, double *w_out, int n, const double *pos) {
    for (int i = 0; i < n; ++i) {
        double tri[6], a;
        for (int v = 0; v < 3; ++v) {
            tri[2 * v] = pos[2 * tri_inds[3 * i + v]];
            tri[2 * v + 1] = pos[2 * tri_inds[3 * i + v] + 1];
        }
        or_tri_geometry(&s[2 * i], tri, &w_out[3 * i], &a);
    }
}

/* gradient rows of the linear interpolant on each triangle, weighted by
 * sqrt|area| / (2 area): two rows (x, y) of three entries per triangle,
 * triangles.c:66-103 */
void oracle_triangle_gradient_csr(int ntri, const double *pts, const int *tris, int *row_inds, int *cols,
                                  double *vals) {
    int i = 0, row = 0;
    for (int t = 0; t < ntri; ++t) {
        int v0 = tris[3 * t], v1 = tris[3 * t + 1], v2 = tris[3 * t + 2];
        double p0x = pts[2 * v0], p0y = pts[2 * v0 + 1];
        double p1x = pts[2 * v1], p1y = pts[2 * v1 + 1];
        double p2x = pts[2 * v2], p2y = pts[2 * v2 + 1];
        double area = 0.5 * (p0y * (p1x - p2x) + p1y * p2x - p1x * p2y + p0x * (-p1y + p2y));
        double weight = 1.0 / (2 * area);
        weight *= sqrt(fabs(area));
        row_inds[row++] = i;
        vals[i] = (-p1y + p2y) * weight; cols[i++] = v0;
        vals[i] = (p0y - p2y) * weight; cols[i++] = v1;
        vals[i] = (-p0y + p1y) * weight; cols[i++] = v2;
        row_inds[row++] = i;
        vals[i] = (p1x - p2x) * weight; cols[i++] = v0;
        vals[i] = (-p0x + p2x) * weight; cols[i++] = v1;
        vals[i] = (p0x - p1x) * weight; cols[i++] = v2;
    }
    row_inds[row] = i;
}

/* lens matrix of a Delaunay source plane, triangles.c:106-145: per image
 * pixel (x-major) nothing when masked out, one entry 1.0 on the next casted
 * vertex for pixels with i % ncast == j % ncast == 0, otherwise the
 * barycentric weights of the pixel's deflected position in its triangle
 * (nothing when it fell outside the triangulation, tri index < 0) */
void oracle_delaunay_lens_matrix_csr(or_image_space im, int ntri, const double *pts, const int *tris, int ncast,
                                     const double *unc_pts, const int *unc_tri, int *row_inds, int *cols,
                                     double *vals) {
    int row = 0, i = 0, cast_idx = 0, unc_idx = 0;
    for (int iim = 0; iim < im.nx; ++iim)
        for (int jim = 0; jim < im.ny; ++jim, ++row) {
            row_inds[row] = i;
            if (!im.mask[row]) continue;
            if (iim % ncast || jim % ncast) {
                const double *s = &unc_pts[2 * unc_idx];
                int t = unc_tri[unc_idx++];
                if (t < 0) continue;
                double tri[6], w[3], a;
                int p[3];
                for (int v = 0; v < 3; ++v) {
                    p[v] = tris[3 * t + v];
                    tri[2 * v] = pts[2 * p[v]];
                    tri[2 * v + 1] = pts[2 * p[v] + 1];
                }
                or_tri_geometry(s, tri, w, &a);
                for (int v = 0; v < 3; ++v) {
                    vals[i] = w[v];
                    cols[i++] = p[v];
                }
            } else {
                vals[i] = 1.0;
                cols[i++] = cast_idx++;
            }
        }
    row_inds[row] = i;
    (void)ntri;
}

/* ------------------------------------------------------------------------- */
/* rasterization                                                              */
/* ------------------------------------------------------------------------- */

#define OR_STACK_CAP 64          /* raster.c:11 */
#define OR_FRAG_CAP 128000       /* raster.c:12 */
#define OR_POP_LIMIT (1 << 22)   /* guard: the reference can loop forever */

/* x86-64 cvttsd2si semantics: out-of-range and NaN give INT_MIN.  This is
 * what the reference's (int)floor(...) compiles to. */
static inline int or_f2i(double v) {
    if (!(v >= -2147483648.0 && v < 2147483648.0)) return INT32_MIN;
    return (int)v;
}

typedef struct {
    double v[3][2];   /* vertex positions in the source plane */
    double b[3][2];   /* barycentric coordinates w.r.t. the parent triangle */
    int lo[2], hi[2]; /* covered source-pixel index range [lo, hi) per axis */
} or_tri;

typedef struct {
    double area;
    int idx;
} or_frag;

/* How a triangle is cut by a grid line, indexed by the 3-bit mask of
 * vertices on the + side (raster.c:135-142).  These tables fix both the
 * geometry and the order fragments are emitted in, so they are kept as the
 * reference has them. */
static const int or_ncut[8] = {0, 2, 2, 2, 2, 2, 2, 0};
static const int or_cut_edge[8][2][2] = {{{0, 0}, {0, 0}}, {{0, 1}, {0, 2}}, {{0, 1}, {1, 2}},
                                         {{0, 2}, {1, 2}}, {{0, 2}, {1, 2}}, {{0, 1}, {1, 2}},
                                         {{0, 1}, {0, 2}}, {{0, 0}, {0, 0}}};
static const int or_nchild[8] = {1, 3, 3, 3, 3, 3, 3, 1};
static const int or_child_vert[8][3][3] = {
    {{0, 1, 2}, {0, 0, 0}, {0, 0, 0}}, {{1, 2, 4}, {1, 4, 3}, {0, 3, 4}},
    {{0, 3, 4}, {0, 4, 2}, {1, 4, 3}}, {{2, 3, 4}, {0, 4, 3}, {0, 1, 4}},
    {{0, 1, 4}, {0, 4, 3}, {2, 3, 4}}, {{1, 4, 3}, {0, 3, 4}, {0, 4, 2}},
    {{0, 3, 4}, {1, 2, 4}, {1, 4, 3}}, {{0, 1, 2}, {0, 0, 0}, {0, 0, 0}}};
static const int or_child_side[8][3] = {{0, 0, 0}, {0, 0, 1}, {0, 0, 1}, {0, 1, 1},
                                        {0, 0, 1}, {0, 1, 1}, {0, 1, 1}, {1, 0, 0}};

typedef struct {
    or_frag *buf;
    int n, cap;
    int status;
} or_fragbuf;

static void or_emit(or_fragbuf *fb, double area, int idx) {
    if (fb->n >= fb->cap) {
        int nc = fb->cap ? 2 * fb->cap : 1024;
        fb->buf = (or_frag *)realloc(fb->buf, (size_t)nc * sizeof(or_frag));
        fb->cap = nc;
    }
    fb->buf[fb->n].area = area;
    fb->buf[fb->n].idx = idx;
    fb->n++;
    if (fb->n >= OR_FRAG_CAP) fb->status |= OR_RASTER_FRAG_OVERFLOW;
}

/* Depth-first recursive split of one source-plane triangle along source
 * grid lines until every piece covers one pixel.  raster.c:120-269 */
static void or_raster_tri(const double *tv, or_image_space g, or_fragbuf *fb) {
    or_tri stk[OR_STACK_CAP];
    or_tri tri;
    const double cell[2] = {g.dx, g.dy};
    const double org[2] = {g.xmin, g.ymin};
    memset(&tri, 0, sizeof(tri));
    tri.lo[0] = g.nx;
    tri.lo[1] = g.ny;
    tri.hi[0] = 0;
    tri.hi[1] = 0;
    tri.b[1][1] = 1.0;
    tri.b[2][0] = 1.0;
    for (int v = 0; v < 3; ++v) {
        tri.v[v][0] = tv[2 * v];
        tri.v[v][1] = tv[2 * v + 1];
        int i = or_f2i(floor((tri.v[v][0] - g.xmin) / g.dx));
        int j = or_f2i(floor((tri.v[v][1] - g.ymin) / g.dy));
        if (i < tri.lo[0]) tri.lo[0] = i;
        if (j < tri.lo[1]) tri.lo[1] = j;
        if (i + 1 > tri.hi[0]) tri.hi[0] = i + 1;
        if (j + 1 > tri.hi[1]) tri.hi[1] = j + 1;
    }
    int ns = 0;
    long pops = 0;
    stk[ns++] = tri;
    while (ns > 0) {
        if (++pops > OR_POP_LIMIT) {
            fb->status |= OR_RASTER_POP_LIMIT;
            return;
        }
        tri = stk[--ns];
        int ax = 0, ext = 0;
        for (int a = 0; a < 2; ++a) {
            int s = tri.hi[a] - tri.lo[a];
            if (s > ext) {
                ext = s;
                ax = a;
            }
        }
        int cut = tri.lo[ax] + ext / 2;
        double cutc = cut * cell[ax] + org[ax];
        if (ext == 1) {
            int i = tri.lo[0], j = tri.lo[1];
            if (i < 0 || i >= g.nx || j < 0 || j >= g.ny) continue;
            double x1 = tri.b[0][0], y1 = tri.b[0][1];
            double x2 = tri.b[1][0], y2 = tri.b[1][1];
            double x3 = tri.b[2][0], y3 = tri.b[2][1];
            or_emit(fb, -0.25 * (x1 * (y2 - y3) + x2 * (y3 - y1) + x3 * (y1 - y2)), g.ny * i + j);
            continue;
        }
        double pv[5][2], pb[5][2], dist[3];
        memset(pv, 0, sizeof(pv));
        memset(pb, 0, sizeof(pb));
        memcpy(pv, tri.v, sizeof(tri.v));
        memcpy(pb, tri.b, sizeof(tri.b));
        int mask = 0;
        for (int v = 0; v < 3; ++v) {
            dist[v] = pv[v][ax] - cutc;
            mask |= ((dist[v] > 0.) << v);
        }
        const int side_idx[3] = {tri.lo[ax], cut, tri.hi[ax]};
        for (int k = 0; k < or_ncut[mask]; ++k) {
            int e0 = or_cut_edge[mask][k][0], e1 = or_cut_edge[mask][k][1];
            /* point on edge (e0, e1) where it crosses the cut line */
            double w0 = dist[e1], w1 = -dist[e0];
            double wn = 1.0 / (w0 + w1);
            double a0 = w0 * wn, a1 = w1 * wn;
            pv[3 + k][0] = a0 * pv[e0][0] + a1 * pv[e1][0];
            pv[3 + k][1] = a0 * pv[e0][1] + a1 * pv[e1][1];
            pb[3 + k][0] = a0 * pb[e0][0] + a1 * pb[e1][0];
            pb[3 + k][1] = a0 * pb[e0][1] + a1 * pb[e1][1];
        }
        for (int c = 0; c < or_nchild[mask]; ++c) {
            int s = or_child_side[mask][c];
            tri.lo[ax] = side_idx[s];
            tri.hi[ax] = side_idx[s + 1];
            if (tri.hi[0] <= 0 || tri.lo[0] >= g.nx) continue;
            if (tri.hi[1] <= 0 || tri.lo[1] >= g.ny) continue;
            for (int v = 0; v < 3; ++v) {
                int src = or_child_vert[mask][c][v];
                tri.v[v][0] = pv[src][0];
                tri.v[v][1] = pv[src][1];
                tri.b[v][0] = pb[src][0];
                tri.b[v][1] = pb[src][1];
            }
            stk[ns++] = tri;
            if (ns >= OR_STACK_CAP) {
                fb->status |= OR_RASTER_STACK_OVERFLOW;
                return;
            }
        }
    }
}

/* stable merge sort of fragments by source-pixel index (glibc qsort is a
 * stable merge sort, which the reference's reduction depends on) */
static void or_sort_frags(or_frag *a, or_frag *tmp, int n) {
    if (n < 2) return;
    int h = n / 2;
    or_sort_frags(a, tmp, h);
    or_sort_frags(a + h, tmp, n - h);
    int i = 0, j = h, k = 0;
    while (i < h && j < n) tmp[k++] = (a[j].idx < a[i].idx) ? a[j++] : a[i++];
    while (i < h) tmp[k++] = a[i++];
    while (j < n) tmp[k++] = a[j++];
    memcpy(a, tmp, (size_t)n * sizeof(or_frag));
}

/* Source positions of the 9-point stencil of image pixel (i, j): corners,
 * edge midpoints and centre, in the order raster.c:319-347 uses. */
static void or_stencil(or_image_space im, int i, int j, double pts[9][2]) {
    double x0 = im.xmin + i * im.dx, x1 = im.xmin + (i + 1) * im.dx;
    double y0 = im.ymin + j * im.dy, y1 = im.ymin + (j + 1) * im.dy;
    double xm = 0.5 * (x0 + x1), ym = 0.5 * (y0 + y1);
    double px[9] = {x0, xm, x1, x1, x1, xm, x0, x0, xm};
    double py[9] = {y0, y0, y0, ym, y1, y1, y1, ym, ym};
    for (int k = 0; k < 9; ++k) {
        pts[k][0] = px[k];
        pts[k][1] = py[k];
    }
}

/* Shared core: rows of the lens matrix from given deflected stencils.
 * defl(i,j,k) supplies source-plane stencil point k of pixel (i,j).
 * If row_inds is NULL only the per-row entry counts are returned in nnz_row.
 * raster.c:272-395 */
typedef void (*or_stencil_fn)(void *ctx, int i, int j, double out[9][2]);

static int or_lens_matrix_core(or_image_space im, or_image_space src, or_stencil_fn fn, void *ctx,
                               int *row_inds, int *cols, double *vals, long *nnz_out, int drop) {
    or_fragbuf fb = {0};
    or_frag *tmp = NULL;
    int tmpcap = 0;
    int stale = 0; /* grididx left in fragbuffer[0] by the last non-empty row */
    long k = 0;
    int row = 0;
    int status = 0;
    for (int i = 0; i < im.nx; ++i) {
        for (int j = 0; j < im.ny; ++j, ++row) {
            if (row_inds) row_inds[row] = (int)k;
            if (!im.mask[row]) continue;
            double sp[9][2];
            fn(ctx, i, j, sp);
            fb.n = 0;
            for (int t = 0; t < 8; ++t) {
                double tv[6] = {sp[(t + 1) % 8][0], sp[(t + 1) % 8][1], sp[t][0], sp[t][1], sp[8][0], sp[8][1]};
                or_raster_tri(tv, src, &fb);
            }
            if (fb.n > tmpcap) {
                tmpcap = fb.n;
                tmp = (or_frag *)realloc(tmp, (size_t)tmpcap * sizeof(or_frag));
            }
            or_sort_frags(fb.buf, tmp, fb.n);
            if (fb.n > 0) stale = fb.buf[0].idx;
            /* reduce: a new column group starts at 0.0 without its first
             * fragment (raster.c:375-391) */
            double val = 0.0;
            int col = stale;
            for (int f = 0; f < fb.n; ++f) {
                if (fb.buf[f].idx != col) {
                    if (cols) {
                        vals[k] = val;
                        cols[k] = col;
                    }
                    k++;
                    val = 0.0;
                    col = fb.buf[f].idx;
                    if (!drop) val += fb.buf[f].area;
                } else {
                    val += fb.buf[f].area;
                }
            }
            if (cols) {
                vals[k] = val;
                cols[k] = col;
            }
            k++;
        }
    }
    if (row_inds) row_inds[row] = (int)k;
    status |= fb.status;
    free(fb.buf);
    free(tmp);
    if (nnz_out) *nnz_out = k;
    return status;
}

typedef struct {
    or_image_space im;
    or_lens_model glm;
    int bad;
} or_defl_ctx;

static void or_stencil_deflect(void *vctx, int i, int j, double out[9][2]) {
    or_defl_ctx *c = (or_defl_ctx *)vctx;
    double pts[9][2];
    or_stencil(c->im, i, j, pts);
    for (int k = 0; k < 9; ++k) c->bad |= or_deflect_one(c->glm, pts[k], out[k]);
}

/* Full lens-matrix build: deflect the 9-point stencil of every unmasked
 * pixel, rasterize 8 fan triangles, reduce to CSR.  Passing NULL for
 * row_inds/cols/vals only counts entries (nnz_out).  Returns status bits. */
int oracle_lens_matrix_csr(or_image_space im, or_image_space src, or_lens_model glm, int *row_inds, int *cols,
                           double *vals, long *nnz_out) {
    or_defl_ctx c = {im, glm, 0};
    int st = or_lens_matrix_core(im, src, or_stencil_deflect, &c, row_inds, cols, vals, nnz_out, 1);
    return st | (c.bad ? OR_RASTER_BAD_LENS : 0);
}

/* Same, but from a precomputed deflected lattice: lattice[(2nx+1)(2ny+1)][2],
 * lattice point (a, b) = stencil coordinates x index a, y index b. */
typedef struct {
    const double *lat;
    int ly;
} or_lat_ctx;

static void or_stencil_lattice(void *vctx, int i, int j, double out[9][2]) {
    or_lat_ctx *c = (or_lat_ctx *)vctx;
    static const int dxo[9] = {0, 1, 2, 2, 2, 1, 0, 0, 1};
    static const int dyo[9] = {0, 0, 0, 1, 2, 2, 2, 1, 1};
    for (int k = 0; k < 9; ++k) {
        long a = 2 * i + dxo[k], b = 2 * j + dyo[k];
        out[k][0] = c->lat[2 * (a * c->ly + b)];
        out[k][1] = c->lat[2 * (a * c->ly + b) + 1];
    }
}

int oracle_lens_matrix_from_lattice(or_image_space im, or_image_space src, const double *lattice, int *row_inds,
                                    int *cols, double *vals, long *nnz_out) {
    or_lat_ctx c = {lattice, 2 * im.ny + 1};
    return or_lens_matrix_core(im, src, or_stencil_lattice, &c, row_inds, cols, vals, nnz_out, 1);
}

/* Same reduction without the reference's dropped first fragment of each
 * later column group: every entry is then the exact clipped area of the
 * fan triangles over that source pixel (x 0.25 per triangle).  Used only to
 * pin the rasterization geometry against brute-force polygon clipping. */
int oracle_lens_matrix_from_lattice_nodrop(or_image_space im, or_image_space src, const double *lattice,
                                           int *row_inds, int *cols, double *vals, long *nnz_out) {
    or_lat_ctx c = {lattice, 2 * im.ny + 1};
    return or_lens_matrix_core(im, src, or_stencil_lattice, &c, row_inds, cols, vals, nnz_out, 0);
}

/* image-plane lattice of stencil points, (2nx+1) x (2ny+1), x-major, with
 * the same expressions the per-pixel stencil uses */
void oracle_stencil_lattice(or_image_space im, double *pts) {
    int lx = 2 * im.nx + 1, ly = 2 * im.ny + 1;
    for (int a = 0; a < lx; ++a) {
        int i = a / 2;
        double x = (a & 1) ? 0.5 * ((im.xmin + i * im.dx) + (im.xmin + (i + 1) * im.dx)) : im.xmin + i * im.dx;
        for (int b = 0; b < ly; ++b) {
            int j = b / 2;
            double y = (b & 1) ? 0.5 * ((im.ymin + j * im.dy) + (im.ymin + (j + 1) * im.dy)) : im.ymin + j * im.dy;
            pts[2 * ((long)a * ly + b)] = x;
            pts[2 * ((long)a * ly + b) + 1] = y;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Kaiser-Bessel gridding                                                     */
/* ------------------------------------------------------------------------- */

/* modified Bessel function I0 by its power series (all terms positive) */
double oracle_bessel_i0(double x) {
    double q = 0.25 * x * x, term = 1.0, sum = 1.0;
    for (int k = 1; k < 500; ++k) {
        term *= q / ((double)k * (double)k);
        sum += term;
        if (term < 1e-17 * sum) break;
    }
    return sum;
}

/* separable KB kernel value.  nufft.c:11-21 */
double oracle_kb_g(double ak, double aq, double beta, double w) {
    if (!(ak >= -w && ak <= w && aq >= -w && aq <= w)) return 0.0;
    double sk = 1.0 - (ak / w) * (ak / w);
    double sq = 1.0 - (aq / w) * (aq / w);
    if (sk < 0.0) sk = 0.0;
    if (sq < 0.0) sq = 0.0;
    double bk = beta * sqrt(sk), bq = beta * sqrt(sq);
    return 1.0 / (4.0 * w * w) * oracle_bessel_i0(bk) * oracle_bessel_i0(bq);
}

/* direction 0: degrid (grid -> vis, accumulated into vis);
 * direction 1: grid (vis -> grid, accumulated into grid).
 * nthreads > 1 parallelises over visibilities with atomics, as the
 * reference does.  nufft.c:42-122 */
void oracle_grid(or_visibility_space vs, double *vis, or_fourier_space gs, double *grid, int wsup, double beta,
                 int direction, int nthreads) {
    const double w = (double)wsup;
    for (int ch = 0; ch < vs.nchannels; ++ch) {
        const double chs = vs.channels[ch];
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(static)
        for (int i = 0; i < vs.nrows; ++i) {
            double u = vs.uv[3 * i + 0] * chs;
            double v = -vs.uv[3 * i + 1] * chs;
            double ph = OR_TWO_PI * (u * gs.gcx + v * gs.gcy);
            double cp = cos(ph), sp = sin(ph);
            int nu0 = (int)round(u / gs.du);
            int nv0 = (int)round(v / gs.dv);
            long vi = (long)vs.nrows * ch + i;
            for (int a = nu0 - wsup; a <= nu0 + wsup; ++a) {
                for (int b = nv0 - wsup; b <= nv0 + wsup; ++b) {
                    double kv = oracle_kb_g(u / gs.du - a, v / gs.dv - b, beta, w);
                    double sgn = 1 - 2 * ((a + b) & 1);
                    double wr = 0.5 * kv * sgn, wi = 0.5 * kv * sgn;
                    int iu = a, iv = b;
                    while (iu < 0) iu += gs.nu;
                    while (iv < 0) iv += gs.nv;
                    while (iu >= gs.nu) iu -= gs.nu;
                    while (iv >= gs.nv) iv -= gs.nv;
                    if (iv >= gs.half_nv) {
                        iu = (gs.nu - iu) % gs.nu;
                        iv = gs.nv - iv;
                        wi *= -1;
                    }
                    long fi = (long)gs.half_nv * iu + iv;
                    if (direction == 0) {
                        double tr = wr * grid[2 * fi], ti = wi * grid[2 * fi + 1];
                        double dr = 2 * (cp * tr + sp * ti);
                        double di = 2 * (-sp * tr + cp * ti);
#pragma omp atomic
                        vis[2 * vi] += dr;
#pragma omp atomic
                        vis[2 * vi + 1] += di;
                    } else {
                        if (iv == 0) {
                            wr *= 2;
                            wi *= 2;
                        }
                        double xr = vis[2 * vi], xi = vis[2 * vi + 1];
                        double gr = wr * (cp * xr - sp * xi);
                        double gi = wi * (sp * xr + cp * xi);
#pragma omp atomic
                        grid[2 * fi] += gr;
#pragma omp atomic
                        grid[2 * fi + 1] += gi;
                    }
                }
            }
        }
    }
}

/* dense DFT as CSR: row 2k = cos, 2k+1 = sin of -2 pi ARCSEC (u x + v y),
 * image points are pixel corners.  nufft.c:128-161 */
void oracle_dft_matrix_csr(or_visibility_space vs, or_image_space im, int *row_inds, int *cols, double *vals) {
    long k = 0;
    int row;
    for (row = 0; row < 2 * vs.nrows; ++row) {
        int vi = row / 2, odd = row % 2;
        row_inds[row] = (int)k;
        double u = vs.uv[3 * vi] * vs.channels[0];
        double v = -vs.uv[3 * vi + 1] * vs.channels[0];
        int col = 0;
        for (int i = 0; i < im.nx; ++i) {
            for (int j = 0; j < im.ny; ++j, ++col) {
                if (!im.mask[col]) continue;
                double x = i * im.dx + im.xmin;
                double y = j * im.dy + im.ymin;
                double arg = -2.0 * M_PI * OR_ARCSEC * (u * x + v * y);
                vals[k] = odd ? sin(arg) : cos(arg);
                cols[k] = col;
                k++;
            }
        }
    }
    row_inds[row] = (int)k;
}

/* 5-point Laplacian; boundary rows keep the -4 diagonal.  nufft.c:197-229 */
void oracle_curv_matrix_csr(or_image_space im, int *row_inds, int *cols, double *vals) {
    long k = 0;
    int row = 0;
    double h = 1.0 / (im.dx * im.dy);
    for (int i = 0; i < im.nx; ++i) {
        for (int j = 0; j < im.ny; ++j, ++row) {
            row_inds[row] = (int)k;
            if (i > 0) { vals[k] = h; cols[k++] = im.ny * (i - 1) + j; }
            if (j > 0) { vals[k] = h; cols[k++] = im.ny * i + (j - 1); }
            vals[k] = -4.0 * h;
            cols[k++] = im.ny * i + j;
            if (j < im.ny - 1) { vals[k] = h; cols[k++] = im.ny * i + (j + 1); }
            if (i < im.nx - 1) { vals[k] = h; cols[k++] = im.ny * (i + 1) + j; }
        }
    }
    row_inds[row] = (int)k;
}

/* forward differences, x block then y block.  nufft.c:163-194 */
void oracle_grad_matrix_csr(or_image_space im, int *row_inds, int *cols, double *vals) {
    long k = 0;
    int row = 0;
    for (int i = 0; i < im.nx; ++i)
        for (int j = 0; j < im.ny; ++j, ++row) {
            row_inds[row] = (int)k;
            if (i < im.nx - 1) {
                vals[k] = -1.0 / im.dx; cols[k++] = im.ny * i + j;
                vals[k] = 1.0 / im.dx; cols[k++] = im.ny * (i + 1) + j;
            }
        }
    for (int i = 0; i < im.nx; ++i)
        for (int j = 0; j < im.ny; ++j, ++row) {
            row_inds[row] = (int)k;
            if (j < im.ny - 1) {
                vals[k] = -1.0 / im.dy; cols[k++] = im.ny * i + j;
                vals[k] = 1.0 / im.dy; cols[k++] = im.ny * i + (j + 1);
            }
        }
    row_inds[row] = (int)k;
}

/* ------------------------------------------------------------------------- */
/* CSR helpers for the CPU baseline                                           */
/* ------------------------------------------------------------------------- */

void oracle_csr_matvec(int nrows, const int *rp, const int *ci, const double *v, const double *x, double *y,
                       int nthreads) {
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(static)
    for (int r = 0; r < nrows; ++r) {
        double s = 0.0;
        for (int k = rp[r]; k < rp[r + 1]; ++k) s += v[k] * x[ci[k]];
        y[r] = s;
    }
}

int oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
