// FP64 lens deflection, written once for device and host (the host build is
// only used by the CPU self-test of the device arithmetic).
//
// Follows lentils/backend/deflect.c:26-57 (lens equation), :70-108 (shear),
// :179-239 (2F1 series, 20 terms) and :243-301 (PEMD).  The complex
// arithmetic is spelled out as gcc lowers C99 complex: products inline,
// real/complex quotients through Smith's method (libgcc __divdc3 up to
// gcc 11), csqrt by glibc's hypot-based algorithm.  Compile with
// -ffp-contract=off: no operation below may be fused.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#include "lt_internal.h"

namespace lt {

#define LT_HD __host__ __device__ inline

// per-lens constants of one mass-model record; returns false for an
// unsupported type.  PEMD: deflect.c:281-290 (per-lens part, host pow so the
// bits match the reference); shear: deflect.c:82-89 with the fpars layout of
// lens_models.py:129-148.
inline bool pack_lens_one(const generic_mass_model &m, DevLens &L) {
    const double *f = m.fpars;
    L = DevLens{};
    L.type = (int)m.type;
    L.x = f[0];
    L.y = f[1];
    if (m.type == 0) {
        const double q = f[2], t = 2 * f[7];
        L.q = q;
        L.qq = q * q;
        L.qt = (1 - q * q) / (q * q);
        L.t = t;
        L.tm2 = t - 2;
        L.b = pow(0.5 * f[6] * pow(q, t - 0.5) * (3.0 - t) / (2.0 - t), 1.0 / t);
        L.bb = L.b * L.b;
        L.cr = f[5];
        L.ci = -f[4];
        return true;
    }
    if (m.type == 1) {
        L.ss = f[2];
        L.c2 = -f[5] * f[5] + f[4] * f[4];
        L.s2 = -2 * f[4] * f[5];
        L.gks = f[6];
        L.s_gka = f[8];
        L.c_gka = f[9];
        L.gss = f[10];
        L.s_gsa = f[12];
        L.c_gsa = f[13];
        return true;
    }
    return false;
}

LT_HD void cmul(double a, double b, double c, double d, double &re, double &im) {
    re = a * c - b * d;
    im = a * d + b * c;
}

// real/complex quotient a / (c + i d) with b = 0 numerator imaginary part:
// Smith's method plus the C99 Annex G recovery of infinities and zeros.
LT_HD void cdiv(double a, double b, double c, double d, double &re, double &im) {
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
            a = copysign(isinf(a) ? 1.0 : 0.0, a);
            b = copysign(isinf(b) ? 1.0 : 0.0, b);
            x = INFINITY * (a * c + b * d);
            y = INFINITY * (b * c - a * d);
        } else if ((isinf(c) || isinf(d)) && isfinite(a) && isfinite(b)) {
            c = copysign(isinf(c) ? 1.0 : 0.0, c);
            d = copysign(isinf(d) ? 1.0 : 0.0, d);
            x = 0.0 * (a * c + b * d);
            y = 0.0 * (b * c - a * d);
        }
    }
    re = x;
    im = y;
}

// principal complex square root, glibc's algorithm for finite arguments
// (exact-zero parts handled separately, cancellation avoided through
// 2 Re Im = Im x).  Overflow/underflow rescaling is not needed for the
// |w2| <= ~1e3 arguments the series sees.
LT_HD void csqrt_d(double x, double y, double &re, double &im) {
    if (y == 0.0) {
        if (x < 0.0) {
            re = 0.0;
            im = copysign(sqrt(-x), y);
        } else {
            re = fabs(sqrt(x));
            im = copysign(0.0, y);
        }
        return;
    }
    if (x == 0.0) {
        double r = sqrt(0.5 * fabs(y));
        re = r;
        im = copysign(r, y);
        return;
    }
    double d = hypot(x, y), r, s;
    if (x > 0.0) {
        r = sqrt(0.5 * (d + x));
        s = 0.5 * (y / r);
    } else {
        s = sqrt(0.5 * (d - x));
        r = fabs(0.5 * (y / s));
    }
    re = r;
    im = copysign(s, y);
}

// x^y for the radial factor (b/rs)^(t-2).  pow is kept even for y = -1:
// the reference calls pow, and glibc's pow(x, -1) differs from 1/x in the
// last bit for ~0.1% of arguments.
LT_HD double pow_radial(double x, double y) { return pow(x, y); }

// 2F1(1, 2-t; 2-t/2; u) by 20 terms of its series.  deflect.c:179-239
LT_HD void hyp2f1_series(const DevLens &L, double zr, double zi, double &fr, double &fi) {
    double x = zr, y = zi;
    double r2 = L.qq * x * x + y * y;
    double zzr, zzi, w2r, w2i;
    cmul(x, y, x, y, zzr, zzi);
    cdiv(L.qt * r2, 0.0, zzr, zzi, w2r, w2i);
    double sr, si;
    csqrt_d(1.0 - w2r, -w2i, sr, si);
    double ur = 0.5 * (1.0 - sr);
    double ui = 0.5 * (-si);
    double acc_r = 0.0, acc_i = 0.0, pr = 1.0, pi = 0.0, an = 1.0;
    const double t = L.t;
#pragma unroll
    for (int n = 0; n < 20; n++) {
        acc_r += an * pr;
        acc_i += an * pi;
        double den = (double)(2 * n + 4) - t;
        double nr, ni;
        cmul(pr, pi, ur, ui, nr, ni);
        pr = nr;
        pi = ni;
        an *= (den - t) / den;
    }
    fr = acc_r;
    fi = acc_i;
}

// PEMD deflection accumulated into (ax, ay).  deflect.c:243-301
LT_HD void pemd_deflection(const DevLens &L, double px, double py, double &ax, double &ay) {
    double zr, zi;
    cmul(px - L.x, py - L.y, L.cr, L.ci, zr, zi);
    double rs2 = L.qq * zr * zr + zi * zi;
    double rs = sqrt(rs2);
    double dr, di;
    cdiv(L.bb, 0.0, L.q * zr, L.q * zi, dr, di);
    double pw = pow_radial(L.b / rs, L.tm2);
    double Ar = dr * pw, Ai = di * pw;
    double Fr, Fi;
    hyp2f1_series(L, zr, zi, Fr, Fi);
    double gr, gi, hr, hi;
    cmul(Ar, Ai, Fr, Fi, gr, gi);
    cmul(gr, gi, L.cr, L.ci, hr, hi);
    ax += hr;
    ay += -hi;
}

// external shear + Keeton (2001) convergence/shear gradient terms.
// deflect.c:70-108
LT_HD void shear_deflection(const DevLens &L, double px, double py, double &ax, double &ay) {
    double sx = px - L.x, sy = py - L.y;
    ax += L.ss * (L.c2 * sx + L.s2 * sy);
    ay += L.ss * (L.s2 * sx - L.c2 * sy);
    double xr = sx * L.c_gka + sy * L.s_gka;
    double yr = -sx * L.s_gka + sy * L.c_gka;
    double ex = -0.25 * L.gks * (3 * xr * xr + yr * yr);
    double ey = -0.5 * L.gks * xr * yr;
    ax += ex * L.c_gka - ey * L.s_gka;
    ay += ex * L.s_gka + ey * L.c_gka;
    xr = sx * L.c_gsa + sy * L.s_gsa;
    yr = -sx * L.s_gsa + sy * L.c_gsa;
    ex = -0.5 * L.gss * (xr * xr - yr * yr);
    ey = 1.0 * L.gss * xr * yr;
    ax += ex * L.c_gsa - ey * L.s_gsa;
    ay += ex * L.s_gsa + ey * L.c_gsa;
}

// single-plane lens equation: every component is evaluated at the unlensed
// position and subtracted in order.  deflect.c:26-57
LT_HD void deflect_point(const DevLens *lenses, int n, double px, double py, double &ox, double &oy) {
    ox = px;
    oy = py;
    for (int i = 0; i < n; ++i) {
        double ax = 0.0, ay = 0.0;
        if (lenses[i].type == 0)
            pemd_deflection(lenses[i], px, py, ax, ay);
        else
            shear_deflection(lenses[i], px, py, ax, ay);
        ox -= ax;
        oy -= ay;
    }
}

// coordinate of stencil lattice index k along one axis: even k are pixel
// corners xmin + (k/2) dx, odd k are midpoints 0.5 (x_i + x_{i+1}), with the
// expressions of raster.c:289-347 so that shared points agree bitwise
LT_HD double lattice_coord(double xmin, double dx, int k) {
    int i = k >> 1;
    if (k & 1) return 0.5 * ((xmin + i * dx) + (xmin + (i + 1) * dx));
    return xmin + i * dx;
}

}  // namespace lt
