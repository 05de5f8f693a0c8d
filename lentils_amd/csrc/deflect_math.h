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

// The series with its derivatives (the want_deriv branch of
// deflect.c:179-239): F, and d[0..7] = dF/dt, dF/dq, dF/dx, dF/dy (complex,
// interleaved; x, y the rotated, centred coordinates).  Mixed real/complex
// products componentwise, complex quotients by cdiv, as gcc evaluates the
// reference.
LT_HD void hyp2f1_series_d(const DevLens &L, double zr, double zi, double &fr, double &fi, double d[8]) {
    const double x = zr, y = zi, q = L.q, t = L.t, qt = L.qt;
    double r2 = L.qq * x * x + y * y;
    double zzr, zzi, w2r, w2i;
    cmul(x, y, x, y, zzr, zzi);
    cdiv(qt * r2, 0.0, zzr, zzi, w2r, w2i);
    double sr, si;
    csqrt_d(1.0 - w2r, -w2i, sr, si);
    double ur = 0.5 * (1.0 - sr);
    double ui = 0.5 * (-si);
    double acc_r = 0.0, acc_i = 0.0, pr = 1.0, pi = 0.0, an = 1.0;
    double dtr = 0.0, dti = 0.0, dur = 0.0, dui = 0.0, da = 0.0, p1r = 0.0, p1i = 0.0;
#pragma unroll
    for (int n = 0; n < 20; n++) {
        acc_r += an * pr;
        acc_i += an * pi;
        double den = (double)(2 * n + 4) - t;
        dtr += da * pr;  // df_dt += da_dt u_n
        dti += da * pi;
        const double ann = an * (double)n;  // df_du += a_n n u_{n-1}
        dur += ann * p1r;
        dui += ann * p1i;
        p1r = pr;
        p1i = pi;
        da = (double)(-2 * (2 + n)) / (den * den) * an + 2.0 * ((double)(2 + n) - t) / den * da;
        double nr, ni;
        cmul(pr, pi, ur, ui, nr, ni);
        pr = nr;
        pi = ni;
        an *= (den - t) / den;
    }
    fr = acc_r;
    fi = acc_i;
    double qr, qi, gr, gi;
    cdiv(1.0, 0.0, 4.0 * sr, 4.0 * si, qr, qi);  // du/dw2 = 1 / (4 sqrt(1 - w2))
    cmul(dur, dui, qr, qi, gr, gi);              // dF/dw2
    const double q3 = q * q * q;
    double ar, ai, br, bi;
    cmul(q3 * zr, q3 * zi, zr, zi, ar, ai);
    cdiv(-2.0 * (q * q * q * q * x * x + y * y), 0.0, ar, ai, br, bi);  // dw2/dq
    double z3r, z3i;
    cmul(zzr, zzi, zr, zi, z3r, z3i);
    const double c1 = -2.0 * qt * y, c2 = -2.0 * qt * x;
    double xr_, xi_, yr_, yi_;
    cdiv(c1 * y, c1 * (-q * q * x), z3r, z3i, xr_, xi_);         // dw2/dx
    cdiv(-(y * c2), (q * q * x) * c2, z3r, z3i, yr_, yi_);       // dw2/dy
    d[0] = dtr;
    d[1] = dti;
    cmul(gr, gi, br, bi, d[2], d[3]);
    cmul(gr, gi, xr_, xi_, d[4], d[5]);
    cmul(gr, gi, yr_, yi_, d[6], d[7]);
}

// PEMD deflection accumulated into (ax, ay) and its derivatives
// d alpha / d(b_eff, qh, q, x, y, th) as (x, y) pairs into dv[12]
// (deflect.c:243-338; b_eff the internal radius of deflect.c:284, held fixed
// for the other five, th in degrees).
LT_HD void pemd_deflection_d(const DevLens &L, double px, double py, double &ax, double &ay, double dv[12]) {
    const double q = L.q, t = L.t, b = L.b, cs = L.cr, sn = -L.ci;
    double zr, zi;
    cmul(px - L.x, py - L.y, L.cr, L.ci, zr, zi);
    const double x = zr, y = zi;
    double rs2 = L.qq * zr * zr + zi * zi;
    double rs = sqrt(rs2);
    double dr, di;
    cdiv(L.bb, 0.0, q * zr, q * zi, dr, di);
    double pw = pow_radial(b / rs, L.tm2);
    double Ar = dr * pw, Ai = di * pw;
    double Fr, Fi, fd[8];
    hyp2f1_series_d(L, zr, zi, Fr, Fi, fd);
    double gr, gi, hr, hi;
    cmul(Ar, Ai, Fr, Fi, gr, gi);
    cmul(gr, gi, L.cr, L.ci, hr, hi);
    const double alr = hr, ali = -hi;
    ax += alr;
    ay += ali;
    const double Abr = Ar * t / b, Abi = Ai * t / b;
    const double lg = log(b / rs);
    const double Atr = Ar * lg, Ati = Ai * lg;
    const double kq = (t - 1.0) * q * q * x * x + y * y, qrs = q * rs2;
    const double Aqr = -Ar * kq / qrs, Aqi = -Ai * kq / qrs;
    const double zrsr = zr * rs2, zrsi = zi * rs2, qqx = q * q * x;
    double nr_, ni_, Axr, Axi, Ayr, Ayi;
    cmul(-Ar, -Ai, y * y + qqx * (x * (t - 1.0)), qqx * ((t - 2.0) * y), nr_, ni_);
    cdiv(nr_, ni_, zrsr, zrsi, Axr, Axi);
    cmul(-Ar, -Ai, y * ((t - 2.0) * x), q * q * x * x + y * ((t - 1.0) * y), nr_, ni_);
    cdiv(nr_, ni_, zrsr, zrsi, Ayr, Ayi);
    double ur, ui, vr, vi, wr, wi;
    cmul(Fr, Fi, Abr, Abi, ur, ui);
    cmul(ur, ui, L.cr, L.ci, wr, wi);
    dv[0] += wr;
    dv[1] += -wi;
    cmul(Fr, Fi, Atr, Ati, ur, ui);
    cmul(Ar, Ai, fd[0], fd[1], vr, vi);
    cmul(ur + vr, ui + vi, L.cr, L.ci, wr, wi);
    dv[2] += 2.0 * wr;
    dv[3] += 2.0 * (-wi);
    cmul(Fr, Fi, Aqr, Aqi, ur, ui);
    cmul(Ar, Ai, fd[2], fd[3], vr, vi);
    cmul(ur + vr, ui + vi, L.cr, L.ci, wr, wi);
    dv[4] += wr;
    dv[5] += -wi;
    cmul(Fr, Fi, Axr, Axi, ur, ui);
    cmul(Ar, Ai, fd[4], fd[5], vr, vi);
    cmul(ur + vr, ui + vi, L.cr, L.ci, wr, wi);
    const double dxsr = wr, dxsi = -wi;
    cmul(Fr, Fi, Ayr, Ayi, ur, ui);
    cmul(Ar, Ai, fd[6], fd[7], vr, vi);
    cmul(ur + vr, ui + vi, L.cr, L.ci, wr, wi);
    const double dysr = wr, dysi = -wi;
    dv[6] += -(dxsr * cs - dysr * sn);
    dv[7] += -(dxsi * cs - dysi * sn);
    dv[8] += -(dxsr * sn + dysr * cs);
    dv[9] += -(dxsi * sn + dysi * cs);
    const double k_deg = M_PI / 180;
    dv[10] += k_deg * (dxsr * y - dysr * x + -ali);
    dv[11] += k_deg * (dxsi * y - dysi * x + alr);
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
