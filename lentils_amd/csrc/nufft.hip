// Kaiser-Bessel NUFFT on gfx950: degridding, atomic-free (in HBM) gridding,
// apodization / zero padding, rocFFT r2c / c2r.
//
// Semantics follow lentils/backend/nufft.c:42-122 (grid_cpu) and
// lentils/operators/fourier_operators.py:12-126:
//   * u = U ch, v = -V ch; NGP by round(); taps ngp +- W; kernel
//     1/(4W^2) I0(beta sqrt(1-(k/W)^2)) I0(...) with k = u/du - i_u;
//     checkerboard sign (-1)^(i_u+i_v) on pre-wrap indices; phase
//     2 pi (u gcx + v gcy) with the truncated 2 pi of common.h.
//   * taps with wrapped i_v >= half_nv fold to ((nu - i_u) % nu, nv - i_v)
//     with the imaginary weight negated; the adjoint doubles final
//     column 0.
// Layout: visibilities are binned once into uv tiles of TU x TV cells (raw,
// centred tap coordinates) and kept in that order on the device.  A
// workgroup takes a chunk of one tile: degridding stages the tile's
// (TU+2W) x (TV+2W) window (fold / wrap / sign applied while staging) in LDS
// and reads taps from it; gridding accumulates the window in LDS with LDS
// atomics and writes it once as a partial.  A reduction gathers the partials
// covering each final grid cell in a fixed order -- no global atomics, and
// the result does not depend on scheduling.
#include <cstring>

#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <cmath>
#include <type_traits>

#include "nufft_plan.h"

namespace lt {

namespace {

constexpr double kTwoPiRef = 6.28318530718;  // lentils/backend/common.h:9
constexpr int kNThreads = 256;

__host__ __device__ inline uint32_t pack_ngp(int a, int b) {
    return (uint32_t)(a + 32768) | ((uint32_t)(b + 32768) << 16);
}
__device__ inline void unpack_ngp(uint32_t p, int &a, int &b) {
    a = (int)(p & 0xffff) - 32768;
    b = (int)(p >> 16) - 32768;
}

struct Geo {
    int nu, nv, half_nv, nrows, W, TU, TV, ntu, ntv, cu, cv, WU, WV, bu, bv;
    double du, dv, gcx, gcy;
};

// pass 1: tile of every channel-expanded visibility, Nyquist-band check
__global__ void k_vis_tile(Geo G, long nvis, const double *__restrict__ uvw, const double *__restrict__ chan,
                           uint32_t *__restrict__ tile, int32_t *__restrict__ idx, int *__restrict__ bad) {
    long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nvis) return;
    const int ch = (int)(e / G.nrows);
    const long r = e - (long)ch * G.nrows;
    const double u = uvw[3 * r] * chan[ch];
    const double v = -uvw[3 * r + 1] * chan[ch];
    const double tu = u / G.du, tv = v / G.dv;
    idx[e] = (int32_t)e;
    if (!(fabs(tu) < 1e9 && fabs(tv) < 1e9)) {
        atomicOr(bad, 1);
        tile[e] = 0;
        return;
    }
    const int a = (int)round(tu), b = (int)round(tv);
    if (abs(a) + G.W > G.bu || abs(b) + G.W > G.bv) {
        atomicOr(bad, 1);
        tile[e] = 0;
        return;
    }
    // tile-major cell key: tile id * (TU*TV) + cell inside the tile, so that
    // a stable sort groups visibilities by tile and, inside it, by cell
    const int ac = a + G.cu, bc = b + G.cv;
    const uint32_t t = (uint32_t)((ac / G.TU) * G.ntv + bc / G.TV);
    tile[e] = t * (uint32_t)(G.TU * G.TV) + (uint32_t)((ac % G.TU) * G.TV + bc % G.TV);
}

// pass 2: per sorted visibility, NGP and exact fractional offsets
template <class T>
__global__ void k_vis_fill(Geo G, long nvis, const double *__restrict__ uvw, const double *__restrict__ chan,
                           const int32_t *__restrict__ perm, int32_t *__restrict__ ngp, T *__restrict__ frac,
                           T *__restrict__ ph) {
    long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nvis) return;
    const long e = perm[p];
    const int ch = (int)(e / G.nrows);
    const long r = e - (long)ch * G.nrows;
    const double u = uvw[3 * r] * chan[ch];
    const double v = -uvw[3 * r + 1] * chan[ch];
    const double tu = u / G.du, tv = v / G.dv;
    const int a = (int)round(tu), b = (int)round(tv);
    ngp[p] = (int32_t)pack_ngp(a, b);
    frac[2 * p] = (T)(tu - a);  // exact in double (Sterbenz)
    frac[2 * p + 1] = (T)(tv - b);
    if (ph) {
        const double arg = kTwoPiRef * (u * G.gcx + v * G.gcy);
        ph[2 * p] = (T)cos(arg);
        ph[2 * p + 1] = (T)sin(arg);
    }
}

__global__ void k_count_tiles(long n, const uint32_t *__restrict__ key, uint32_t tt, int32_t *__restrict__ cnt) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicAdd(&cnt[key[i] / tt], 1);
}

template <class T>
__device__ inline T kb_value(const KbPoly &P, T arg, T W) {
    if (!(arg >= -W && arg <= W)) return (T)0;
    T q = arg / W;
    T s = (T)1 - q * q;
    if (s < (T)0) s = (T)0;
    T acc = (T)P.c[P.deg];
    for (int m = P.deg - 1; m >= 0; --m) acc = acc * s + (T)P.c[m];
    return acc;
}

// the 2W+1 kernel values of one axis at fractional offset f (|f| <= 1/2)
// from the per-tap polynomials kc[k][m]
template <class T, int W>
__device__ inline void kb_taps(const T *__restrict__ kc, T f, T (&out)[2 * W + 1]) {
    constexpr int NT = 2 * W + 1;
    constexpr int D = TapDeg<T, W>::D;
#pragma unroll
    for (int k = 0; k < NT; ++k) {
        T acc = kc[k * (D + 1) + D];
#pragma unroll
        for (int m = D - 1; m >= 0; --m) acc = acc * f + kc[k * (D + 1) + m];
        out[k] = acc;
    }
    if (f > (T)0) out[0] = (T)0;
    if (f < (T)0) out[NT - 1] = (T)0;
}

// same as fetch_folded, from a band of the half-plane spectrum stored
// [v][u] for v < nbv (columns beyond the band are zero)
template <class T>
__device__ inline Cplx<T> fetch_band(const Geo &G, const Cplx<T> *__restrict__ band, int a, int b, int nbv) {
    Cplx<T> z{(T)0, (T)0};
    if (abs(a) > G.bu || abs(b) > G.bv || abs(b) >= nbv) return z;
    const T sg = ((a + b) & 1) ? (T)-1 : (T)1;
    if (b >= 0) {
        const int iu = a < 0 ? a + G.nu : a;
        z = band[(long)b * G.nu + iu];
        z.re *= sg;
        z.im *= sg;
    } else {
        const int iu = -a < 0 ? -a + G.nu : -a;
        z = band[(long)(-b) * G.nu + iu];
        z.re *= sg;
        z.im *= -sg;
    }
    return z;
}

// value of raw tap (a, b) of the half-plane grid, with the reference's
// wrap/fold and the checkerboard sign applied; zero outside the band
template <class T>
__device__ inline Cplx<T> fetch_folded(const Geo &G, const Cplx<T> *__restrict__ grid, int a, int b) {
    Cplx<T> z{(T)0, (T)0};
    if (abs(a) > G.bu || abs(b) > G.bv) return z;
    const T sg = ((a + b) & 1) ? (T)-1 : (T)1;
    if (b >= 0) {
        const int iu = a < 0 ? a + G.nu : a;
        z = grid[(long)iu * G.half_nv + b];
        z.re *= sg;
        z.im *= sg;
    } else {
        const int iu = -a < 0 ? -a + G.nu : -a;
        z = grid[(long)iu * G.half_nv - b];
        z.re *= sg;
        z.im *= -sg;
    }
    return z;
}

template <class T, int W>
__global__ void __launch_bounds__(kNThreads) k_degrid(Geo G, const T *__restrict__ kc, const Cplx<T> *__restrict__ grid,
                                                      const int32_t *__restrict__ ck_tile,
                                                      const int32_t *__restrict__ ck_beg,
                                                      const int32_t *__restrict__ ck_end,
                                                      const int32_t *__restrict__ ngp, const T *__restrict__ frac,
                                                      const T *__restrict__ ph, const double *__restrict__ wts,
                                                      Cplx<T> *__restrict__ vis, int accumulate) {
    extern __shared__ unsigned char smem[];
    Cplx<T> *win = reinterpret_cast<Cplx<T> *>(smem);
    constexpr int NT = 2 * W + 1;
    const int c = blockIdx.x;
    const int tile = ck_tile[c];
    const int tu = tile / G.ntv, tv = tile - tu * G.ntv;
    const int a0 = tu * G.TU - W - G.cu, b0 = tv * G.TV - W - G.cv;  // raw coords of window (0, 0)
    const int nwin = G.WU * G.WV;
    for (int i = threadIdx.x; i < nwin; i += kNThreads) {
        const int wa = i / G.WV, wb = i - wa * G.WV;
        win[i] = fetch_folded(G, grid, a0 + wa, b0 + wb);
    }
    __syncthreads();
    const int beg = ck_beg[c], end = ck_end[c];
    const T c0 = (T)(1.0 / (4.0 * W * W));
    for (int p = beg + threadIdx.x; p < end; p += kNThreads) {
        int na, nb;
        unpack_ngp((uint32_t)ngp[p], na, nb);
        const T fu = frac[2 * (long)p], fv = frac[2 * (long)p + 1];
        T ku[NT], kv[NT];
        kb_taps<T, W>(kc, fu, ku);
        kb_taps<T, W>(kc, fv, kv);
        const int r0 = na - a0 - W, s0 = nb - b0 - W;  // window row/col of tap (0, 0)
        T ar = 0, ai = 0;
#pragma unroll
        for (int k = 0; k < NT; ++k) {
            const Cplx<T> *row = win + (r0 + k) * G.WV + s0;
            T rr = 0, ri = 0;
#pragma unroll
            for (int l = 0; l < NT; ++l) {
                const Cplx<T> g = row[l];
                rr += kv[l] * g.re;
                ri += kv[l] * g.im;
            }
            ar += ku[k] * rr;
            ai += ku[k] * ri;
        }
        T cs = (T)1, sn = (T)0;
        if (ph) {
            cs = ph[2 * (long)p];
            sn = ph[2 * (long)p + 1];
        }
        T outr = c0 * (cs * ar + sn * ai);
        T outi = c0 * (-sn * ar + cs * ai);
        if (wts) {
            const T w = (T)wts[p];
            outr *= w;
            outi *= w;
        }
        if (accumulate) {
            outr += vis[p].re;
            outi += vis[p].im;
        }
        vis[p] = Cplx<T>{outr, outi};
    }
}

// The 2W+1 taps from the W+1 polynomials of taps 0..W, split into even and
// odd parts in f: P_k(f) = E_k(f^2) + f O_k(f^2).  The kernel is even and
// tap k sits at argument f + W - k, so tap 2W-k at f equals tap k at -f:
// P_{2W-k}(f) = E_k(f^2) - f O_k(f^2).  Half the polynomial work of kb_taps;
// the fit error of tap k at -f bounds the error of tap 2W-k.  No zeroing.
template <class T, int W>
__device__ inline void kb_taps_sym(const T *__restrict__ kc, T f, T (&out)[2 * W + 1]) {
    constexpr int D = TapDeg<T, W>::D;
    constexpr int DE = D / 2 * 2;              // highest even power
    constexpr int DO = (D % 2) ? D : D - 1;    // highest odd power
    const T s = f * f;
#pragma unroll
    for (int k = 0; k <= W; ++k) {
        const T *c = kc + k * (D + 1);
        T e = c[DE];
#pragma unroll
        for (int m = DE - 2; m >= 0; m -= 2) e = e * s + c[m];
        T o = c[DO];
#pragma unroll
        for (int m = DO - 2; m >= 1; m -= 2) o = o * s + c[m];
        out[k] = e + f * o;
        if (k < W) out[2 * W - k] = e - f * o;
    }
}

// The 2W taps of the window that starts W - 1 cells below floor(x), from
// the W polynomials of taps 0..W-1 in s = x - floor(x) - 1/2, split into
// even and odd parts: tap 2W-1-j at s is tap j at -s (the kernel is even),
// so P_{2W-1-j}(s) = E_j(s^2) - s O_j(s^2).  Every tap of this window is
// inside the support for any s, so there is nothing to select or zero.
template <class T, int W>
__device__ inline void kb_taps_half(const T *__restrict__ kc, T s, T (&out)[2 * W]) {
    constexpr int D = TapDeg<T, W>::D;
    constexpr int DE = D / 2 * 2;
    constexpr int DO = (D % 2) ? D : D - 1;
    const T q = s * s;
#pragma unroll
    for (int j = 0; j < W; ++j) {
        const T *c = kc + j * (D + 1);
        T e = c[DE];
#pragma unroll
        for (int m = DE - 2; m >= 0; m -= 2) e = e * q + c[m];
        T o = c[DO];
#pragma unroll
        for (int m = DO - 2; m >= 1; m -= 2) o = o * q + c[m];
        out[j] = e + s * o;
        out[2 * W - 1 - j] = e - s * o;
    }
}

template <class T>
struct Vec2;
template <>
struct Vec2<float> {
    using type = float2;
};
template <>
struct Vec2<double> {
    using type = double2;
};

// Degridder for tiles of exactly 32 cells in v (the window row stride WV =
// 32 + 2W is then a compile-time constant, so every window read of a
// visibility is one LDS base address plus an immediate offset).  A
// visibility with f != 0 on both axes has exactly 2W nonzero taps per axis
// (tap 0 is outside the support for f > 0, tap 2W for f < 0): those are
// selected and the 2W x 2W window product is formed.  Waves holding a
// visibility with f == 0 on an axis (all 2W+1 taps inside the support) take
// the full (2W+1)^2 product.  Same results as k_degrid up to rounding.
template <class T, int W>
__global__ void __launch_bounds__(kNThreads) k_degrid32(Geo G, const T *__restrict__ kc,
                                                        const T *__restrict__ kch,
                                                        const Cplx<T> *__restrict__ grid,
                                                        const int32_t *__restrict__ ck_tile,
                                                        const int32_t *__restrict__ ck_beg,
                                                        const int32_t *__restrict__ ck_end,
                                                        const int32_t *__restrict__ ngp, const T *__restrict__ frac,
                                                        const T *__restrict__ ph, const double *__restrict__ wts,
                                                        Cplx<T> *__restrict__ vis, int accumulate) {
    extern __shared__ __align__(16) unsigned char smem_d32[];
    Cplx<T> *win = reinterpret_cast<Cplx<T> *>(smem_d32);
    // window reads as one aligned 2-vector (one LDS instruction with an
    // immediate offset per tap)
    using V = typename Vec2<T>::type;
    const V *wv = reinterpret_cast<const V *>(smem_d32);
    constexpr int NT = 2 * W + 1;
    constexpr int NF = 2 * W;
    constexpr int WV = 32 + 2 * W;
    const int c = blockIdx.x;
    const int tile = ck_tile[c];
    const int tu = tile / G.ntv, tv = tile - tu * G.ntv;
    const int a0 = tu * G.TU - W - G.cu, b0 = tv * 32 - W - G.cv;
    const int nwin = G.WU * WV;
    for (int i = threadIdx.x; i < nwin; i += kNThreads) {
        const int wa = i / WV, wb = i - wa * WV;
        win[i] = fetch_folded(G, grid, a0 + wa, b0 + wb);
    }
    __syncthreads();
    const int beg = ck_beg[c], end = ck_end[c];
    const T c0 = (T)(1.0 / (4.0 * W * W));
    for (int p0 = beg; p0 < end; p0 += kNThreads) {
        const int p = p0 + threadIdx.x;
        const bool live = p < end;
        const int q = live ? p : beg;  // idle lanes read a valid entry, store nothing
        int na, nb;
        unpack_ngp((uint32_t)ngp[q], na, nb);
        const T fu = frac[2 * (long)q], fv = frac[2 * (long)q + 1];
        const int r0 = na - a0 - W, s0 = nb - b0 - W;
        T ar = 0, ai = 0;
        if (__all(fu != (T)0 && fv != (T)0)) {
            const int du = fu > (T)0 ? 1 : 0, dv = fv > (T)0 ? 1 : 0;
            T ku6[NF], kv6[NF];
            if (kch) {
                // the window starts W - 1 cells below floor(x): r0 + du
                kb_taps_half<T, W>(kch, du ? fu - (T)0.5 : fu + (T)0.5, ku6);
                kb_taps_half<T, W>(kch, dv ? fv - (T)0.5 : fv + (T)0.5, kv6);
            } else {
                T ku[NT], kv[NT];
                kb_taps_sym<T, W>(kc, fu, ku);
                kb_taps_sym<T, W>(kc, fv, kv);
#pragma unroll
                for (int j = 0; j < NF; ++j) {
                    ku6[j] = du ? ku[j + 1] : ku[j];
                    kv6[j] = dv ? kv[j + 1] : kv[j];
                }
            }
            const V *base = wv + (r0 + du) * WV + s0 + dv;
#pragma unroll
            for (int k = 0; k < NF; ++k) {
                T rr = 0, ri = 0;
#pragma unroll
                for (int l = 0; l < NF; ++l) {
                    const V g = base[k * WV + l];
                    rr += kv6[l] * g.x;
                    ri += kv6[l] * g.y;
                }
                ar += ku6[k] * rr;
                ai += ku6[k] * ri;
            }
        } else {
            T ku[NT], kv[NT];
            kb_taps_sym<T, W>(kc, fu, ku);
            kb_taps_sym<T, W>(kc, fv, kv);
            if (fu > (T)0) ku[0] = (T)0;
            if (fu < (T)0) ku[NT - 1] = (T)0;
            if (fv > (T)0) kv[0] = (T)0;
            if (fv < (T)0) kv[NT - 1] = (T)0;
            const V *base = wv + r0 * WV + s0;
#pragma unroll
            for (int k = 0; k < NT; ++k) {
                T rr = 0, ri = 0;
#pragma unroll
                for (int l = 0; l < NT; ++l) {
                    const V g = base[k * WV + l];
                    rr += kv[l] * g.x;
                    ri += kv[l] * g.y;
                }
                ar += ku[k] * rr;
                ai += ku[k] * ri;
            }
        }
        if (!live) continue;
        T cs = (T)1, sn = (T)0;
        if (ph) {
            cs = ph[2 * (long)p];
            sn = ph[2 * (long)p + 1];
        }
        T outr = c0 * (cs * ar + sn * ai);
        T outi = c0 * (-sn * ar + cs * ai);
        if (wts) {
            const T w = (T)wts[p];
            outr *= w;
            outi *= w;
        }
        if (accumulate) {
            outr += vis[p].re;
            outi += vis[p].im;
        }
        vis[p] = Cplx<T>{outr, outi};
    }
}

__device__ inline int gcd_i(int a, int b) {
    while (b) {
        int t = a % b;
        a = b;
        b = t;
    }
    return a;
}

__device__ inline int coprime_stride(int n) {
    if (n <= 2) return 1;
    int s = (int)(((long)n * 40503L) >> 16);
    if (s < 1) s = 1;
    while (gcd_i(s, n) != 1) ++s;
    return s;
}

template <class T>
__device__ inline void lds_add(T *addr, T v) {
    atomicAdd(addr, v);
}

template <class T, int W>
__global__ void __launch_bounds__(kNThreads) k_grid(Geo G, const T *__restrict__ kc, const Cplx<T> *__restrict__ vis,
                                                    const double *__restrict__ wts,
                                                    const int32_t *__restrict__ ck_tile,
                                                    const int32_t *__restrict__ ck_beg,
                                                    const int32_t *__restrict__ ck_end,
                                                    const int32_t *__restrict__ ngp, const T *__restrict__ frac,
                                                    const T *__restrict__ ph, Cplx<T> *__restrict__ partial) {
    extern __shared__ unsigned char smem[];
    T *win = reinterpret_cast<T *>(smem);  // interleaved re, im
    constexpr int NT = 2 * W + 1;
    const int c = blockIdx.x;
    const int tile = ck_tile[c];
    const int tu = tile / G.ntv, tv = tile - tu * G.ntv;
    const int a0 = tu * G.TU - W - G.cu, b0 = tv * G.TV - W - G.cv;
    const int nwin = G.WU * G.WV;
    for (int i = threadIdx.x; i < 2 * nwin; i += kNThreads) win[i] = (T)0;
    __syncthreads();
    const int beg = ck_beg[c], end = ck_end[c];
    const int n = end - beg;
    // a stride coprime to n (near the golden ratio of n) spreads consecutive
    // threads over the chunk so that a wavefront's LDS atomics rarely meet on
    // one address; q -> q*stride mod n is a permutation of the chunk
    const int stride = coprime_stride(n);
    for (int q = threadIdx.x; q < n; q += kNThreads) {
        const int p = beg + (int)(((long)q * stride) % n);
        int na, nb;
        unpack_ngp((uint32_t)ngp[p], na, nb);
        const T fu = frac[2 * (long)p], fv = frac[2 * (long)p + 1];
        Cplx<T> V = vis[p];
        if (wts) {
            const T w = (T)wts[p];
            V.re *= w;
            V.im *= w;
        }
        T cs = (T)1, sn = (T)0;
        if (ph) {
            cs = ph[2 * (long)p];
            sn = ph[2 * (long)p + 1];
        }
        const T rr = (T)0.5 * (cs * V.re - sn * V.im);
        const T ri = (T)0.5 * (sn * V.re + cs * V.im);
        T ku[NT], kv[NT];
        kb_taps<T, W>(kc, fu, ku);
        kb_taps<T, W>(kc, fv, kv);
        const int r0 = na - a0 - W, s0 = nb - b0 - W;
#pragma unroll
        for (int k = 0; k < NT; ++k) {
            const T tr = ku[k] * rr, ti = ku[k] * ri;
            T *row = win + 2 * ((r0 + k) * G.WV + s0);
#pragma unroll
            for (int l = 0; l < NT; ++l) {
                lds_add(row + 2 * l, tr * kv[l]);
                lds_add(row + 2 * l + 1, ti * kv[l]);
            }
        }
    }
    __syncthreads();
    // sign (-1)^(a+b) and 1/(4W^2) applied once per window cell
    const T c0 = (T)(1.0 / (4.0 * W * W));
    Cplx<T> *out = partial + (long)c * nwin;
    for (int i = threadIdx.x; i < nwin; i += kNThreads) {
        const int wa = i / G.WV, wb = i - wa * G.WV;
        const T sg = (((a0 + wa) + (b0 + wb)) & 1) ? -c0 : c0;
        out[i] = Cplx<T>{win[2 * i] * sg, win[2 * i + 1] * sg};
    }
}

// sum of the partial windows of every chunk whose window covers raw cell
// (a, b), in tile / chunk order
template <class T>
__device__ inline Cplx<T> gather_raw(const Geo &G, int a, int b, const int32_t *__restrict__ tile_ck0,
                                     const int32_t *__restrict__ tile_ck1, const Cplx<T> *__restrict__ partial) {
    Cplx<T> s{(T)0, (T)0};
    const int ac = a + G.cu, bc = b + G.cv;
    const int nwin = G.WU * G.WV;
    const int W = G.W;
    int tu_lo = (ac + W - G.WU + 1);
    tu_lo = tu_lo <= 0 ? 0 : (tu_lo + G.TU - 1) / G.TU;
    const int tu_hi = min((ac + W) / G.TU, G.ntu - 1);
    int tv_lo = (bc + W - G.WV + 1);
    tv_lo = tv_lo <= 0 ? 0 : (tv_lo + G.TV - 1) / G.TV;
    const int tv_hi = min((bc + W) / G.TV, G.ntv - 1);
    for (int tu = tu_lo; tu <= tu_hi; ++tu) {
        const int wa = ac - (tu * G.TU - W);
        for (int tv = tv_lo; tv <= tv_hi; ++tv) {
            const int wb = bc - (tv * G.TV - W);
            const int t = tu * G.ntv + tv;
            const int k0 = tile_ck0[t], k1 = tile_ck1[t];
            for (int k = k0; k < k1; ++k) {
                const Cplx<T> v = partial[(long)k * nwin + wa * G.WV + wb];
                s.re += v.re;
                s.im += v.im;
            }
        }
    }
    return s;
}

template <class T>
__global__ void __launch_bounds__(kNThreads) k_grid_reduce(Geo G, const int32_t *__restrict__ tile_ck0,
                                                           const int32_t *__restrict__ tile_ck1,
                                                           const Cplx<T> *__restrict__ partial,
                                                           Cplx<T> *__restrict__ grid) {
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)G.nu * G.half_nv) return;
    const int iu = (int)(g / G.half_nv), iv = (int)(g - (long)iu * G.half_nv);
    const int a = iu <= G.bu ? iu : iu - G.nu;
    Cplx<T> tot{(T)0, (T)0};
    if (abs(a) <= G.bu && iv <= G.bv) {
        tot = gather_raw(G, a, iv, tile_ck0, tile_ck1, partial);
        if (iv >= 1 && -a >= -G.bu && -a <= G.bu) {
            const Cplx<T> f = gather_raw(G, -a, -iv, tile_ck0, tile_ck1, partial);
            tot.re += f.re;
            tot.im -= f.im;
        }
        if (iv == 0) {
            tot.re *= (T)2;
            tot.im *= (T)2;
        }
    }
    grid[g] = tot;
}

// zero-padding * apodization into the real FFT input (whole buffer written)
template <class T>
__global__ void k_apod_pad(int nx, int ny, int padx, int pady, int nxp, int nyp, const double *__restrict__ img,
                           const double *__restrict__ apod, T *__restrict__ out) {
    long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)nxp * nyp) return;
    const int i = (int)(g / nyp) - padx, j = (int)(g % nyp) - pady;
    T v = (T)0;
    if (i >= 0 && i < nx && j >= 0 && j < ny) v = (T)(img[(long)i * ny + j] * apod[(long)i * ny + j]);
    out[g] = v;
}

// crop * apodization of the c2r output: img = alpha * crop(x) * apod (+ img)
template <class T>
__global__ void k_crop_apod(int nx, int ny, int padx, int pady, int nyp, const T *__restrict__ in,
                            const double *__restrict__ apod, double alpha, int accumulate, double *__restrict__ img) {
    long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)nx * ny) return;
    const int i = (int)(g / ny), j = (int)(g % ny);
    double v = alpha * ((double)in[(long)(i + padx) * nyp + (j + pady)] * apod[g]);
    img[g] = accumulate ? img[g] + v : v;
}

// c2r treats columns 0 and nv/2 as Hermitian: make them so explicitly,
// X[k] <- (X[k] + conj(X[-k])) / 2, which is what taking the real part of
// the inverse transform along the first axis does (numpy/pocketfft)
template <class T>
__global__ void k_hermitian_cols(int nu, int half_nv, int col, Cplx<T> *__restrict__ X) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > nu / 2) return;
    const int m = (nu - k) % nu;
    Cplx<T> a = X[(long)k * half_nv + col], b = X[(long)m * half_nv + col];
    Cplx<T> na{(T)0.5 * (a.re + b.re), (T)0.5 * (a.im - b.im)};
    Cplx<T> nb{(T)0.5 * (b.re + a.re), (T)0.5 * (b.im - a.im)};
    X[(long)k * half_nv + col] = na;
    if (m != k) X[(long)m * half_nv + col] = nb;
}

template <class T>
__global__ void k_to_sorted(long n, const int32_t *__restrict__ perm, const double *__restrict__ src,
                            Cplx<T> *__restrict__ dst) {
    long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const long e = perm[p];
    dst[p] = Cplx<T>{(T)src[2 * e], (T)src[2 * e + 1]};
}

template <class T>
__global__ void k_from_sorted(long n, const int32_t *__restrict__ perm, const Cplx<T> *__restrict__ src,
                              double *__restrict__ dst, int accumulate) {
    long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const long e = perm[p];
    double r = (double)src[p].re, i = (double)src[p].im;
    if (accumulate) {
        r += dst[2 * e];
        i += dst[2 * e + 1];
    }
    dst[2 * e] = r;
    dst[2 * e + 1] = i;
}

template <class T>
__global__ void k_c128_to_T(long n, const double *__restrict__ a, T *__restrict__ b) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (T)a[i];
}

template <class T>
__global__ void k_T_add_c128(long n, const T *__restrict__ a, double *__restrict__ b) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] += (double)a[i];
}


// Fused D^T C D in the spectral domain for one chunk of one sample: degrid
// the staged input window, weight, and grid back into an LDS accumulator.
// The phase factor cancels (e^{-i phi} then e^{+i phi}) and the kernel
// weights are computed once per visibility for both directions.
template <class T, int W>
__global__ void __launch_bounds__(kNThreads) k_normal(Geo G, const T *__restrict__ kc, const Cplx<T> *__restrict__ spec_in,
                                                      long sstride, const int32_t *__restrict__ ck_tile,
                                                      const int32_t *__restrict__ ck_beg,
                                                      const int32_t *__restrict__ ck_end,
                                                      const int32_t *__restrict__ ngp, const T *__restrict__ frac,
                                                      const double *__restrict__ wts, Cplx<T> *__restrict__ partial,
                                                      long pstride, const int *__restrict__ active) {
    extern __shared__ unsigned char smem[];
    constexpr int NT = 2 * W + 1;
    const int nwin = G.WU * G.WV;
    Cplx<T> *win = reinterpret_cast<Cplx<T> *>(smem);
    T *acc = reinterpret_cast<T *>(smem + (size_t)nwin * sizeof(Cplx<T>));
    const int smp = active ? active[blockIdx.y] : (int)blockIdx.y;
    const Cplx<T> *grid = spec_in + (long)smp * sstride;
    const int c = blockIdx.x;
    const int tile = ck_tile[c];
    const int tu = tile / G.ntv, tv = tile - tu * G.ntv;
    const int a0 = tu * G.TU - W - G.cu, b0 = tv * G.TV - W - G.cv;
    for (int i = threadIdx.x; i < nwin; i += kNThreads) {
        const int wa = i / G.WV, wb = i - wa * G.WV;
        win[i] = fetch_folded(G, grid, a0 + wa, b0 + wb);
        acc[2 * i] = (T)0;
        acc[2 * i + 1] = (T)0;
    }
    __syncthreads();
    const int beg = ck_beg[c], end = ck_end[c];
    const int n = end - beg;
    const T c0h = (T)(0.5 / (4.0 * W * W));
    const int stride = coprime_stride(n);
    for (int q = threadIdx.x; q < n; q += kNThreads) {
        const int p = beg + (int)(((long)q * stride) % n);
        int na, nb;
        unpack_ngp((uint32_t)ngp[p], na, nb);
        const T fu = frac[2 * (long)p], fv = frac[2 * (long)p + 1];
        T ku[NT], kv[NT];
        kb_taps<T, W>(kc, fu, ku);
        kb_taps<T, W>(kc, fv, kv);
        const int r0 = na - a0 - W, s0 = nb - b0 - W;
        T ar = 0, ai = 0;
#pragma unroll
        for (int k = 0; k < NT; ++k) {
            const Cplx<T> *row = win + (r0 + k) * G.WV + s0;
            T rr = 0, ri = 0;
#pragma unroll
            for (int l = 0; l < NT; ++l) {
                const Cplx<T> g = row[l];
                rr += kv[l] * g.re;
                ri += kv[l] * g.im;
            }
            ar += ku[k] * rr;
            ai += ku[k] * ri;
        }
        const T wv = (T)wts[p] * c0h;
        ar *= wv;
        ai *= wv;
#pragma unroll
        for (int k = 0; k < NT; ++k) {
            const T tr = ku[k] * ar, ti = ku[k] * ai;
            T *row = acc + 2 * ((r0 + k) * G.WV + s0);
#pragma unroll
            for (int l = 0; l < NT; ++l) {
                lds_add(row + 2 * l, tr * kv[l]);
                lds_add(row + 2 * l + 1, ti * kv[l]);
            }
        }
    }
    __syncthreads();
    const T c0 = (T)(1.0 / (4.0 * W * W));
    Cplx<T> *out = partial + (long)smp * pstride + (long)c * nwin;
    for (int i = threadIdx.x; i < nwin; i += kNThreads) {
        const int wa = i / G.WV, wb = i - wa * G.WV;
        const T sg = (((a0 + wa) + (b0 + wb)) & 1) ? -c0 : c0;
        out[i] = Cplx<T>{acc[2 * i] * sg, acc[2 * i + 1] * sg};
    }
}

template <class T>
__global__ void __launch_bounds__(kNThreads) k_grid_reduce_b(Geo G, const int32_t *__restrict__ tile_ck0,
                                                             const int32_t *__restrict__ tile_ck1,
                                                             const Cplx<T> *__restrict__ partial, long pstride,
                                                             Cplx<T> *__restrict__ grid, long gstride,
                                                             const int *__restrict__ active) {
    const int smp = active ? active[blockIdx.y] : (int)blockIdx.y;
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)G.nu * G.half_nv) return;
    const Cplx<T> *pp = partial + (long)smp * pstride;
    const int iu = (int)(g / G.half_nv), iv = (int)(g - (long)iu * G.half_nv);
    const int a = iu <= G.bu ? iu : iu - G.nu;
    Cplx<T> tot{(T)0, (T)0};
    if (abs(a) <= G.bu && iv <= G.bv) {
        tot = gather_raw(G, a, iv, tile_ck0, tile_ck1, pp);
        if (iv >= 1) {
            const Cplx<T> f = gather_raw(G, -a, -iv, tile_ck0, tile_ck1, pp);
            tot.re += f.re;
            tot.im -= f.im;
        }
        if (iv == 0) {
            tot.re *= (T)2;
            tot.im *= (T)2;
        }
    }
    grid[(long)smp * gstride + g] = tot;
}

// batched zero padding (x apod when apod != nullptr)
// slot y of the padded batch <- sample active[y] (or y) of the images
template <class T, class TI = double>
__global__ void k_pad_b(int nx, int ny, int padx, int pady, int nxp, int nyp, const TI *__restrict__ img,
                        const double *__restrict__ apod, T *__restrict__ out, const int *__restrict__ active) {
    const int slot = blockIdx.y;
    const int smp = active ? active[slot] : slot;
    long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)nxp * nyp) return;
    const int i = (int)(g / nyp) - padx, j = (int)(g % nyp) - pady;
    T v = (T)0;
    if (i >= 0 && i < nx && j >= 0 && j < ny) {
        const long k = (long)i * ny + j;
        const TI x = img[(long)smp * nx * ny + k];
        v = (T)(apod ? x * (TI)apod[k] : x);
    }
    out[(long)slot * nxp * nyp + g] = v;
}

template <class T, class TO = double>
__global__ void k_crop_b(int nx, int ny, int padx, int pady, int nxp, int nyp, const T *__restrict__ in,
                         const double *__restrict__ apod, TO alpha, int accumulate, TO *__restrict__ img,
                         const int *__restrict__ active) {
    const int slot = blockIdx.y;
    const int smp = active ? active[slot] : slot;
    long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)nx * ny) return;
    const int i = (int)(g / ny), j = (int)(g % ny);
    const TO x = (TO)in[(long)slot * nxp * nyp + (long)(i + padx) * nyp + (j + pady)];
    const TO v = alpha * (apod ? x * (TO)apod[g] : x);
    TO *o = img + (long)smp * nx * ny + g;
    *o = accumulate ? *o + v : v;
}

template <class T>
__global__ void k_hermitian_cols_b(int nu, int half_nv, int col, Cplx<T> *__restrict__ X, long stride,
                                   const int *__restrict__ active) {
    const int smp = active ? active[blockIdx.y] : (int)blockIdx.y;
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > nu / 2) return;
    Cplx<T> *Y = X + (long)smp * stride;
    const int m = (nu - k) % nu;
    Cplx<T> a = Y[(long)k * half_nv + col], b = Y[(long)m * half_nv + col];
    Y[(long)k * half_nv + col] = Cplx<T>{(T)0.5 * (a.re + b.re), (T)0.5 * (a.im - b.im)};
    if (m != k) Y[(long)m * half_nv + col] = Cplx<T>{(T)0.5 * (b.re + a.re), (T)0.5 * (b.im - a.im)};
}

template <class T>
__global__ void k_spec_mul_b(long n, const double *__restrict__ mul, Cplx<T> *__restrict__ X, long stride,
                             const int *__restrict__ active) {
    const int smp = active ? active[blockIdx.y] : (int)blockIdx.y;
    long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    Cplx<T> *Y = X + (long)smp * stride;
    const double mr = mul[2 * g], mi = mul[2 * g + 1];
    const double xr = (double)Y[g].re, xi = (double)Y[g].im;
    Y[g] = Cplx<T>{(T)(xr * mr - xi * mi), (T)(xr * mi + xi * mr)};
}

// ---------------------------------------------------------------------------
// precomputed normal stencil
// ---------------------------------------------------------------------------

__device__ inline void key_to_cell(const Geo &G, uint32_t key, int &ac, int &bc) {
    const uint32_t tt = (uint32_t)(G.TU * G.TV);
    const uint32_t t = key / tt, c = key - t * tt;
    const int tu = (int)(t / G.ntv), tv = (int)(t - (uint32_t)tu * G.ntv);
    ac = tu * G.TU + (int)(c / G.TV);
    bc = tv * G.TV + (int)(c % G.TV);
}

__global__ void k_cell_flags(long n, const uint32_t *__restrict__ key, int64_t *__restrict__ flag) {
    long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) flag[p] = (p == 0 || key[p] != key[p - 1]) ? 1 : 0;
}

// cell list: first visibility of each cell, its centred raw coordinates,
// and the dense cell map
__global__ void k_cell_list(Geo G, long n, const uint32_t *__restrict__ key, const int64_t *__restrict__ idx,
                            int32_t *__restrict__ cell_beg, int32_t *__restrict__ cellmap) {
    long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    if (p == 0 || key[p] != key[p - 1]) {
        const long c = idx[p];
        cell_beg[c] = (int32_t)p;
        int ac, bc;
        key_to_cell(G, key[p], ac, bc);
        cellmap[(long)ac * G.nv + bc] = (int32_t)c;
    }
}

// T_c = sum_{v in cell c} w_v z_v z_v^T with z_v = ku (x) kv (NT^2 values);
// one wavefront-sized block per cell, thread i owns row i in registers
template <int W>
__global__ void __launch_bounds__(128) k_cell_T(Geo G, KbPoly P, long ncell, const int32_t *__restrict__ cell_beg,
                                                const int32_t *__restrict__ frac_dummy, const double *__restrict__ frac,
                                                const double *__restrict__ wts, double *__restrict__ Tout) {
    constexpr int NT = 2 * W + 1, N2 = NT * NT, BATCH = 8;
    __shared__ double ks[BATCH][2][NT];
    __shared__ double zs[BATCH][N2];
    __shared__ double ws[BATCH];
    const long c = blockIdx.x;
    const int beg = cell_beg[c], end = cell_beg[c + 1];
    const int i = threadIdx.x;
    double acc[N2];
#pragma unroll
    for (int j = 0; j < N2; ++j) acc[j] = 0.0;
    const double Wt = (double)W;
    for (int b0 = beg; b0 < end; b0 += BATCH) {
        const int nb = min(BATCH, end - b0);
        // kernel values: thread t -> (vis t / (2 NT), axis, tap)
        for (int t = threadIdx.x; t < nb * 2 * NT; t += blockDim.x) {
            const int v = t / (2 * NT), rem = t - v * 2 * NT, ax = rem / NT, k = rem - ax * NT;
            const double f = frac[2 * (long)(b0 + v) + ax];
            ks[v][ax][k] = kb_value<double>(P, f + (double)(W - k), Wt);
        }
        for (int t = threadIdx.x; t < nb; t += blockDim.x) ws[t] = wts[b0 + t];
        __syncthreads();
        for (int t = threadIdx.x; t < nb * N2; t += blockDim.x) {
            const int v = t / N2, j = t - v * N2;
            zs[v][j] = ks[v][0][j / NT] * ks[v][1][j % NT];
        }
        __syncthreads();
        if (i < N2) {
            for (int v = 0; v < nb; ++v) {
                const double zi = ws[v] * zs[v][i];
#pragma unroll
                for (int j = 0; j < N2; ++j) acc[j] += zi * zs[v][j];
            }
        }
        __syncthreads();
    }
    if (i < N2) {
        double *o = Tout + c * (long)N2 * N2 + (long)i * N2;
#pragma unroll
        for (int j = 0; j < N2; ++j) o[j] = acc[j];
    }
}

// mark raw cells within W of a visibility cell
__global__ void k_mark_rows(Geo G, long ncell, const int32_t *__restrict__ cell_beg, const uint32_t *__restrict__ key,
                            int64_t *__restrict__ mark) {
    long c = blockIdx.x;
    if (c >= ncell) return;
    int ac, bc;
    key_to_cell(G, key[cell_beg[c]], ac, bc);
    const int NT = 2 * G.W + 1;
    for (int t = threadIdx.x; t < NT * NT; t += blockDim.x) {
        const int a = ac - G.W + t / NT, b = bc - G.W + t % NT;
        if (a >= 0 && a < G.nu && b >= 0 && b < G.nv) mark[(long)a * G.nv + b] = 1;
    }
}

__global__ void k_rows(long n, const int64_t *__restrict__ mark, const int64_t *__restrict__ off,
                       int32_t *__restrict__ rowmap, int32_t *__restrict__ row_cell) {
    long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const bool on = (off[g + 1] - off[g]) != 0;
    rowmap[g] = on ? (int32_t)off[g] : -1;
    if (on) row_cell[off[g]] = (int32_t)g;
    (void)mark;
}

// coef[o][row] = 0.5 c0^2 sum over cells c with both r' and r' + o in c's
// footprint of T_c[r' - c][r' + o - c], summed in a fixed cell order
template <int W>
__global__ void k_coef(Geo G, long nrow, const int32_t *__restrict__ row_cell, const int32_t *__restrict__ cellmap,
                       const double *__restrict__ T, double scale, double *__restrict__ coef) {
    constexpr int NT = 2 * W + 1, N2 = NT * NT, R = 2 * W, NR = 2 * R + 1;
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nrow * NR * NR) return;
    const int o = (int)(g / nrow);
    const long row = g - (long)o * nrow;
    const int ou = o / NR - R, ov = o % NR - R;
    const int rc = row_cell[row];
    const int ru = rc / G.nv, rv = rc % G.nv;
    const int su = ru + ou, sv = rv + ov;
    double acc = 0.0;
    const int u0 = max(ru, su) - W, u1 = min(ru, su) + W;
    const int v0 = max(rv, sv) - W, v1 = min(rv, sv) + W;
    for (int cu = u0; cu <= u1; ++cu) {
        if (cu < 0 || cu >= G.nu) continue;
        for (int cv = v0; cv <= v1; ++cv) {
            if (cv < 0 || cv >= G.nv) continue;
            const int c = cellmap[(long)cu * G.nv + cv];
            if (c < 0) continue;
            const int i = (ru - cu + W) * NT + (rv - cv + W);
            const int j = (su - cu + W) * NT + (sv - cv + W);
            acc += T[(long)c * N2 * N2 + (long)i * N2 + j];
        }
    }
    coef[g] = scale * acc;
}

// Hermitian fold of the stencil's rows.  The stencil's input S (the
// checkerboard-signed spectrum on raw cells) satisfies S(-p) = conj(S(p)),
// and the final half-plane spectrum at (a, b >= 0) is the raw row at (a, b)
// plus the conjugate of the raw row at (-a, -b) (for b = 0 the Hermitian
// pair that the c2r transform takes).  With real coefficients
//   conj(R(-p)) = sign(p) sum_o c_(-p)[o] S(p - o),
// so the two rows are one stencil row at p with coefficients
//   c_p[o] + c_(-p)[-o]  (offset -o is index NO - 1 - o).
// Folded rows: cells (a, b >= 0) inside the band where either raw row exists.
__device__ inline int mirror_cell(const Geo &G, long g) {
    const int a = (int)(g / G.nv) - G.cu, b = (int)(g % G.nv) - G.cv;
    const int ma = -a + G.cu, mb = -b + G.cv;
    return (ma >= 0 && ma < G.nu && mb >= 0 && mb < G.nv) ? ma * G.nv + mb : -1;
}

__global__ void k_fold_mark(Geo G, const int32_t *__restrict__ rowmap, int64_t *__restrict__ mark) {
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long ng = (long)G.nu * G.nv;
    if (g >= ng) return;
    const int a = (int)(g / G.nv) - G.cu, b = (int)(g % G.nv) - G.cv;
    int m = 0;
    if (b >= 0 && b <= G.bv && abs(a) <= G.bu) {
        const int mc = mirror_cell(G, g);
        m = (rowmap[g] >= 0 || (mc >= 0 && rowmap[mc] >= 0)) ? 1 : 0;
    }
    mark[g] = m;
}

// new rowmap / row_cell, the raw rows of each folded row (r1: the cell's,
// r2: its mirror's; -1: none) and its band cell b * nu + (a mod nu)
__global__ void k_fold_rows(Geo G, const int64_t *__restrict__ off, const int32_t *__restrict__ rowmap_raw,
                            int32_t *__restrict__ rowmap, int32_t *__restrict__ row_cell, int32_t *__restrict__ r1,
                            int32_t *__restrict__ r2, int32_t *__restrict__ dst) {
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long ng = (long)G.nu * G.nv;
    if (g >= ng) return;
    if (off[g + 1] == off[g]) {
        rowmap[g] = -1;
        return;
    }
    const int32_t fr = (int32_t)off[g];
    rowmap[g] = fr;
    row_cell[fr] = (int32_t)g;
    const int mc = mirror_cell(G, g);
    r1[fr] = rowmap_raw[g];
    r2[fr] = mc >= 0 ? rowmap_raw[mc] : -1;
    const int a = (int)(g / G.nv) - G.cu, b = (int)(g % G.nv) - G.cv;
    dst[fr] = b * G.nu + (a < 0 ? a + G.nu : a);
}

__global__ void k_fold_coef(long nrow, int NO, const int32_t *__restrict__ r1, const int32_t *__restrict__ r2,
                            long nrow_raw, const double *__restrict__ coef_raw, double *__restrict__ coef) {
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nrow * NO) return;
    const int o = (int)(g / nrow);
    const long row = g - (long)o * nrow;
    const int a = r1[row], b = r2[row];
    double c = 0.0;
    if (a >= 0) c = coef_raw[(long)o * nrow_raw + a];
    if (b >= 0) c += coef_raw[(long)(NO - 1 - o) * nrow_raw + b];
    coef[g] = c;
}

__global__ void k_stencil_tiles(long nrow, const int32_t *__restrict__ row_cell, int nv, int st, int ntv,
                                int64_t *__restrict__ tmark) {
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrow) return;
    const int rc = row_cell[r];
    const int ru = rc / nv, rv = rc % nv;
    tmark[(long)(ru / st) * ntv + rv / st] = 1;
}

__global__ void k_compact(long n, const int64_t *__restrict__ off, int32_t *__restrict__ out) {
    long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n && off[g + 1] != off[g]) out[off[g]] = (int32_t)g;
}

// R[s][row] = sign(r') sum_o coef[o][row] S_s(r' + o) for NB samples per
// block, written to out[s * ostride + (odst ? odst[row] : row)] (the rows'
// band cells, or a compact per-row buffer); S windows of the tile staged
// in LDS with fold / wrap / sign.  The
// workgroups of one tile (one per group of NB samples) run back to back on
// one XCD, so the tile's coefficients come from that XCD's L2 after the
// first group.  RT > 0: the radius R compiled in (the offset loop unrolled,
// its coefficient loads issued together); RT = 0: R at run time.
template <class T, int NB, class TC, bool BAND = false, int RT = 0>
__global__ void __launch_bounds__(256) k_stencil_apply(Geo G, int R_, int ST, int ntv_st, const int32_t *__restrict__ tiles,
                                                       int ngroup, const int32_t *__restrict__ rowmap, long nrow,
                                                       const TC *__restrict__ coef, const Cplx<T> *__restrict__ spec,
                                                       long sstride, Cplx<T> *__restrict__ out,
                                                       const int32_t *__restrict__ odst, long ostride,
                                                       const int *__restrict__ active, int nactive, int nbv) {
    extern __shared__ unsigned char smem[];
    const int R = RT > 0 ? RT : R_;
    const int WS = ST + 2 * R;
    // the window as separate real and imaginary planes of NB x WS rows of
    // WP (odd) values: the staging writes (stride WP) and the reads (16
    // consecutive values per lane row) are free of LDS bank conflicts
    const int WP = WS | 1;
    T *wre = reinterpret_cast<T *>(smem);
    T *wim = wre + NB * WS * WP;
    const long lb = xcd_block(blockIdx.x, gridDim.x);
    const int grp = (int)(lb % ngroup);
    const int tile = tiles[lb / ngroup];
    const int tu = tile / ntv_st, tv = tile - tu * ntv_st;
    const int u0 = tu * ST - R, v0 = tv * ST - R;  // centred coords of window (0, 0)
    int smp[NB];
    int ns = 0;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        const int q = grp * NB + k;
        smp[k] = q < nactive ? (active ? active[q] : q) : -1;
        if (smp[k] >= 0) ns = k + 1;
    }
    for (int k = 0; k < ns; ++k) {
        const Cplx<T> *grid = spec + (long)smp[k] * sstride;
        for (int i = threadIdx.x; i < WS * WS; i += blockDim.x) {
            int wa, wb;
            Cplx<T> c;
            if (BAND) {  // the band is stored [v][u]: consecutive threads take consecutive u
                wb = i / WS;
                wa = i - wb * WS;
                c = fetch_band(G, grid, u0 + wa - G.cu, v0 + wb - G.cv, nbv);
            } else {
                wa = i / WS;
                wb = i - wa * WS;
                c = fetch_folded(G, grid, u0 + wa - G.cu, v0 + wb - G.cv);
            }
            wre[k * WS * WP + wa * WP + wb] = c.re;
            wim[k * WS * WP + wa * WP + wb] = c.im;
        }
    }
    __syncthreads();
    const int ly = threadIdx.x / ST, lx = threadIdx.x - ly * ST;
    const int ru = tu * ST + ly, rv = tv * ST + lx;
    if (ru >= G.nu || rv >= G.nv) return;
    const int row = rowmap[(long)ru * G.nv + rv];
    if (row < 0) return;
    T ar[NB], ai[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) ar[k] = ai[k] = (T)0;
    const int NR = 2 * R + 1;
    for (int ou = 0; ou < NR; ++ou) {
#pragma unroll
        for (int ov = 0; ov < NR; ++ov) {
            const T cf = (T)coef[(long)(ou * NR + ov) * nrow + row];
            const int wi = (ly + ou) * WP + (lx + ov);
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                ar[k] += cf * wre[k * WS * WP + wi];
                ai[k] += cf * wim[k * WS * WP + wi];
            }
        }
    }
    const T sg = (((ru - G.cu) + (rv - G.cv)) & 1) ? (T)-1 : (T)1;
    const long d = odst ? (long)odst[row] : (long)row;
    for (int k = 0; k < ns; ++k) out[(long)smp[k] * ostride + d] = Cplx<T>{sg * ar[k], sg * ai[k]};
}

// --- FP64 normal stencil on the matrix cores ------------------------------
// A 4 x 4 tile of stencil rows against its 16 x 16 window of grid cells is a
// dense 16 x 256 matrix (each row's 169 coefficients placed at their cells,
// zeros elsewhere) times the window's values of 8 samples (16 real columns:
// real and imaginary parts, the coefficients being real).
// v_mfma_f64_16x16x4f64 takes it in 64 steps of 4 window cells.  Operand
// layout of that instruction: lane l holds A[l % 16][l / 16] and
// B[l / 16][l % 16], and element r of its result D[l / 16 + 4 r][l % 16].
// The matrix is packed once per stencil in that layout and read from
// L2 while a workgroup runs over its sample groups; one wave per tile of a
// block of kMfBTU x kMfBTV tiles, the block's window staged in LDS per group
// of 8 samples and shared by its tiles.
constexpr int kMfT = 4;                        // tile edge (rows)
constexpr int kMfBTU = 2, kMfBTV = 4;          // tiles per block along u and v (one wave each)
constexpr int kMfNT = kMfBTU * kMfBTV;         // tiles per block
constexpr int kMfBU = kMfBTU * kMfT, kMfBV = kMfBTV * kMfT;  // block edges (rows)
constexpr int kMfR = 6;                        // stencil radius
constexpr int kMfWU = kMfBU + 2 * kMfR, kMfWV = kMfBV + 2 * kMfR;  // block window (cells)
// staged window: cells in pairs along u (u even, u + 1), each column's two
// values side by side, so that one 16-byte LDS read gives a lane its values
// for two matrix steps; rows of pairs padded by 4 doubles
constexpr int kMfR2 = kMfWV * 32 + 4;
static_assert(kMfWU % 2 == 0 && kMfT % 2 == 0, "pairs along u need even window and tile edges");
typedef double lt_double4 __attribute__((ext_vector_type(4)));

__global__ void k_mf_blocks(long nrow, const int32_t *__restrict__ row_cell, int nv, int ntv_b,
                            int64_t *__restrict__ bmark) {
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrow) return;
    const int rc = row_cell[r];
    const int ru = rc / nv, rv = rc % nv;
    bmark[(long)(ru / kMfBU) * ntv_b + rv / kMfBV] = 1;
}

// A[((block * kMfNT + tile) * 32 + pair) * 64 + lane][half] with step
// 4 w + q = pair (w / 2) * 4 + q, half w % 2, and the band cells of the
// tiles' rows (odst; -1 where a cell is no stencil row)
template <class TC>
__global__ void k_mf_pack(Geo G, long nblock, const int32_t *__restrict__ blocks, int ntv_b,
                          const int32_t *__restrict__ rowmap, long nrow, const TC *__restrict__ coef,
                          const int32_t *__restrict__ odst, TC *__restrict__ A, int32_t *__restrict__ rows) {
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nblock * kMfNT * 64 * 64) return;
    const int l = (int)(g & 63), c = (int)((g >> 6) & 63);
    const long bt = g >> 12;
    const long b = bt / kMfNT;
    const int t = (int)(bt - b * kMfNT);
    const int blk = blocks[b];
    const int bu = blk / ntv_b, bv = blk - bu * ntv_b;
    const int i = l & 15, k = 4 * c + (l >> 4);
    const int iu = i >> 2, iv = i & 3;   // row within the tile
    const int wu = k >> 4, wv = k & 15;  // cell within the tile's window
    const int ru = bu * kMfBU + (t / kMfBTV) * kMfT + iu, rv = bv * kMfBV + (t % kMfBTV) * kMfT + iv;
    int row = -1;
    if (ru < G.nu && rv < G.nv) row = rowmap[(long)ru * G.nv + rv];
    const int ou = wu - iu, ov = wv - iv;  // offset index in [0, 2R]
    constexpr int NR = 2 * kMfR + 1;
    TC a = (TC)0;
    if (row >= 0 && ou >= 0 && ou < NR && ov >= 0 && ov < NR) a = coef[(long)(ou * NR + ov) * nrow + row];
    // steps in pairs (4 w + q, 4 (w + 1) + q), w even: one 16-byte load per two
    A[((bt * 32 + (c >> 3) * 4 + (c & 3)) * 64 + l) * 2 + ((c >> 2) & 1)] = a;
    if (c == 0 && l < 16) rows[bt * 16 + i] = row >= 0 ? odst[row] : -1;
}

// R[s][row] = sign(r') sum_o coef[o][row] S_s(r' + o) on the matrix cores,
// written to the rows' band cells out[s * ostride + cell].  Workgroup =
// (block, chunk of GPW sample groups): one wave per tile of the block; the
// waves stage the block's window for a group of 8 samples in LDS
// together, then each runs its tile's 64 matrix steps with the packed
// matrix read from L2 as it goes (the chunks of one block run back to back
// on one XCD) and the window's values from LDS.
// one group's window: per cell the band position and the fold's signs once,
// then the 8 samples' values (consecutive threads take consecutive u)
template <class T>
__device__ inline void mf_stage(const Geo &G, const Cplx<T> *__restrict__ spec, long sstride,
                                const int *__restrict__ active, int nactive, int g, int u0, int v0, int nbv,
                                int t, int nt, T *__restrict__ win) {
    long soff[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const int q = g * 8 + s;
        soff[s] = q < nactive ? (long)(active ? active[q] : q) * sstride : -1;
    }
    for (int ci = t; ci < kMfWU * kMfWV; ci += nt) {
        const int wv = ci / kMfWU, wu = ci - wv * kMfWU;
        const int a = u0 + wu - G.cu, bb = v0 + wv - G.cv;
        T *cell = win + (wu >> 1) * kMfR2 + wv * 32 + (wu & 1);
        if (abs(a) > G.bu || abs(bb) > G.bv || abs(bb) >= nbv) {
#pragma unroll
            for (int s = 0; s < 16; ++s) cell[2 * s] = (T)0;
            continue;
        }
        const T sg = ((a + bb) & 1) ? (T)-1 : (T)1;
        long off;
        T sim;
        if (bb >= 0) {
            off = (long)bb * G.nu + (a < 0 ? a + G.nu : a);
            sim = sg;
        } else {
            off = (long)(-bb) * G.nu + (-a < 0 ? -a + G.nu : -a);
            sim = -sg;
        }
        Cplx<T> z[8];
#pragma unroll
        for (int s = 0; s < 8; ++s) z[s] = soff[s] >= 0 ? spec[soff[s] + off] : Cplx<T>{(T)0, (T)0};
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            cell[4 * s] = sg * z[s].re;
            cell[4 * s + 2] = sim * z[s].im;
        }
    }
}

template <int GPW>
__global__ void __launch_bounds__(64 * kMfNT) k_stencil_mfma(Geo G, const int32_t *__restrict__ blocks, int ntv_b, int ngc,
                                                      const double *__restrict__ A, const int32_t *__restrict__ rows,
                                                      const Cplx<double> *__restrict__ spec, long sstride,
                                                      Cplx<double> *__restrict__ outp, long ostride,
                                                      const int *__restrict__ active, int nactive, int nbv) {
    extern __shared__ unsigned char smem[];
    double *win = reinterpret_cast<double *>(smem);  // [kMfWU / 2][kMfR2]: cell pairs along u
    const long lb = xcd_block(blockIdx.x, gridDim.x);
    const long b = lb / ngc;
    const int gc = (int)(lb - b * ngc);
    const int blk = blocks[b];
    const int bu = blk / ntv_b, bv = blk - bu * ntv_b;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int tu = w / kMfBTV, tv = w - tu * kMfBTV;
    const double2 *Aw = reinterpret_cast<const double2 *>(A) + (b * kMfNT + w) * 32 * 64 + l;
    int orow[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) orow[r] = rows[(b * kMfNT + w) * 16 + (l >> 4) + 4 * r];
    const int j = l & 15;  // column: sample j / 2 of the group, real (even) or imaginary (odd) part
    const int u0 = bu * kMfBU - kMfR, v0 = bv * kMfBV - kMfR;  // raw cell of window (0, 0)
    const double2 *wl =
        reinterpret_cast<const double2 *>(win + (tu * kMfT / 2) * kMfR2 + (tv * kMfT + (l >> 4)) * 32 + 2 * j);
    const int ngroups = (nactive + 7) / 8;
    const int g0 = gc * GPW, g1 = min(ngroups, g0 + GPW);
    for (int g = g0; g < g1; ++g) {
        __syncthreads();  // the previous group's reads are done
        mf_stage(G, spec, sstride, active, nactive, g, u0, v0, nbv, threadIdx.x, blockDim.x, win);
        __syncthreads();
        lt_double4 acc = {0.0, 0.0, 0.0, 0.0};
        // pairs of window rows (w, w + 1): steps 4 w + q, then 4 (w + 1) + q,
        // in the order of the vector form
#pragma unroll 2
        for (int pr = 0; pr < 8; ++pr) {
            double2 av[4], bv[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                av[q] = Aw[(pr * 4 + q) * 64];
                bv[q] = wl[(pr * kMfR2 + q * 4 * 32) / 2];
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[q].x, bv[q].x, acc, 0, 0, 0);
#pragma unroll
            for (int q = 0; q < 4; ++q) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[q].y, bv[q].y, acc, 0, 0, 0);
        }
        const int q = g * 8 + (j >> 1);
        if (q < nactive) {
            const int smp = active ? active[q] : q;
            double *out = reinterpret_cast<double *>(outp) + (long)smp * ostride * 2;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int cell = orow[r];
                if (cell < 0) continue;
                const int i = (l >> 4) + 4 * r;
                const int ru = bu * kMfBU + tu * kMfT + (i >> 2), rv = bv * kMfBV + tv * kMfT + (i & 3);
                const double sg = (((ru - G.cu) + (rv - G.cv)) & 1) ? -1.0 : 1.0;
                out[(long)cell * 2 + (j & 1)] = sg * acc[r];
            }
        }
    }
}

// The FP32 stencil of the precision form on the matrix cores: the same
// tiles, packed matrices (of the FP32 coefficients) and staged windows as
// k_stencil_mfma, with v_mfma_f32_16x16x4f32 (twice the FP64 matrix rate on
// this chip, half the operand bytes).  Its result layout differs from the
// FP64 instruction's: element r of D sits at row 4 (l / 16) + r, column
// l % 16 (checked on the GPU with a probe of known operands).  FP32
// accumulation: the results agree with the vector form to FP32 rounding.
typedef float lt_float4 __attribute__((ext_vector_type(4)));

template <int GPW>
__global__ void __launch_bounds__(64 * kMfNT) k_stencil_mfma32(Geo G, const int32_t *__restrict__ blocks, int ntv_b,
                                                        int ngc, const float *__restrict__ A,
                                                        const int32_t *__restrict__ rows,
                                                        const Cplx<float> *__restrict__ spec, long sstride,
                                                        Cplx<float> *__restrict__ outp, long ostride,
                                                        const int *__restrict__ active, int nactive, int nbv) {
    extern __shared__ unsigned char smem[];
    float *win = reinterpret_cast<float *>(smem);  // [kMfWU / 2][kMfR2]: cell pairs along u
    const long lb = xcd_block(blockIdx.x, gridDim.x);
    const long b = lb / ngc;
    const int gc = (int)(lb - b * ngc);
    const int blk = blocks[b];
    const int bu = blk / ntv_b, bv = blk - bu * ntv_b;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int tu = w / kMfBTV, tv = w - tu * kMfBTV;
    const float2 *Aw = reinterpret_cast<const float2 *>(A) + (b * kMfNT + w) * 32 * 64 + l;
    int orow[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) orow[r] = rows[(b * kMfNT + w) * 16 + 4 * (l >> 4) + r];
    const int j = l & 15;  // column: sample j / 2 of the group, real (even) or imaginary (odd) part
    const int u0 = bu * kMfBU - kMfR, v0 = bv * kMfBV - kMfR;
    const float2 *wl =
        reinterpret_cast<const float2 *>(win + (tu * kMfT / 2) * kMfR2 + (tv * kMfT + (l >> 4)) * 32 + 2 * j);
    const int ngroups = (nactive + 7) / 8;
    const int g0 = gc * GPW, g1 = min(ngroups, g0 + GPW);
    for (int g = g0; g < g1; ++g) {
        __syncthreads();  // the previous group's reads are done
        mf_stage(G, spec, sstride, active, nactive, g, u0, v0, nbv, threadIdx.x, blockDim.x, win);
        __syncthreads();
        lt_float4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
        for (int pr = 0; pr < 8; ++pr) {
            float2 av[4], bq[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                av[q] = Aw[(pr * 4 + q) * 64];
                bq[q] = wl[(pr * kMfR2 + q * 4 * 32) / 2];
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av[q].x, bq[q].x, acc, 0, 0, 0);
#pragma unroll
            for (int q = 0; q < 4; ++q) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av[q].y, bq[q].y, acc, 0, 0, 0);
        }
        const int q = g * 8 + (j >> 1);
        if (q < nactive) {
            const int smp = active ? active[q] : q;
            float *out = reinterpret_cast<float *>(outp) + (long)smp * ostride * 2;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int cell = orow[r];
                if (cell < 0) continue;
                const int i = 4 * (l >> 4) + r;
                const int ru = bu * kMfBU + tu * kMfT + (i >> 2), rv = bv * kMfBV + tv * kMfT + (i & 3);
                const float sg = (((ru - G.cu) + (rv - G.cv)) & 1) ? -1.f : 1.f;
                out[(long)cell * 2 + (j & 1)] = sg * acc[r];
            }
        }
    }
}

// the full-spectrum form: the half-plane spectrum from the folded rows'
// compact outputs (the cells without a row are zero)
template <class T>
__global__ void k_stencil_fold(Geo G, const int32_t *__restrict__ rowmap, long nrow, const Cplx<T> *__restrict__ rbuf,
                               Cplx<T> *__restrict__ spec, long sstride, const int *__restrict__ active) {
    const int smp = active ? active[blockIdx.y] : (int)blockIdx.y;
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)G.nu * G.half_nv) return;
    const int iu = (int)(g / G.half_nv), iv = (int)(g - (long)iu * G.half_nv);
    const int a = iu <= G.bu ? iu : iu - G.nu;
    Cplx<T> tot{(T)0, (T)0};
    if (abs(a) <= G.bu && iv <= G.bv) {
        const int r = rowmap[(long)(a + G.cu) * G.nv + (iv + G.cv)];
        if (r >= 0) tot = rbuf[(long)smp * nrow + r];
    }
    spec[(long)smp * sstride + g] = tot;
}

// Hermitian symmetrisation of band row v = 0 (what c2r assumes of column 0)
template <class T>
__global__ void k_herm_band0(int nu, long sstride, Cplx<T> *__restrict__ band) {
    Cplx<T> *Y = band + (long)blockIdx.y * sstride;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > nu / 2) return;
    const int m = (nu - k) % nu;
    const Cplx<T> a = Y[k], b = Y[m];
    Y[k] = Cplx<T>{(T)0.5 * (a.re + b.re), (T)0.5 * (a.im - b.im)};
    if (m != k) Y[m] = Cplx<T>{(T)0.5 * (b.re + a.re), (T)0.5 * (b.im - a.im)};
}

// slot rows of the zero-padded, apodised image: R[slot][i][k], k < nyp
template <class T, class TI>
__global__ void k_band_pad(int nx, int ny, int pady, int nyp, const TI *__restrict__ img,
                           const double *__restrict__ apod, const int *__restrict__ active, T *__restrict__ R) {
    const int slot = blockIdx.y;
    const int smp = active ? active[slot] : slot;
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)nx * nyp) return;
    const int i = (int)(g / nyp), k = (int)(g - (long)i * nyp);
    const int j = k - pady;
    T v = (T)0;
    if (j >= 0 && j < ny) {
        const long q = (long)i * ny + j;
        v = (T)(img[(long)smp * nx * ny + q] * (TI)apod[q]);
    }
    R[(long)slot * nx * nyp + g] = v;
}

// B[slot][v][u] = C[slot][u - padx][v] (0 outside the image rows), v < nbv:
// a 32 x 32 tile transpose through LDS
template <class T>
__global__ void __launch_bounds__(256) k_band_gather(int nx, int padx, int nu, int half_nv, int nbv,
                                                     const Cplx<T> *__restrict__ C, Cplx<T> *__restrict__ B) {
    __shared__ Cplx<T> tile[32][33];
    const int slot = blockIdx.z;
    const int u0 = blockIdx.x * 32, v0 = blockIdx.y * 32;
    const Cplx<T> *Cs = C + (long)slot * nx * half_nv;
    Cplx<T> *Bs = B + (long)slot * nbv * nu;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int r = ty; r < 32; r += 8) {  // read rows u (image rows), columns v
        const int u = u0 + r, v = v0 + tx;
        const int i = u - padx;
        Cplx<T> z{(T)0, (T)0};
        if (i >= 0 && i < nx && v < nbv) z = Cs[(long)i * half_nv + v];
        tile[r][tx] = z;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {  // write rows v, columns u
        const int v = v0 + r, u = u0 + tx;
        if (v < nbv && u < nu) Bs[(long)v * nu + u] = tile[tx][r];
    }
}

// C[slot][i][v] = F[slot][v][i + padx] for v < nbv, 0 for nbv <= v < half_nv
template <class T>
__global__ void __launch_bounds__(256) k_band_scatter(int nx, int padx, int nu, int half_nv, int nbv,
                                                      const Cplx<T> *__restrict__ F, Cplx<T> *__restrict__ C) {
    __shared__ Cplx<T> tile[32][33];
    const int slot = blockIdx.z;
    const int i0 = blockIdx.x * 32, v0 = blockIdx.y * 32;
    const Cplx<T> *Fs = F + (long)slot * nbv * nu;
    Cplx<T> *Cs = C + (long)slot * nx * half_nv;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int r = ty; r < 32; r += 8) {  // read rows v, columns i
        const int v = v0 + r, i = i0 + tx;
        Cplx<T> z{(T)0, (T)0};
        if (v < nbv && i < nx) z = Fs[(long)v * nu + i + padx];
        tile[r][tx] = z;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {  // write rows i, columns v
        const int i = i0 + r, v = v0 + tx;
        if (i < nx && v < half_nv) Cs[(long)i * half_nv + v] = tile[tx][r];
    }
}

// out[smp] (+)= alpha * apod * R[slot][i][j + pady]
template <class T, class TO>
__global__ void k_band_crop(int nx, int ny, int pady, int nyp, const T *__restrict__ R, const double *__restrict__ apod,
                            TO alpha, int accumulate, const int *__restrict__ active, TO *__restrict__ out) {
    const int slot = blockIdx.y;
    const int smp = active ? active[slot] : slot;
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)nx * ny) return;
    const int i = (int)(g / ny), j = (int)(g - (long)i * ny);
    const TO x = (TO)R[(long)slot * nx * nyp + (long)i * nyp + j + pady];
    const TO v = alpha * (x * (TO)apod[g]);
    TO *o = out + (long)smp * nx * ny + g;
    *o = accumulate ? *o + v : v;
}

__global__ void k_max_abs_b(long nrow, const int32_t *__restrict__ row_cell, int nv, int cv, int *__restrict__ out) {
    const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrow) return;
    const int b = (int)(row_cell[r] % nv) - cv;
    atomicMax(out, b < 0 ? -b : b);
}

// ---------------------------------------------------------------------------
// Fused row transforms of the band normal operator.  Each workgroup takes
// 2*PAIRS consecutive image rows of one sample; every pair (a, b) is
// transformed as one complex sequence a + i b of length N = nyp (the rows
// zero-padded along y) by an unnormalised mixed-radix (4, then 2) Stockham
// FFT in LDS, and split with the two-real-sequences identity.
//   forward: apodize + pad + r2c of the rows + pick the band's nbv columns,
//            written as the band's column-major input [slot][v][u];
//   inverse: Hermitian extension of the band rows + c2r + crop + apodize
//            (+ alpha, accumulate), written to the output images.
// They replace k_band_pad + rows r2c + k_band_gather and k_band_scatter +
// rows c2r + k_band_crop with one pass over the data each way.
// twd[e] = exp(-2 pi i e / N) (e < N); the inverse uses its conjugate.
// ---------------------------------------------------------------------------
template <class T, bool INV>
__device__ inline void dft4(Cplx<T> (&v)[4]) {
    const Cplx<T> t0{v[0].re + v[2].re, v[0].im + v[2].im}, t1{v[0].re - v[2].re, v[0].im - v[2].im};
    const Cplx<T> t2{v[1].re + v[3].re, v[1].im + v[3].im}, t3{v[1].re - v[3].re, v[1].im - v[3].im};
    v[0] = Cplx<T>{t0.re + t2.re, t0.im + t2.im};
    v[2] = Cplx<T>{t0.re - t2.re, t0.im - t2.im};
    if (!INV) {  // -i t3 = (t3.im, -t3.re)
        v[1] = Cplx<T>{t1.re + t3.im, t1.im - t3.re};
        v[3] = Cplx<T>{t1.re - t3.im, t1.im + t3.re};
    } else {
        v[1] = Cplx<T>{t1.re - t3.im, t1.im + t3.re};
        v[3] = Cplx<T>{t1.re + t3.im, t1.im - t3.re};
    }
}

// a times the twiddle w (its conjugate for the inverse), each part one
// explicit fma over one rounded product, so that every transform using it
// rounds the same way whatever the compiler's contraction choices
template <class T, bool INV>
__device__ inline Cplx<T> tw_mul(Cplx<T> a, Cplx<T> w) {
    const T wi = INV ? -w.im : w.im;
    return Cplx<T>{fma(a.re, w.re, -(a.im * wi)), fma(a.re, wi, a.im * w.re)};
}

// a b with the same explicit fma form: the radix-4 passes read w = w^k
// from their table and form w^2 = w w and w^3 = w w^2 (one table read per
// butterfly instead of three; every transform derives them alike)
template <class T>
__device__ inline Cplx<T> tw_sq(Cplx<T> a, Cplx<T> b) {
    return Cplx<T>{fma(a.re, b.re, -(a.im * b.im)), fma(a.re, b.im, a.im * b.re)};
}

// one Stockham stage of radix R over PAIRS sequences of length N in buf
template <class T, int N, int R, int PAIRS, bool INV>
__device__ inline void stockham_stage(Cplx<T> (*buf)[N], int Ns, const Cplx<T> *__restrict__ twd) {
    constexpr int NB = N / R * PAIRS;  // butterflies of the stage over all sequences
    constexpr int PER = (NB + 255) / 256;
    Cplx<T> v[PER][R];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int g = threadIdx.x + 256 * q;
        if (g < NB) {
            const int pr = g / (N / R), j = g - pr * (N / R), k = j % Ns;
            const int step = N / (Ns * R);
            Cplx<T> w[R];
            w[0] = twd[0];
            if (R > 1) w[1] = twd[k * step];
            if (R > 2) w[2] = tw_sq(w[1], w[1]);
            if (R > 3) w[3] = tw_sq(w[1], w[2]);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                v[q][r] = tw_mul<T, INV>(buf[pr][j + r * (N / R)], w[r]);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int g = threadIdx.x + 256 * q;
        if (g < NB) {
            const int pr = g / (N / R), j = g - pr * (N / R), k = j % Ns;
            if constexpr (R == 4) {
                dft4<T, INV>(v[q]);
                const int d = (j / Ns) * Ns * 4 + k;
#pragma unroll
                for (int r = 0; r < 4; ++r) buf[pr][d + r * Ns] = v[q][r];
            } else {
                const Cplx<T> b0{v[q][0].re + v[q][1].re, v[q][0].im + v[q][1].im};
                const Cplx<T> b1{v[q][0].re - v[q][1].re, v[q][0].im - v[q][1].im};
                const int d = (j / Ns) * Ns * 2 + k;
                buf[pr][d] = b0;
                buf[pr][d + Ns] = b1;
            }
        }
    }
    __syncthreads();
}

template <class T, int N, int PAIRS, bool INV>
__device__ inline void fft_rows(Cplx<T> (*buf)[N], const Cplx<T> *__restrict__ twd) {
    int Ns = 1;
#pragma unroll
    for (int n = N; n >= 4 && n % 4 == 0; n /= 4) {
        stockham_stage<T, N, 4, PAIRS, INV>(buf, Ns, twd);
        Ns *= 4;
    }
    if (Ns < N) stockham_stage<T, N, 2, PAIRS, INV>(buf, Ns, twd);
}

template <class T, int N, int PAIRS, class TI>
__global__ void __launch_bounds__(256) k_rows_fwd(int nx, int ny, int pady, int padx, int nu, int nbv,
                                                  const TI *__restrict__ img, const double *__restrict__ apod,
                                                  const int *__restrict__ active, const uint8_t *__restrict__ rowmask,
                                                  const Cplx<T> *__restrict__ twd, Cplx<T> *__restrict__ band) {
    __shared__ Cplx<T> buf[PAIRS][N];
    const int slot = blockIdx.y;
    const int smp = active ? active[slot] : slot;
    const int row0 = blockIdx.x * 2 * PAIRS;
    const TI *im = img + (long)smp * nx * ny;
    // rows known to be zero (rowmask 0) are not read; a block of such rows
    // only writes its zero band columns
    const uint8_t *rm = rowmask ? rowmask + (long)smp * nx : nullptr;
    bool any = false;
    for (int rr = 0; rr < 2 * PAIRS; ++rr) any = any || (row0 + rr < nx && (!rm || rm[row0 + rr]));
    Cplx<T> *B = band + (long)slot * nbv * nu;
    if (!any) {
        for (int q = threadIdx.x; q < nbv * 2 * PAIRS; q += 256) {
            const int v = q / (2 * PAIRS), rr = q - v * 2 * PAIRS;
            if (row0 + rr < nx) B[(long)v * nu + padx + row0 + rr] = Cplx<T>{(T)0, (T)0};
        }
        return;
    }
    for (int q = threadIdx.x; q < PAIRS * N; q += 256) {
        const int pr = q / N, k = q - pr * N, j = k - pady;
        const int ia = row0 + 2 * pr, ib = ia + 1;
        T re = (T)0, imv = (T)0;
        if (j >= 0 && j < ny) {
            if (ia < nx && (!rm || rm[ia])) re = (T)(im[(long)ia * ny + j] * (TI)apod[(long)ia * ny + j]);
            if (ib < nx && (!rm || rm[ib])) imv = (T)(im[(long)ib * ny + j] * (TI)apod[(long)ib * ny + j]);
        }
        buf[pr][k] = Cplx<T>{re, imv};
    }
    __syncthreads();
    fft_rows<T, N, PAIRS, false>(buf, twd);
    const T h = (T)0.5;
    for (int q = threadIdx.x; q < nbv * 2 * PAIRS; q += 256) {
        const int v = q / (2 * PAIRS), rr = q - v * 2 * PAIRS, pr = rr >> 1;
        if (row0 + rr >= nx) continue;
        const Cplx<T> z = buf[pr][v], zm = buf[pr][(N - v) & (N - 1)];
        const Cplx<T> x = (rr & 1) ? Cplx<T>{h * (z.im + zm.im), h * (zm.re - z.re)}
                                   : Cplx<T>{h * (z.re + zm.re), h * (z.im - zm.im)};
        B[(long)v * nu + padx + row0 + rr] = x;
    }
}

template <class T, int N, int PAIRS, class TO>
__global__ void __launch_bounds__(256) k_rows_inv(int nx, int ny, int pady, int padx, int nu, int nbv,
                                                  const Cplx<T> *__restrict__ band, const double *__restrict__ apod,
                                                  TO alpha, int accumulate, const int *__restrict__ active,
                                                  const uint8_t *__restrict__ rowmask,
                                                  const Cplx<T> *__restrict__ twd, TO *__restrict__ out) {
    __shared__ Cplx<T> buf[PAIRS][N];
    const int slot = blockIdx.y;
    const int smp = active ? active[slot] : slot;
    const int row0 = blockIdx.x * 2 * PAIRS;
    // output rows with rowmask 0 are not needed: not written, and a block of
    // only such rows does nothing
    const uint8_t *rm = rowmask ? rowmask + (long)smp * nx : nullptr;
    bool any = false;
    for (int rr = 0; rr < 2 * PAIRS; ++rr) any = any || (row0 + rr < nx && (!rm || rm[row0 + rr]));
    if (!any) return;
    const Cplx<T> *B = band + (long)slot * nbv * nu;
    constexpr int H = N / 2;
    // pass a: the a rows' Hermitian spectra; pass b: + i times the b rows'
    for (int pass = 0; pass < 2; ++pass) {
        for (int q = threadIdx.x; q < (H + 1) * PAIRS; q += 256) {
            const int v = q / PAIRS, pr = q - v * PAIRS;
            const int row = row0 + 2 * pr + pass;
            Cplx<T> c{(T)0, (T)0};
            // rows not needed are taken as zero (not read)
            if (v < nbv && row < nx && (!rm || rm[row])) c = B[(long)v * nu + padx + row];
            if (v == 0 || v == H) c.im = (T)0;
            // value at v and its conjugate at N - v
            if (pass == 0) {
                buf[pr][v] = c;
                if (v > 0 && v < H) buf[pr][N - v] = Cplx<T>{c.re, -c.im};
            } else {  // += i c at v, += i conj(c) at N - v
                Cplx<T> &z = buf[pr][v];
                z = Cplx<T>{z.re - c.im, z.im + c.re};
                if (v > 0 && v < H) {
                    Cplx<T> &zm = buf[pr][N - v];
                    zm = Cplx<T>{zm.re + c.im, zm.im + c.re};
                }
            }
        }
        __syncthreads();
    }
    fft_rows<T, N, PAIRS, true>(buf, twd);
    TO *o = out + (long)smp * nx * ny;
    for (int q = threadIdx.x; q < 2 * PAIRS * ny; q += 256) {
        const int rr = q / ny, j = q - rr * ny;
        const int row = row0 + rr;
        if (row >= nx || (rm && !rm[row])) continue;
        const Cplx<T> z = buf[rr >> 1][pady + j];
        const long g = (long)row * ny + j;
        const TO val = alpha * ((TO)((rr & 1) ? z.im : z.re) * (TO)apod[g]);
        o[g] = accumulate ? o[g] + val : val;
    }
}

// FP64 row transforms of length 1024 with the first and the last radix-4
// pass in registers: one row pair per 256-thread workgroup, thread j doing
// butterfly j of each of the five Stockham passes (the same butterflies,
// twiddles and operation order as k_rows_fwd / k_rows_inv, so the same
// bits).  The forward transform feeds its first pass from the image in
// global memory and writes back only the outputs the band needs; the
// inverse builds its first pass's inputs from the band's Hermitian
// extension in global memory and writes the image from its last pass.  The
// LDS exchanges between passes use 16-byte units swizzled within blocks of
// 16 (low 4 bits XORed with a function of the next two), which makes the
// writes after the passes of span 1 and 4 free of bank conflicts; reads of
// 16 consecutive units stay conflict-free under any such swizzle.
constexpr int kR4N = 1024;

// Workgroup order of the row-pair transforms over the 8 XCDs (workgroup b
// runs on XCD b % 8): quads of 4 consecutive row pairs, whose band values
// share 128-byte lines, run back to back on one XCD, and the quads (sample
// major) are dealt to the XCDs in turn, so that each XCD reads the same
// eighth of the apodization rows for every sample.  Grid:
// 32 * ceil(nslot * ceil(P / 4) / 8) workgroups for P row pairs per sample;
// false for the padding workgroups.
__device__ inline bool rowpair_of(long b, int P, int nslot, int &slot, int &rpair) {
    const int P4 = (P + 3) / 4;
    const long x = b & 7, i = b >> 3;
    const long Q = 8 * (i >> 2) + x;
    if (Q >= (long)nslot * P4) return false;
    slot = (int)(Q / P4);
    rpair = (int)(Q - (long)slot * P4) * 4 + (int)(i & 3);
    return rpair < P;
}
static unsigned rowpair_grid(int P, int nslot) {
    const long quads = (long)nslot * ((P + 3) / 4);
    return (unsigned)(32 * ((quads + 7) / 8));
}

template <int SW>
__device__ inline int r4swz(int i) {
    return SW == 1 ? (i ^ ((i >> 4) & 3)) : SW == 2 ? (i ^ (((i >> 4) & 3) << 2)) : i;
}

// The twiddles in LDS: the first quarter period tq[e] = w^e (e < 256, its
// position swizzled like the exchanges so that strided reads spread over the
// banks) and w^(e + 256 q) = (-i)^q w^e by exact quarter turns (the global
// table is built the same way, so the values are its bits).
__device__ inline int r4tsw(int e) { return e ^ ((e >> 4) & 15); }
// (-i)^q t is written as a swap (odd q) and two sign selects; a four-way
// select of the rotated pairs compiled to a wrong real part for q = 3 in
// FP32 (a probe of every m against the full table showed it), this form
// matches the table in both precisions.
template <class T>
__device__ inline Cplx<T> r4tw(const Cplx<T> *tq, int m) {  // m < 1024
    const Cplx<T> t = tq[r4tsw(m & 255)];
    const int q = m >> 8;
    const bool sw = (q & 1) != 0;
    const T x = sw ? t.im : t.re, y = sw ? t.re : t.im;
    return Cplx<T>{(q & 2) ? -x : x, ((q + 1) & 2) ? -y : y};
}

// a middle pass (span NS = 4, 16, 64): read with swizzle SI, write with SO
template <class T, bool INV, int NS, int SI, int SO>
__device__ inline void r4_pass(Cplx<T> *buf, const Cplx<T> *tq) {
    const int j = threadIdx.x, k = j & (NS - 1);
    constexpr int step = kR4N / (NS * 4);
    __syncthreads();  // the previous pass's writes (and the twiddle table's)
    Cplx<T> w[3];  // w^k, w^2k, w^3k
    w[0] = r4tw(tq, k * step);
    w[1] = tw_sq(w[0], w[0]);
    w[2] = tw_sq(w[0], w[1]);
    Cplx<T> v[4];
    v[0] = buf[r4swz<SI>(j)];
#pragma unroll
    for (int r = 1; r < 4; ++r) v[r] = tw_mul<T, INV>(buf[r4swz<SI>(j + 256 * r)], w[r - 1]);
    __syncthreads();  // every read before the writes
    dft4<T, INV>(v);
    const int d = (j / NS) * NS * 4 + k;
#pragma unroll
    for (int r = 0; r < 4; ++r) buf[r4swz<SO>(d + r * NS)] = v[r];
}

// the last pass (span 256): its four outputs j + 256 r in v
template <class T, bool INV>
__device__ inline void r4_last(Cplx<T> (&v)[4], const Cplx<T> *buf, const Cplx<T> *tq) {
    const int j = threadIdx.x;
    __syncthreads();
    Cplx<T> w[3];  // w^k, w^2k, w^3k
    w[0] = r4tw(tq, j);
    w[1] = tw_sq(w[0], w[0]);
    w[2] = tw_sq(w[0], w[1]);
    v[0] = buf[j];
#pragma unroll
    for (int r = 1; r < 4; ++r) v[r] = tw_mul<T, INV>(buf[j + 256 * r], w[r - 1]);
    dft4<T, INV>(v);
}

template <class T, class TI>
__global__ void __launch_bounds__(256) k_rows_fwd_r(int nslot, int nx, int ny, int pady, int padx, int nu, int nbv,
                                                    const TI *__restrict__ img, const double *__restrict__ apod,
                                                    const int *__restrict__ active, const uint8_t *__restrict__ rowmask,
                                                    const Cplx<T> *__restrict__ twd,
                                                    Cplx<T> *__restrict__ band) {
    constexpr int N = kR4N;
    __shared__ Cplx<T> buf[N];
    __shared__ Cplx<T> tq[256];
    int slot, rpair;
    if (!rowpair_of(blockIdx.x, (nx + 1) / 2, nslot, slot, rpair)) return;
    const int smp = active ? active[slot] : slot;
    const int row0 = rpair * 2, ia = row0, ib = row0 + 1;
    const TI *im = img + (long)smp * nx * ny;
    const uint8_t *rm = rowmask ? rowmask + (long)smp * nx : nullptr;
    const bool ua = ia < nx && (!rm || rm[ia]), ub = ib < nx && (!rm || rm[ib]);
    Cplx<T> *B = band + (long)slot * nbv * nu;
    if (!ua && !ub) {
        for (int q = threadIdx.x; q < nbv * 2; q += 256) {
            const int v = q >> 1, rr = q & 1;
            if (row0 + rr < nx) B[(long)v * nu + padx + row0 + rr] = Cplx<T>{(T)0, (T)0};
        }
        return;
    }
    const int j = threadIdx.x;
    tq[r4tsw(j)] = twd[j];
    // pass 1 (span 1, unit twiddles) from the apodized rows x[j + 256 r]
    Cplx<T> v[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int c = j + 256 * r - pady;
        T re = (T)0, imv = (T)0;
        if (c >= 0 && c < ny) {
            if (ua) re = (T)(im[(long)ia * ny + c] * (TI)apod[(long)ia * ny + c]);
            if (ub) imv = (T)(im[(long)ib * ny + c] * (TI)apod[(long)ib * ny + c]);
        }
        v[r] = Cplx<T>{re, imv};
    }
    dft4<T, false>(v);
#pragma unroll
    for (int r = 0; r < 4; ++r) buf[r4swz<1>(4 * j + r)] = v[r];
    r4_pass<T, false, 4, 1, 2>(buf, tq);
    r4_pass<T, false, 16, 2, 0>(buf, tq);
    r4_pass<T, false, 64, 0, 0>(buf, tq);
    r4_last<T, false>(v, buf, tq);
    __syncthreads();  // every read of the last pass before its writes
    // the outputs the band needs: z[o] for o < nbv and o > N - nbv
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int o = j + 256 * r;
        if (o < nbv || N - o < nbv) buf[o] = v[r];
    }
    __syncthreads();
    const T h = (T)0.5;
    for (int q = threadIdx.x; q < nbv * 2; q += 256) {
        const int vv = q >> 1, rr = q & 1;
        if (row0 + rr >= nx) continue;
        const Cplx<T> z = buf[vv], zm = buf[(N - vv) & (N - 1)];
        const Cplx<T> x = rr ? Cplx<T>{h * (z.im + zm.im), h * (zm.re - z.re)}
                                  : Cplx<T>{h * (z.re + zm.re), h * (z.im - zm.im)};
        B[(long)vv * nu + padx + row0 + rr] = x;
    }
}

template <class T, class TO>
__global__ void __launch_bounds__(256) k_rows_inv_r(int nslot, int nx, int ny, int pady, int padx, int nu, int nbv,
                                                    const Cplx<T> *__restrict__ band,
                                                    const double *__restrict__ apod, TO alpha, int accumulate,
                                                    const int *__restrict__ active,
                                                    const uint8_t *__restrict__ rowmask,
                                                    const Cplx<T> *__restrict__ twd, TO *__restrict__ out) {
    constexpr int N = kR4N, H = N / 2;
    __shared__ Cplx<T> buf[N];
    __shared__ Cplx<T> tq[256];
    int slot, rpair;
    if (!rowpair_of(blockIdx.x, (nx + 1) / 2, nslot, slot, rpair)) return;
    const int smp = active ? active[slot] : slot;
    const int row0 = rpair * 2, ia = row0, ib = row0 + 1;
    const uint8_t *rm = rowmask ? rowmask + (long)smp * nx : nullptr;
    const bool ua = ia < nx && (!rm || rm[ia]), ub = ib < nx && (!rm || rm[ib]);
    if (!ua && !ub) return;  // no output row needed
    const Cplx<T> *B = band + (long)slot * nbv * nu;
    const int j = threadIdx.x;
    tq[r4tsw(j)] = twd[j];
    // pass 1 inputs: z[k] = A[k] + i Bb[k] of the band rows' Hermitian
    // extension (conjugates above H; imaginary parts of columns 0 and H
    // dropped; rows not needed taken as zero)
    Cplx<T> v[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int k = j + 256 * r;
        const int vv = k <= H ? k : N - k;
        Cplx<T> ca{(T)0, (T)0}, cb{(T)0, (T)0};
        if (vv < nbv) {
            if (ua) ca = B[(long)vv * nu + padx + ia];
            if (ub) cb = B[(long)vv * nu + padx + ib];
        }
        if (vv == 0 || vv == H) {
            ca.im = (T)0;
            cb.im = (T)0;
        }
        v[r] = k <= H ? Cplx<T>{ca.re - cb.im, ca.im + cb.re} : Cplx<T>{ca.re + cb.im, -ca.im + cb.re};
    }
    dft4<T, true>(v);
#pragma unroll
    for (int r = 0; r < 4; ++r) buf[r4swz<1>(4 * j + r)] = v[r];
    r4_pass<T, true, 4, 1, 2>(buf, tq);
    r4_pass<T, true, 16, 2, 0>(buf, tq);
    r4_pass<T, true, 64, 0, 0>(buf, tq);
    r4_last<T, true>(v, buf, tq);
    TO *o = out + (long)smp * nx * ny;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int c = j + 256 * r - pady;
        if (c < 0 || c >= ny) continue;
        if (ua) {
            const long g = (long)ia * ny + c;
            const TO val = alpha * ((TO)v[r].re * (TO)apod[g]);
            o[g] = accumulate ? o[g] + val : val;
        }
        if (ub) {
            const long g = (long)ib * ny + c;
            const TO val = alpha * ((TO)v[r].im * (TO)apod[g]);
            o[g] = accumulate ? o[g] + val : val;
        }
    }
}

// The band's column transforms along u (length 1024, complex, one
// contiguous column per workgroup: slot s, band row v), with the same
// passes as the row transforms.  Forward: only the image rows u in
// [padx, padx + nx) hold values (the band's other cells are zero), so they
// alone are read; the last pass writes all 1024 outputs from registers.
// Inverse: all 1024 inputs; the last pass writes only the image rows, the
// only cells the row transforms read back.
template <class T>
__global__ void __launch_bounds__(256) k_cols_fwd_r(int nbv, int nu, int padx, int nx, long bsz,
                                                    const Cplx<T> *__restrict__ in, Cplx<T> *__restrict__ out,
                                                    const Cplx<T> *__restrict__ twd) {
    __shared__ Cplx<T> buf[kR4N];
    __shared__ Cplx<T> tq[256];
    const int slot = (int)(blockIdx.x / nbv), vv = (int)(blockIdx.x - (long)slot * nbv);
    const Cplx<T> *X = in + (long)slot * bsz + (long)vv * nu;
    Cplx<T> *Y = out + (long)slot * bsz + (long)vv * nu;
    const int j = threadIdx.x;
    tq[r4tsw(j)] = twd[j];
    Cplx<T> v[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int u = j + 256 * r;
        v[r] = (u >= padx && u < padx + nx) ? X[u] : Cplx<T>{(T)0, (T)0};
    }
    dft4<T, false>(v);
#pragma unroll
    for (int r = 0; r < 4; ++r) buf[r4swz<1>(4 * j + r)] = v[r];
    r4_pass<T, false, 4, 1, 2>(buf, tq);
    r4_pass<T, false, 16, 2, 0>(buf, tq);
    r4_pass<T, false, 64, 0, 0>(buf, tq);
    r4_last<T, false>(v, buf, tq);
#pragma unroll
    for (int r = 0; r < 4; ++r) Y[j + 256 * r] = v[r];
}

template <class T>
__global__ void __launch_bounds__(256) k_cols_inv_r(int nbv, int nu, int padx, int nx, long bsz,
                                                    const Cplx<T> *__restrict__ in, Cplx<T> *__restrict__ out,
                                                    const Cplx<T> *__restrict__ twd) {
    __shared__ Cplx<T> buf[kR4N];
    __shared__ Cplx<T> tq[256];
    const int slot = (int)(blockIdx.x / nbv), vv = (int)(blockIdx.x - (long)slot * nbv);
    const Cplx<T> *X = in + (long)slot * bsz + (long)vv * nu;
    Cplx<T> *Y = out + (long)slot * bsz + (long)vv * nu;
    const int j = threadIdx.x;
    tq[r4tsw(j)] = twd[j];
    Cplx<T> v[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = X[j + 256 * r];
    dft4<T, true>(v);
#pragma unroll
    for (int r = 0; r < 4; ++r) buf[r4swz<1>(4 * j + r)] = v[r];
    r4_pass<T, true, 4, 1, 2>(buf, tq);
    r4_pass<T, true, 16, 2, 0>(buf, tq);
    r4_pass<T, true, 64, 0, 0>(buf, tq);
    r4_last<T, true>(v, buf, tq);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int u = j + 256 * r;
        if (u >= padx && u < padx + nx) Y[u] = v[r];
    }
}

Geo make_geo(const NufftPlan &P) {
    Geo G;
    G.nu = P.nu;
    G.nv = P.nv;
    G.half_nv = P.half_nv;
    G.nrows = P.nrows;
    G.W = P.W;
    G.TU = P.TU;
    G.TV = P.TV;
    G.ntu = P.ntu;
    G.ntv = P.ntv;
    G.cu = P.cu;
    G.cv = P.cv;
    G.WU = P.WU;
    G.WV = P.WV;
    G.bu = (P.nu - 1) / 2;
    G.bv = (P.nv - 1) / 2;
    G.du = P.du;
    G.dv = P.dv;
    G.gcx = P.gcx;
    G.gcy = P.gcy;
    return G;
}

lt_status fft_check(rocfft_status s, const char *what) {
    if (s != rocfft_status_success) {
        set_error("rocFFT %s failed (status %d)", what, (int)s);
        return LT_ERR_FFT;
    }
    return LT_OK;
}

lt_status kb_poly(double beta, int prec, KbPoly &P) {
    // c_m = (beta^2/4)^m / (m!)^2; keep terms until the tail is below the
    // target relative to I0(beta) (the value at s = 1)
    const double q = 0.25 * beta * beta;
    double c = 1.0, sum = 1.0;
    P.c[0] = 1.0;
    int m = 1;
    for (; m < kMaxPoly; ++m) {
        c *= q / ((double)m * (double)m);
        P.c[m] = c;
        sum += c;
    }
    const double eps = (prec == LT_F64) ? 1e-18 : 1e-9;
    double tail = 0.0;
    int deg = kMaxPoly - 1;
    for (int k = kMaxPoly - 1; k >= 1; --k) {
        tail += P.c[k];
        if (tail > eps * sum) {
            deg = k;
            break;
        }
    }
    if (deg >= kMaxPoly - 1) {
        set_error("kb_beta %.3g too large for the kernel polynomial", beta);
        return LT_ERR_ARG;
    }
    P.deg = deg;
    return LT_OK;
}

}  // namespace

// g(x) = I0(beta sqrt(1 - (x/W)^2)) continued analytically to |x| > W
// (where it is J0(beta sqrt((x/W)^2 - 1))): the entire power series in s
double kb_entire(double beta, int W, double x) {
    const double s = 1.0 - (x / W) * (x / W);
    const double q = 0.25 * beta * beta * s;
    double term = 1.0, acc = 1.0;
    for (int m = 1; m < 400; ++m) {
        term *= q / ((double)m * (double)m);
        acc += term;
        if (m > std::fabs(q) && std::fabs(term) < 1e-19 * std::fabs(acc)) break;
    }
    return acc;
}

namespace {

// Chebyshev interpolation at n = D+1 nodes of y(t), t in [-1, 1] (the
// samples y[j] at t_j = cos(pi (j + 1/2) / n)), converted to monomials in
// f = t / 2: mono[e], e = 0..D
void cheb_to_mono(const std::vector<double> &y, int n, double *mono) {
    std::vector<double> cheb(n), tk(n), tkm1(n), tkp1(n), m(n, 0.0);
    for (int i = 0; i < n; ++i) {
        double a = 0.0;
        for (int j = 0; j < n; ++j) a += y[j] * std::cos(M_PI * i * (j + 0.5) / n);
        cheb[i] = (i == 0 ? 1.0 : 2.0) * a / n;
    }
    // sum_i cheb_i T_i(t) as monomials in t, with
    // T_0 = 1, T_1 = t, T_{i+1} = 2 t T_i - T_{i-1}
    std::fill(tkm1.begin(), tkm1.end(), 0.0);
    std::fill(tk.begin(), tk.end(), 0.0);
    tkm1[0] = 1.0;
    for (int e = 0; e < n; ++e) m[e] += cheb[0] * tkm1[e];
    if (n > 1) {
        tk[1] = 1.0;
        for (int e = 0; e < n; ++e) m[e] += cheb[1] * tk[e];
    }
    for (int i = 2; i < n; ++i) {
        for (int e = 0; e < n; ++e) tkp1[e] = (e > 0 ? 2.0 * tk[e - 1] : 0.0) - tkm1[e];
        for (int e = 0; e < n; ++e) m[e] += cheb[i] * tkp1[e];
        tkm1.swap(tk);
        tk.swap(tkp1);
    }
    double sc = 1.0;
    for (int e = 0; e < n; ++e, sc *= 2.0) mono[e] = m[e] * sc;
}

// fits of the taps at distances f + c_k, f in [-1/2, 1/2] (tap k), and the
// max error relative to the peak over a dense sample; edge_zero: tap 0 is
// skipped for f > 0 and the last tap for f < 0 (outside the support there)
double fit_taps(double beta, int W, int D, int NT, double c0, bool edge_zero, std::vector<double> &mono) {
    const int n = D + 1;
    mono.assign((size_t)NT * n, 0.0);
    std::vector<double> y(n);
    for (int k = 0; k < NT; ++k) {
        for (int j = 0; j < n; ++j) y[j] = kb_entire(beta, W, 0.5 * std::cos(M_PI * (j + 0.5) / n) + c0 - k);
        cheb_to_mono(y, n, &mono[(size_t)k * n]);
    }
    const double peak = kb_entire(beta, W, 0.0);
    double err = 0.0;
    for (int k = 0; k < NT; ++k) {
        for (int i = 0; i <= 512; ++i) {
            const double f = -0.5 + i / 512.0;
            if (edge_zero && ((k == 0 && f > 0) || (k == NT - 1 && f < 0))) continue;
            double acc = mono[(size_t)k * n + D];
            for (int e = D - 1; e >= 0; --e) acc = acc * f + mono[(size_t)k * n + e];
            err = std::max(err, std::fabs(acc - kb_entire(beta, W, f + c0 - k)) / peak);
        }
    }
    return err;
}

}  // namespace

// Chebyshev interpolation of every tap on f in [-1/2, 1/2] at D+1 nodes,
// converted to monomials in f; returns the max error over a dense sample
// (edge taps only where they are nonzero) relative to the peak I0(beta)
double kb_taps_fit(double beta, int W, int D, std::vector<double> &mono) {
    return fit_taps(beta, W, D, 2 * W + 1, (double)W, true, mono);
}

// The 2W taps of the window starting W - 1 cells below floor(x), x = u/du:
// tap j at distance s + W - 1/2 - j with s = x - floor(x) - 1/2 in
// [-1/2, 1/2], all inside the support (|distance| <= W)
double kb_taps_fit_half(double beta, int W, int D, std::vector<double> &mono) {
    return fit_taps(beta, W, D, 2 * W, W - 0.5, false, mono);
}

namespace {

template <class T>
const T *taps_of(const NufftPlan &P);
template <>
const float *taps_of<float>(const NufftPlan &P) {
    return P.taps32.p;
}
template <>
const double *taps_of<double>(const NufftPlan &P) {
    return P.taps64.p;
}

template <class T>
const T *taps_half_of(const NufftPlan &P);
template <>
const float *taps_half_of<float>(const NufftPlan &P) {
    return P.taps32h.p;
}
template <>
const double *taps_half_of<double>(const NufftPlan &P) {
    return P.taps64h.p;
}

template <class T>
lt_status launch_degrid(NufftPlan &P, const void *grid, void *vis, const double *wts, bool acc) {
    if (P.nchunks == 0) return LT_OK;
    Geo G = make_geo(P);
    const T *kc = taps_of<T>(P);
    const T *kch = P.degrid_half ? taps_half_of<T>(P) : nullptr;
    const size_t sh = (size_t)P.WU * P.WV * sizeof(Cplx<T>);
    const T *ph = P.phase ? (P.prec == LT_F64 ? (const T *)P.ph64.p : (const T *)P.ph32.p) : nullptr;
    const T *fr = P.prec == LT_F64 ? (const T *)P.frac64.p : (const T *)P.frac32.p;
#define LT_DEGRID_CASE(WW)                                                                                        \
    case WW:                                                                                                      \
        if (P.TV == 32)                                                                                           \
            hipLaunchKernelGGL((k_degrid32<T, WW>), dim3((unsigned)P.nchunks), dim3(kNThreads), sh, stream(), G, \
                               kc, kch, (const Cplx<T> *)grid, P.ck_tile.p, P.ck_beg.p, P.ck_end.p, P.ngp.p, fr, ph,   \
                               wts, (Cplx<T> *)vis, acc ? 1 : 0);                                                 \
        else                                                                                                      \
            hipLaunchKernelGGL((k_degrid<T, WW>), dim3((unsigned)P.nchunks), dim3(kNThreads), sh, stream(), G,   \
                               kc, (const Cplx<T> *)grid, P.ck_tile.p, P.ck_beg.p, P.ck_end.p, P.ngp.p, fr, ph,   \
                               wts, (Cplx<T> *)vis, acc ? 1 : 0);                                                 \
        break;
    switch (P.W) {
        LT_DEGRID_CASE(1)
        LT_DEGRID_CASE(2)
        LT_DEGRID_CASE(3)
        LT_DEGRID_CASE(4)
        LT_DEGRID_CASE(5)
        LT_DEGRID_CASE(6)
        LT_DEGRID_CASE(7)
        LT_DEGRID_CASE(8)
        default:
            set_error("kernel_support %d not supported (1..8)", P.W);
            return LT_ERR_UNSUPPORTED;
    }
#undef LT_DEGRID_CASE
    LT_CHECK_LAUNCH();
    return LT_OK;
}

template <class T>
lt_status launch_grid(NufftPlan &P, const void *vis, const double *wts, void *grid_out) {
    Geo G = make_geo(P);
    const T *kc = taps_of<T>(P);
    const T *ph = P.phase ? (P.prec == LT_F64 ? (const T *)P.ph64.p : (const T *)P.ph32.p) : nullptr;
    const T *fr = P.prec == LT_F64 ? (const T *)P.frac64.p : (const T *)P.frac32.p;
    if (P.nchunks > 0) {
        const size_t sh = (size_t)P.WU * P.WV * sizeof(Cplx<T>);
#define LT_GRID_CASE(WW)                                                                                        \
    case WW:                                                                                                    \
        hipLaunchKernelGGL((k_grid<T, WW>), dim3((unsigned)P.nchunks), dim3(kNThreads), sh, stream(), G, kc, \
                           (const Cplx<T> *)vis, wts, P.ck_tile.p, P.ck_beg.p, P.ck_end.p, P.ngp.p, fr, ph,      \
                           (Cplx<T> *)P.partial.p);                                                             \
        break;
        switch (P.W) {
            LT_GRID_CASE(1)
            LT_GRID_CASE(2)
            LT_GRID_CASE(3)
            LT_GRID_CASE(4)
            LT_GRID_CASE(5)
            LT_GRID_CASE(6)
            LT_GRID_CASE(7)
            LT_GRID_CASE(8)
            default:
                set_error("kernel_support %d not supported (1..8)", P.W);
                return LT_ERR_UNSUPPORTED;
        }
#undef LT_GRID_CASE
        LT_CHECK_LAUNCH();
    }
    const long ng = (long)P.nu * P.half_nv;
    hipLaunchKernelGGL((k_grid_reduce<T>), dim3(blocks_for(ng, kNThreads)), dim3(kNThreads), 0, stream(), G,
                       P.tile_ck0.p, P.tile_ck1.p, (const Cplx<T> *)P.partial.p, (Cplx<T> *)grid_out);
    LT_CHECK_LAUNCH();
    return LT_OK;
}

}  // namespace

NufftPlan::~NufftPlan() {
    if (fwd) rocfft_plan_destroy(fwd);
    if (inv) rocfft_plan_destroy(inv);
    if (fwd_info) rocfft_execution_info_destroy(fwd_info);
    if (inv_info) rocfft_execution_info_destroy(inv_info);
}

lt_status NufftPlan::create(const visibility_space &vs, const fourier_space &fs, int wsup, double kb_beta,
                            int precision) {
    LT_TRY(ensure_init());
    if (wsup < 1 || wsup > 8) {
        set_error("kernel_support must be in 1..8 (got %d)", wsup);
        return LT_ERR_UNSUPPORTED;
    }
    if (precision != LT_F64 && precision != LT_F32) {
        set_error("precision must be LT_F64 or LT_F32");
        return LT_ERR_ARG;
    }
    if (fs.nu <= 2 * wsup + 2 || fs.nv <= 2 * wsup + 2 || fs.half_nv != fs.nv / 2 + 1 || fs.nu > 65000 ||
        fs.nv > 65000) {
        set_error("bad fourier space %d x %d (half %d)", fs.nu, fs.nv, fs.half_nv);
        return LT_ERR_ARG;
    }
    prec = precision;
    W = wsup;
    beta = kb_beta;
    nu = fs.nu;
    nv = fs.nv;
    half_nv = fs.half_nv;
    du = fs.du;
    dv = fs.dv;
    gcx = fs.gcx;
    gcy = fs.gcy;
    phase = (gcx != 0.0 || gcy != 0.0);
    nrows = vs.nrows;
    nch = vs.nchannels;
    nvis = (long)nrows * nch;
    uvw_host.assign(vs.uv, vs.uv + 3L * nrows);
    chan_host.assign(vs.channels, vs.channels + nch);
    if (nvis >= INT32_MAX) {
        set_error("too many visibilities (%ld)", nvis);
        return LT_ERR_CAPACITY;
    }
    LT_TRY(kb_poly(beta, prec, poly));
    {
        std::vector<double> mono;
        const int D = tap_degree(prec == LT_F64, W);
        taps_err = kb_taps_fit(beta, W, D, mono);
        const double tol = (prec == LT_F64) ? 1e-13 : 1e-6;
        if (!(taps_err <= tol)) {
            set_error("kb_beta %.4g with kernel_support %d: kernel polynomial error %.2g above %.1g", beta, W,
                      taps_err, tol);
            return LT_ERR_ARG;
        }
        if (prec == LT_F64) {
            LT_TRY(taps64.upload(mono.data(), (long)mono.size()));
        } else {
            std::vector<float> m32(mono.begin(), mono.end());
            LT_TRY(taps32.upload(m32.data(), (long)m32.size()));
        }
        std::vector<double> half;
        if (kb_taps_fit_half(beta, W, D, half) <= tol) {
            if (prec == LT_F64) {
                LT_TRY(taps64h.upload(half.data(), (long)half.size()));
            } else {
                std::vector<float> h32(half.begin(), half.end());
                LT_TRY(taps32h.upload(h32.data(), (long)h32.size()));
            }
        }
    }
    cu = nu / 2;
    cv = nv / 2;
    TU = std::min(TU, nu);
    TV = std::min(TV, nv);
    ntu = (nu + TU - 1) / TU;
    ntv = (nv + TV - 1) / TV;
    WU = TU + 2 * W;
    WV = TV + 2 * W;
    const long ntiles = (long)ntu * ntv;
    Geo G = make_geo(*this);
    DevBuf<double> duvw, dch;
    DevBuf<uint32_t> tile;
    DevBuf<int32_t> idx;
    DevBuf<int> bad;
    DevBuf<int32_t> cnt;
    LT_TRY(duvw.upload(vs.uv, 3L * nrows));
    LT_TRY(dch.upload(vs.channels, nch));
    LT_TRY(tile.alloc(std::max<long>(nvis, 1)));
    LT_TRY(key_sorted.alloc(std::max<long>(nvis, 1)));
    LT_TRY(idx.alloc(std::max<long>(nvis, 1)));
    LT_TRY(perm.alloc(std::max<long>(nvis, 1)));
    LT_TRY(bad.alloc(1));
    LT_TRY(bad.zero());
    LT_TRY(cnt.alloc(ntiles));
    LT_TRY(cnt.zero());
    if (nvis > 0) {
        hipLaunchKernelGGL(k_vis_tile, dim3(blocks_for(nvis, 256)), dim3(256), 0, stream(), G, nvis, duvw.p, dch.p,
                           tile.p, idx.p, bad.p);
        LT_CHECK_LAUNCH();
        int hbad = 0;
        LT_TRY(bad.download(&hbad, 1));
        LT_TRY(sync());
        if (hbad) {
            set_error("visibility taps fall outside the Nyquist band of the %d x %d grid", nu, nv);
            return LT_ERR_NYQUIST;
        }
        // stable radix sort by (tile, cell): inside a cell, visibilities keep
        // the caller's order
        unsigned bits = 1;
        while ((1L << bits) < ntiles * TU * TV) ++bits;
        size_t tmpb = 0;
        LT_HIP(rocprim::radix_sort_pairs(nullptr, tmpb, tile.p, key_sorted.p, idx.p, perm.p, (size_t)nvis, 0u, bits,
                                         stream()));
        DevBuf<uint8_t> tmp;
        LT_TRY(tmp.alloc(std::max<size_t>(tmpb, 1)));
        LT_HIP(rocprim::radix_sort_pairs(tmp.p, tmpb, tile.p, key_sorted.p, idx.p, perm.p, (size_t)nvis, 0u, bits,
                                         stream()));
        hipLaunchKernelGGL(k_count_tiles, dim3(blocks_for(nvis, 256)), dim3(256), 0, stream(), nvis, tile.p,
                           (uint32_t)(TU * TV), cnt.p);
        LT_CHECK_LAUNCH();
    }
    LT_TRY(ngp.alloc(std::max<long>(nvis, 1)));
    if (prec == LT_F64) {
        LT_TRY(frac64.alloc(2 * std::max<long>(nvis, 1)));
        if (phase) LT_TRY(ph64.alloc(2 * std::max<long>(nvis, 1)));
    } else {
        LT_TRY(frac32.alloc(2 * std::max<long>(nvis, 1)));
        if (phase) LT_TRY(ph32.alloc(2 * std::max<long>(nvis, 1)));
    }
    if (nvis > 0) {
        if (prec == LT_F64)
            hipLaunchKernelGGL(k_vis_fill<double>, dim3(blocks_for(nvis, 256)), dim3(256), 0, stream(), G, nvis,
                               duvw.p, dch.p, perm.p, ngp.p, frac64.p, phase ? ph64.p : nullptr);
        else
            hipLaunchKernelGGL(k_vis_fill<float>, dim3(blocks_for(nvis, 256)), dim3(256), 0, stream(), G, nvis,
                               duvw.p, dch.p, perm.p, ngp.p, frac32.p, phase ? ph32.p : nullptr);
        LT_CHECK_LAUNCH();
    }
    // chunks: each tile's range split into near-equal pieces of <= chunk
    std::vector<int32_t> hc(ntiles);
    LT_TRY(cnt.download(hc.data(), ntiles));
    LT_TRY(sync());
    std::vector<int32_t> ct, cb, ce, t0(ntiles), t1(ntiles);
    long pos = 0;
    for (long t = 0; t < ntiles; ++t) {
        t0[t] = (int32_t)ct.size();
        const long n = hc[t];
        if (n > 0) {
            const long k = (n + chunk - 1) / chunk;
            for (long j = 0; j < k; ++j) {
                ct.push_back((int32_t)t);
                cb.push_back((int32_t)(pos + n * j / k));
                ce.push_back((int32_t)(pos + n * (j + 1) / k));
            }
        }
        pos += n;
        t1[t] = (int32_t)ct.size();
    }
    nchunks = (long)ct.size();
    if (nchunks) {
        LT_TRY(ck_tile.upload(ct.data(), nchunks));
        LT_TRY(ck_beg.upload(cb.data(), nchunks));
        LT_TRY(ck_end.upload(ce.data(), nchunks));
    }
    LT_TRY(tile_ck0.upload(t0.data(), ntiles));
    LT_TRY(tile_ck1.upload(t1.data(), ntiles));
    LT_TRY(partial.alloc(std::max<long>(nchunks, 1) * WU * WV * 2 * tsize()));
    LT_TRY(spec.alloc((long)nu * half_nv * 2 * tsize()));
    LT_TRY(visbuf.alloc(std::max<long>(nvis, 1) * 2 * tsize()));
    return sync();
}

lt_status NufftPlan::attach_image(const image_space &im, int pad) {
    if (pad < 2) {
        set_error("pad_factor must be >= 2 (the reference's zero padding breaks for pad = 1)");
        return LT_ERR_ARG;
    }
    has_image = true;
    nx = im.nx;
    ny = im.ny;
    padx = ((pad - 1) * nx + 1) / 2;  // fourier_operators.py:34-35
    pady = ((pad - 1) * ny + 1) / 2;
    nxp = nx + 2 * padx;
    nyp = ny + 2 * pady;
    if (nxp != nu || nyp != nv) {
        set_error("padded image %d x %d does not match the grid %d x %d", nxp, nyp, nu, nv);
        return LT_ERR_ARG;
    }
    // ApodizationCorrectionOperator, fourier_operators.py:54-68
    std::vector<double> ap((long)nx * ny), sx(nx), sy(ny);
    for (int i = 0; i < nx; ++i) {
        const double c = (i - 0.5 * nx) / nxp;
        const double a = std::sqrt(beta * beta - (2 * M_PI * W * c) * (2 * M_PI * W * c));
        sx[i] = std::sinh(a) / a;
    }
    for (int j = 0; j < ny; ++j) {
        const double c = (j - 0.5 * ny) / nyp;
        const double a = std::sqrt(beta * beta - (2 * M_PI * W * c) * (2 * M_PI * W * c));
        sy[j] = std::sinh(a) / a;
    }
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) ap[(long)i * ny + j] = 1.0 / (sx[i] * sy[j]);
    LT_TRY(apod.upload(ap.data(), ap.size()));
    LT_TRY(realbuf.alloc((long)nxp * nyp * tsize()));
    LT_TRY(imgbuf.alloc((long)nx * ny));
    // rocFFT plans: row-major (nxp, nyp) real <-> (nxp, nyp/2+1) complex
    LT_TRY(fft_check(rocfft_setup(), "setup"));
    const size_t len[2] = {(size_t)nyp, (size_t)nxp};
    const rocfft_precision rp = prec == LT_F64 ? rocfft_precision_double : rocfft_precision_single;
    LT_TRY(fft_check(rocfft_plan_create(&fwd, rocfft_placement_notinplace, rocfft_transform_type_real_forward, rp, 2,
                                        len, 1, nullptr),
                     "plan (r2c)"));
    LT_TRY(fft_check(rocfft_plan_create(&inv, rocfft_placement_notinplace, rocfft_transform_type_real_inverse, rp, 2,
                                        len, 1, nullptr),
                     "plan (c2r)"));
    size_t wf = 0, wi = 0;
    LT_TRY(fft_check(rocfft_plan_get_work_buffer_size(fwd, &wf), "work size"));
    LT_TRY(fft_check(rocfft_plan_get_work_buffer_size(inv, &wi), "work size"));
    LT_TRY(fft_check(rocfft_execution_info_create(&fwd_info), "info"));
    LT_TRY(fft_check(rocfft_execution_info_create(&inv_info), "info"));
    if (wf) {
        LT_TRY(fwd_work.alloc(wf));
        LT_TRY(fft_check(rocfft_execution_info_set_work_buffer(fwd_info, fwd_work.p, wf), "work buffer"));
    }
    if (wi) {
        LT_TRY(inv_work.alloc(wi));
        LT_TRY(fft_check(rocfft_execution_info_set_work_buffer(inv_info, inv_work.p, wi), "work buffer"));
    }
    LT_TRY(fft_check(rocfft_execution_info_set_stream(fwd_info, stream()), "stream"));
    LT_TRY(fft_check(rocfft_execution_info_set_stream(inv_info, stream()), "stream"));
    return sync();
}

lt_status NufftPlan::degrid(const void *grid_in, void *vis, const double *w, bool acc) {
    return prec == LT_F64 ? launch_degrid<double>(*this, grid_in, vis, w, acc)
                          : launch_degrid<float>(*this, grid_in, vis, w, acc);
}

lt_status NufftPlan::grid(const void *vis, const double *w, void *grid_out) {
    return prec == LT_F64 ? launch_grid<double>(*this, vis, w, grid_out) : launch_grid<float>(*this, vis, w, grid_out);
}

lt_status NufftPlan::forward(const double *img, void *vis, const double *w) {
    if (!has_image) {
        set_error("NUFFT plan has no image space");
        return LT_ERR_ARG;
    }
    const long n = (long)nxp * nyp;
    if (prec == LT_F64)
        hipLaunchKernelGGL(k_apod_pad<double>, dim3(blocks_for(n, 256)), dim3(256), 0, stream(), nx, ny, padx, pady,
                           nxp, nyp, img, apod.p, (double *)realbuf.p);
    else
        hipLaunchKernelGGL(k_apod_pad<float>, dim3(blocks_for(n, 256)), dim3(256), 0, stream(), nx, ny, padx, pady,
                           nxp, nyp, img, apod.p, (float *)realbuf.p);
    LT_CHECK_LAUNCH();
    void *in[1] = {realbuf.p};
    void *out[1] = {spec.p};
    LT_TRY(fft_check(rocfft_execute(fwd, in, out, fwd_info), "execute r2c"));
    return degrid(spec.p, vis, w, false);
}

lt_status NufftPlan::adjoint(const void *vis, const double *w, double *img, double alpha, bool acc) {
    if (!has_image) {
        set_error("NUFFT plan has no image space");
        return LT_ERR_ARG;
    }
    LT_TRY(grid(vis, w, spec.p));
    const int cols[2] = {0, nv / 2};
    const int ncol = (nv % 2 == 0) ? 2 : 1;
    for (int k = 0; k < ncol; ++k) {
        if (prec == LT_F64)
            hipLaunchKernelGGL(k_hermitian_cols<double>, dim3(blocks_for(nu / 2 + 1, 256)), dim3(256), 0, stream(), nu,
                               half_nv, cols[k], (Cplx<double> *)spec.p);
        else
            hipLaunchKernelGGL(k_hermitian_cols<float>, dim3(blocks_for(nu / 2 + 1, 256)), dim3(256), 0, stream(), nu,
                               half_nv, cols[k], (Cplx<float> *)spec.p);
        LT_CHECK_LAUNCH();
    }
    void *in[1] = {spec.p};
    void *out[1] = {realbuf.p};
    LT_TRY(fft_check(rocfft_execute(inv, in, out, inv_info), "execute c2r"));
    const long n = (long)nx * ny;
    if (prec == LT_F64)
        hipLaunchKernelGGL(k_crop_apod<double>, dim3(blocks_for(n, 256)), dim3(256), 0, stream(), nx, ny, padx, pady,
                           nyp, (const double *)realbuf.p, apod.p, alpha, acc ? 1 : 0, img);
    else
        hipLaunchKernelGGL(k_crop_apod<float>, dim3(blocks_for(n, 256)), dim3(256), 0, stream(), nx, ny, padx, pady,
                           nyp, (const float *)realbuf.p, apod.p, alpha, acc ? 1 : 0, img);
    LT_CHECK_LAUNCH();
    return LT_OK;
}


NufftPlan::BatchFFT::~BatchFFT() {
    if (fwd) rocfft_plan_destroy(fwd);
    if (inv) rocfft_plan_destroy(inv);
    if (fi) rocfft_execution_info_destroy(fi);
    if (ii) rocfft_execution_info_destroy(ii);
}

lt_status NufftPlan::batch_fft(int n, BatchFFT *&out, int precision) {
    if (precision < 0) precision = prec;
    for (auto &b : bfft)
        if (b->n == n && b->prec == precision) {
            out = b.get();
            return LT_OK;
        }
    auto b = std::make_unique<BatchFFT>();
    b->n = n;
    b->prec = precision;
    const size_t len[2] = {(size_t)nyp, (size_t)nxp};
    const rocfft_precision rp = precision == LT_F64 ? rocfft_precision_double : rocfft_precision_single;
    LT_TRY(fft_check(rocfft_plan_create(&b->fwd, rocfft_placement_notinplace, rocfft_transform_type_real_forward, rp,
                                        2, len, (size_t)n, nullptr),
                     "batched plan (r2c)"));
    LT_TRY(fft_check(rocfft_plan_create(&b->inv, rocfft_placement_notinplace, rocfft_transform_type_real_inverse, rp,
                                        2, len, (size_t)n, nullptr),
                     "batched plan (c2r)"));
    size_t wf = 0, wi = 0;
    LT_TRY(fft_check(rocfft_plan_get_work_buffer_size(b->fwd, &wf), "work size"));
    LT_TRY(fft_check(rocfft_plan_get_work_buffer_size(b->inv, &wi), "work size"));
    LT_TRY(fft_check(rocfft_execution_info_create(&b->fi), "info"));
    LT_TRY(fft_check(rocfft_execution_info_create(&b->ii), "info"));
    if (wf) {
        LT_TRY(b->fw.alloc(wf));
        LT_TRY(fft_check(rocfft_execution_info_set_work_buffer(b->fi, b->fw.p, wf), "work buffer"));
    }
    if (wi) {
        LT_TRY(b->iw.alloc(wi));
        LT_TRY(fft_check(rocfft_execution_info_set_work_buffer(b->ii, b->iw.p, wi), "work buffer"));
    }
    LT_TRY(fft_check(rocfft_execution_info_set_stream(b->fi, stream()), "stream"));
    LT_TRY(fft_check(rocfft_execution_info_set_stream(b->ii, stream()), "stream"));
    out = b.get();
    bfft.push_back(std::move(b));
    return LT_OK;
}

lt_status NufftPlan::ensure_batch(int nb) {
    if (nb <= batch) return LT_OK;
    if (!has_image) {
        set_error("NUFFT plan has no image space");
        return LT_ERR_ARG;
    }
    LT_TRY(realB.alloc((long)nb * nxp * nyp * tsize()));
    LT_TRY(specB.alloc((long)nb * nu * half_nv * 2 * tsize()));
    LT_TRY(specB2.alloc((long)nb * nu * half_nv * 2 * tsize()));
    LT_TRY(partialB.alloc((long)nb * std::max<long>(nchunks, 1) * WU * WV * 2 * tsize()));
    batch = nb;
    return LT_OK;
}

template <class T>
static lt_status normal_impl(NufftPlan &P, const double *imgs, double *out, const double *w, const int *active,
                             int nactive, double alpha, bool acc, const double *spec_mul, NormalStencil *st) {
    // active samples are packed into slots 0..nactive-1 (image side reads and
    // writes through `active`, the spectral side works on slots)
    if (nactive <= 0) return LT_OK;
    if (nactive > P.batch) {
        set_error("normal operator: %d samples exceed the batch %d", nactive, P.batch);
        return LT_ERR_ARG;
    }
    const int nrun = std::min(P.batch, (nactive + NufftPlan::kBatchStep - 1) / NufftPlan::kBatchStep *
                                           NufftPlan::kBatchStep);
    NufftPlan::BatchFFT *bf = nullptr;
    LT_TRY(P.batch_fft(nrun, bf));
    const long nreal = (long)P.nxp * P.nyp, nspec = (long)P.nu * P.half_nv, npix = (long)P.nx * P.ny;
    const bool bcb = spec_mul != nullptr;
    hipLaunchKernelGGL((k_pad_b<T>), dim3(blocks_for(nreal, 256), nactive), dim3(256), 0, stream(), P.nx, P.ny,
                       P.padx, P.pady, P.nxp, P.nyp, imgs, bcb ? nullptr : P.apod.p, (T *)P.realB.p, active);
    LT_CHECK_LAUNCH();
    void *in[1] = {P.realB.p};
    void *o1[1] = {P.specB.p};
    LT_TRY(fft_check(rocfft_execute(bf->fwd, in, o1, bf->fi), "execute batched r2c"));
    Cplx<T> *spec_out = (Cplx<T> *)P.specB2.p;
    const int *slots = nullptr;  // spectral kernels index slots directly
    {
        Geo G = make_geo(P);
        if (bcb) {
            spec_out = (Cplx<T> *)P.specB.p;
            hipLaunchKernelGGL((k_spec_mul_b<T>), dim3(blocks_for(nspec, 256), nactive), dim3(256), 0, stream(), nspec,
                               spec_mul, spec_out, nspec, slots);
            LT_CHECK_LAUNCH();
        } else if (st) {
            LT_TRY(st->apply(P, P.specB.p, nspec, spec_out, slots, nactive));
        } else {
            const long pstride = std::max<long>(P.nchunks, 1) * P.WU * P.WV;
            if (P.nchunks > 0) {
                const size_t sh = (size_t)P.WU * P.WV * sizeof(Cplx<T>) * 2;
                const T *fr = P.prec == LT_F64 ? (const T *)P.frac64.p : (const T *)P.frac32.p;
#define LT_NORMAL_CASE(WW)                                                                                         \
    case WW:                                                                                                       \
        hipLaunchKernelGGL((k_normal<T, WW>), dim3((unsigned)P.nchunks, nactive), dim3(kNThreads), sh, stream(), G, \
                           taps_of<T>(P), (const Cplx<T> *)P.specB.p, nspec, P.ck_tile.p, P.ck_beg.p, P.ck_end.p, P.ngp.p, \
                           fr, w, (Cplx<T> *)P.partialB.p, pstride, slots);                                        \
        break;
                switch (P.W) {
                    LT_NORMAL_CASE(1)
                    LT_NORMAL_CASE(2)
                    LT_NORMAL_CASE(3)
                    LT_NORMAL_CASE(4)
                    LT_NORMAL_CASE(5)
                    LT_NORMAL_CASE(6)
                    LT_NORMAL_CASE(7)
                    LT_NORMAL_CASE(8)
                    default:
                        set_error("kernel_support %d not supported (1..8)", P.W);
                        return LT_ERR_UNSUPPORTED;
                }
#undef LT_NORMAL_CASE
                LT_CHECK_LAUNCH();
            }
            hipLaunchKernelGGL((k_grid_reduce_b<T>), dim3(blocks_for(nspec, kNThreads), nactive), dim3(kNThreads), 0,
                               stream(), G, P.tile_ck0.p, P.tile_ck1.p, (const Cplx<T> *)P.partialB.p, pstride,
                               spec_out, nspec, slots);
            LT_CHECK_LAUNCH();
        }
        const int cols[2] = {0, P.nv / 2};
        const int ncol = (P.nv % 2 == 0) ? 2 : 1;
        for (int k = 0; k < ncol; ++k) {
            hipLaunchKernelGGL((k_hermitian_cols_b<T>), dim3(blocks_for(P.nu / 2 + 1, 256), nactive), dim3(256), 0,
                               stream(), P.nu, P.half_nv, cols[k], spec_out, nspec, slots);
            LT_CHECK_LAUNCH();
        }
    }
    void *i2[1] = {(void *)spec_out};
    void *o2[1] = {P.realB.p};
    LT_TRY(fft_check(rocfft_execute(bf->inv, i2, o2, bf->ii), "execute batched c2r"));
    hipLaunchKernelGGL((k_crop_b<T>), dim3(blocks_for(npix, 256), nactive), dim3(256), 0, stream(), P.nx, P.ny,
                       P.padx, P.pady, P.nxp, P.nyp, (const T *)P.realB.p, bcb ? nullptr : P.apod.p, alpha,
                       acc ? 1 : 0, out, active);
    LT_CHECK_LAUNCH();
    return LT_OK;
}

NufftPlan::BandFFT::~BandFFT() {
    for (int k = 0; k < 4; ++k) {
        if (plan[k]) rocfft_plan_destroy(plan[k]);
        if (info[k]) rocfft_execution_info_destroy(info[k]);
    }
}

lt_status NufftPlan::band_fft(int n, int precision, int nbv, BandFFT *&out) {
    for (auto &b : bandfft)
        if (b->n == n && b->prec == precision && b->nbv == nbv) {
            out = b.get();
            return LT_OK;
        }
    auto b = std::make_unique<BandFFT>();
    b->n = n;
    b->prec = precision;
    b->nbv = nbv;
    const rocfft_precision rp = precision == LT_F64 ? rocfft_precision_double : rocfft_precision_single;
    const size_t ly = (size_t)nyp, lu = (size_t)nu;
    // rows along y (r2c / c2r) and columns along u (c2c), all out of place
    LT_TRY(fft_check(rocfft_plan_create(&b->plan[0], rocfft_placement_notinplace, rocfft_transform_type_real_forward,
                                        rp, 1, &ly, (size_t)n * nx, nullptr),
                     "band plan (rows r2c)"));
    LT_TRY(fft_check(rocfft_plan_create(&b->plan[1], rocfft_placement_notinplace, rocfft_transform_type_complex_forward,
                                        rp, 1, &lu, (size_t)n * nbv, nullptr),
                     "band plan (columns forward)"));
    LT_TRY(fft_check(rocfft_plan_create(&b->plan[2], rocfft_placement_notinplace, rocfft_transform_type_complex_inverse,
                                        rp, 1, &lu, (size_t)n * nbv, nullptr),
                     "band plan (columns inverse)"));
    LT_TRY(fft_check(rocfft_plan_create(&b->plan[3], rocfft_placement_notinplace, rocfft_transform_type_real_inverse,
                                        rp, 1, &ly, (size_t)n * nx, nullptr),
                     "band plan (rows c2r)"));
    for (int k = 0; k < 4; ++k) {
        size_t ws = 0;
        LT_TRY(fft_check(rocfft_plan_get_work_buffer_size(b->plan[k], &ws), "work size"));
        LT_TRY(fft_check(rocfft_execution_info_create(&b->info[k]), "info"));
        if (ws) {
            LT_TRY(b->work[k].alloc(ws));
            LT_TRY(fft_check(rocfft_execution_info_set_work_buffer(b->info[k], b->work[k].p, ws), "work buffer"));
        }
        LT_TRY(fft_check(rocfft_execution_info_set_stream(b->info[k], stream()), "stream"));
    }
    out = b.get();
    bandfft.push_back(std::move(b));
    return LT_OK;
}

static lt_status ensure_twiddles(NufftPlan &P) {
    const int N = P.nyp;
    if (P.twd64.n == 2 * (size_t)N && P.twd32.n == 2 * (size_t)N) return LT_OK;
    std::vector<double> t64(2 * (size_t)N);
    std::vector<float> t32(2 * (size_t)N);
    // the first quarter period from cos / sin, the others by exact quarter
    // turns (times -i), so that a kernel holding only the first quarter
    // rebuilds every entry bit for bit
    for (int e = 0; e < N; ++e) {
        if (e < N / 4) {
            const double a = 2.0 * M_PI * (double)e / (double)N;
            t64[2 * e] = std::cos(a);
            t64[2 * e + 1] = -std::sin(a);
        } else {
            t64[2 * e] = t64[2 * (e - N / 4) + 1];
            t64[2 * e + 1] = -t64[2 * (e - N / 4)];
        }
        t32[2 * e] = (float)t64[2 * e];
        t32[2 * e + 1] = (float)t64[2 * e + 1];
    }
    LT_TRY(P.twd64.upload(t64.data(), t64.size()));
    LT_TRY(P.twd32.upload(t32.data(), t32.size()));
    return sync();
}

static bool fused_rows_ok(const NufftPlan &P, int nbv) {
    const int N = P.nyp;
    return P.fused_rows && N >= 64 && N <= 2048 && (N & (N - 1)) == 0 && P.nv == N && P.half_nv == N / 2 + 1 &&
           nbv >= 1 && nbv <= N / 2 + 1 && P.pady >= 0 && P.pady + P.ny <= N && P.padx >= 0 &&
           P.padx + P.nx <= P.nu;
}

template <class T>
static const Cplx<T> *twiddles_of(const NufftPlan &P);
template <>
const Cplx<double> *twiddles_of<double>(const NufftPlan &P) {
    return (const Cplx<double> *)P.twd64.p;
}
template <>
const Cplx<float> *twiddles_of<float>(const NufftPlan &P) {
    return (const Cplx<float> *)P.twd32.p;
}

// sequence pairs per workgroup: up to 16 KB of LDS, at most 4
template <class T, int N>
constexpr int rows_pairs() {
    constexpr int per = N * 2 * (int)sizeof(T);
    return per >= 16384 ? 1 : (16384 / per > 4 ? 4 : 16384 / per);
}

// writes the image rows of B; its padding rows stay as they are (zero)
template <class T, class TI>
static lt_status rows_forward(NufftPlan &P, int nbv, const TI *imgs, const int *active, int nactive, Cplx<T> *B,
                              const uint8_t *rowmask) {
    const Cplx<T> *tw = twiddles_of<T>(P);
    {
        if (P.nyp == kR4N && P.wave_rows) {
            hipLaunchKernelGGL((k_rows_fwd_r<T, TI>), dim3(rowpair_grid((P.nx + 1) / 2, nactive)), dim3(256), 0, stream(),
                               nactive, P.nx, P.ny, P.pady, P.padx, P.nu, nbv, imgs, P.apod.p, active, rowmask, tw, B);
            LT_CHECK_LAUNCH();
            return LT_OK;
        }
    }
#define LT_ROWS_FWD(NN)                                                                                            \
    case NN: {                                                                                                     \
        constexpr int PR = rows_pairs<T, NN>();                                                                    \
        hipLaunchKernelGGL((k_rows_fwd<T, NN, PR, TI>), dim3((P.nx + 2 * PR - 1) / (2 * PR), nactive), dim3(256), \
                           0, stream(), P.nx, P.ny, P.pady, P.padx, P.nu, nbv, imgs, P.apod.p, active, rowmask, tw, \
                           B);                                                                                     \
        break;                                                                                                     \
    }
    switch (P.nyp) {
        LT_ROWS_FWD(64)
        LT_ROWS_FWD(128)
        LT_ROWS_FWD(256)
        LT_ROWS_FWD(512)
        LT_ROWS_FWD(1024)
        LT_ROWS_FWD(2048)
        default:
            set_error("fused row transform: unsupported length %d", P.nyp);
            return LT_ERR_ARG;
    }
#undef LT_ROWS_FWD
    LT_CHECK_LAUNCH();
    return LT_OK;
}

template <class T, class TO>
static lt_status rows_inverse(NufftPlan &P, int nbv, const Cplx<T> *F, const int *active, int nactive, TO alpha,
                              bool acc, TO *out, const uint8_t *rowmask) {
    const Cplx<T> *tw = twiddles_of<T>(P);
    {
        if (P.nyp == kR4N && P.wave_rows) {
            hipLaunchKernelGGL((k_rows_inv_r<T, TO>), dim3(rowpair_grid((P.nx + 1) / 2, nactive)), dim3(256), 0, stream(),
                               nactive, P.nx, P.ny, P.pady, P.padx, P.nu, nbv, F, P.apod.p, alpha, acc ? 1 : 0, active,
                               rowmask, tw, out);
            LT_CHECK_LAUNCH();
            return LT_OK;
        }
    }
#define LT_ROWS_INV(NN)                                                                                            \
    case NN: {                                                                                                     \
        constexpr int PR = rows_pairs<T, NN>();                                                                    \
        hipLaunchKernelGGL((k_rows_inv<T, NN, PR, TO>), dim3((P.nx + 2 * PR - 1) / (2 * PR), nactive), dim3(256), \
                           0, stream(), P.nx, P.ny, P.pady, P.padx, P.nu, nbv, F, P.apod.p, alpha, acc ? 1 : 0,    \
                           active, rowmask, tw, out);                                                              \
        break;                                                                                                     \
    }
    switch (P.nyp) {
        LT_ROWS_INV(64)
        LT_ROWS_INV(128)
        LT_ROWS_INV(256)
        LT_ROWS_INV(512)
        LT_ROWS_INV(1024)
        LT_ROWS_INV(2048)
        default:
            set_error("fused row transform: unsupported length %d", P.nyp);
            return LT_ERR_ARG;
    }
#undef LT_ROWS_INV
    LT_CHECK_LAUNCH();
    return LT_OK;
}

// D^T C D through the stencil with a band-pruned separable transform: rows of
// the padded image are transformed along y (only the image's rows exist),
// only the stencil's band of nbv spectral columns is carried to the
// transform along u, and back the same way.  Equals the full 2-D path.
template <class T, class TI>
static lt_status normal_band_impl(NufftPlan &P, NormalStencil &St, const TI *imgs, TI *out, const int *active,
                                  int nactive, TI alpha, bool acc, const uint8_t *rowmask = nullptr) {
    if (nactive <= 0) return LT_OK;
    if (nactive > P.batch) {
        set_error("normal operator: %d samples exceed the batch %d", nactive, P.batch);
        return LT_ERR_ARG;
    }
    const int prec = sizeof(T) == 8 ? LT_F64 : LT_F32;
    const int nbv = St.nbv;
    const int nrun = std::min(P.batch, (nactive + NufftPlan::kBatchStep - 1) / NufftPlan::kBatchStep *
                                           NufftPlan::kBatchStep);
    NufftPlan::BandFFT *bf = nullptr;
    LT_TRY(P.band_fft(nrun, prec, nbv, bf));
    const long rsz = (long)P.nx * P.nyp, csz = (long)P.nx * P.half_nv, bsz = (long)nbv * P.nu;
    const size_t need_r = (size_t)P.batch * rsz * sizeof(T), need_c = (size_t)P.batch * csz * 2 * sizeof(T),
                 need_b = (size_t)P.batch * bsz * 2 * sizeof(T);
    const bool fused = fused_rows_ok(P, nbv);
    if (!fused && P.bandR.n < need_r) LT_TRY(P.bandR.alloc(need_r));
    if (!fused && P.bandC.n < need_c) LT_TRY(P.bandC.alloc(need_c));
    // the band buffers the row transforms and the stencil write hold zeros
    // outside the cells they write (from allocation on); the column
    // transforms write their outputs to the plan's scratch bandB / bandF
    const int pi = prec == LT_F64 ? 0 : 1;
    DevBuf<uint8_t> &bin = St.band_in[pi], &bout = St.band_out[pi];
    if (bin.n < need_b) {
        LT_TRY(bin.alloc(need_b));
        LT_TRY(bin.zero());
    }
    if (bout.n < need_b) {
        LT_TRY(bout.alloc(need_b));
        LT_TRY(bout.zero());
    }
    if (P.bandB.n < need_b) LT_TRY(P.bandB.alloc(need_b));
    if (P.bandF.n < need_b) LT_TRY(P.bandF.alloc(need_b));
    T *R = (T *)P.bandR.p;
    Cplx<T> *Cb = (Cplx<T> *)P.bandC.p, *Bin = (Cplx<T> *)bin.p, *Bf = (Cplx<T> *)P.bandB.p,
            *Fin = (Cplx<T> *)bout.p, *Fo = (Cplx<T> *)P.bandF.p;
    if (fused) {
        LT_TRY(ensure_twiddles(P));
        LT_TRY((rows_forward<T, TI>(P, nbv, imgs, active, nactive, Bin, rowmask)));
    } else {
        hipLaunchKernelGGL((k_band_pad<T, TI>), dim3(blocks_for(rsz, 256), nactive), dim3(256), 0, stream(), P.nx,
                           P.ny, P.pady, P.nyp, imgs, P.apod.p, active, R);
        LT_CHECK_LAUNCH();
        void *in[1] = {R};
        void *o[1] = {Cb};
        LT_TRY(fft_check(rocfft_execute(bf->plan[0], in, o, bf->info[0]), "execute band rows r2c"));
        hipLaunchKernelGGL((k_band_gather<T>), dim3((P.nu + 31) / 32, (nbv + 31) / 32, nactive), dim3(256), 0,
                           stream(), P.nx, P.padx, P.nu, P.half_nv, nbv, Cb, Bin);
        LT_CHECK_LAUNCH();
    }
    // own column transforms for columns of 1024 (the row transforms' twiddle
    // table), rocFFT otherwise
    const bool own_cols = fused && P.own_cols && P.nu == kR4N && P.nyp == kR4N;
    if (own_cols) {
        hipLaunchKernelGGL((k_cols_fwd_r<T>), dim3((unsigned)((long)nactive * nbv)), dim3(256), 0, stream(), nbv, P.nu,
                           P.padx, P.nx, bsz, Bin, Bf, twiddles_of<T>(P));
        LT_CHECK_LAUNCH();
    } else {
        void *in[1] = {Bin};
        void *o[1] = {Bf};
        LT_TRY(fft_check(rocfft_execute(bf->plan[1], in, o, bf->info[1]), "execute band columns forward"));
    }
    // the stencil writes its (folded) rows' band cells of Fin directly
    if (prec == LT_F64)
        LT_TRY((stencil_apply_t<double, double>(St, P, St.coef.p, Bf, bsz, Fin, nullptr, nactive, true)));
    else
        LT_TRY((stencil_apply_t<float, float>(St, P, St.coef32.p, Bf, bsz, Fin, nullptr, nactive, true)));
    // row v = 0 exactly Hermitian along u (its two halves are computed as
    // separate rows) for rocFFT's c2r; the fused inverse row transform takes
    // the real part of column 0 after the column transform, which equals the
    // transform of the Hermitian part (Re IFFT(X) = IFFT((X + conj X(-u)) / 2)),
    // so it needs no such step
    if (!fused) {
        hipLaunchKernelGGL((k_herm_band0<T>), dim3(blocks_for(P.nu / 2 + 1, 256), nactive), dim3(256), 0, stream(),
                           P.nu, bsz, Fin);
        LT_CHECK_LAUNCH();
    }
    if (own_cols) {
        hipLaunchKernelGGL((k_cols_inv_r<T>), dim3((unsigned)((long)nactive * nbv)), dim3(256), 0, stream(), nbv, P.nu,
                           P.padx, P.nx, bsz, Fin, Fo, twiddles_of<T>(P));
        LT_CHECK_LAUNCH();
    } else {
        void *in[1] = {Fin};
        void *o[1] = {Fo};
        LT_TRY(fft_check(rocfft_execute(bf->plan[2], in, o, bf->info[2]), "execute band columns inverse"));
    }
    if (fused) return rows_inverse<T, TI>(P, nbv, Fo, active, nactive, alpha, acc, out, rowmask);
    hipLaunchKernelGGL((k_band_scatter<T>), dim3((P.nx + 31) / 32, (P.half_nv + 31) / 32, nactive), dim3(256), 0,
                       stream(), P.nx, P.padx, P.nu, P.half_nv, nbv, Fo, Cb);
    LT_CHECK_LAUNCH();
    {
        void *in[1] = {Cb};
        void *o[1] = {R};
        LT_TRY(fft_check(rocfft_execute(bf->plan[3], in, o, bf->info[3]), "execute band rows c2r"));
    }
    hipLaunchKernelGGL((k_band_crop<T, TI>), dim3(blocks_for((long)P.nx * P.ny, 256), nactive), dim3(256), 0, stream(),
                       P.nx, P.ny, P.pady, P.nyp, R, P.apod.p, alpha, acc ? 1 : 0, active, out);
    LT_CHECK_LAUNCH();
    return LT_OK;
}

// Autotuner of the stencil's form: samples per workgroup of the vector form
// (1, 2, 4, 8) and sample groups per workgroup of the matrix-core form (1, 2,
// 4, 8); times the whole band normal operator at the full batch for each and
// keeps the fastest.  imgs / out: nb images of scratch.
lt_status stencil_autotune(NufftPlan &P, NormalStencil &St, const double *imgs, double *out, const int *all, int nb) {
    if (!St.use_band || nb <= 0) return LT_OK;
    hipEvent_t e0, e1;
    LT_HIP(hipEventCreate(&e0));
    LT_HIP(hipEventCreate(&e1));
    // three rounds over the candidates, each candidate's best round kept (so
    // that clocks settling during the first timings favour none of them);
    // negative: the matrix-core form with that many groups of 8 samples per
    // workgroup (radius 6 only)
    const int cands[8] = {1, 2, 4, 8, -1, -2, -4, -8};
    const int ncand = St.R == kMfR ? 8 : 4;
    float tmin[8];
    for (int c = 0; c < 8; ++c) tmin[c] = 1e30f;
    for (int round = 0; round < 3; ++round)
        for (int c = 0; c < ncand; ++c) {
            St.nb_sel = cands[c];
            LT_TRY((normal_band_impl<double, double>(P, St, imgs, out, all, nb, 1.0, false)));  // warm-up
            LT_HIP(hipEventRecord(e0, stream()));
            for (int r = 0; r < 3; ++r)
                LT_TRY((normal_band_impl<double, double>(P, St, imgs, out, all, nb, 1.0, false)));
            LT_HIP(hipEventRecord(e1, stream()));
            LT_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            LT_HIP(hipEventElapsedTime(&ms, e0, e1));
            tmin[c] = std::min(tmin[c], ms);
        }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    int best = 0;
    for (int c = 1; c < ncand; ++c)
        if (tmin[c] < tmin[best]) best = c;
    St.nb_sel = cands[best];
    return LT_OK;
}

// The FP32 stencil's form (precision form): the vector form with 4 samples
// per workgroup against the matrix-core form with 1, 2 or 4 groups of 8
// samples per workgroup (radius 6 only), timed like stencil_autotune.
lt_status stencil_autotune32(NufftPlan &P, NormalStencil &St, const float *imgs, float *out, const int *all, int nb) {
    if (!St.use_band || nb <= 0) return LT_OK;
    if (St.R != kMfR) {
        St.nb_sel32 = 4;
        return LT_OK;
    }
    if (!St.coef32.p) LT_TRY(St.apply32(P, nullptr, 0, nullptr, nullptr, 0));  // makes coef32
    const int cands[4] = {4, -1, -2, -4};
    float tmin[4];
    for (int c = 0; c < 4; ++c) tmin[c] = 1e30f;
    hipEvent_t e0, e1;
    LT_HIP(hipEventCreate(&e0));
    LT_HIP(hipEventCreate(&e1));
    auto rounds = [&]() -> lt_status {
        for (int round = 0; round < 3; ++round)
            for (int c = 0; c < 4; ++c) {
                St.nb_sel32 = cands[c];
                LT_TRY((normal_band_impl<float, float>(P, St, imgs, out, all, nb, 1.f, false)));  // warm-up
                LT_HIP(hipEventRecord(e0, stream()));
                for (int r = 0; r < 3; ++r)
                    LT_TRY((normal_band_impl<float, float>(P, St, imgs, out, all, nb, 1.f, false)));
                LT_HIP(hipEventRecord(e1, stream()));
                LT_HIP(hipEventSynchronize(e1));
                float ms = 0.f;
                LT_HIP(hipEventElapsedTime(&ms, e0, e1));
                tmin[c] = std::min(tmin[c], ms);
            }
        return LT_OK;
    };
    const int forced = P.force_nb;  // a forced shape would stand in for every candidate
    P.force_nb = 0;
    const lt_status st = rounds();
    P.force_nb = forced;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (st != LT_OK) {
        St.nb_sel32 = 0;
        return st;
    }
    int best = 0;
    for (int c = 1; c < 4; ++c)
        if (tmin[c] < tmin[best]) best = c;
    St.nb_sel32 = cands[best];
    return LT_OK;
}

lt_status NufftPlan::normal_batched32(const float *imgs, float *out, const int *active, int nactive, float alpha,
                                      bool acc, NormalStencil &st, const uint8_t *rowmask) {
    if (st.use_band) {
        if (!st.coef32.p) LT_TRY(st.apply32(*this, nullptr, 0, nullptr, nullptr, 0));  // makes coef32
        return normal_band_impl<float, float>(*this, st, imgs, out, active, nactive, alpha, acc, rowmask);
    }
    if (nactive <= 0) return LT_OK;
    if (nactive > batch) {
        set_error("normal operator: %d samples exceed the batch %d", nactive, batch);
        return LT_ERR_ARG;
    }
    const int nrun = std::min(batch, (nactive + kBatchStep - 1) / kBatchStep * kBatchStep);
    BatchFFT *bf = nullptr;
    LT_TRY(batch_fft(nrun, bf, LT_F32));
    const long nreal = (long)nxp * nyp, nspec = (long)nu * half_nv, npix = (long)nx * ny;
    hipLaunchKernelGGL((k_pad_b<float, float>), dim3(blocks_for(nreal, 256), nactive), dim3(256), 0, stream(), nx, ny,
                       padx, pady, nxp, nyp, imgs, apod.p, (float *)realB.p, active);
    LT_CHECK_LAUNCH();
    void *in[1] = {realB.p};
    void *o1[1] = {specB.p};
    LT_TRY(fft_check(rocfft_execute(bf->fwd, in, o1, bf->fi), "execute batched r2c (FP32)"));
    LT_TRY(st.apply32(*this, specB.p, nspec, specB2.p, nullptr, nactive));
    const int cols[2] = {0, nv / 2};
    const int ncol = (nv % 2 == 0) ? 2 : 1;
    for (int k = 0; k < ncol; ++k) {
        hipLaunchKernelGGL((k_hermitian_cols_b<float>), dim3(blocks_for(nu / 2 + 1, 256), nactive), dim3(256), 0,
                           stream(), nu, half_nv, cols[k], (Cplx<float> *)specB2.p, nspec, nullptr);
        LT_CHECK_LAUNCH();
    }
    void *i2[1] = {specB2.p};
    void *o2[1] = {realB.p};
    LT_TRY(fft_check(rocfft_execute(bf->inv, i2, o2, bf->ii), "execute batched c2r (FP32)"));
    hipLaunchKernelGGL((k_crop_b<float, float>), dim3(blocks_for(npix, 256), nactive), dim3(256), 0, stream(), nx, ny,
                       padx, pady, nxp, nyp, (const float *)realB.p, apod.p, alpha, acc ? 1 : 0, out, active);
    LT_CHECK_LAUNCH();
    return LT_OK;
}

lt_status NufftPlan::normal_batched(const double *imgs, double *out, const double *w, const int *active, int nactive,
                                    double alpha, bool acc, NormalStencil *st, const uint8_t *rowmask) {
    if (st && st->use_band && prec == LT_F64)
        return normal_band_impl<double, double>(*this, *st, imgs, out, active, nactive, alpha, acc, rowmask);
    return prec == LT_F64 ? normal_impl<double>(*this, imgs, out, w, active, nactive, alpha, acc, nullptr, st)
                          : normal_impl<float>(*this, imgs, out, w, active, nactive, alpha, acc, nullptr, st);
}

lt_status NufftPlan::convolve_batched(const double *imgs, double *out, const double *mul, const int *active,
                                      int nactive, double alpha, bool acc) {
    return prec == LT_F64 ? normal_impl<double>(*this, imgs, out, nullptr, active, nactive, alpha, acc, mul, nullptr)
                          : normal_impl<float>(*this, imgs, out, nullptr, active, nactive, alpha, acc, mul, nullptr);
}

lt_status NormalStencil::build(NufftPlan &P, const double *w_sorted) {
    if (P.prec != LT_F64 || P.W > 4 || P.nvis == 0) {
        set_error("normal stencil: needs an FP64 plan with kernel_support <= 4");
        return LT_ERR_UNSUPPORTED;
    }
    Geo G = make_geo(P);
    W = P.W;
    R = 2 * W;
    NO = (2 * R + 1) * (2 * R + 1);
    const long n = P.nvis, ng = (long)P.nu * P.nv;
    const int NT = 2 * W + 1, N2 = NT * NT;
    DevBuf<int64_t> flag, idx;
    LT_TRY(flag.alloc(n + 1));
    LT_TRY(idx.alloc(n + 1));
    hipLaunchKernelGGL(k_cell_flags, dim3(blocks_for(n, 256)), dim3(256), 0, stream(), n, P.key_sorted.p, flag.p);
    LT_CHECK_LAUNCH();
    LT_TRY(exclusive_scan_i64(flag.p, idx.p, n));
    int64_t nc = 0;
    LT_TRY(idx.download(&nc, 1, n));
    LT_TRY(sync());
    ncell = nc;
    DevBuf<int32_t> cell_beg, cellmap;
    LT_TRY(cell_beg.alloc(ncell + 1));
    LT_TRY(cellmap.alloc(ng));
    LT_HIP(hipMemsetAsync(cellmap.p, 0xff, ng * sizeof(int32_t), stream()));
    hipLaunchKernelGGL(k_cell_list, dim3(blocks_for(n, 256)), dim3(256), 0, stream(), G, n, P.key_sorted.p, idx.p,
                       cell_beg.p, cellmap.p);
    LT_CHECK_LAUNCH();
    const int32_t nn = (int32_t)n;
    LT_HIP(hipMemcpyAsync(cell_beg.p + ncell, &nn, sizeof(int32_t), hipMemcpyHostToDevice, stream()));
    LT_TRY(sync());
    // per-cell outer-product sums
    DevBuf<double> T;
    LT_TRY(T.alloc(ncell * (long)N2 * N2));
    const int thr = N2 <= 64 ? 64 : 128;
#define LT_CELLT_CASE(WW)                                                                                    \
    case WW:                                                                                                 \
        hipLaunchKernelGGL((k_cell_T<WW>), dim3((unsigned)ncell), dim3(thr), 0, stream(), G, P.poly, ncell, \
                           cell_beg.p, (const int32_t *)nullptr, P.frac64.p, w_sorted, T.p);                 \
        break;
    switch (W) {
        LT_CELLT_CASE(1)
        LT_CELLT_CASE(2)
        LT_CELLT_CASE(3)
        LT_CELLT_CASE(4)
    }
#undef LT_CELLT_CASE
    LT_CHECK_LAUNCH();
    // stencil rows
    DevBuf<int64_t> mark, off;
    LT_TRY(mark.alloc(ng + 1));
    LT_TRY(mark.zero());
    LT_TRY(off.alloc(ng + 1));
    hipLaunchKernelGGL(k_mark_rows, dim3((unsigned)ncell), dim3(64), 0, stream(), G, ncell, cell_beg.p,
                       P.key_sorted.p, mark.p);
    LT_CHECK_LAUNCH();
    LT_TRY(exclusive_scan_i64(mark.p, off.p, ng));
    int64_t nr = 0;
    LT_TRY(off.download(&nr, 1, ng));
    LT_TRY(sync());
    nrow = nr;
    LT_TRY(rowmap.alloc(ng));
    LT_TRY(row_cell.alloc(nrow));
    hipLaunchKernelGGL(k_rows, dim3(blocks_for(ng, 256)), dim3(256), 0, stream(), ng, mark.p, off.p, rowmap.p,
                       row_cell.p);
    LT_CHECK_LAUNCH();
    LT_TRY(coef.alloc((long)NO * nrow));
    const double c0 = 1.0 / (4.0 * W * W);
    const long ncoef = (long)NO * nrow;
#define LT_COEF_CASE(WW)                                                                                          \
    case WW:                                                                                                      \
        hipLaunchKernelGGL((k_coef<WW>), dim3(blocks_for(ncoef, 256)), dim3(256), 0, stream(), G, nrow, row_cell.p, \
                           cellmap.p, T.p, 0.5 * c0 * c0, coef.p);                                               \
        break;
    switch (W) {
        LT_COEF_CASE(1)
        LT_COEF_CASE(2)
        LT_COEF_CASE(3)
        LT_COEF_CASE(4)
    }
#undef LT_COEF_CASE
    LT_CHECK_LAUNCH();
    // Hermitian fold of the rows (k_fold_mark): half the rows, and the
    // stencil's outputs are the final band cells
    {
        DevBuf<int32_t> rowmap_raw = std::move(rowmap);
        DevBuf<double> coef_raw = std::move(coef);
        nrow_raw = nrow;
        hipLaunchKernelGGL(k_fold_mark, dim3(blocks_for(ng, 256)), dim3(256), 0, stream(), G, rowmap_raw.p, mark.p);
        LT_CHECK_LAUNCH();
        LT_TRY(exclusive_scan_i64(mark.p, off.p, ng));
        int64_t nf = 0;
        LT_TRY(off.download(&nf, 1, ng));
        LT_TRY(sync());
        nrow = nf;
        DevBuf<int32_t> r1, r2;
        LT_TRY(rowmap.alloc(ng));
        LT_TRY(row_cell.alloc(std::max<long>(nrow, 1)));
        LT_TRY(row_dst.alloc(std::max<long>(nrow, 1)));
        LT_TRY(r1.alloc(std::max<long>(nrow, 1)));
        LT_TRY(r2.alloc(std::max<long>(nrow, 1)));
        hipLaunchKernelGGL(k_fold_rows, dim3(blocks_for(ng, 256)), dim3(256), 0, stream(), G, off.p, rowmap_raw.p,
                           rowmap.p, row_cell.p, r1.p, r2.p, row_dst.p);
        LT_CHECK_LAUNCH();
        LT_TRY(coef.alloc(std::max<long>((long)NO * nrow, 1)));
        if (nrow > 0) {
            hipLaunchKernelGGL(k_fold_coef, dim3(blocks_for((long)NO * nrow, 256)), dim3(256), 0, stream(), nrow, NO,
                               r1.p, r2.p, nrow_raw, coef_raw.p, coef.p);
            LT_CHECK_LAUNCH();
        }
        LT_TRY(sync());
    }
    // stencil tiles
    const int ntu_st = (P.nu + ST - 1) / ST, ntv_st = (P.nv + ST - 1) / ST;
    const long nts = (long)ntu_st * ntv_st;
    DevBuf<int64_t> tmark, toff;
    LT_TRY(tmark.alloc(nts + 1));
    LT_TRY(tmark.zero());
    LT_TRY(toff.alloc(nts + 1));
    hipLaunchKernelGGL(k_stencil_tiles, dim3(blocks_for(nrow, 256)), dim3(256), 0, stream(), nrow, row_cell.p, P.nv,
                       ST, ntv_st, tmark.p);
    LT_CHECK_LAUNCH();
    LT_TRY(exclusive_scan_i64(tmark.p, toff.p, nts));
    int64_t nt = 0;
    LT_TRY(toff.download(&nt, 1, nts));
    LT_TRY(sync());
    ntile = nt;
    LT_TRY(tiles.alloc(std::max<long>(ntile, 1)));
    hipLaunchKernelGGL(k_compact, dim3(blocks_for(nts, 256)), dim3(256), 0, stream(), nts, toff.p, tiles.p);
    LT_CHECK_LAUNCH();
    // spectral band the stencil reads: |v| up to the rows' largest + R
    {
        DevBuf<int> mb;
        LT_TRY(mb.alloc(1));
        LT_TRY(mb.zero());
        hipLaunchKernelGGL(k_max_abs_b, dim3(blocks_for(nrow, 256)), dim3(256), 0, stream(), nrow, row_cell.p, P.nv,
                           P.cv, mb.p);
        LT_CHECK_LAUNCH();
        int hb = 0;
        LT_TRY(mb.download(&hb, 1));
        LT_TRY(sync());
        nbv = std::min(P.half_nv, hb + R + 1);
    }
    return sync();
}

lt_status NormalStencil::mfma_build(NufftPlan &P) {
    if (mf_nblock >= 0) return LT_OK;
    if (R != kMfR || !use_band) {
        set_error("normal stencil: the matrix-core form needs the band form with radius %d", kMfR);
        return LT_ERR_UNSUPPORTED;
    }
    Geo G = make_geo(P);
    const int ntu_b = (P.nu + kMfBU - 1) / kMfBU;
    mf_ntv = (P.nv + kMfBV - 1) / kMfBV;
    const long nbs = (long)ntu_b * mf_ntv;
    DevBuf<int64_t> bmark, boff;
    LT_TRY(bmark.alloc(nbs + 1));
    LT_TRY(bmark.zero());
    LT_TRY(boff.alloc(nbs + 1));
    hipLaunchKernelGGL(k_mf_blocks, dim3(blocks_for(nrow, 256)), dim3(256), 0, stream(), nrow, row_cell.p, P.nv,
                       mf_ntv, bmark.p);
    LT_CHECK_LAUNCH();
    LT_TRY(exclusive_scan_i64(bmark.p, boff.p, nbs));
    int64_t nb = 0;
    LT_TRY(boff.download(&nb, 1, nbs));
    LT_TRY(sync());
    LT_TRY(mf_blocks.alloc(std::max<long>(nb, 1)));
    hipLaunchKernelGGL(k_compact, dim3(blocks_for(nbs, 256)), dim3(256), 0, stream(), nbs, boff.p, mf_blocks.p);
    LT_CHECK_LAUNCH();
    LT_TRY(mf_A.alloc(std::max<long>(nb, 1) * kMfNT * 64 * 64));
    LT_TRY(mf_rows.alloc(std::max<long>(nb, 1) * kMfNT * 16));
    const long npk = nb * kMfNT * 64 * 64;
    if (npk > 0) {
        hipLaunchKernelGGL(k_mf_pack<double>, dim3(blocks_for(npk, 256)), dim3(256), 0, stream(), G, (long)nb, mf_blocks.p,
                           mf_ntv, rowmap.p, nrow, coef.p, row_dst.p, mf_A.p, mf_rows.p);
        LT_CHECK_LAUNCH();
    }
    mf_nblock = nb;
    return sync();
}

lt_status NormalStencil::mfma_build32(NufftPlan &P) {
    if (mf_A32.p) return LT_OK;
    if (!coef32.p) {
        set_error("normal stencil: FP32 coefficients missing");
        return LT_ERR_ARG;
    }
    LT_TRY(mfma_build(P));
    Geo G = make_geo(P);
    const long nb = mf_nblock;
    LT_TRY(mf_A32.alloc(std::max<long>(nb, 1) * kMfNT * 64 * 64));
    const long npk = nb * kMfNT * 64 * 64;
    if (npk > 0) {
        hipLaunchKernelGGL(k_mf_pack<float>, dim3(blocks_for(npk, 256)), dim3(256), 0, stream(), G, nb, mf_blocks.p,
                           mf_ntv, rowmap.p, nrow, coef32.p, row_dst.p, mf_A32.p, mf_rows.p);
        LT_CHECK_LAUNCH();
    }
    return sync();
}

// band: spec_in is a [v][u] band of nbv rows and the stencil writes its
// rows' cells of the band spec_out (same stride); otherwise the rows go to
// the compact St.rbuf and k_stencil_fold makes the half-plane spectrum
template <class T, class TC>
static lt_status stencil_apply_t(NormalStencil &St, NufftPlan &P, const TC *coef, const void *spec_in, long sstride,
                                 void *spec_out, const int *active, int nactive, bool band = false) {
    Geo G = make_geo(P);
    if (!band) {
        const size_t need = (size_t)std::max(P.batch, 1) * St.nrow * sizeof(Cplx<double>);
        if (St.rbuf.n < need) LT_TRY(St.rbuf.alloc(need));
    }
    Cplx<T> *ob = band ? (Cplx<T> *)spec_out : (Cplx<T> *)St.rbuf.p;
    const int32_t *odst = band ? St.row_dst.p : nullptr;
    const long ostride = band ? sstride : St.nrow;
    const int WS = St.ST + 2 * St.R;
    const int ntv_st = (P.nv + St.ST - 1) / St.ST;
    const bool f32 = std::is_same<T, float>::value && std::is_same<TC, float>::value;
    const int nb_sel = P.force_nb != 0 ? P.force_nb : (f32 ? (St.nb_sel32 != 0 ? St.nb_sel32 : 4) : St.nb_sel);
    if (nb_sel < 0 && f32 && band && St.R == kMfR && nactive > 0) {
        // FP32 matrix-core form: -nb_sel groups of 8 samples per workgroup
        LT_TRY(St.mfma_build32(P));
        const int gpw = -nb_sel;
        const int ngroups = (nactive + 7) / 8;
        const int ngc = (ngroups + gpw - 1) / gpw;
        const size_t sh = (size_t)(kMfWU / 2) * kMfR2 * sizeof(float);
        const dim3 grid((unsigned)(St.mf_nblock * ngc));
#define LT_STENCIL_MF32(GG)                                                                                       \
    hipLaunchKernelGGL((k_stencil_mfma32<GG>), grid, dim3(64 * kMfNT), sh, stream(), G, St.mf_blocks.p,          \
                       St.mf_ntv, ngc, St.mf_A32.p, St.mf_rows.p, (const Cplx<float> *)spec_in, sstride,         \
                       (Cplx<float> *)ob, ostride, active, nactive, St.nbv);
        if (St.mf_nblock > 0) {
            if (gpw == 1)
                LT_STENCIL_MF32(1)
            else if (gpw == 2)
                LT_STENCIL_MF32(2)
            else if (gpw == 4)
                LT_STENCIL_MF32(4)
            else
                LT_STENCIL_MF32(8)
            LT_CHECK_LAUNCH();
        }
#undef LT_STENCIL_MF32
        return LT_OK;
    }
    if (nb_sel < 0 && std::is_same<T, double>::value && std::is_same<TC, double>::value && band &&
        St.R == kMfR && nactive > 0) {
        // matrix-core form: -nb_sel groups of 8 samples per workgroup
        LT_TRY(St.mfma_build(P));
        const int gpw = -nb_sel;
        const int ngroups = (nactive + 7) / 8;
        const int ngc = (ngroups + gpw - 1) / gpw;
        const size_t sh = (size_t)(kMfWU / 2) * kMfR2 * sizeof(double);
        const dim3 grid((unsigned)(St.mf_nblock * ngc));
#define LT_STENCIL_MF(GG)                                                                                          \
    hipLaunchKernelGGL((k_stencil_mfma<GG>), grid, dim3(64 * kMfNT), sh, stream(), G, St.mf_blocks.p, St.mf_ntv,   \
                       ngc,                                                                                         \
                       St.mf_A.p, St.mf_rows.p, (const Cplx<double> *)spec_in, sstride, (Cplx<double> *)ob,       \
                       ostride, active, nactive, St.nbv);
        if (St.mf_nblock > 0) {
            if (gpw == 1)
                LT_STENCIL_MF(1)
            else if (gpw == 2)
                LT_STENCIL_MF(2)
            else if (gpw == 4)
                LT_STENCIL_MF(4)
            else
                LT_STENCIL_MF(8)
            LT_CHECK_LAUNCH();
        }
#undef LT_STENCIL_MF
        return LT_OK;
    }
    if (St.ntile > 0 && nactive > 0) {
        // NB samples share one staged coefficient read (autotuned: 1, 2, 4 or 8)
#define LT_STENCIL_NB(NBV)                                                                                        \
    {                                                                                                             \
        const size_t sh = (size_t)NBV * WS * (WS | 1) * sizeof(Cplx<T>);                                          \
        const int ngroup = (nactive + NBV - 1) / NBV;                                                             \
        const dim3 grid((unsigned)(St.ntile * ngroup));                                                           \
        if (band && St.R == 6)                                                                                    \
            hipLaunchKernelGGL((k_stencil_apply<T, NBV, TC, true, 6>), grid, dim3(St.ST * St.ST), sh, stream(), G, \
                               St.R, St.ST, ntv_st, St.tiles.p, ngroup, St.rowmap.p, St.nrow, coef,               \
                               (const Cplx<T> *)spec_in, sstride, ob, odst, ostride, active, nactive, St.nbv); \
        else if (band)                                                                                            \
            hipLaunchKernelGGL((k_stencil_apply<T, NBV, TC, true>), grid, dim3(St.ST * St.ST), sh, stream(), G,   \
                               St.R, St.ST, ntv_st, St.tiles.p, ngroup, St.rowmap.p, St.nrow, coef,               \
                               (const Cplx<T> *)spec_in, sstride, ob, odst, ostride, active, nactive, St.nbv); \
        else                                                                                                      \
            hipLaunchKernelGGL((k_stencil_apply<T, NBV, TC, false>), grid, dim3(St.ST * St.ST), sh, stream(), G,  \
                               St.R, St.ST, ntv_st, St.tiles.p, ngroup, St.rowmap.p, St.nrow, coef,               \
                               (const Cplx<T> *)spec_in, sstride, ob, odst, ostride, active, nactive, St.nbv); \
    }
        if (nb_sel == 1)
            LT_STENCIL_NB(1)
        else if (nb_sel == 2)
            LT_STENCIL_NB(2)
        else if (nb_sel == 8)
            LT_STENCIL_NB(8)
        else
            LT_STENCIL_NB(4)
#undef LT_STENCIL_NB
        LT_CHECK_LAUNCH();
    }
    const long nspec = (long)P.nu * P.half_nv;
    if (nactive > 0 && !band) {
        hipLaunchKernelGGL((k_stencil_fold<T>), dim3(blocks_for(nspec, 256), nactive), dim3(256), 0, stream(), G,
                           St.rowmap.p, St.nrow, (const Cplx<T> *)St.rbuf.p, (Cplx<T> *)spec_out, sstride, active);
        LT_CHECK_LAUNCH();
    }
    return LT_OK;
}

__global__ void k_d2f(long n, const double *__restrict__ a, float *__restrict__ b) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

lt_status NormalStencil::apply(NufftPlan &P, const void *spec_in, long sstride, void *spec_out, const int *active,
                               int nactive) {
    if (P.prec != LT_F64) {
        set_error("normal stencil: FP64 plans only");
        return LT_ERR_UNSUPPORTED;
    }
    return stencil_apply_t<double, double>(*this, P, coef.p, spec_in, sstride, spec_out, active, nactive);
}

__global__ void k_absmax_part(long n, const double *__restrict__ a, double *__restrict__ out) {
    __shared__ double sh[256];
    double m = 0.0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) m = fmax(m, fabs(a[i]));
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

__global__ void k_d2f_scaled(long n, const double *__restrict__ a, double inv, float *__restrict__ b) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)(a[i] * inv);
}

lt_status NormalStencil::apply32(NufftPlan &P, const void *spec_in, long sstride, void *spec_out, const int *active,
                                 int nactive) {
    if (!coef32.p) {
        // FP32 coefficients normalised by their largest magnitude (the FP64
        // ones carry the visibility weights, ~1e15 and more); the caller
        // multiplies the result by coef32_scale
        const long n = (long)NO * nrow;
        DevBuf<double> part;
        LT_TRY(part.alloc(256));
        hipLaunchKernelGGL(k_absmax_part, dim3(256), dim3(256), 0, stream(), n, coef.p, part.p);
        std::vector<double> h(256);
        LT_TRY(part.download(h.data(), 256));
        LT_TRY(sync());
        double m = 0.0;
        for (double v : h) m = std::max(m, v);
        coef32_scale = m > 0.0 ? m : 1.0;
        LT_TRY(coef32.alloc(std::max<long>(n, 1)));
        hipLaunchKernelGGL(k_d2f_scaled, dim3(blocks_for(n, 256)), dim3(256), 0, stream(), n, coef.p,
                           1.0 / coef32_scale, coef32.p);
        LT_CHECK_LAUNCH();
    }
    return stencil_apply_t<float, float>(*this, P, coef32.p, spec_in, sstride, spec_out, active, nactive);
}

__global__ void k_gather_weights(long n, const int32_t *__restrict__ perm, const double *__restrict__ w_in,
                                 double *__restrict__ w_out) {
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) w_out[q] = w_in[perm[q]];
}

// Times the normal stencil's kernel as the solver launches it (band form,
// the plan's resident stencil, nb samples, the tuned samples per workgroup)
// with HIP events on the library stream: mean duration per launch over reps
// launches.  Algorithmic work per launch: every stencil row's (the folded
// rows: half-plane cells) NO FP64 coefficients once (8 B each), and per
// sample its output (16 B per row) and
// the distinct complex band values its windows read (16 B each: raw cells
// within R of a row, a cell and its Hermitian mirror counted once); 4 flops
// (a complex value times a real coefficient, added) per row, offset and
// sample.
lt_status stencil_time(NufftPlan &P, int nb, int reps, double *ms, double *bytes, double *flops, long *rows,
                       long *win_cells) {
    NormalStencil *St = P.wstencil.get();
    if (!St || !St->use_band || P.prec != LT_F64) {
        set_error("stencil timing: the plan has no resident FP64 band stencil (lt_nufft_set_weights)");
        return LT_ERR_ARG;
    }
    if (nb <= 0 || reps <= 0) {
        set_error("stencil timing: bad batch or repetition count");
        return LT_ERR_ARG;
    }
    LT_TRY(P.ensure_batch(nb));
    const long npix = (long)P.nx * P.ny;
    DevBuf<double> imgs, out;
    DevBuf<int> all;
    LT_TRY(imgs.alloc(npix * nb));
    LT_TRY(out.alloc(npix * nb));
    LT_TRY(imgs.zero());
    std::vector<int> ids(nb);
    for (int k = 0; k < nb; ++k) ids[k] = k;
    LT_TRY(all.upload(ids.data(), nb));
    // first application allocates the band buffers; the tuner picks the shape
    LT_TRY((normal_band_impl<double, double>(P, *St, imgs.p, out.p, all.p, nb, 1.0, false)));
    if (P.wstencil_tuned_batch != nb) {
        LT_TRY(stencil_autotune(P, *St, imgs.p, out.p, all.p, nb));
        P.wstencil_tuned_batch = nb;
    }
    const long bsz = (long)St->nbv * P.nu;
    hipEvent_t e0, e1;
    LT_HIP(hipEventCreate(&e0));
    LT_HIP(hipEventCreate(&e1));
    LT_TRY((stencil_apply_t<double, double>(*St, P, St->coef.p, P.bandB.p, bsz, St->band_out[0].p, nullptr, nb, true)));
    LT_HIP(hipEventRecord(e0, stream()));
    for (int r = 0; r < reps; ++r)
        LT_TRY((stencil_apply_t<double, double>(*St, P, St->coef.p, P.bandB.p, bsz, St->band_out[0].p, nullptr, nb, true)));
    LT_HIP(hipEventRecord(e1, stream()));
    LT_HIP(hipEventSynchronize(e1));
    float t = 0.f;
    LT_HIP(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms) *ms = (double)t / reps;
    // distinct band values the windows read
    const Geo G = make_geo(P);
    const long ng = (long)P.nu * P.nv;
    std::vector<int32_t> rm(ng);
    LT_TRY(St->rowmap.download(rm.data(), ng));
    LT_TRY(sync());
    std::vector<uint8_t> m(ng, 0), mu(ng, 0);
    const int R = St->R;
    for (long g = 0; g < ng; ++g)  // dilate along v
        if (rm[g] >= 0) {
            const long a = g / P.nv, b = g % P.nv;
            for (long q = std::max(0L, b - R); q <= std::min((long)P.nv - 1, b + R); ++q) mu[a * P.nv + q] = 1;
        }
    for (long g = 0; g < ng; ++g)  // then along u
        if (mu[g]) {
            const long a = g / P.nv, b = g % P.nv;
            for (long q = std::max(0L, a - R); q <= std::min((long)P.nu - 1, a + R); ++q) m[q * P.nv + b] = 1;
        }
    long nwin = 0;
    for (long g = 0; g < ng; ++g) {
        if (!m[g]) continue;
        const long a = g / P.nv - G.cu, b = g % P.nv - G.cv;  // centred raw coordinates
        const long ma = -a + G.cu, mb = -b + G.cv;
        const bool mirror_in = ma >= 0 && ma < P.nu && mb >= 0 && mb < P.nv && m[ma * P.nv + mb];
        // a pair counts once: keep the member with the larger index
        if (!mirror_in || g >= ma * P.nv + mb) ++nwin;
    }
    if (rows) *rows = St->nrow;
    if (win_cells) *win_cells = nwin;
    if (bytes) *bytes = 8.0 * St->NO * St->nrow + 16.0 * nb * (double)(St->nrow + nwin);
    if (flops) *flops = 4.0 * St->NO * (double)St->nrow * nb;
    return LT_OK;
}

lt_status NufftPlan::set_weights(const double *weights) {
    if (!weights && nvis > 0) {
        set_error("set_weights: no weights");
        return LT_ERR_ARG;
    }
    LT_TRY(wplan.alloc(std::max<long>(nvis, 1)));
    if (nvis > 0) {
        DevBuf<double> raw;
        LT_TRY(raw.upload(weights, nvis));
        hipLaunchKernelGGL(k_gather_weights, dim3(blocks_for(nvis, 256)), dim3(256), 0, stream(), nvis, perm.p, raw.p,
                           wplan.p);
        LT_CHECK_LAUNCH();
        LT_TRY(sync());
    }
    wstencil.reset();
    wstencil_tuned_batch = -1;
    if (has_image && prec == LT_F64 && W <= 4 && nvis > 0) {
        auto st = std::make_unique<NormalStencil>();
        LT_TRY(st->build(*this, wplan.p));
        wstencil = std::move(st);
    }
    return sync();
}

lt_status NufftPlan::to_sorted(const double *src, void *dst) {
    if (nvis == 0) return LT_OK;
    if (prec == LT_F64)
        hipLaunchKernelGGL(k_to_sorted<double>, dim3(blocks_for(nvis, 256)), dim3(256), 0, stream(), nvis, perm.p, src,
                           (Cplx<double> *)dst);
    else
        hipLaunchKernelGGL(k_to_sorted<float>, dim3(blocks_for(nvis, 256)), dim3(256), 0, stream(), nvis, perm.p, src,
                           (Cplx<float> *)dst);
    LT_CHECK_LAUNCH();
    return LT_OK;
}

lt_status NufftPlan::from_sorted(const void *src, double *dst, bool acc) {
    if (nvis == 0) return LT_OK;
    if (prec == LT_F64)
        hipLaunchKernelGGL(k_from_sorted<double>, dim3(blocks_for(nvis, 256)), dim3(256), 0, stream(), nvis, perm.p,
                           (const Cplx<double> *)src, dst, acc ? 1 : 0);
    else
        hipLaunchKernelGGL(k_from_sorted<float>, dim3(blocks_for(nvis, 256)), dim3(256), 0, stream(), nvis, perm.p,
                           (const Cplx<float> *)src, dst, acc ? 1 : 0);
    LT_CHECK_LAUNCH();
    return LT_OK;
}

// grid (complex128 host layout) <-> plan precision
lt_status grid_c128_to_plan(NufftPlan &P, const double *src, void *dst) {
    const long n = 2L * P.nu * P.half_nv;
    if (P.prec == LT_F64) {
        LT_HIP(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, stream()));
    } else {
        hipLaunchKernelGGL(k_c128_to_T<float>, dim3(blocks_for(n, 256)), dim3(256), 0, stream(), n, src, (float *)dst);
        LT_CHECK_LAUNCH();
    }
    return LT_OK;
}

lt_status grid_plan_add_c128(NufftPlan &P, const void *src, double *dst) {
    const long n = 2L * P.nu * P.half_nv;
    if (P.prec == LT_F64)
        hipLaunchKernelGGL(k_T_add_c128<double>, dim3(blocks_for(n, 256)), dim3(256), 0, stream(), n,
                           (const double *)src, dst);
    else
        hipLaunchKernelGGL(k_T_add_c128<float>, dim3(blocks_for(n, 256)), dim3(256), 0, stream(), n,
                           (const float *)src, dst);
    LT_CHECK_LAUNCH();
    return LT_OK;
}

}  // namespace lt
