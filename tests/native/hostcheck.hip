// Host build of the device arithmetic (deflect_math.h, raster_core.h): lets
// the CPU test suite compare the exact code the kernels run against the
// oracle, bit for bit, without a GPU.  Test infrastructure only.
#include <vector>

#include "../../lentils_amd/csrc/deflect_math.h"
#include "../../lentils_amd/csrc/raster_core.h"

using namespace lt;

namespace {
struct VecSink {
    std::vector<double> area;
    std::vector<int> idx;
    void operator()(double a, int i) {
        area.push_back(a);
        idx.push_back(i);
    }
    bool stop() const { return false; }
};
struct StoreEmit {
    int *cols;
    double *vals;
    long base;
    void operator()(int k, int col, double val) {
        if (cols) {
            cols[base + k] = col;
            vals[base + k] = val;
        }
    }
};
}  // namespace

extern "C" {

int hc_deflect_points(const generic_mass_model *lenses, int nl, const double *pin, long n, double *pout) {
    std::vector<DevLens> L(nl);
    for (int k = 0; k < nl; ++k)
        if (!pack_lens_one(lenses[k], L[k])) return 1;
    for (long p = 0; p < n; ++p) deflect_point(L.data(), nl, pin[2 * p], pin[2 * p + 1], pout[2 * p], pout[2 * p + 1]);
    return 0;
}

// deflection with the PEMD parameter derivatives, per point summed over
// the PEMD components (the device kernel's per-component values, added in
// component order)
int hc_deflect_points_deriv(const generic_mass_model *lenses, int nl, const double *pin, long n, double *pout,
                            double *dv) {
    std::vector<DevLens> L(nl);
    for (int k = 0; k < nl; ++k)
        if (!pack_lens_one(lenses[k], L[k])) return 1;
    for (long p = 0; p < n; ++p) {
        const double px = pin[2 * p], py = pin[2 * p + 1];
        double ox = px, oy = py, acc[12] = {0};
        for (int i = 0; i < nl; ++i) {
            double ax = 0.0, ay = 0.0;
            if (L[i].type == 0) {
                double d[12] = {0};
                pemd_deflection_d(L[i], px, py, ax, ay, d);
                for (int k = 0; k < 12; ++k) acc[k] += d[k];
            } else {
                shear_deflection(L[i], px, py, ax, ay);
            }
            ox -= ax;
            oy -= ay;
        }
        pout[2 * p] = ox;
        pout[2 * p + 1] = oy;
        for (int k = 0; k < 12; ++k) dv[12 * p + k] = acc[k];
    }
    return 0;
}

// lens matrix from a lattice with the kernels' per-row code path; cols /
// vals may be NULL (count only)
long hc_lens_matrix_from_lattice(int nx, int ny, const char *mask, int snx, int sny, double sxmin, double symin,
                                 double sdx, double sdy, const double *lat, int *rp, int *cols, double *vals,
                                 int *status) {
    SrcGrid g{snx, sny, sxmin, symin, sdx, sdy};
    long k = 0;
    int stale = 0, st = 0;
    const int ly = 2 * ny + 1;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
            const long r = (long)i * ny + j;
            rp[r] = (int)k;
            if (mask && !mask[r]) continue;
            VecSink s;
            st |= raster_pixel(lat, ly, i, j, g, s);
            StoreEmit e{cols, vals, k};
            const int n = (int)s.area.size();
            if (n < 64) {  // reduce_fragments reads short lists in blocks up to 64 entries
                s.area.resize(64, 0.0);
                s.idx.resize(64, 0);
            }
            int first = 0;
            k += reduce_fragments(s.area.data(), s.idx.data(), n, stale, e, &first);
            if (n > 0) stale = first;
        }
    rp[(long)nx * ny] = (int)k;
    if (status) *status = st;
    return k;
}

// Fragments of image pixel (i, j) in emission order.  mode 0: the
// sequential depth-first walk (raster_pixel); mode 1: the long-row
// decomposition the GPU uses (breadth-first frontier of at most `front`
// items per fan triangle, then one depth-first walk per frontier item).
// Returns the fragment count (at most cap are stored), or -1 when mode 1
// meets a stack overflow / pop limit (the GPU then falls back to mode 0).
long hc_pixel_frags(const double *lat, int ny, int i, int j, int snx, int sny, double sxmin, double symin, double sdx,
                    double sdy, int mode, int front, double *area, int *idx, long cap, int *levels_out) {
    SrcGrid g{snx, sny, sxmin, symin, sdx, sdy};
    const int ly = 2 * ny + 1;
    VecSink s;
    int levels_max = 0;
    if (mode == 0) {
        raster_pixel(lat, ly, i, j, g, s);
    } else {
        for (int t = 0; t < 8; ++t) {
            double tv[6];
            fan_vertices(lat, ly, i, j, t, tv);
            std::vector<RItem> cur(1), nxt;
            raster_root(tv, g, cur[0].tri);
            cur[0].depth = 0;
            cur[0].kind = 0;
            int level = 0;
            for (; level < kLongLevels; ++level) {
                std::vector<int> cnt(cur.size() + 1, 0);
                int ovf = 0, any_sub = 0;
                for (size_t q = 0; q < cur.size(); ++q) {
                    const bool ex = ritem_splits(cur[q]);
                    cnt[q + 1] = cnt[q] + (ex ? ritem_count(cur[q], g, &ovf) : 1);
                    any_sub |= ex;
                }
                if (ovf) return -1;
                if (!any_sub || cnt.back() > front) break;
                nxt.assign(cnt.back(), RItem{});
                for (size_t q = 0; q < cur.size(); ++q) {
                    if (ritem_splits(cur[q]))
                        ritem_write(cur[q], g, nxt.data() + cnt[q]);
                    else
                        nxt[cnt[q]] = cur[q];
                }
                cur.swap(nxt);
            }
            if (level > levels_max) levels_max = level;
            for (const RItem &it : cur) {
                if (it.kind == 1) {
                    s(it.area, it.idx);
                } else if (raster_dfs<true>(it.tri, it.depth, g, s)) {
                    return -1;
                }
            }
        }
    }
    const long n = (long)s.area.size();
    for (long k = 0; k < n && k < cap; ++k) {
        area[k] = s.area[k];
        idx[k] = s.idx[k];
    }
    if (levels_out) *levels_out = levels_max;
    return n;
}

}  // extern "C"
