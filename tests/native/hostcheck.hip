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
            k += reduce_fragments(s.area.data(), s.idx.data(), n, stale, e);
            if (n > 0) stale = s.idx[0];
        }
    rp[(long)nx * ny] = (int)k;
    if (status) *status = st;
    return k;
}

}  // extern "C"
