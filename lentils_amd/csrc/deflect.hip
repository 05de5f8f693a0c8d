// FP64 deflection kernels: arbitrary points, and the stencil lattice of an
// image grid for a batch of lens-parameter samples.
//
// The per-lens constants (including the pow() of deflect.c:284) are computed
// once on the host; each thread evaluates one point of one sample with the
// arithmetic of deflect_math.h.  Work is pure FP64 ALU (~420 flop/point), so
// the kernels are sized for occupancy: 256 threads, one point per thread,
// lens records read through the scalar cache (uniform across the wave).
#include <cmath>

#include "deflect_math.h"
#include "lt_internal.h"

namespace lt {

lt_status pack_lenses(const generic_mass_model *in, int n, std::vector<DevLens> &out) {
    out.clear();
    for (int k = 0; k < n; ++k) {
        DevLens L;
        if (!pack_lens_one(in[k], L)) {
            set_error("lens type %u is not supported (0 = PEMD, 1 = external shear)", in[k].type);
            return LT_ERR_UNSUPPORTED;
        }
        out.push_back(L);
    }
    return LT_OK;
}

__global__ void __launch_bounds__(256) k_deflect_points(const DevLens *__restrict__ lenses, int nl,
                                                         const double *__restrict__ pin, double *__restrict__ pout,
                                                         long n) {
    long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    double ox, oy;
    deflect_point(lenses, nl, pin[2 * p], pin[2 * p + 1], ox, oy);
    pout[2 * p] = ox;
    pout[2 * p + 1] = oy;
}

// points with their PEMD parameter derivatives: the deflected point, and per
// PEMD component c (in the model's order) the 12 values
// d alpha / d(b_eff, qh, q, x, y, th) at dv[(p * npemd + c) * 12 + k]
// (deflect.c:243-338); external shear adds its deflection only (its
// derivative block is disabled in the reference, deflect.c:111-122)
__global__ void __launch_bounds__(256) k_deflect_points_deriv(const DevLens *__restrict__ lenses, int nl, int npemd,
                                                               const double *__restrict__ pin,
                                                               double *__restrict__ pout, double *__restrict__ dv,
                                                               long n) {
    long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double px = pin[2 * p], py = pin[2 * p + 1];
    double ox = px, oy = py;
    int c = 0;
    for (int i = 0; i < nl; ++i) {
        double ax = 0.0, ay = 0.0;
        if (lenses[i].type == 0) {
            double d[12];
#pragma unroll
            for (int k = 0; k < 12; ++k) d[k] = 0.0;
            pemd_deflection_d(lenses[i], px, py, ax, ay, d);
            double *o = dv + (p * npemd + c) * 12;
#pragma unroll
            for (int k = 0; k < 12; ++k) o[k] = d[k];
            ++c;
        } else {
            shear_deflection(lenses[i], px, py, ax, ay);
        }
        ox -= ax;
        oy -= ay;
    }
    pout[2 * p] = ox;
    pout[2 * p + 1] = oy;
}

lt_status deflect_points_deriv_device(const std::vector<DevLens> &lens, const double *pin_dev, double *pout_dev,
                                      double *dv_dev, long n) {
    int npemd = 0;
    for (const DevLens &L : lens) npemd += (L.type == 0);
    DevBuf<DevLens> dl;
    LT_TRY(dl.upload(lens.data(), lens.size()));
    if (n > 0) {
        hipLaunchKernelGGL(k_deflect_points_deriv, dim3(blocks_for(n, 256)), dim3(256), 0, stream(), dl.p,
                           (int)lens.size(), npemd, pin_dev, pout_dev, dv_dev, n);
        LT_CHECK_LAUNCH();
    }
    return sync();
}

// lattice point (a, b) of sample s; lattice is x-major, (2nx+1)*(2ny+1)
// points per sample, 2 doubles each
__global__ void __launch_bounds__(256) k_deflect_lattice(const DevLens *__restrict__ lenses,
                                                          const int *__restrict__ lens_off, int nsamples, int lx,
                                                          int ly, double xmin, double dx, double ymin, double dy,
                                                          double *__restrict__ lat) {
    const long per = (long)lx * ly;
    long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= per * nsamples) return;
    const int s = (int)(g / per);
    const long p = g - (long)s * per;
    const int a = (int)(p / ly), b = (int)(p - (long)a * ly);
    const double px = lattice_coord(xmin, dx, a);
    const double py = lattice_coord(ymin, dy, b);
    const int l0 = lens_off[s], l1 = lens_off[s + 1];
    double ox, oy;
    deflect_point(lenses + l0, l1 - l0, px, py, ox, oy);
    lat[2 * g] = ox;
    lat[2 * g + 1] = oy;
}

lt_status deflect_points_device(const std::vector<DevLens> &lens, const double *pin_dev, double *pout_dev, long n) {
    DevBuf<DevLens> dl;
    LT_TRY(dl.upload(lens.data(), lens.size()));
    if (n > 0) {
        hipLaunchKernelGGL(k_deflect_points, dim3(blocks_for(n, 256)), dim3(256), 0, stream(), dl.p,
                           (int)lens.size(), pin_dev, pout_dev, n);
        LT_CHECK_LAUNCH();
    }
    return sync();
}

lt_status deflect_lattice_device(const DevLens *lens_dev, const int *lens_off_dev, int nsamples, const image_space &im,
                                 double *lat_dev) {
    const int lx = 2 * im.nx + 1, ly = 2 * im.ny + 1;
    const long n = (long)lx * ly * nsamples;
    if (n == 0) return LT_OK;
    hipLaunchKernelGGL(k_deflect_lattice, dim3(blocks_for(n, 256)), dim3(256), 0, stream(), lens_dev, lens_off_dev,
                       nsamples, lx, ly, im.xmin, im.dx, im.ymin, im.dy, lat_dev);
    LT_CHECK_LAUNCH();
    return LT_OK;
}

}  // namespace lt
