// C ABI of liblentils_hip.so: library state, drop-in symbols with the
// reference's signatures, and the lt_* deflection / lens-matrix entry points.
#include <cstdarg>
#include <cstring>
#include <map>
#include <mutex>

#include "lt_internal.h"
#include "raster_batch.h"

namespace lt {

static thread_local char g_err[1024] = "";
static int g_device = -1;
static hipStream_t g_stream = nullptr;
static int g_cus = 0;
static std::mutex g_init_mu;

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

const char *last_error() { return g_err; }

lt_status init_device(int device) {
    std::lock_guard<std::mutex> lk(g_init_mu);
    if (g_stream && device == g_device) return LT_OK;
    int n = 0;
    LT_HIP(hipGetDeviceCount(&n));
    if (n <= 0) {
        set_error("no HIP device visible");
        return LT_ERR_HIP;
    }
    if (device < 0 || device >= n) {
        set_error("device %d out of range (%d visible)", device, n);
        return LT_ERR_ARG;
    }
    LT_HIP(hipSetDevice(device));
    if (g_stream) {
        (void)hipStreamDestroy(g_stream);
        g_stream = nullptr;
    }
    LT_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    hipDeviceProp_t prop;
    LT_HIP(hipGetDeviceProperties(&prop, device));
    g_cus = prop.multiProcessorCount;
    LT_TRY(noalloc_setup(device));
    g_device = device;
    return LT_OK;
}

lt_status ensure_init() {
    if (g_stream) {
        LT_HIP(hipSetDevice(g_device));
        return LT_OK;
    }
    return init_device(0);
}

hipStream_t stream() { return g_stream; }
int device_cu_count() { return g_cus; }

// The pool: hipMalloc / hipFree cost tens of microseconds each and hipFree
// synchronizes the device, while a lens batch and its solver allocate a few
// hundred buffers per step, of sizes that recur from step to step (up to
// the nonzero counts: hence size classes of a quarter octave).  All device
// work of the library runs on its one stream, so a block's earlier users
// are ordered before its next ones.  Blocks beyond kPoolCap pooled bytes
// are freed.
namespace {
std::mutex g_pool_mu;
std::map<std::pair<int, size_t>, std::vector<void *>> g_pool;
size_t g_pool_bytes = 0;
constexpr size_t kPoolCap = size_t(64) << 30;

size_t pool_class(size_t bytes) {
    if (bytes <= 4096) return (bytes + 255) / 256 * 256;
    int k = 63 - __builtin_clzll((unsigned long long)bytes);
    const size_t step = (size_t)1 << (k - 2);
    return (bytes + step - 1) / step * step;
}
}  // namespace

hipError_t pool_get(void **p, size_t bytes) {
    const size_t cls = pool_class(bytes);
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    {
        std::lock_guard<std::mutex> lk(g_pool_mu);
        auto it = g_pool.find({dev, cls});
        if (it != g_pool.end() && !it->second.empty()) {
            *p = it->second.back();
            it->second.pop_back();
            g_pool_bytes -= cls;
            return hipSuccess;
        }
    }
    return hipMalloc(p, cls);
}

void pool_put(void *p, size_t bytes) {
    const size_t cls = pool_class(bytes);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        (void)hipFree(p);
        return;
    }
    std::lock_guard<std::mutex> lk(g_pool_mu);
    if (g_pool_bytes + cls > kPoolCap) {
        (void)hipFree(p);
        return;
    }
    g_pool[{dev, cls}].push_back(p);
    g_pool_bytes += cls;
}

lt_status build_curvature(int nx, int ny, double dx, double dy, int32_t coloff, DevCSR &C);
lt_status build_gradient(int nx, int ny, double dx, double dy, DevCSR &G);
lt_status build_dft(const visibility_space &vs, const image_space &im, int *row_inds, int *cols, double *vals);
lt_status deflect_points_device(const std::vector<DevLens> &lens, const double *pin_dev, double *pout_dev, long n);
lt_status deflect_lattice_device(const DevLens *lens_dev, const int *lens_off_dev, int nsamples, const image_space &im,
                                 double *lat_dev);

static lt_status csr_download(const DevCSR &A, int *rp, int *ci, double *va) {
    LT_TRY(A.rp.download(rp, A.nrows + 1));
    if (A.nnz) {
        LT_TRY(A.ci.download(ci, A.nnz));
        LT_TRY(A.va.download(va, A.nnz));
    }
    return sync();
}

}  // namespace lt

using namespace lt;

extern "C" {

lt_status lt_init(int device) { return init_device(device); }

const char *lt_last_error_message(void) { return last_error(); }

const char *lt_version(void) { return "lentils_hip 0.1 gfx950"; }

lt_status lt_deflect_points(const global_lens_model *glm, const double *p_in, long npoints, double *p_out) {
    LT_TRY(ensure_init());
    if (!glm || npoints < 0 || (npoints > 0 && (!p_in || !p_out))) {
        set_error("lt_deflect_points: bad arguments");
        return LT_ERR_ARG;
    }
    std::vector<DevLens> lens;
    LT_TRY(pack_lenses(glm->lenses, glm->num_lenses, lens));
    if (npoints == 0) return LT_OK;
    DevBuf<double> din, dout;
    LT_TRY(din.upload(p_in, 2 * npoints));
    LT_TRY(dout.alloc(2 * npoints));
    LT_TRY(deflect_points_device(lens, din.p, dout.p, npoints));
    LT_TRY(dout.download(p_out, 2 * npoints));
    return sync();
}

lt_status lt_deflect_points_deriv(const global_lens_model *glm, const double *p_in, long npoints, double *p_out,
                                  double *dalpha) {
    LT_TRY(ensure_init());
    if (!glm || npoints < 0 || (npoints > 0 && (!p_in || !p_out))) {
        set_error("lt_deflect_points_deriv: bad arguments");
        return LT_ERR_ARG;
    }
    std::vector<DevLens> lens;
    LT_TRY(pack_lenses(glm->lenses, glm->num_lenses, lens));
    long npemd = 0;
    for (const DevLens &L : lens) npemd += (L.type == 0);
    if (npoints == 0) return LT_OK;
    if (npemd > 0 && !dalpha) {
        set_error("lt_deflect_points_deriv: dalpha is NULL");
        return LT_ERR_ARG;
    }
    DevBuf<double> din, dout, dd;
    LT_TRY(din.upload(p_in, 2 * npoints));
    LT_TRY(dout.alloc(2 * npoints));
    LT_TRY(dd.alloc(std::max(1L, 12 * npemd * npoints)));
    LT_TRY(deflect_points_deriv_device(lens, din.p, dout.p, dd.p, npoints));
    LT_TRY(dout.download(p_out, 2 * npoints));
    if (npemd > 0) LT_TRY(dd.download(dalpha, 12 * npemd * npoints));
    return sync();
}

// want_deriv != 0: the reference adds every point's PEMD derivatives into
// the same deriv[0..11] (deflect.c:26-66 passes one pointer to every point),
// points in order and components in order; the per-point values come from
// the GPU in chunks and are added here in that order
void deflect_points(global_lens_model glm, double *p_in, int npoints, double *p_out, int want_deriv, double *deriv) {
    if (!want_deriv) {
        if (lt_deflect_points(&glm, p_in, npoints, p_out) != LT_OK)
            fprintf(stderr, "lentils_hip: %s\n", last_error());
        return;
    }
    long npemd = 0;
    for (int i = 0; i < glm.num_lenses; ++i) npemd += (glm.lenses[i].type == 0);
    const long chunk = 1L << 20;
    std::vector<double> dv((size_t)std::max(1L, 12 * npemd * std::min<long>(chunk, std::max(npoints, 1))));
    for (long p0 = 0; p0 < npoints; p0 += chunk) {
        const long n = std::min<long>(chunk, npoints - p0);
        if (lt_deflect_points_deriv(&glm, p_in + 2 * p0, n, p_out + 2 * p0, dv.data()) != LT_OK) {
            fprintf(stderr, "lentils_hip: %s\n", last_error());
            return;
        }
        if (!deriv) continue;
        for (long q = 0; q < n * npemd; ++q)
            for (int k = 0; k < 12; ++k) deriv[k] += dv[q * 12 + k];
    }
}

lt_status lt_lens_matrix_nnz(image_space im, image_space src, global_lens_model glm, long *nnz, int *status) {
    LensBatch lb;
    int cnt = glm.num_lenses;
    LT_TRY(lb.build(im, &src, 1, glm.lenses, &cnt, nullptr));
    if (nnz) *nnz = lb.L.nnz;
    if (status) *status = lb.status_host[0];
    return LT_OK;
}

static lt_status lens_matrix_common(LensBatch &lb, int *row_inds, int *cols, double *vals, long capacity, long *nnz,
                                    int *status) {
    if (nnz) *nnz = lb.L.nnz;
    if (status) *status = lb.status_host[0];
    if (capacity >= 0 && lb.L.nnz > capacity) {
        set_error("lens matrix has %ld nonzeros, caller capacity is %ld", lb.L.nnz, capacity);
        return LT_ERR_CAPACITY;
    }
    return lb.download(0, row_inds, cols, vals);
}

lt_status lt_lens_matrix_csr(image_space im, image_space src, global_lens_model glm, int *row_inds, int *cols,
                             double *vals, long capacity, long *nnz, int *status) {
    LensBatch lb;
    int cnt = glm.num_lenses;
    LT_TRY(lb.build(im, &src, 1, glm.lenses, &cnt, nullptr));
    return lens_matrix_common(lb, row_inds, cols, vals, capacity, nnz, status);
}

lt_status lt_lens_matrix_from_lattice(image_space im, image_space src, const double *lattice, int *row_inds, int *cols,
                                      double *vals, long capacity, long *nnz, int *status) {
    LensBatch lb;
    int cnt = 0;
    LT_TRY(lb.build(im, &src, 1, nullptr, &cnt, lattice));
    return lens_matrix_common(lb, row_inds, cols, vals, capacity, nnz, status);
}

lt_status lt_deflect_lattice(image_space im, global_lens_model glm, double *p_out) {
    LT_TRY(ensure_init());
    std::vector<DevLens> lens;
    LT_TRY(pack_lenses(glm.lenses, glm.num_lenses, lens));
    if (lens.empty()) lens.push_back(DevLens{});
    const int off[2] = {0, glm.num_lenses};
    const long n = (long)(2 * im.nx + 1) * (2 * im.ny + 1) * 2;
    DevBuf<DevLens> dl;
    DevBuf<int> doff;
    DevBuf<double> lat;
    LT_TRY(dl.upload(lens.data(), lens.size()));
    LT_TRY(doff.upload(off, 2));
    LT_TRY(lat.alloc(n));
    LT_TRY(deflect_lattice_device(dl.p, doff.p, 1, im, lat.p));
    LT_TRY(lat.download(p_out, n));
    return sync();
}

void manifold_lens_matrix_csr(image_space im, image_space src, global_lens_model glm, int *row_inds, int *cols,
                              double *vals) {
    if (lt_lens_matrix_csr(im, src, glm, row_inds, cols, vals, -1, nullptr, nullptr) != LT_OK)
        fprintf(stderr, "lentils_hip: %s\n", last_error());
}

void curv_matrix_findiff_csr(image_space im, int *row_inds, int *cols, double *vals) {
    DevCSR C;
    lt_status s = ensure_init();
    if (s == LT_OK) s = build_curvature(im.nx, im.ny, im.dx, im.dy, 0, C);
    if (s == LT_OK) s = csr_download(C, row_inds, cols, vals);
    if (s != LT_OK) fprintf(stderr, "lentils_hip: %s\n", last_error());
}

void grad_matrix_findiff_csr(image_space im, int *row_inds, int *cols, double *vals) {
    DevCSR G;
    lt_status s = ensure_init();
    if (s == LT_OK) s = build_gradient(im.nx, im.ny, im.dx, im.dy, G);
    if (s == LT_OK) s = csr_download(G, row_inds, cols, vals);
    if (s != LT_OK) fprintf(stderr, "lentils_hip: %s\n", last_error());
}

void dft_matrix_csr(visibility_space vs, image_space im, int *row_inds, int *cols, double *vals) {
    lt_status s = ensure_init();
    if (s == LT_OK) s = build_dft(vs, im, row_inds, cols, vals);
    if (s != LT_OK) fprintf(stderr, "lentils_hip: %s\n", last_error());
}

// --- lens batch handles ------------------------------------------------------

struct lt_lensbatch_s {
    LensBatch lb;
};

lt_status lt_lensbatch_create(image_space im, const image_space *srcs, int ns, const generic_mass_model *lenses,
                              const int *lens_counts, lt_lensbatch *out) {
    if (!out || !srcs || !lens_counts || ns <= 0) {
        set_error("lt_lensbatch_create: bad arguments");
        return LT_ERR_ARG;
    }
    auto *h = new lt_lensbatch_s();
    lt_status s = h->lb.build(im, srcs, ns, lenses, lens_counts, nullptr);
    if (s != LT_OK) {
        delete h;
        return s;
    }
    *out = h;
    return LT_OK;
}

lt_status lt_lensbatch_info(lt_lensbatch h, int sample, long *nnz, int *status) {
    if (!h || sample < 0 || sample >= h->lb.nsamples) {
        set_error("lt_lensbatch_info: bad handle or sample");
        return LT_ERR_ARG;
    }
    if (nnz) *nnz = h->lb.nnzoff_host[sample + 1] - h->lb.nnzoff_host[sample];
    if (status) *status = h->lb.status_host[sample];
    return LT_OK;
}

lt_status lt_lensbatch_download(lt_lensbatch h, int sample, int *row_inds, int *cols, double *vals) {
    if (!h) {
        set_error("lt_lensbatch_download: null handle");
        return LT_ERR_ARG;
    }
    return h->lb.download(sample, row_inds, cols, vals);
}

lt_status lt_lensbatch_destroy(lt_lensbatch h) {
    delete h;
    return LT_OK;
}

}  // extern "C"

namespace lt {
LensBatch *lensbatch_of(lt_lensbatch h) { return h ? &h->lb : nullptr; }
}  // namespace lt
