// Lens-matrix construction on the GPU for a batch of lens samples.
//
// Pipeline (one launch each, all samples at once):
//   1. deflect the stencil lattice of the image grid        (deflect.hip)
//   2. count pass: one thread per image pixel rasterizes its 8 fan
//      triangles depth-first (raster_core.h), keeps up to kShort fragments
//      in registers/scratch, sorts and reduces them to the CSR entry count;
//      pixels with more fragments are marked long
//   3. long rows: one thread re-emits the fragments to HBM, a segmented sort
//      orders them by (source index, emission position), one thread reduces
//   4. exclusive scan of entry counts -> row pointers; running max of the
//      last row with fragments -> the stale column of empty rows
//   5. fill pass: short rows are rasterized again and written in place, long
//      rows copied from their staging area
// The result is a block-diagonal CSR (global row / column indices) whose
// per-sample blocks are bit-identical to manifold_lens_matrix_csr
// (lentils/backend/raster.c:272-395) given the same deflections.
#include <algorithm>
#include <numeric>

#include "lt_internal.h"
#include "raster_core.h"
#include "raster_batch.h"

namespace lt {

lt_status deflect_lattice_device(const DevLens *lens_dev, const int *lens_off_dev, int nsamples, const image_space &im,
                                 double *lat_dev);

namespace {

constexpr int kShort = 32;  // fragments a short row may hold
constexpr int kRasterThreads = 128;

struct LocalFrags {
    double area[kShort];
    int idx[kShort];
    int n = 0;
    __device__ void operator()(double a, int i) {
        if (n < kShort) {
            area[n] = a;
            idx[n] = i;
        }
        ++n;
    }
};

struct CountEmit {
    __device__ void operator()(int, int, double) {}
};

struct StoreEmit {
    int32_t *cols;
    double *vals;
    int32_t coloff;
    __device__ void operator()(int k, int col, double val) {
        cols[k] = coloff + col;
        vals[k] = val;
    }
};

struct BatchGeom {
    int nsamples, nx, ny, lx, ly;
    long npix;
};

__global__ void __launch_bounds__(kRasterThreads) k_raster_count(BatchGeom G, const SrcGrid *__restrict__ src,
                                                                  const double *__restrict__ lat,
                                                                  const uint8_t *__restrict__ mask,
                                                                  int64_t *__restrict__ nnz_row,
                                                                  int32_t *__restrict__ nfrag,
                                                                  int32_t *__restrict__ minidx,
                                                                  int32_t *__restrict__ lastne, int *__restrict__ status) {
    const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= G.npix * G.nsamples) return;
    const int s = (int)(r / G.npix);
    const long p = r - (long)s * G.npix;
    if (!mask[p]) {
        nnz_row[r] = 0;
        nfrag[r] = 0;
        minidx[r] = -1;
        lastne[r] = -1;
        return;
    }
    const int i = (int)(p / G.ny), j = (int)(p - (long)i * G.ny);
    LocalFrags fr;
    const double *L = lat + (long)s * G.lx * G.ly * 2;
    int st = raster_pixel(L, G.ly, i, j, src[s], fr);
    if (st) atomicOr(&status[s], st);
    nfrag[r] = fr.n;
    lastne[r] = fr.n > 0 ? (int32_t)r : -1;
    if (fr.n > kShort) {  // long row: counted by the long-row path
        nnz_row[r] = 0;
        minidx[r] = -1;
        return;
    }
    CountEmit ce;
    int k = reduce_fragments(fr.area, fr.idx, fr.n, 0, ce);
    nnz_row[r] = k;
    minidx[r] = fr.n > 0 ? fr.idx[0] : -1;
}

struct GlobalFrags {
    double *area;
    uint64_t *key;
    long base;
    int n = 0;
    __device__ void operator()(double a, int i) {
        area[base + n] = a;
        key[base + n] = ((uint64_t)(uint32_t)i << 32) | (uint32_t)n;
        ++n;
    }
};

__global__ void k_long_emit(BatchGeom G, const SrcGrid *__restrict__ src, const double *__restrict__ lat,
                            const int64_t *__restrict__ rows, const int64_t *__restrict__ foff, int nlong,
                            double *__restrict__ area, uint64_t *__restrict__ key) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nlong) return;
    const long r = rows[l];
    const int s = (int)(r / G.npix);
    const long p = r - (long)s * G.npix;
    const int i = (int)(p / G.ny), j = (int)(p - (long)i * G.ny);
    GlobalFrags gf{area, key, foff[l]};
    raster_pixel(lat + (long)s * G.lx * G.ly * 2, G.ly, i, j, src[s], gf);
}

// sequential reduction of one sorted long row (raster.c:375-391); entries
// staged at the row's fragment offset
__global__ void k_long_reduce(const int64_t *__restrict__ rows, const int64_t *__restrict__ foff, int nlong,
                              const double *__restrict__ area, const uint64_t *__restrict__ key,
                              int32_t *__restrict__ scol, double *__restrict__ sval, int64_t *__restrict__ nnz_row,
                              int32_t *__restrict__ minidx) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nlong) return;
    const long b = foff[l], e = foff[l + 1];
    const long r = rows[l];
    int col = (int)(key[b] >> 32);
    minidx[r] = col;
    double val = 0.0;
    long k = 0;
    for (long f = b; f < e; ++f) {
        const int idx = (int)(key[f] >> 32);
        if (idx != col) {
            scol[b + k] = col;
            sval[b + k] = val;
            ++k;
            val = 0.0;
            col = idx;
        } else {
            val += area[b + (long)(uint32_t)key[f]];
        }
    }
    scol[b + k] = col;
    sval[b + k] = val;
    ++k;
    nnz_row[r] = k;
}

__global__ void __launch_bounds__(kRasterThreads) k_raster_fill(BatchGeom G, const SrcGrid *__restrict__ src,
                                                                 const int64_t *__restrict__ srcoff,
                                                                 const double *__restrict__ lat,
                                                                 const uint8_t *__restrict__ mask,
                                                                 const int64_t *__restrict__ rowoff,
                                                                 const int32_t *__restrict__ nfrag,
                                                                 const int32_t *__restrict__ minidx,
                                                                 const int32_t *__restrict__ lastne,
                                                                 int32_t *__restrict__ cols, double *__restrict__ vals) {
    const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= G.npix * G.nsamples) return;
    const int s = (int)(r / G.npix);
    const long p = r - (long)s * G.npix;
    if (!mask[p] || nfrag[r] > kShort) return;
    const int i = (int)(p / G.ny), j = (int)(p - (long)i * G.ny);
    int stale = 0;
    if (nfrag[r] == 0) {
        const int32_t q = lastne[r];
        if (q >= 0 && (long)q >= (long)s * G.npix) stale = minidx[q];
    }
    LocalFrags fr;
    if (nfrag[r] > 0) raster_pixel(lat + (long)s * G.lx * G.ly * 2, G.ly, i, j, src[s], fr);
    StoreEmit se{cols + rowoff[r], vals + rowoff[r], (int32_t)srcoff[s]};
    reduce_fragments(fr.area, fr.idx, fr.n, stale, se);
}

__global__ void k_long_copy(const int64_t *__restrict__ rows, const int64_t *__restrict__ foff, int nlong,
                            const int64_t *__restrict__ rowoff, const int64_t *__restrict__ nnz_row,
                            const int32_t *__restrict__ scol, const double *__restrict__ sval,
                            const int64_t *__restrict__ srcoff, long npix, int32_t *__restrict__ cols,
                            double *__restrict__ vals) {
    const int l = blockIdx.x;
    if (l >= nlong) return;
    const long r = rows[l];
    const int s = (int)(r / npix);
    const long b = foff[l], n = nnz_row[r], o = rowoff[r];
    for (long k = threadIdx.x; k < n; k += blockDim.x) {
        cols[o + k] = (int32_t)srcoff[s] + scol[b + k];
        vals[o + k] = sval[b + k];
    }
}

__global__ void k_long_flags(long n, const int32_t *__restrict__ nfrag, int64_t *__restrict__ flag) {
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) flag[r] = nfrag[r] > kShort ? 1 : 0;
}

__global__ void k_long_compact(long n, const int64_t *__restrict__ pos, const int32_t *__restrict__ nfrag,
                               int64_t *__restrict__ rows, int32_t *__restrict__ cnt) {
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n && pos[r + 1] != pos[r]) {
        rows[pos[r]] = r;
        cnt[pos[r]] = nfrag[r];
    }
}

__global__ void k_sample_bounds(int ns, long npix, const int64_t *__restrict__ rowoff, int64_t *__restrict__ out) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s <= ns) out[s] = rowoff[(long)s * npix];
}

__global__ void k_rowptr(const int64_t *__restrict__ rowoff, long n, int32_t *__restrict__ rp) {
    long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r <= n) rp[r] = (int32_t)rowoff[r];
}

}  // namespace

lt_status LensBatch::build(const image_space &im_, const image_space *srcs, int ns, const generic_mass_model *lenses,
                           const int *lens_counts, const double *lattice_host) {
    LT_TRY(ensure_init());
    if (ns <= 0 || im_.nx <= 0 || im_.ny <= 0) {
        set_error("lens batch: empty image grid or no samples");
        return LT_ERR_ARG;
    }
    im = im_;
    nsamples = ns;
    npix = (long)im.nx * im.ny;
    src_host.resize(ns);
    srcoff_host.assign(ns + 1, 0);
    for (int s = 0; s < ns; ++s) {
        const image_space &g = srcs[s];
        if (g.nx <= 0 || g.ny <= 0) {
            set_error("lens batch: sample %d has an empty source grid", s);
            return LT_ERR_ARG;
        }
        src_host[s] = SrcGrid{g.nx, g.ny, g.xmin, g.ymin, g.dx, g.dy};
        srcoff_host[s + 1] = srcoff_host[s] + (int64_t)g.nx * g.ny;
    }
    status_host.assign(ns, 0);
    // mask (x-major bytes); a NULL mask means every pixel is used
    std::vector<uint8_t> mask_host(npix, 1);
    if (im.mask)
        for (long p = 0; p < npix; ++p) mask_host[p] = im.mask[p] ? 1 : 0;
    LT_TRY(mask.upload(mask_host.data(), npix));
    LT_TRY(src.upload(src_host.data(), ns));
    LT_TRY(srcoff.upload(srcoff_host.data(), ns + 1));
    const int lx = 2 * im.nx + 1, ly = 2 * im.ny + 1;
    const long latn = (long)lx * ly * 2;
    LT_TRY(lattice.alloc(latn * ns));
    if (lattice_host) {
        LT_TRY(lattice.upload(lattice_host, latn * ns));
    } else {
        std::vector<DevLens> packed;
        std::vector<int> off(ns + 1, 0);
        std::vector<DevLens> one;
        for (int s = 0; s < ns; ++s) {
            LT_TRY(pack_lenses(lenses + off[s], lens_counts[s], one));
            packed.insert(packed.end(), one.begin(), one.end());
            off[s + 1] = off[s] + lens_counts[s];
        }
        DevBuf<DevLens> dl;
        DevBuf<int> doff;
        if (packed.empty()) packed.push_back(DevLens{});
        LT_TRY(dl.upload(packed.data(), packed.size()));
        LT_TRY(doff.upload(off.data(), off.size()));
        LT_TRY(deflect_lattice_device(dl.p, doff.p, ns, im, lattice.p));
        LT_TRY(sync());
    }
    const long R = npix * ns;
    DevBuf<int64_t> nnz_row, rowoff;
    DevBuf<int32_t> nfrag, minidx, lastne;
    DevBuf<int> dstatus;
    LT_TRY(nnz_row.alloc(R));
    LT_TRY(rowoff.alloc(R + 1));
    LT_TRY(nfrag.alloc(R));
    LT_TRY(minidx.alloc(R));
    LT_TRY(lastne.alloc(R));
    LT_TRY(dstatus.alloc(ns));
    LT_TRY(dstatus.zero());
    BatchGeom G{ns, im.nx, im.ny, lx, ly, npix};
    hipLaunchKernelGGL(k_raster_count, dim3(blocks_for(R, kRasterThreads)), dim3(kRasterThreads), 0, stream(), G,
                       src.p, lattice.p, mask.p, nnz_row.p, nfrag.p, minidx.p, lastne.p, dstatus.p);
    LT_CHECK_LAUNCH();
    // long rows
    // long rows: compacted on the device; only their (few) ids and fragment
    // counts come back to size the staging area
    DevBuf<int64_t> lflag, lpos;
    LT_TRY(lflag.alloc(R + 1));
    LT_TRY(lpos.alloc(R + 1));
    hipLaunchKernelGGL(k_long_flags, dim3(blocks_for(R, 256)), dim3(256), 0, stream(), R, nfrag.p, lflag.p);
    LT_CHECK_LAUNCH();
    LT_TRY(exclusive_scan_i64(lflag.p, lpos.p, R));
    int64_t nl64 = 0;
    LT_TRY(lpos.download(&nl64, 1, R));
    LT_TRY(sync());
    std::vector<int64_t> lrows(nl64), loff(nl64 + 1, 0);
    if (nl64 > 0) {
        DevBuf<int64_t> lr;
        DevBuf<int32_t> ln;
        LT_TRY(lr.alloc(nl64));
        LT_TRY(ln.alloc(nl64));
        hipLaunchKernelGGL(k_long_compact, dim3(blocks_for(R, 256)), dim3(256), 0, stream(), R, lpos.p, nfrag.p, lr.p,
                           ln.p);
        LT_CHECK_LAUNCH();
        std::vector<int32_t> hn(nl64);
        LT_TRY(lr.download(lrows.data(), nl64));
        LT_TRY(ln.download(hn.data(), nl64));
        LT_TRY(sync());
        for (long k = 0; k < nl64; ++k) loff[k + 1] = loff[k] + hn[k];
    }
    const int nlong = (int)lrows.size();
    DevBuf<int64_t> drows, dfoff;
    DevBuf<double> farea, sval;
    DevBuf<uint64_t> fkey;
    DevBuf<int32_t> scol;
    if (nlong) {
        const long F = loff.back();
        LT_TRY(drows.upload(lrows.data(), nlong));
        LT_TRY(dfoff.upload(loff.data(), nlong + 1));
        LT_TRY(farea.alloc(F));
        LT_TRY(fkey.alloc(F));
        LT_TRY(scol.alloc(F));
        LT_TRY(sval.alloc(F));
        hipLaunchKernelGGL(k_long_emit, dim3(blocks_for(nlong, 64)), dim3(64), 0, stream(), G, src.p, lattice.p,
                           drows.p, dfoff.p, nlong, farea.p, fkey.p);
        LT_CHECK_LAUNCH();
        LT_TRY(segmented_sort_u64(fkey.p, dfoff.p, loff));
        hipLaunchKernelGGL(k_long_reduce, dim3(blocks_for(nlong, 64)), dim3(64), 0, stream(), drows.p, dfoff.p,
                           nlong, farea.p, fkey.p, scol.p, sval.p, nnz_row.p, minidx.p);
        LT_CHECK_LAUNCH();
    }
    LT_TRY(exclusive_scan_i64(nnz_row.p, rowoff.p, R));
    LT_TRY(inclusive_max_i32(lastne.p, lastne.p, R));
    int64_t total = 0;
    LT_TRY(rowoff.download(&total, 1, R));
    LT_TRY(sync());
    if (total >= INT32_MAX) {
        set_error("lens batch: %lld nonzeros exceed the int32 index range; use fewer samples per batch",
                  (long long)total);
        return LT_ERR_CAPACITY;
    }
    L.nrows = R;
    L.ncols = srcoff_host[ns];
    L.nnz = total;
    LT_TRY(L.rp.alloc(R + 1));
    LT_TRY(L.ci.alloc(std::max<int64_t>(total, 1)));
    LT_TRY(L.va.alloc(std::max<int64_t>(total, 1)));
    hipLaunchKernelGGL(k_rowptr, dim3(blocks_for(R + 1, 256)), dim3(256), 0, stream(), rowoff.p, R, L.rp.p);
    LT_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_raster_fill, dim3(blocks_for(R, kRasterThreads)), dim3(kRasterThreads), 0, stream(), G,
                       src.p, srcoff.p, lattice.p, mask.p, rowoff.p, nfrag.p, minidx.p, lastne.p, L.ci.p, L.va.p);
    LT_CHECK_LAUNCH();
    if (nlong) {
        hipLaunchKernelGGL(k_long_copy, dim3(nlong), dim3(256), 0, stream(), drows.p, dfoff.p, nlong, rowoff.p,
                           nnz_row.p, scol.p, sval.p, srcoff.p, npix, L.ci.p, L.va.p);
        LT_CHECK_LAUNCH();
    }
    LT_TRY(dstatus.download(status_host.data(), ns));
    // per-sample nnz offsets (row pointer at every sample boundary)
    nnzoff_host.assign(ns + 1, 0);
    {
        DevBuf<int64_t> bnd;
        LT_TRY(bnd.alloc(ns + 1));
        hipLaunchKernelGGL(k_sample_bounds, dim3(blocks_for(ns + 1, 64)), dim3(64), 0, stream(), ns, npix, rowoff.p,
                           bnd.p);
        LT_CHECK_LAUNCH();
        LT_TRY(bnd.download(nnzoff_host.data(), ns + 1));
        LT_TRY(sync());
    }
    nlong_rows = nlong;
    return csr_prepare(L);
}

lt_status LensBatch::download(int s, int *row_inds, int *cols, double *vals) const {
    if (s < 0 || s >= nsamples) {
        set_error("sample %d out of range", s);
        return LT_ERR_ARG;
    }
    std::vector<int32_t> rp(npix + 1);
    LT_TRY(L.rp.download(rp.data(), npix + 1, (size_t)s * npix));
    const long n = nnzoff_host[s + 1] - nnzoff_host[s];
    std::vector<int32_t> c(n);
    if (n) {
        LT_TRY(L.ci.download(c.data(), n, nnzoff_host[s]));
        LT_TRY(L.va.download(vals, n, nnzoff_host[s]));
    }
    LT_TRY(sync());
    for (long r = 0; r <= npix; ++r) row_inds[r] = rp[r] - (int32_t)nnzoff_host[s];
    for (long k = 0; k < n; ++k) cols[k] = c[k] - (int32_t)srcoff_host[s];
    return LT_OK;
}

}  // namespace lt
