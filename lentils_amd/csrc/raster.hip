// Lens-matrix construction on the GPU for a batch of lens samples.
//
// Pipeline (one launch each, all samples at once):
//   1. deflect the stencil lattice of the image grid        (deflect.hip)
//   2. count pass: one thread per image pixel rasterizes its 8 fan
//      triangles depth-first (raster_core.h), keeps up to kShort fragments
//      in registers/scratch, sorts and reduces them to the CSR entry count;
//      pixels with more fragments are marked long
//   3. long rows (the pixels around the lens centre): each fan triangle's
//      split tree is expanded breadth first by a workgroup into a frontier
//      kept in the reference's depth-first order, every frontier subtree is
//      walked by its own thread (count, scan, emit to HBM), a segmented radix
//      sort orders the fragments by (source index, emission position) and
//      the reduction runs in parallel over groups of equal source index
//   4. exclusive scan of entry counts -> row pointers; running max of the
//      last row with fragments -> the stale column of empty rows
//   5. fill pass: short rows are rasterized again and written in place, long
//      rows copied from their staging area
// The result is a block-diagonal CSR (global row / column indices) whose
// per-sample blocks are bit-identical to manifold_lens_matrix_csr
// (lentils/backend/raster.c:272-395) given the same deflections.
#include <cstring>

#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <numeric>

#include "lt_internal.h"
#include "raster_core.h"
#include "raster_batch.h"

namespace lt {

lt_status deflect_lattice_device(const DevLens *lens_dev, const int *lens_off_dev, int nsamples, const image_space &im,
                                 double *lat_dev);

namespace {

constexpr int kShort = 64;  // fragments a short row may hold
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
    // a row with more than kShort fragments is long: stop counting it here
    __device__ bool stop() const { return n > kShort; }
};

struct CountSink {
    long n = 0;
    __device__ void operator()(double, int) { ++n; }
    __device__ bool stop() const { return false; }
};

struct CountEmit {
    __device__ void operator()(int, int, double) {}
};

// The count pass keeps the first kStage CSR entries of every short row here
// (row r at r * kStage); the fill pass copies rows of at most kStage entries
// instead of rasterizing them again.
constexpr int kStage = 8;

struct StageEmit {
    int32_t *cols;
    double *vals;
    __device__ void operator()(int k, int col, double val) {
        if (k < kStage) {
            cols[k] = col;
            vals[k] = val;
        }
    }
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
                                                                  int32_t *__restrict__ lastne, int *__restrict__ status,
                                                                  int32_t *__restrict__ stcol,
                                                                  double *__restrict__ stval) {
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
    StageEmit se{stcol + r * kStage, stval + r * kStage};
    int first = 0;
    int k = reduce_fragments(fr.area, fr.idx, fr.n, 0, se, &first);
    nnz_row[r] = k;
    minidx[r] = fr.n > 0 ? first : -1;
}

// fragment n of a long row goes to base + n; its sort key is
// (source index, emission position)
struct GlobalFrags {
    double *area;
    uint64_t *key;
    long base;
    long n = 0;
    __device__ void operator()(double a, int i) {
        area[base + n] = a;
        key[base + n] = ((uint64_t)(uint32_t)i << 32) | (uint32_t)n;
        ++n;
    }
    __device__ bool stop() const { return false; }
};

// Long rows with at most kMedium fragments are walked by one thread per fan
// triangle; longer ("huge") rows take the frontier decomposition below.
// The first pass counts every long row up to kMedium + 1.
constexpr long kMedium = 512;

// count of one fan triangle's walk that also adds to its row's count in
// LDS (shared by the row's 8 lanes) and stops once the row is past kMedium
struct RowCappedCount {
    int *row;
    long n = 0;
    __device__ void operator()(double, int) {
        ++n;
        atomicAdd(row, 1);
    }
    __device__ bool stop() const { return *(volatile int *)row > kMedium; }
};

// The long rows' counts, each row's 8 fan triangles walked in parallel by 8
// consecutive lanes (each triangle's walk starts from an empty stack, as in
// raster_pixel; all 8 stop once the row is past kMedium fragments): a row of
// at most kMedium fragments gets its exact count, and tcnt its triangles'
// counts for the emission; a longer row gets kMedium + 1, as the one-walk
// count gives, and triangle counts that add up to more than kMedium.
__global__ void __launch_bounds__(64) k_long_count8(BatchGeom G, const SrcGrid *__restrict__ src,
                                                    const double *__restrict__ lat, const int64_t *__restrict__ rows,
                                                    int nlong, int32_t *__restrict__ tcnt, int64_t *__restrict__ cnt,
                                                    int *__restrict__ status) {
    __shared__ int rowc[8];  // the workgroup's 8 rows
    if (threadIdx.x < 8) rowc[threadIdx.x] = 0;
    __syncthreads();
    const long gt = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long l = gt >> 3;
    const int t = (int)(gt & 7);
    int n = 0, st = 0, s = 0;
    if (l < nlong) {
        const long r = rows[l];
        s = (int)(r / G.npix);
        const long p = r - (long)s * G.npix;
        const int i = (int)(p / G.ny), j = (int)(p - (long)i * G.ny);
        double tv[6];
        fan_vertices(lat + (long)s * G.lx * G.ly * 2, G.ly, i, j, t, tv);
        RowCappedCount cc{&rowc[threadIdx.x >> 3]};
        st = raster_triangle<true>(tv, src[s], cc);
        n = (int)cc.n;
        tcnt[gt] = n;
    }
    int sum = n, sto = st;  // the row's 8 lanes share one wave
    for (int m = 1; m < 8; m <<= 1) {
        sum += __shfl_xor(sum, m);
        sto |= __shfl_xor(sto, m);
    }
    if (l < nlong && t == 0) {
        if (sum <= kMedium && sto) atomicOr(&status[s], sto);
        cnt[l] = sum > kMedium ? kMedium + 1 : sum;
    }
}

// Huge rows, step 1: one workgroup per (huge row, fan triangle) expands the
// triangle's split tree breadth first in depth-first order (ritem_count /
// ritem_write) while the frontier fits kLongFrontier items.  front holds two
// kLongFrontier buffers per workgroup; the final frontier ends in the first.
// A row whose expansion would overflow the reference's stack is flagged for
// the sequential path.
constexpr int kExpandThreads = 256;

__global__ void __launch_bounds__(kExpandThreads) k_long_expand(BatchGeom G, const SrcGrid *__restrict__ src,
                                                                 const double *__restrict__ lat,
                                                                 const int64_t *__restrict__ rows, int nhuge,
                                                                 RItem *__restrict__ front,
                                                                 int32_t *__restrict__ nfront,
                                                                 int *__restrict__ fallback) {
    __shared__ int cnt[kLongFrontier + 1];
    __shared__ int sh_n, sh_sel, sh_ovf, sh_sub, sh_stop;
    const int b = blockIdx.x;
    const int l = b >> 3, t = b & 7;
    if (l >= nhuge) return;
    const long r = rows[l];
    const int s = (int)(r / G.npix);
    const long p = r - (long)s * G.npix;
    const int i = (int)(p / G.ny), j = (int)(p - (long)i * G.ny);
    const SrcGrid g = src[s];
    RItem *bufA = front + (long)b * 2 * kLongFrontier;
    RItem *bufB = bufA + kLongFrontier;
    if (threadIdx.x == 0) {
        double tv[6];
        fan_vertices(lat + (long)s * G.lx * G.ly * 2, G.ly, i, j, t, tv);
        RItem it;
        raster_root(tv, g, it.tri);
        it.depth = 0;
        it.kind = 0;
        it.area = 0.0;
        it.idx = 0;
        bufA[0] = it;
        sh_n = 1;
        sh_sel = 0;
    }
    __syncthreads();
    for (int level = 0; level < kLongLevels; ++level) {
        const int n = sh_n;
        const RItem *cur = sh_sel ? bufB : bufA;
        RItem *nxt = sh_sel ? bufA : bufB;
        if (threadIdx.x == 0) {
            sh_ovf = 0;
            sh_sub = 0;
        }
        __syncthreads();
        for (int q = threadIdx.x; q < n; q += kExpandThreads) {
            int ovf = 0;
            const bool ex = ritem_splits(cur[q]);
            cnt[q] = ex ? ritem_count(cur[q], g, &ovf) : 1;
            if (ovf) atomicOr(&sh_ovf, 1);
            if (ex) atomicOr(&sh_sub, 1);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int acc = 0;
            for (int q = 0; q < n; ++q) {
                const int c = cnt[q];
                cnt[q] = acc;
                acc += c;
            }
            cnt[n] = acc;
            sh_stop = (sh_ovf || !sh_sub || acc > kLongFrontier) ? 1 : 0;
        }
        __syncthreads();
        if (sh_stop) break;
        for (int q = threadIdx.x; q < n; q += kExpandThreads) {
            if (ritem_splits(cur[q]))
                ritem_write(cur[q], g, nxt + cnt[q]);
            else
                nxt[cnt[q]] = cur[q];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            sh_n = cnt[n];
            sh_sel ^= 1;
        }
        __syncthreads();
    }
    const int n = sh_n;
    if (sh_sel)
        for (int q = threadIdx.x; q < n; q += kExpandThreads) bufA[q] = bufB[q];
    if (threadIdx.x == 0) {
        nfront[b] = n;
        if (sh_ovf) atomicOr(&fallback[l], 1);
    }
}

// Huge rows, step 2 (count): fragments of every frontier item, each subtree
// walked depth first from the stack depth the reference has at its root.
// A subtree that overflows that stack or hits the pop limit flags its row
// for the sequential path.
__global__ void __launch_bounds__(kRasterThreads) k_long_sub_count(const int64_t *__restrict__ rows, long npix,
                                                                   const SrcGrid *__restrict__ src,
                                                                   const RItem *__restrict__ front,
                                                                   const int32_t *__restrict__ nfront, int nb,
                                                                   const int32_t *__restrict__ order,
                                                                   int *__restrict__ fallback,
                                                                   int64_t *__restrict__ cnt) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nb * kLongFrontier) return;
    const long gid = order[t];  // items in decreasing size: similar walks share a wave
    const int b = (int)(gid / kLongFrontier), q = (int)(gid - (long)b * kLongFrontier);
    if (q >= nfront[b]) {
        cnt[gid] = 0;
        return;
    }
    const int l = b >> 3;
    const RItem &it = front[(long)b * 2 * kLongFrontier + q];
    if (it.kind == 1) {
        cnt[gid] = 1;
        return;
    }
    const SrcGrid g = src[(int)(rows[l] / npix)];
    CountSink cs;
    if (raster_dfs<true>(it.tri, it.depth, g, cs)) atomicOr(&fallback[l], 1);
    cnt[gid] = cs.n;
}

// walk-order keys of the frontier slots: larger source-cell extent first,
// empty slots last
__global__ void k_item_keys(long nitems, const RItem *__restrict__ front, const int32_t *__restrict__ nfront,
                            uint32_t *__restrict__ keys, int32_t *__restrict__ ids) {
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= nitems) return;
    const int b = (int)(gid / kLongFrontier), q = (int)(gid - (long)b * kLongFrontier);
    uint32_t k = 0xffffffffu;
    if (q < nfront[b]) {
        const RItem &it = front[(long)b * 2 * kLongFrontier + q];
        long ext = 0;
        if (it.kind == 0) ext = (long)(it.tri.hi[0] - it.tri.lo[0]) * (long)(it.tri.hi[1] - it.tri.lo[1]);
        if (ext < 0) ext = 0;
        if (ext > 0x7fffffffL) ext = 0x7fffffffL;
        k = 0x7fffffffu - (uint32_t)ext;
    }
    keys[gid] = k;
    ids[gid] = (int32_t)gid;
}

// fragment count of every huge row: from the frontier counts, or by the
// sequential walk for flagged rows
__global__ void k_long_rowtot(BatchGeom G, const SrcGrid *__restrict__ src, const double *__restrict__ lat,
                              const int64_t *__restrict__ rows, int nhuge, const int *__restrict__ fallback,
                              const int64_t *__restrict__ pos, int64_t *__restrict__ tot, int *__restrict__ status) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nhuge) return;
    const long r = rows[l];
    const int s = (int)(r / G.npix);
    long n;
    if (!fallback[l]) {
        n = pos[(long)(l + 1) * 8 * kLongFrontier] - pos[(long)l * 8 * kLongFrontier];
    } else {
        const long p = r - (long)s * G.npix;
        const int i = (int)(p / G.ny), j = (int)(p - (long)i * G.ny);
        CountSink cs;
        const int st = raster_pixel<false>(lat + (long)s * G.lx * G.ly * 2, G.ly, i, j, src[s], cs);
        if (st) atomicOr(&status[s], st);
        n = cs.n;
    }
    // the reference's fragment buffer holds kRasterFragCap entries and is
    // overrun beyond that (raster.c:204-212); every fragment is kept here
    if (n >= kRasterFragCap) atomicOr(&status[s], 2);  // LT_RASTER_FRAG_OVERFLOW
    tot[l] = n;
}

// Huge rows, step 3 (emit): every frontier item writes its fragments at its
// row's staging offset (foff, indexed by long-row number hmap[h]) + its
// emission position
__global__ void __launch_bounds__(kRasterThreads) k_long_sub_emit(const int64_t *__restrict__ rows,
                                                                  const int32_t *__restrict__ hmap, long npix,
                                                                  const SrcGrid *__restrict__ src,
                                                                  const RItem *__restrict__ front,
                                                                  const int32_t *__restrict__ nfront, int nb,
                                                                  const int32_t *__restrict__ order,
                                                                  const int *__restrict__ fallback,
                                                                  const int64_t *__restrict__ pos,
                                                                  const int64_t *__restrict__ foff,
                                                                  double *__restrict__ area,
                                                                  uint64_t *__restrict__ key) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nb * kLongFrontier) return;
    const long gid = order[t];
    const int b = (int)(gid / kLongFrontier), q = (int)(gid - (long)b * kLongFrontier);
    const int l = b >> 3;
    if (q >= nfront[b] || fallback[l]) return;
    const long n0 = pos[gid] - pos[(long)l * 8 * kLongFrontier];
    const RItem &it = front[(long)b * 2 * kLongFrontier + q];
    GlobalFrags gf{area, key, foff[hmap[l]], n0};
    if (it.kind == 1) {
        gf(it.area, it.idx);
        return;
    }
    const SrcGrid g = src[(int)(rows[l] / npix)];
    raster_dfs<true>(it.tri, it.depth, g, gf);
}

// Emission of the long rows flagged for it (medium rows, and huge rows the
// frontier path gave back): the 8 fan triangles of a row of at most kMedium
// fragments walked in parallel, each from its position in the row's
// sequence (the counts of the triangles before it); a huge row given back by
// the frontier path takes the one sequential walk (lane t = 0).
__global__ void __launch_bounds__(64) k_long_emit8(BatchGeom G, const SrcGrid *__restrict__ src,
                                                   const double *__restrict__ lat, const int64_t *__restrict__ rows,
                                                   const int64_t *__restrict__ foff, int nlong,
                                                   const int *__restrict__ seq, const int32_t *__restrict__ tcnt,
                                                   double *__restrict__ area, uint64_t *__restrict__ key) {
    const long gt = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long l = gt >> 3;
    const int t = (int)(gt & 7);
    if (l >= nlong || !seq[l]) return;
    int n0 = 0, sum = 0;
    for (int q = 0; q < 8; ++q) {
        const int c = tcnt[8 * l + q];
        n0 += q < t ? c : 0;
        sum += c;
    }
    const long r = rows[l];
    const int s = (int)(r / G.npix);
    const long p = r - (long)s * G.npix;
    const int i = (int)(p / G.ny), j = (int)(p - (long)i * G.ny);
    const double *ls = lat + (long)s * G.lx * G.ly * 2;
    if (sum > kMedium) {
        if (t == 0) {
            GlobalFrags gf{area, key, foff[l]};
            raster_pixel<true>(ls, G.ly, i, j, src[s], gf);
        }
        return;
    }
    double tv[6];
    fan_vertices(ls, G.ly, i, j, t, tv);
    GlobalFrags gf{area, key, foff[l], n0};
    raster_triangle<true>(tv, src[s], gf);
}

// Reduction of the sorted long rows (raster.c:375-391) in parallel over
// groups of equal source index: a group starts where the index changes; its
// value is the in-order sum of its fragments, without the first one unless
// the group opens the row (the reference's dropped fragment).
__device__ inline int row_of(const int64_t *__restrict__ foff, int nlong, long f) {
    int lo = 0, hi = nlong;  // foff[lo] <= f < foff[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (foff[mid] <= f) lo = mid;
        else hi = mid;
    }
    return lo;
}

// The long rows' fragments sorted in one radix sort instead of per row
// (a segmented sort gives each row one workgroup, and the lens-centre rows
// of 1e5 fragments then run alone): the key (source index << 32 | emission
// position) of row l becomes l << (bi + bn) | index << bn | position, which
// orders by row first, then as the per-row key; keys are unique, so the
// order is the segmented sort's.  *bad = 1 if a field does not fit.
__global__ void k_frag_pack(const int64_t *__restrict__ foff, int nlong, uint64_t *__restrict__ key, int bi, int bn,
                            int *__restrict__ bad) {
    const long F = foff[nlong];
    const long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    const int l = row_of(foff, nlong, f);
    const uint64_t k = key[f];
    const uint64_t i = k >> 32, n = k & 0xffffffffull;
    if ((i >> bi) != 0 || (n >> bn) != 0) *bad = 1;
    key[f] = ((uint64_t)l << (bi + bn)) | (i << bn) | n;
}

__global__ void k_frag_unpack(long F, uint64_t *__restrict__ key, int bi, int bn) {
    const long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    const uint64_t k = key[f];
    const uint64_t n = k & ((1ull << bn) - 1), i = (k >> bn) & ((1ull << bi) - 1);
    key[f] = (i << 32) | n;
}

__global__ void k_group_flags(const int64_t *__restrict__ foff, int nlong, const uint64_t *__restrict__ key,
                              int64_t *__restrict__ flag) {
    const long F = foff[nlong];
    const long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    const int l = row_of(foff, nlong, f);
    flag[f] = (f == foff[l] || (key[f] >> 32) != (key[f - 1] >> 32)) ? 1 : 0;
}

__global__ void k_group_sum(const int64_t *__restrict__ foff, int nlong, const uint64_t *__restrict__ key,
                            const double *__restrict__ area, const int64_t *__restrict__ gpos,
                            int32_t *__restrict__ scol, double *__restrict__ sval) {
    const long F = foff[nlong];
    const long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F || gpos[f + 1] == gpos[f]) return;  // not a group start
    const int l = row_of(foff, nlong, f);
    const long b = foff[l], e = foff[l + 1];
    const uint32_t idx = (uint32_t)(key[f] >> 32);
    double val = 0.0;
    for (long k = (f == b) ? f : f + 1; k < e && (uint32_t)(key[k] >> 32) == idx; ++k)
        val += area[b + (long)(uint32_t)key[k]];
    scol[gpos[f]] = (int32_t)idx;
    sval[gpos[f]] = val;
}

__global__ void k_group_rows(const int64_t *__restrict__ rows, const int64_t *__restrict__ foff, int nlong,
                             const uint64_t *__restrict__ key, const int64_t *__restrict__ gpos,
                             int64_t *__restrict__ nnz_row, int32_t *__restrict__ minidx,
                             int64_t *__restrict__ ebase) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nlong) return;
    const long r = rows[l];
    nnz_row[r] = gpos[foff[l + 1]] - gpos[foff[l]];
    minidx[r] = (int32_t)(key[foff[l]] >> 32);
    ebase[l] = gpos[foff[l]];
}

__global__ void __launch_bounds__(kRasterThreads) k_raster_fill(BatchGeom G, const SrcGrid *__restrict__ src,
                                                                 const int64_t *__restrict__ srcoff,
                                                                 const double *__restrict__ lat,
                                                                 const uint8_t *__restrict__ mask,
                                                                 const int64_t *__restrict__ rowoff,
                                                                 const int32_t *__restrict__ nfrag,
                                                                 const int32_t *__restrict__ minidx,
                                                                 const int32_t *__restrict__ lastne,
                                                                 const int32_t *__restrict__ stcol,
                                                                 const double *__restrict__ stval,
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
    const long o = rowoff[r];
    const int k = (int)(rowoff[r + 1] - o);
    if (nfrag[r] > 0 && k <= kStage) {  // entries kept by the count pass
        for (int e = 0; e < k; ++e) {
            cols[o + e] = (int32_t)srcoff[s] + stcol[r * kStage + e];
            vals[o + e] = stval[r * kStage + e];
        }
        return;
    }
    LocalFrags fr;
    if (nfrag[r] > 0) raster_pixel(lat + (long)s * G.lx * G.ly * 2, G.ly, i, j, src[s], fr);
    StoreEmit se{cols + o, vals + o, (int32_t)srcoff[s]};
    int first = 0;
    reduce_fragments(fr.area, fr.idx, fr.n, stale, se, &first);
}

__global__ void k_long_copy(const int64_t *__restrict__ rows, const int64_t *__restrict__ ebase, int nlong,
                            const int64_t *__restrict__ rowoff, const int64_t *__restrict__ nnz_row,
                            const int32_t *__restrict__ scol, const double *__restrict__ sval,
                            const int64_t *__restrict__ srcoff, long npix, int32_t *__restrict__ cols,
                            double *__restrict__ vals) {
    const int l = blockIdx.x;
    if (l >= nlong) return;
    const long r = rows[l];
    const int s = (int)(r / npix);
    const long b = ebase[l], n = nnz_row[r], o = rowoff[r];
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
    DevBuf<int32_t> stcol;
    DevBuf<double> stval;
    LT_TRY(stcol.alloc(R * kStage));
    LT_TRY(stval.alloc(R * kStage));
    hipLaunchKernelGGL(k_raster_count, dim3(blocks_for(R, kRasterThreads)), dim3(kRasterThreads), 0, stream(), G,
                       src.p, lattice.p, mask.p, nnz_row.p, nfrag.p, minidx.p, lastne.p, dstatus.p, stcol.p, stval.p);
    LT_CHECK_LAUNCH();
    // long rows: compacted on the device; their fragments are produced by the
    // frontier decomposition (k_long_expand / k_long_sub_*), sorted by
    // (source index, emission position) and reduced group-parallel
    DevBuf<int64_t> lflag, lpos;
    LT_TRY(lflag.alloc(R + 1));
    LT_TRY(lpos.alloc(R + 1));
    hipLaunchKernelGGL(k_long_flags, dim3(blocks_for(R, 256)), dim3(256), 0, stream(), R, nfrag.p, lflag.p);
    LT_CHECK_LAUNCH();
    LT_TRY(exclusive_scan_i64(lflag.p, lpos.p, R));
    int64_t nl64 = 0;
    LT_TRY(lpos.download(&nl64, 1, R));
    LT_TRY(sync());
    const int nlong = (int)nl64;
    DevBuf<int64_t> drows, dfoff, ebase;
    DevBuf<double> sval;
    DevBuf<int32_t> scol;
    if (nlong) {
        LT_TRY(drows.alloc(nlong));
        {
            DevBuf<int32_t> ln;
            LT_TRY(ln.alloc(nlong));
            hipLaunchKernelGGL(k_long_compact, dim3(blocks_for(R, 256)), dim3(256), 0, stream(), R, lpos.p, nfrag.p,
                               drows.p, ln.p);
            LT_CHECK_LAUNCH();
        }
        // exact counts of medium rows; rows beyond kMedium are huge
        DevBuf<int64_t> lcnt;
        DevBuf<int32_t> tcnt;  // per fan triangle
        LT_TRY(lcnt.alloc(nlong));
        LT_TRY(tcnt.alloc(8L * nlong));
        hipLaunchKernelGGL(k_long_count8, dim3(blocks_for(8L * nlong, 64)), dim3(64), 0, stream(), G, src.p,
                           lattice.p, drows.p, nlong, tcnt.p, lcnt.p, dstatus.p);
        LT_CHECK_LAUNCH();
        std::vector<int64_t> hcnt(nlong), hrows_all(nlong);
        LT_TRY(lcnt.download(hcnt.data(), nlong));
        LT_TRY(drows.download(hrows_all.data(), nlong));
        LT_TRY(sync());
        std::vector<int32_t> hmap;
        std::vector<int64_t> hrow;
        for (int l = 0; l < nlong; ++l)
            if (hcnt[l] > kMedium) {
                hmap.push_back(l);
                hrow.push_back(hrows_all[l]);
            }
        const int nhuge = (int)hmap.size();
        std::vector<int> seq(nlong, 1);
        std::vector<int64_t> loff(nlong + 1, 0);
        DevBuf<int32_t> dhmap, nfront;
        DevBuf<int64_t> dhrow, cnt, pos, tot;
        DevBuf<RItem> front;
        DevBuf<int32_t> order;  // walk order of the frontier slots
        DevBuf<int> hfb;
        const int nb = nhuge * 8;
        const long nitems = (long)nb * kLongFrontier;
        if (nhuge) {
            LT_TRY(dhmap.upload(hmap.data(), nhuge));
            LT_TRY(dhrow.upload(hrow.data(), nhuge));
            LT_TRY(front.alloc(2 * nitems));
            LT_TRY(nfront.alloc(nb));
            LT_TRY(hfb.alloc(nhuge));
            LT_TRY(hfb.zero());
            LT_TRY(cnt.alloc(nitems));
            LT_TRY(pos.alloc(nitems + 1));
            LT_TRY(tot.alloc(nhuge));
            hipLaunchKernelGGL(k_long_expand, dim3(nb), dim3(kExpandThreads), 0, stream(), G, src.p, lattice.p,
                               dhrow.p, nhuge, front.p, nfront.p, hfb.p);
            LT_CHECK_LAUNCH();
            {  // walk order: items by decreasing extent (a permutation of the slots)
                DevBuf<uint32_t> k1, k2;
                DevBuf<int32_t> i1;
                LT_TRY(k1.alloc(nitems));
                LT_TRY(k2.alloc(nitems));
                LT_TRY(i1.alloc(nitems));
                LT_TRY(order.alloc(nitems));
                hipLaunchKernelGGL(k_item_keys, dim3(blocks_for(nitems, 256)), dim3(256), 0, stream(), nitems,
                                   front.p, nfront.p, k1.p, i1.p);
                LT_CHECK_LAUNCH();
                size_t tb = 0;
                LT_HIP(rocprim::radix_sort_pairs(nullptr, tb, k1.p, k2.p, i1.p, order.p, (size_t)nitems, 0, 32,
                                                 stream()));
                DevBuf<uint8_t> tmp;
                LT_TRY(tmp.alloc(std::max<size_t>(tb, 1)));
                LT_HIP(rocprim::radix_sort_pairs(tmp.p, tb, k1.p, k2.p, i1.p, order.p, (size_t)nitems, 0, 32,
                                                 stream()));
                LT_TRY(sync());
            }
            hipLaunchKernelGGL(k_long_sub_count, dim3(blocks_for(nitems, kRasterThreads)), dim3(kRasterThreads), 0,
                               stream(), dhrow.p,
                               npix, src.p, front.p, nfront.p, nb, order.p, hfb.p, cnt.p);
            LT_CHECK_LAUNCH();
            LT_TRY(exclusive_scan_i64(cnt.p, pos.p, nitems));
            hipLaunchKernelGGL(k_long_rowtot, dim3(blocks_for(nhuge, 64)), dim3(64), 0, stream(), G, src.p,
                               lattice.p, dhrow.p, nhuge, hfb.p, pos.p, tot.p, dstatus.p);
            LT_CHECK_LAUNCH();
            std::vector<int64_t> htot(nhuge);
            std::vector<int> hf(nhuge);
            LT_TRY(tot.download(htot.data(), nhuge));
            LT_TRY(hfb.download(hf.data(), nhuge));
            LT_TRY(sync());
            for (int h = 0; h < nhuge; ++h) {
                hcnt[hmap[h]] = htot[h];
                seq[hmap[h]] = hf[h] ? 1 : 0;
            }
        }
        for (int l = 0; l < nlong; ++l) loff[l + 1] = loff[l] + hcnt[l];
        const long F = loff.back();
        if (F >= (long)UINT32_MAX) {
            set_error("lens batch: %ld long-row fragments exceed the sort range", F);
            return LT_ERR_CAPACITY;
        }
        DevBuf<int> dseq;
        LT_TRY(dseq.upload(seq.data(), nlong));
        LT_TRY(dfoff.upload(loff.data(), nlong + 1));
        DevBuf<double> farea;
        DevBuf<uint64_t> fkey, fkey2;
        LT_TRY(farea.alloc(F));
        LT_TRY(fkey.alloc(F));
        LT_TRY(fkey2.alloc(F));
        if (nhuge) {
            hipLaunchKernelGGL(k_long_sub_emit, dim3(blocks_for(nitems, kRasterThreads)), dim3(kRasterThreads), 0,
                               stream(), dhrow.p,
                               dhmap.p, npix, src.p, front.p, nfront.p, nb, order.p, hfb.p, pos.p, dfoff.p, farea.p,
                               fkey.p);
            LT_CHECK_LAUNCH();
        }
        hipLaunchKernelGGL(k_long_emit8, dim3(blocks_for(8L * nlong, 64)), dim3(64), 0, stream(), G, src.p,
                           lattice.p, drows.p, dfoff.p, nlong, dseq.p, tcnt.p, farea.p, fkey.p);
        LT_CHECK_LAUNCH();
        {
            auto bits = [](uint64_t v) {  // bits holding 0 .. v - 1
                int b = 0;
                while (b < 64 && (v - 1) >> b) ++b;
                return std::max(b, 1);
            };
            int64_t nsrc_max = 1, cnt_max = 1;
            for (int q = 0; q < ns; ++q) nsrc_max = std::max(nsrc_max, srcoff_host[q + 1] - srcoff_host[q]);
            for (int l = 0; l < nlong; ++l) cnt_max = std::max(cnt_max, hcnt[l]);
            const int bl = bits((uint64_t)nlong), bi = bits((uint64_t)nsrc_max), bn = bits((uint64_t)cnt_max);
            bool packed = bl + bi + bn <= 64 && F > 0;
            if (packed) {
                DevBuf<int> bad;
                LT_TRY(bad.alloc(1));
                LT_TRY(bad.zero());
                hipLaunchKernelGGL(k_frag_pack, dim3(blocks_for(F, 256)), dim3(256), 0, stream(), dfoff.p, nlong,
                                   fkey.p, bi, bn, bad.p);
                LT_CHECK_LAUNCH();
                int hb = 1;
                LT_TRY(bad.download(&hb, 1));
                LT_TRY(sync());
                if (hb) {  // an index outside its sample's grid: unpack and sort per row
                    hipLaunchKernelGGL(k_frag_unpack, dim3(blocks_for(F, 256)), dim3(256), 0, stream(), F, fkey.p,
                                       bi, bn);
                    LT_CHECK_LAUNCH();
                    packed = false;
                }
            }
            size_t tb = 0;
            if (packed) {
                LT_HIP(rocprim::radix_sort_keys(nullptr, tb, fkey.p, fkey2.p, (size_t)F, 0, bl + bi + bn, stream()));
            } else if (rocprim::segmented_radix_sort_keys(nullptr, tb, fkey.p, fkey2.p, (unsigned)F, (unsigned)nlong,
                                                          dfoff.p, dfoff.p + 1, 0, 64, stream()) != hipSuccess) {
                set_error("segmented sort sizing failed");
                return LT_ERR_HIP;
            }
            DevBuf<uint8_t> tmp;
            LT_TRY(tmp.alloc(std::max<size_t>(tb, 1)));
            if (packed) {
                LT_HIP(rocprim::radix_sort_keys(tmp.p, tb, fkey.p, fkey2.p, (size_t)F, 0, bl + bi + bn, stream()));
                hipLaunchKernelGGL(k_frag_unpack, dim3(blocks_for(F, 256)), dim3(256), 0, stream(), F, fkey2.p, bi,
                                   bn);
                LT_CHECK_LAUNCH();
            } else {
                LT_HIP(rocprim::segmented_radix_sort_keys(tmp.p, tb, fkey.p, fkey2.p, (unsigned)F, (unsigned)nlong,
                                                          dfoff.p, dfoff.p + 1, 0, 64, stream()));
            }
        }
        DevBuf<int64_t> gflag, gpos;
        LT_TRY(gflag.alloc(F));
        LT_TRY(gpos.alloc(F + 1));
        hipLaunchKernelGGL(k_group_flags, dim3(blocks_for(F, 256)), dim3(256), 0, stream(), dfoff.p, nlong, fkey2.p,
                           gflag.p);
        LT_CHECK_LAUNCH();
        LT_TRY(exclusive_scan_i64(gflag.p, gpos.p, F));
        int64_t ngroups = 0;
        LT_TRY(gpos.download(&ngroups, 1, F));
        LT_TRY(sync());
        LT_TRY(scol.alloc(std::max<int64_t>(ngroups, 1)));
        LT_TRY(sval.alloc(std::max<int64_t>(ngroups, 1)));
        LT_TRY(ebase.alloc(nlong));
        hipLaunchKernelGGL(k_group_sum, dim3(blocks_for(F, 256)), dim3(256), 0, stream(), dfoff.p, nlong, fkey2.p,
                           farea.p, gpos.p, scol.p, sval.p);
        LT_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_group_rows, dim3(blocks_for(nlong, 64)), dim3(64), 0, stream(), drows.p, dfoff.p, nlong,
                           fkey2.p, gpos.p, nnz_row.p, minidx.p, ebase.p);
        LT_CHECK_LAUNCH();
        LT_TRY(sync());
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
                       src.p, srcoff.p, lattice.p, mask.p, rowoff.p, nfrag.p, minidx.p, lastne.p, stcol.p, stval.p,
                       L.ci.p, L.va.p);
    LT_CHECK_LAUNCH();
    if (nlong) {
        hipLaunchKernelGGL(k_long_copy, dim3(nlong), dim3(256), 0, stream(), drows.p, ebase.p, nlong, rowoff.p,
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
