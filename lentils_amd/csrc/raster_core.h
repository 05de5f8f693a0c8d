// Exact triangle rasterization of one image pixel into source-plane
// fragments, written once for device and host.
//
// Follows lentils/backend/raster.c:120-269 (rasterize_triangle) and
// :319-363 (9-point stencil, 8 fan triangles): a triangle is split
// depth-first along source grid lines until each piece covers one source
// pixel; each such piece emits {area, source index}.  The emission order is
// part of the result (the reduction drops the first fragment of each column
// group, raster.c:375-391), so the split tables and the LIFO order are kept.
// Compile with -ffp-contract=off -fwrapv.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace lt {

#define LT_HD __host__ __device__ inline

constexpr int kRasterStack = 64;        // raster.c:11
constexpr int kRasterFragCap = 128000;  // raster.c:12 (reported, not enforced)
constexpr int kRasterPopLimit = 1 << 20;

struct SrcGrid {
    int nx, ny;
    double xmin, ymin, dx, dy;
};

// (int) conversion with x86-64 cvttsd2si semantics (NaN / out of range ->
// INT_MIN), which is what the reference's int = floor(...) executes.
LT_HD int f2i_x86(double v) {
    if (!(v >= -2147483648.0 && v < 2147483648.0)) return INT32_MIN;
    return (int)v;
}

struct RTri {
    double v[3][2];
    double b[3][2];
    int lo[2], hi[2];
};

// cut tables, raster.c:135-142 (indexed by the mask of vertices on the +
// side of the cut line)
struct CutTables {
    int ncut[8];
    int edge[8][2][2];
    int nchild[8];
    int vert[8][3][3];
    int side[8][3];
};

#define LT_CUT_TABLES_INIT                                                                     \
    {                                                                                          \
        {0, 2, 2, 2, 2, 2, 2, 0},                                                              \
        {{{0, 0}, {0, 0}}, {{0, 1}, {0, 2}}, {{0, 1}, {1, 2}}, {{0, 2}, {1, 2}},              \
         {{0, 2}, {1, 2}}, {{0, 1}, {1, 2}}, {{0, 1}, {0, 2}}, {{0, 0}, {0, 0}}},             \
        {1, 3, 3, 3, 3, 3, 3, 1},                                                              \
        {{{0, 1, 2}, {0, 0, 0}, {0, 0, 0}}, {{1, 2, 4}, {1, 4, 3}, {0, 3, 4}},                \
         {{0, 3, 4}, {0, 4, 2}, {1, 4, 3}}, {{2, 3, 4}, {0, 4, 3}, {0, 1, 4}},                \
         {{0, 1, 4}, {0, 4, 3}, {2, 3, 4}}, {{1, 4, 3}, {0, 3, 4}, {0, 4, 2}},                \
         {{0, 3, 4}, {1, 2, 4}, {1, 4, 3}}, {{0, 1, 2}, {0, 0, 0}, {0, 0, 0}}},               \
        {{0, 0, 0}, {0, 0, 1}, {0, 0, 1}, {0, 1, 1}, {0, 0, 1}, {0, 1, 1}, {0, 1, 1}, {1, 0, 0}} \
    }

static __constant__ CutTables kCutTablesDev = LT_CUT_TABLES_INIT;
static const CutTables kCutTablesHost = LT_CUT_TABLES_INIT;

__host__ __device__ inline const CutTables &cut_tables() {
#ifdef __HIP_DEVICE_COMPILE__
    return kCutTablesDev;
#else
    return kCutTablesHost;
#endif
}

// The tables packed, per fixed (cut, child, vertex) position, into one
// integer over the 8 masks, so that a lookup by a per-thread mask is a shift
// of an immediate (no memory access).
constexpr CutTables kCutTablesCx = LT_CUT_TABLES_INIT;
constexpr unsigned cut_pack_edge(int k, int j) {
    unsigned r = 0;
    for (int m = 0; m < 8; ++m) r |= (unsigned)kCutTablesCx.edge[m][k][j] << (2 * m);
    return r;
}
constexpr unsigned cut_pack_vert(int c, int v) {
    unsigned r = 0;
    for (int m = 0; m < 8; ++m) r |= (unsigned)kCutTablesCx.vert[m][c][v] << (3 * m);
    return r;
}
constexpr unsigned cut_pack_side(int c) {
    unsigned r = 0;
    for (int m = 0; m < 8; ++m) r |= (unsigned)kCutTablesCx.side[m][c] << m;
    return r;
}
constexpr unsigned cut_pack_ncut() {
    unsigned r = 0;
    for (int m = 0; m < 8; ++m) r |= (unsigned)kCutTablesCx.ncut[m] << (2 * m);
    return r;
}
constexpr unsigned cut_pack_nchild() {
    unsigned r = 0;
    for (int m = 0; m < 8; ++m) r |= (unsigned)kCutTablesCx.nchild[m] << (2 * m);
    return r;
}

// a[s] for s in 0..N-1 without indexing the array by a per-thread value
template <int N>
LT_HD double pick(const double (&a)[N], int s) {
    double r = a[0];
#pragma unroll
    for (int i = 1; i < N; ++i) r = s == i ? a[i] : r;
    return r;
}

// Root node of one fan triangle (raster.c:144-170): its three source-plane
// vertices, barycentric corners (0,0), (0,1), (1,0) and the index range of
// the source cells it touches.
LT_HD void raster_root(const double *tv, const SrcGrid &g, RTri &tri) {
    tri.lo[0] = g.nx;
    tri.lo[1] = g.ny;
    tri.hi[0] = 0;
    tri.hi[1] = 0;
    tri.b[0][0] = 0.0; tri.b[0][1] = 0.0;
    tri.b[1][0] = 0.0; tri.b[1][1] = 1.0;
    tri.b[2][0] = 1.0; tri.b[2][1] = 0.0;
    for (int v = 0; v < 3; ++v) {
        tri.v[v][0] = tv[2 * v];
        tri.v[v][1] = tv[2 * v + 1];
        int i = f2i_x86(floor((tri.v[v][0] - g.xmin) / g.dx));
        int j = f2i_x86(floor((tri.v[v][1] - g.ymin) / g.dy));
        if (i < tri.lo[0]) tri.lo[0] = i;
        if (j < tri.lo[1]) tri.lo[1] = j;
        if (i + 1 > tri.hi[0]) tri.hi[0] = i + 1;
        if (j + 1 > tri.hi[1]) tri.hi[1] = j + 1;
    }
}

// What the reference does with a node it pops (raster.c:180-268).  A node of
// extent 1 is a leaf: returns -1 and sets emit (inside the grid) with the
// fragment's area and source index.  Otherwise the node is cut at the middle
// grid line of its larger extent (ties -> x); the pieces that overlap the
// grid, in the order the reference pushes them, are the kids[c] with ok[c]
// set; returns their number (the reference processes them last to first).
// Every table lookup and intermediate point stays in registers (lookups are
// shifts of packed immediates, the cut points are picked by selects): same
// arithmetic as raster_expand below, fewer memory round trips (for walks of
// many short rows side by side) at the price of more instructions and
// registers (raster_expand serves the long serial walks).
LT_HD int expand_node(const RTri &tri, const SrcGrid &g, RTri (&kids)[3], bool (&ok)[3], double &area, int &idx,
                      bool &emit) {
    int ax = 0, ext = 0;
    for (int a = 0; a < 2; ++a) {
        int s = tri.hi[a] - tri.lo[a];
        if (s > ext) {
            ext = s;
            ax = a;
        }
    }
    const int lo_ax = ax ? tri.lo[1] : tri.lo[0], hi_ax = ax ? tri.hi[1] : tri.hi[0];
    const int cut = lo_ax + ext / 2;
    const double cell = ax ? g.dy : g.dx;
    const double org = ax ? g.ymin : g.xmin;
    const double cutc = cut * cell + org;
    ok[0] = ok[1] = ok[2] = false;
    if (ext == 1) {
        const int i = tri.lo[0], j = tri.lo[1];
        emit = !(i < 0 || i >= g.nx || j < 0 || j >= g.ny);
        if (emit) {
            const double x1 = tri.b[0][0], y1 = tri.b[0][1];
            const double x2 = tri.b[1][0], y2 = tri.b[1][1];
            const double x3 = tri.b[2][0], y3 = tri.b[2][1];
            area = -0.25 * (x1 * (y2 - y3) + x2 * (y3 - y1) + x3 * (y1 - y2));
            idx = g.ny * i + j;
        }
        return -1;
    }
    double px[5], py[5], bx[5], by[5], dist[3];
    int mask = 0;
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        px[v] = tri.v[v][0];
        py[v] = tri.v[v][1];
        bx[v] = tri.b[v][0];
        by[v] = tri.b[v][1];
        dist[v] = (ax ? py[v] : px[v]) - cutc;
        mask |= ((dist[v] > 0.) << v);
    }
    px[3] = px[4] = py[3] = py[4] = bx[3] = bx[4] = by[3] = by[4] = 0.0;
    const int ncut = (int)((cut_pack_ncut() >> (2 * mask)) & 3u);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        if (k < ncut) {
            const int e0 = (int)((cut_pack_edge(k, 0) >> (2 * mask)) & 3u);
            const int e1 = (int)((cut_pack_edge(k, 1) >> (2 * mask)) & 3u);
            const double w0 = pick(dist, e1), w1 = -pick(dist, e0);
            const double wn = 1.0 / (w0 + w1);
            const double a0 = w0 * wn, a1 = w1 * wn;
            px[3 + k] = a0 * pick(px, e0) + a1 * pick(px, e1);
            py[3 + k] = a0 * pick(py, e0) + a1 * pick(py, e1);
            bx[3 + k] = a0 * pick(bx, e0) + a1 * pick(bx, e1);
            by[3 + k] = a0 * pick(by, e0) + a1 * pick(by, e1);
        }
    }
    const int nchild = (int)((cut_pack_nchild() >> (2 * mask)) & 3u);
    int m = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (c >= nchild) continue;
        const int s = (int)((cut_pack_side(c) >> mask) & 1u);
        int lo0 = tri.lo[0], lo1 = tri.lo[1], hi0 = tri.hi[0], hi1 = tri.hi[1];
        const int lo_c = s ? cut : lo_ax, hi_c = s ? hi_ax : cut;
        if (ax) {
            lo1 = lo_c;
            hi1 = hi_c;
        } else {
            lo0 = lo_c;
            hi0 = hi_c;
        }
        if (hi0 <= 0 || lo0 >= g.nx) continue;
        if (hi1 <= 0 || lo1 >= g.ny) continue;
        RTri &k = kids[c];
        k.lo[0] = lo0;
        k.lo[1] = lo1;
        k.hi[0] = hi0;
        k.hi[1] = hi1;
#pragma unroll
        for (int v = 0; v < 3; ++v) {
            const int src = (int)((cut_pack_vert(c, v) >> (3 * mask)) & 7u);
            k.v[v][0] = pick(px, src);
            k.v[v][1] = pick(py, src);
            k.b[v][0] = pick(bx, src);
            k.b[v][1] = pick(by, src);
        }
        ok[c] = true;
        ++m;
    }
    return m;
}

// What the reference does with a node it pops, with the tables in constant
// memory and the points in a local array, writing the children straight to
// kids[0..m-1] (onto the walk's stack in raster_dfs<false>).
LT_HD int raster_expand(const RTri &tri, const SrcGrid &g, RTri *kids, double &area, int &idx, bool &emit) {
    const CutTables &T = cut_tables();
    int ax = 0, ext = 0;
    for (int a = 0; a < 2; ++a) {
        int s = tri.hi[a] - tri.lo[a];
        if (s > ext) {
            ext = s;
            ax = a;
        }
    }
    const int cut = tri.lo[ax] + ext / 2;
    const double cell = ax ? g.dy : g.dx;
    const double org = ax ? g.ymin : g.xmin;
    const double cutc = cut * cell + org;
    if (ext == 1) {
        const int i = tri.lo[0], j = tri.lo[1];
        emit = !(i < 0 || i >= g.nx || j < 0 || j >= g.ny);
        if (emit) {
            const double x1 = tri.b[0][0], y1 = tri.b[0][1];
            const double x2 = tri.b[1][0], y2 = tri.b[1][1];
            const double x3 = tri.b[2][0], y3 = tri.b[2][1];
            area = -0.25 * (x1 * (y2 - y3) + x2 * (y3 - y1) + x3 * (y1 - y2));
            idx = g.ny * i + j;
        }
        return -1;
    }
    double pv[5][2], pb[5][2], dist[3];
    int mask = 0;
    for (int v = 0; v < 3; ++v) {
        pv[v][0] = tri.v[v][0];
        pv[v][1] = tri.v[v][1];
        pb[v][0] = tri.b[v][0];
        pb[v][1] = tri.b[v][1];
        dist[v] = pv[v][ax] - cutc;
        mask |= ((dist[v] > 0.) << v);
    }
    const int sidx[3] = {tri.lo[ax], cut, tri.hi[ax]};
    for (int k = 0; k < T.ncut[mask]; ++k) {
        const int e0 = T.edge[mask][k][0], e1 = T.edge[mask][k][1];
        const double w0 = dist[e1], w1 = -dist[e0];
        const double wn = 1.0 / (w0 + w1);
        const double a0 = w0 * wn, a1 = w1 * wn;
        pv[3 + k][0] = a0 * pv[e0][0] + a1 * pv[e1][0];
        pv[3 + k][1] = a0 * pv[e0][1] + a1 * pv[e1][1];
        pb[3 + k][0] = a0 * pb[e0][0] + a1 * pb[e1][0];
        pb[3 + k][1] = a0 * pb[e0][1] + a1 * pb[e1][1];
    }
    int m = 0;
    for (int c = 0; c < T.nchild[mask]; ++c) {
        const int s = T.side[mask][c];
        int lo[2] = {tri.lo[0], tri.lo[1]}, hi[2] = {tri.hi[0], tri.hi[1]};
        lo[ax] = sidx[s];
        hi[ax] = sidx[s + 1];
        if (hi[0] <= 0 || lo[0] >= g.nx) continue;
        if (hi[1] <= 0 || lo[1] >= g.ny) continue;
        RTri &k = kids[m++];
        k.lo[0] = lo[0];
        k.lo[1] = lo[1];
        k.hi[0] = hi[0];
        k.hi[1] = hi[1];
        for (int v = 0; v < 3; ++v) {
            const int src = T.vert[mask][c][v];
            k.v[v][0] = pv[src][0];
            k.v[v][1] = pv[src][1];
            k.b[v][0] = pb[src][0];
            k.b[v][1] = pb[src][1];
        }
    }
    return m;
}

// Depth-first rasterization from node `root`, which sits on the reference's
// 64-entry stack above `base` other entries (0 for a whole triangle).
// sink(area, idx) is called per fragment in emission order; a sink whose
// stop() turns true ends the walk early (count passes that only need to know
// that a row is long).  Returns LT_RASTER_* status bits: the reference gives
// up on the triangle when a push fills its stack (raster.c:262-266).
// REG (many short walks side by side): expand_node in registers and the
// last child kept as the current node instead of pushed and popped; else
// (long serial walks, where registers cost occupancy) raster_expand onto the
// stack.  Same nodes in the same order either way.
template <bool REG = true, class Sink>
__host__ __device__ int raster_dfs(const RTri &root, int base, const SrcGrid &g, Sink &sink) {
    if constexpr (!REG) {
        RTri stk[kRasterStack + 3];
        int ns = 0, pops = 0;
        stk[ns++] = root;
        while (ns > 0) {
            if (++pops > kRasterPopLimit) return 4;  // LT_RASTER_POP_LIMIT
            const RTri tri = stk[--ns];
            double area = 0.0;
            int idx = 0;
            bool emit = false;
            const int m = raster_expand(tri, g, &stk[ns], area, idx, emit);
            if (m < 0) {
                if (emit) {
                    sink(area, idx);
                    if (sink.stop()) return 0;
                }
                continue;
            }
            // the reference pushes one piece at a time and stops as soon as its
            // stack holds kRasterStack entries
            if (m > 0 && ns + base + m >= kRasterStack) return 1;  // LT_RASTER_STACK_OVERFLOW
            ns += m;
        }
        return 0;
    } else {
        // the reference pushes every child and pops the last one next; here the
        // last child is kept as the current node and only the others go to the
        // stack (same order, same stack depths)
        // The stack's top entry is kept in registers (top, valid when
        // has_top): a leaf's pop then takes the entry pushed last without a
        // round trip through memory; ns counts every entry, nm those in stk.
        RTri stk[kRasterStack + 3];
        RTri top;
        bool has_top = false;
        int ns = 0, nm = 0, pops = 0;
        RTri cur = root;
        for (;;) {
            if (++pops > kRasterPopLimit) return 4;  // LT_RASTER_POP_LIMIT
            RTri kids[3];
            bool ok[3];
            double area = 0.0;
            int idx = 0;
            bool emit = false;
            const int m = expand_node(cur, g, kids, ok, area, idx, emit);
            if (m < 0) {
                if (emit) {
                    sink(area, idx);
                    if (sink.stop()) return 0;
                }
            } else if (m > 0) {
                // the reference pushes one piece at a time and stops as soon as
                // its stack holds kRasterStack entries
                if (ns + base + m >= kRasterStack) return 1;  // LT_RASTER_STACK_OVERFLOW
                int left = m;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    if (!ok[c]) continue;
                    if (--left == 0) {
                        cur = kids[c];
                    } else {
                        if (has_top) stk[nm++] = top;
                        top = kids[c];
                        has_top = true;
                        ++ns;
                    }
                }
                continue;
            }
            if (ns == 0) break;
            --ns;
            if (has_top) {
                cur = top;
                has_top = false;
            } else {
                cur = stk[--nm];
            }
        }
        return 0;
    }
}

// Rasterize one triangle (vertices tv[0..5]).
template <bool REG = true, class Sink>
__host__ __device__ int raster_triangle(const double *tv, const SrcGrid &g, Sink &sink) {
    RTri tri;
    raster_root(tv, g, tri);
    return raster_dfs<REG>(tri, 0, g, sink);
}

// ---------------------------------------------------------------------------
// Rows with very many fragments (the pixels around the lens centre fan over
// much of the source plane) are rasterized in parallel without changing the
// emission order: the split tree of each fan triangle is expanded breadth
// first, keeping every level in the reference's depth-first order (children
// last-pushed first, fragments in place) and splitting only the nodes that
// still span more than kLongSplitCells source cells, until none does or the
// frontier would exceed kLongFrontier items; each remaining subtree is then
// walked depth first by its own thread, starting with the stack depth the
// reference would have at that node.  Concatenating the frontier's outputs
// in order reproduces the reference's fragment sequence exactly (any cut of
// the split tree listed in depth-first order does).
constexpr int kLongFrontier = 4096;
constexpr int kLongLevels = 64;
constexpr long kLongSplitCells = 32;

struct RItem {
    RTri tri;
    int depth;  // entries below this node on the reference's stack
    int kind;   // 0: subtree to expand, 1: fragment
    double area;
    int idx;
};

// whether the breadth-first expansion splits this frontier item further
LT_HD bool ritem_splits(const RItem &it) {
    return it.kind == 0 && (long)(it.tri.hi[0] - it.tri.lo[0]) * (long)(it.tri.hi[1] - it.tri.lo[1]) > kLongSplitCells;
}

// number of frontier items `it` turns into at the next level; sets *ovf
// when expanding it would overflow the reference's stack
LT_HD int ritem_count(const RItem &it, const SrcGrid &g, int *ovf) {
    if (it.kind == 1) return 1;
    RTri kids[3];
    double area = 0.0;
    int idx = 0;
    bool emit = false;
    const int m = raster_expand(it.tri, g, kids, area, idx, emit);
    if (m < 0) return emit ? 1 : 0;
    if (m > 0 && it.depth + m >= kRasterStack) *ovf = 1;
    return m;
}

// writes the items `it` turns into at out[0..], in depth-first order
LT_HD void ritem_write(const RItem &it, const SrcGrid &g, RItem *out) {
    if (it.kind == 1) {
        out[0] = it;
        return;
    }
    RTri kids[3];
    double area = 0.0;
    int idx = 0;
    bool emit = false;
    const int m = raster_expand(it.tri, g, kids, area, idx, emit);
    if (m < 0) {
        if (emit) {
            out[0].kind = 1;
            out[0].depth = it.depth;
            out[0].area = area;
            out[0].idx = idx;
        }
        return;
    }
    for (int k = 0; k < m; ++k) {  // pushed k-th, processed (m-1-k)-th
        RItem &o = out[m - 1 - k];
        o.tri = kids[k];
        o.depth = it.depth + k;
        o.kind = 0;
        o.area = 0.0;
        o.idx = 0;
    }
}

// lattice offsets of the 9 stencil points of a pixel: corners, edge
// midpoints, centre, in the order raster.c:319-347 lists them
LT_HD void stencil_offsets(int k, int &ox, int &oy) {
    const int dxo[9] = {0, 1, 2, 2, 2, 1, 0, 0, 1};
    const int dyo[9] = {0, 0, 0, 1, 2, 2, 2, 1, 1};
    ox = dxo[k];
    oy = dyo[k];
}

// All fragments of image pixel (i, j) from the deflected lattice (x-major,
// ly = 2 ny + 1 points per column): 8 fan triangles
// (p[t+1], p[t], p[8]).  raster.c:351-363
template <bool REG = true, class Sink>
__host__ __device__ int raster_pixel(const double *lat, int ly, int i, int j, const SrcGrid &g, Sink &sink) {
    double sp[9][2];
    for (int k = 0; k < 9; ++k) {
        int ox, oy;
        stencil_offsets(k, ox, oy);
        const long a = 2L * i + ox, b = 2L * j + oy;
        sp[k][0] = lat[2 * (a * ly + b)];
        sp[k][1] = lat[2 * (a * ly + b) + 1];
    }
    int st = 0;
    for (int t = 0; t < 8 && !sink.stop(); ++t) {
        const int t1 = (t + 1) & 7;
        const double tv[6] = {sp[t1][0], sp[t1][1], sp[t][0], sp[t][1], sp[8][0], sp[8][1]};
        st |= raster_triangle<REG>(tv, g, sink);
    }
    return st;
}

// vertices of fan triangle t of pixel (i, j), as raster_pixel forms them
LT_HD void fan_vertices(const double *lat, int ly, int i, int j, int t, double tv[6]) {
    const int t1 = (t + 1) & 7;
    const int pts[3] = {t1, t, 8};
    for (int v = 0; v < 3; ++v) {
        int ox, oy;
        stencil_offsets(pts[v], ox, oy);
        const long a = 2L * i + ox, b = 2L * j + oy;
        tv[2 * v] = lat[2 * (a * ly + b)];
        tv[2 * v + 1] = lat[2 * (a * ly + b) + 1];
    }
}

// The reference's reduction of a row's fragments (raster.c:372-391): a
// stable sort by source index, then one CSR entry per distinct index in
// ascending order, its value the in-order sum of the group's areas -- all
// of them for the first group, all but the first for the others (the
// reference starts a new group without adding its first area).  A row
// without fragments gives one entry (stale, 0).  Calls emit(k, col, val)
// per entry and returns the entry count; *first_col is the first entry's
// column.  Rows of up to 64 fragments take one pass over the list per
// entry (the next larger index found while summing the current one), so
// the lists are only read -- in blocks of 8, so area and idx must then hold
// 64 entries whatever n is; longer ones are sorted in place first.
template <class Emit>
__host__ __device__ int reduce_fragments(double *area, int *idx, int n, int stale, Emit &emit, int *first_col) {
    if (n == 0) {
        emit(0, stale, 0.0);
        *first_col = stale;
        return 1;
    }
    if (n <= 64) {
        // the lists hold 64 entries (the short rows' buffers): each pass reads
        // them in blocks of 8, all loads of a block issued together (entries
        // beyond n are read but not used)
        int col = idx[0];
        for (int f0 = 0; f0 < n; f0 += 8) {
            int iv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) iv[u] = idx[f0 + u];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (f0 + u < n) col = iv[u] < col ? iv[u] : col;
        }
        *first_col = col;
        int k = 0;
        for (bool first = true;; first = false) {
            double val = 0.0;
            bool skip = !first, more = false;
            int next = 0;
            for (int f0 = 0; f0 < n; f0 += 8) {
                int iv[8];
                double av[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    iv[u] = idx[f0 + u];
                    av[u] = area[f0 + u];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    if (f0 + u >= n) break;
                    const int id = iv[u];
                    if (id == col) {
                        if (skip)
                            skip = false;
                        else
                            val += av[u];
                    } else if (id > col && (!more || id < next)) {
                        next = id;
                        more = true;
                    }
                }
            }
            emit(k++, col, val);
            if (!more) return k;
            col = next;
        }
    }
    for (int a = 1; a < n; ++a) {
        const double ar = area[a];
        const int id = idx[a];
        int b = a - 1;
        while (b >= 0 && idx[b] > id) {
            area[b + 1] = area[b];
            idx[b + 1] = idx[b];
            --b;
        }
        area[b + 1] = ar;
        idx[b + 1] = id;
    }
    *first_col = idx[0];
    double val = 0.0;
    int col = idx[0];
    int k = 0;
    for (int f = 0; f < n; ++f) {
        if (idx[f] != col) {
            emit(k++, col, val);
            val = 0.0;
            col = idx[f];
        } else {
            val += area[f];
        }
    }
    emit(k++, col, val);
    return k;
}

}  // namespace lt
