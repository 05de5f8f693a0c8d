// Device-wide scans and a segmented sort of 64-bit keys.
//
// Scans: three launches (per-tile totals, scan of totals in one workgroup,
// per-tile scan + carry).  Deterministic; `in` may alias `out`.
// Segmented sort: segments of <= 32 keys are insertion-sorted by one thread,
// longer ones by a workgroup bitonic network (in LDS up to 4096 keys,
// otherwise in a global scratch area).  Keys are unique where it matters
// (the low 32 bits carry a position), so the order is deterministic.
#include <algorithm>

#include "lt_internal.h"

namespace lt {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

struct OpAdd {
    typedef int64_t T;
    __device__ static T identity() { return 0; }
    __device__ static T apply(T a, T b) { return a + b; }
};
struct OpMax {
    typedef int32_t T;
    __device__ static T identity() { return INT32_MIN; }
    __device__ static T apply(T a, T b) { return a > b ? a : b; }
};

// inclusive scan of one value per thread across the block (Hillis-Steele in LDS)
template <class Op>
__device__ typename Op::T block_inclusive(typename Op::T v, typename Op::T *sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        typename Op::T o = (t >= off) ? sh[t - off] : Op::identity();
        __syncthreads();
        v = Op::apply(o, v);
        sh[t] = v;
        __syncthreads();
    }
    return v;
}

template <class Op>
__global__ void __launch_bounds__(kScanThreads) k_tile_total(const typename Op::T *in, long n, typename Op::T *tot) {
    __shared__ typename Op::T sh[kScanThreads];
    const long base = (long)blockIdx.x * kScanTile + (long)threadIdx.x * kScanItems;
    typename Op::T acc = Op::identity();
    for (int k = 0; k < kScanItems; ++k) {
        long i = base + k;
        if (i < n) acc = Op::apply(acc, in[i]);
    }
    acc = block_inclusive<Op>(acc, sh);
    if (threadIdx.x == kScanThreads - 1) tot[blockIdx.x] = acc;
}

// exclusive (ADD) / inclusive (MAX) scan of the tile totals, one block
template <class Op, bool kExclusive>
__global__ void __launch_bounds__(kScanThreads) k_scan_totals(typename Op::T *tot, long m) {
    __shared__ typename Op::T sh[kScanThreads];
    typename Op::T carry = Op::identity();
    for (long base = 0; base < m; base += kScanThreads) {
        long i = base + threadIdx.x;
        typename Op::T v = (i < m) ? tot[i] : Op::identity();
        typename Op::T inc = block_inclusive<Op>(v, sh);
        typename Op::T ex = (threadIdx.x > 0) ? sh[threadIdx.x - 1] : Op::identity();
        __syncthreads();
        if (i < m) tot[i] = Op::apply(carry, kExclusive ? ex : inc);
        carry = Op::apply(carry, sh[kScanThreads - 1]);
        __syncthreads();
    }
}

template <class Op, bool kExclusive>
__global__ void __launch_bounds__(kScanThreads) k_tile_scan(const typename Op::T *in, typename Op::T *out, long n,
                                                            const typename Op::T *tot, typename Op::T *total_out) {
    __shared__ typename Op::T sh[kScanThreads];
    const long base = (long)blockIdx.x * kScanTile + (long)threadIdx.x * kScanItems;
    typename Op::T vals[kScanItems];
    typename Op::T acc = Op::identity();
    for (int k = 0; k < kScanItems; ++k) {
        long i = base + k;
        vals[k] = (i < n) ? in[i] : Op::identity();
        acc = Op::apply(acc, vals[k]);
    }
    typename Op::T inc = block_inclusive<Op>(acc, sh);
    typename Op::T run = (threadIdx.x > 0) ? sh[threadIdx.x - 1] : Op::identity();
    // tile prefix from the scanned totals
    typename Op::T pre;
    if (kExclusive)
        pre = tot[blockIdx.x];
    else
        pre = (blockIdx.x > 0) ? tot[blockIdx.x - 1] : Op::identity();
    run = Op::apply(pre, run);
    for (int k = 0; k < kScanItems; ++k) {
        long i = base + k;
        typename Op::T nv = Op::apply(run, vals[k]);
        if (i < n) out[i] = kExclusive ? run : nv;
        run = nv;
    }
    if (total_out && blockIdx.x == gridDim.x - 1 && threadIdx.x == kScanThreads - 1) *total_out = run;
    (void)inc;
}

template <class Op, bool kExclusive>
lt_status run_scan(const typename Op::T *in, typename Op::T *out, long n, typename Op::T *total_out) {
    if (n <= 0) return LT_OK;
    long tiles = (n + kScanTile - 1) / kScanTile;
    DevBuf<typename Op::T> tot;
    LT_TRY(tot.alloc(tiles));
    hipLaunchKernelGGL((k_tile_total<Op>), dim3((unsigned)tiles), dim3(kScanThreads), 0, stream(), in, n, tot.p);
    LT_CHECK_LAUNCH();
    hipLaunchKernelGGL((k_scan_totals<Op, kExclusive>), dim3(1), dim3(kScanThreads), 0, stream(), tot.p, tiles);
    LT_CHECK_LAUNCH();
    hipLaunchKernelGGL((k_tile_scan<Op, kExclusive>), dim3((unsigned)tiles), dim3(kScanThreads), 0, stream(), in,
                       out, n, tot.p, total_out);
    LT_CHECK_LAUNCH();
    return sync();  // `tot` is freed on return
}

}  // namespace

lt_status exclusive_scan_i64(const int64_t *in, int64_t *out, long n) {
    // out has n + 1 entries; out[n] receives the total
    return run_scan<OpAdd, true>(in, out, n, out + n);
}

lt_status inclusive_max_i32(const int32_t *in, int32_t *out, long n) {
    return run_scan<OpMax, false>(in, out, n, nullptr);
}


}  // namespace lt
