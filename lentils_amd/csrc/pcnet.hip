// Learned preconditioner for the source inversion: a small convolutional
// network on the source grid, BF16 operands on MFMA (FP32 accumulate),
// trained on the GPU by the kernels in this file.  No reference counterpart
// (the reference preconditions with a sparse LU, test_all.py:63-72).
//
// Network (H x W source grid, H and W multiples of 8), 3x3 convolutions with
// zero padding, average-pool down, nearest up, skip connections:
//   e1 = relu(conv(x0; 2->16))            H
//   e2 = relu(conv(pool e1; 16->32))      H/2
//   e3 = relu(conv(pool e2; 32->64))      H/4
//   e4 = relu(conv(pool e3; 64->128))     H/8
//   d3 = relu(conv(up e4; 128->64)) + e3  H/4
//   d2 = relu(conv(up d3; 64->32)) + e2   H/2
//   d1 = relu(conv(up d2; 32->16)) + e1   H
//   y  = conv(d1; 16->1)                  H
// x0 = (residual / rms(residual), diag(A) / mean(diag(A))) per sample; the
// preconditioned direction is z = y * rms(residual) / mean(diag(A)).
//
// Convolutions are implicit GEMMs on v_mfma_f32_32x32x16_bf16: output
// channels x pixels, reduction over (input channel, tap).  A wavefront owns
// 32 pixels and every output channel; operands are rounded to BF16 as they
// are loaded.  The weight gradient is a GEMM over pixels, split into chunks
// whose partial sums are added in a fixed order (deterministic, no atomics).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <vector>

#include "lt_internal.h"
#include "pcnet.h"

namespace lt {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int kConvThreads = 256;  // 4 wavefronts, 32 pixels each
constexpr int kDwChunk = 2048;     // pixels per weight-gradient partial

__device__ inline __bf16 to_bf16(float f) { return (__bf16)f; }

struct ConvArgs {
    const float *x;     // [nb][ci][H][W]
    const __bf16 *w;    // [Cp][Kp]
    const float *bias;  // [co] or null
    const float *skip;  // [nb][co][H][W] or null (added after the activation)
    float *act;         // [nb][co][H][W] or null: activation before the skip
    float *y;           // [nb][co][H][W]
    int nb, ci, co, H, W, K, Kp, relu;
};

// value of column k (input channel k / 9, tap k % 9) of the implicit im2col
// matrix at pixel p; zero outside the image and for k >= K
__device__ inline float im2col(const float *__restrict__ x, int ci_n, int H, int W, int K, long P, int k, long p) {
    if (k >= K || p >= P) return 0.f;
    const int HW = H * W;
    const long n = p / HW;
    const int rem = (int)(p - n * HW);
    const int yy = rem / W, xx = rem - yy * W;
    const int c = k / 9, tap = k - c * 9;
    const int sy = yy + tap / 3 - 1, sx = xx + tap % 3 - 1;
    if (sy < 0 || sy >= H || sx < 0 || sx >= W) return 0.f;
    return x[((n * ci_n + c) * (long)H + sy) * W + sx];
}

template <int MB>
__global__ void __launch_bounds__(kConvThreads) k_conv(ConvArgs a) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long P = (long)a.nb * a.H * a.W;
    const long p0 = ((long)blockIdx.x * (kConvThreads / 64) + wave) * 32;
    if (p0 >= P) return;
    const int j = lane & 31, half = lane >> 5;
    const long p = p0 + j;
    floatx16 acc[MB];
#pragma unroll
    for (int m = 0; m < MB; ++m)
        for (int v = 0; v < 16; ++v) acc[m][v] = 0.f;
    for (int kc = 0; kc < a.Kp; kc += 16) {
        bf16x8 bv;
#pragma unroll
        for (int q = 0; q < 8; ++q) bv[q] = to_bf16(im2col(a.x, a.ci, a.H, a.W, a.K, P, kc + 8 * half + q, p));
#pragma unroll
        for (int m = 0; m < MB; ++m) {
            const bf16x8 av = *reinterpret_cast<const bf16x8 *>(a.w + (long)(m * 32 + j) * a.Kp + kc + 8 * half);
            acc[m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc[m], 0, 0, 0);
        }
    }
    if (p >= P) return;
    const int HW = a.H * a.W;
    const long n = p / HW;
    const int rem = (int)(p - n * HW);
#pragma unroll
    for (int m = 0; m < MB; ++m) {
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int o = m * 32 + 8 * (v / 4) + 4 * half + (v % 4);
            if (o >= a.co) continue;
            float val = acc[m][v] + (a.bias ? a.bias[o] : 0.f);
            if (a.relu) val = val > 0.f ? val : 0.f;
            const long off = (n * a.co + o) * (long)HW + rem;
            if (a.act) a.act[off] = val;
            if (a.skip) val += a.skip[off];
            a.y[off] = val;
        }
    }
}

struct DwArgs {
    const float *dz;  // [nb][co][H][W]  gradient at the pre-activation
    const float *x;   // [nb][ci][H][W]  layer input
    float *part;      // [nsplit][co][K + 1]  (column K: bias gradient)
    int nb, ci, co, H, W, K;
};

// dW[o][k] = sum_p dz[o][p] col[k][p]; column K is the constant 1 (bias).
// Grid: (ceil((K+1)/32), nsplit); one wavefront per block.
template <int MB>
__global__ void __launch_bounds__(64) k_conv_dw(DwArgs a) {
    const int lane = threadIdx.x;
    const int j = lane & 31, half = lane >> 5;
    const int kb = blockIdx.x;
    const long P = (long)a.nb * a.H * a.W;
    const long pbeg = (long)blockIdx.y * kDwChunk;
    const long pend = pbeg + kDwChunk < P ? pbeg + kDwChunk : P;
    const int HW = a.H * a.W;
    const int k = kb * 32 + j;  // this lane's column of B
    floatx16 acc[MB];
#pragma unroll
    for (int m = 0; m < MB; ++m)
        for (int v = 0; v < 16; ++v) acc[m][v] = 0.f;
    for (long pc = pbeg; pc < pend; pc += 16) {
        bf16x8 bv;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const long p = pc + 8 * half + q;
            float val;
            if (p >= pend) val = 0.f;
            else if (k == a.K) val = 1.f;
            else val = im2col(a.x, a.ci, a.H, a.W, a.K, P, k, p);
            bv[q] = to_bf16(val);
        }
#pragma unroll
        for (int m = 0; m < MB; ++m) {
            const int o = m * 32 + j;
            bf16x8 av;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const long p = pc + 8 * half + q;
                float val = 0.f;
                if (o < a.co && p < pend) {
                    const long n = p / HW;
                    val = a.dz[(n * a.co + o) * (long)HW + (p - n * HW)];
                }
                av[q] = to_bf16(val);
            }
            acc[m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc[m], 0, 0, 0);
        }
    }
    // D[o][k]: row o = m*32 + 8*(v/4) + 4*half + v%4, column = this lane's k
    if (k > a.K) return;
    float *out = a.part + (long)blockIdx.y * a.co * (a.K + 1);
#pragma unroll
    for (int m = 0; m < MB; ++m)
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int o = m * 32 + 8 * (v / 4) + 4 * half + (v % 4);
            if (o < a.co) out[(long)o * (a.K + 1) + k] = acc[m][v];
        }
}

// sum of the weight-gradient partials in split order into the parameter
// gradient (weights [co][K], then bias [co])
__global__ void k_dw_reduce(const float *__restrict__ part, int nsplit, int co, int K, float *__restrict__ gw,
                            float *__restrict__ gb) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = co * (K + 1);
    if (e >= n) return;
    float s = 0.f;
    for (int t = 0; t < nsplit; ++t) s += part[(long)t * n + e];
    const int o = e / (K + 1), k = e - o * (K + 1);
    if (k < K) gw[(long)o * K + k] = s;
    else gb[o] = s;
}

// forward weights [co][ci*9] (FP32) -> BF16 [Cp][Kp], zero padded
__global__ void k_pack_w(const float *__restrict__ w, int co, int K, int Cp, int Kp, __bf16 *__restrict__ out) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)Cp * Kp) return;
    const int o = (int)(e / Kp), k = (int)(e - (long)o * Kp);
    out[e] = to_bf16((o < co && k < K) ? w[(long)o * K + k] : 0.f);
}

// weights of the data-gradient convolution: Wf[c][o*9 + 8 - t] = W[o][c*9 + t]
// (channels swapped, taps flipped), BF16 [Cip][Kp2], zero padded
__global__ void k_pack_wflip(const float *__restrict__ w, int co, int ci, int Cip, int Kp2, __bf16 *__restrict__ out) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)Cip * Kp2) return;
    const int c = (int)(e / Kp2), k2 = (int)(e - (long)c * Kp2);
    float v = 0.f;
    if (c < ci && k2 < co * 9) {
        const int o = k2 / 9, tf = k2 - o * 9;
        v = w[(long)o * ci * 9 + c * 9 + (8 - tf)];
    }
    out[e] = to_bf16(v);
}

// 2x2 average pooling [nc][H][W] -> [nc][H/2][W/2]
__global__ void k_pool(long nc, int H, int W, const float *__restrict__ x, float *__restrict__ y) {
    const int h = H / 2, w = W / 2;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nc * h * w) return;
    const long c = e / (h * w);
    const int r = (int)(e - c * h * w), yy = r / w, xx = r - yy * w;
    const float *s = x + c * H * W + (long)(2 * yy) * W + 2 * xx;
    y[e] = 0.25f * (s[0] + s[1] + s[W] + s[W + 1]);
}

// gradient of the pooling, added into dx
__global__ void k_pool_bwd(long nc, int H, int W, const float *__restrict__ dy, float *__restrict__ dx) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nc * H * W) return;
    const long c = e / ((long)H * W);
    const int r = (int)(e - c * H * W), yy = r / W, xx = r - yy * W;
    dx[e] += 0.25f * dy[c * (H / 2) * (W / 2) + (long)(yy / 2) * (W / 2) + xx / 2];
}

// nearest 2x upsampling [nc][h][w] -> [nc][2h][2w]
__global__ void k_up(long nc, int h, int w, const float *__restrict__ x, float *__restrict__ y) {
    const int H = 2 * h, W = 2 * w;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nc * H * W) return;
    const long c = e / ((long)H * W);
    const int r = (int)(e - c * H * W), yy = r / W, xx = r - yy * W;
    y[e] = x[c * h * w + (long)(yy / 2) * w + xx / 2];
}

// gradient of the upsampling: dx = sum of the 2x2 block of dy
__global__ void k_up_bwd(long nc, int h, int w, const float *__restrict__ dy, float *__restrict__ dx) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nc * h * w) return;
    const long c = e / ((long)h * w);
    const int r = (int)(e - c * h * w), yy = r / w, xx = r - yy * w;
    const int W = 2 * w;
    const float *s = dy + c * 4L * h * w + (long)(2 * yy) * W + 2 * xx;
    dx[e] = s[0] + s[1] + s[W] + s[W + 1];
}

// dz = g * (a > 0)
__global__ void k_relu_bwd(long n, const float *__restrict__ g, const float *__restrict__ a, float *__restrict__ dz) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) dz[e] = a[e] > 0.f ? g[e] : 0.f;
}

// per-sample relative squared error: stats[s] = (|y-t|^2, |t|^2); one block
// per sample
__global__ void k_loss_stats(int HW, const float *__restrict__ y, const float *__restrict__ t,
                             double *__restrict__ stats) {
    __shared__ double s0[256], s1[256];
    const int s = blockIdx.x;
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < HW; i += 256) {
        const double d = (double)y[(long)s * HW + i] - (double)t[(long)s * HW + i];
        a += d * d;
        b += (double)t[(long)s * HW + i] * t[(long)s * HW + i];
    }
    s0[threadIdx.x] = a;
    s1[threadIdx.x] = b;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            s0[threadIdx.x] += s0[threadIdx.x + h];
            s1[threadIdx.x] += s1[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        stats[2 * s] = s0[0];
        stats[2 * s + 1] = s1[0];
    }
}

// dL/dy for L = mean_s |y_s - t_s|^2 / |t_s|^2
__global__ void k_loss_grad(int nb, int HW, const float *__restrict__ y, const float *__restrict__ t,
                            const double *__restrict__ stats, float *__restrict__ dy) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)nb * HW) return;
    const int s = (int)(e / HW);
    const double den = stats[2 * s + 1] > 0 ? stats[2 * s + 1] : 1.0;
    dy[e] = (float)(2.0 * ((double)y[e] - (double)t[e]) / (den * nb));
}

__global__ void k_adam(long n, float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2, float clip) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    float gg = g[e] * clip;
    const float mm = b1 * m[e] + (1.f - b1) * gg;
    const float vv = b2 * v[e] + (1.f - b2) * gg * gg;
    m[e] = mm;
    v[e] = vv;
    p[e] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

__global__ void k_sumsq_f32(long n, const float *__restrict__ g, double *__restrict__ out) {
    __shared__ double sh[256];
    double a = 0.0;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long)gridDim.x * 256) a += (double)g[e] * g[e];
    sh[threadIdx.x] = a;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

inline int round_up(int v, int m) { return (v + m - 1) / m * m; }

}  // namespace

// ---------------------------------------------------------------------------

const int PcNet::kCin[PcNet::kLayers] = {2, 16, 32, 64, 128, 64, 32, 16};
const int PcNet::kCout[PcNet::kLayers] = {16, 32, 64, 128, 64, 32, 16, 1};
const int PcNet::kLevel[PcNet::kLayers] = {0, 1, 2, 3, 2, 1, 0, 0};

long PcNet::param_count() {
    long n = 0;
    for (int l = 0; l < kLayers; ++l) n += (long)kCout[l] * kCin[l] * 9 + kCout[l];
    return n;
}

lt_status PcNet::create(int H_, int W_, int cap) {
    if (H_ <= 0 || W_ <= 0 || H_ % 8 || W_ % 8 || cap <= 0) {
        set_error("pcnet: source grid %d x %d must be a positive multiple of 8", H_, W_);
        return LT_ERR_ARG;
    }
    H = H_;
    W = W_;
    nbcap = cap;
    const long np = param_count();
    long off = 0;
    for (int l = 0; l < kLayers; ++l) {
        woff[l] = off;
        off += (long)kCout[l] * kCin[l] * 9;
        boff[l] = off;
        off += kCout[l];
    }
    LT_TRY(params.alloc(np));
    LT_TRY(grads.alloc(np));
    LT_TRY(adam_m.alloc(np));
    LT_TRY(adam_v.alloc(np));
    LT_TRY(adam_m.zero());
    LT_TRY(adam_v.zero());
    for (int l = 0; l < kLayers; ++l) {
        const int K = kCin[l] * 9;
        LT_TRY(wbf[l].alloc((long)round_up(kCout[l], 32) * round_up(K, 16)));
        LT_TRY(wflip[l].alloc((long)round_up(kCin[l], 32) * round_up(kCout[l] * 9, 16)));
    }
    // activations: index by the tensor they hold
    const long hw0 = (long)H * W;
    auto sz = [&](int c, int lev) { return (long)cap * c * (hw0 >> (2 * lev)); };
    const long sizes[kTensors] = {
        sz(2, 0),                        // X0
        sz(16, 0), sz(16, 1),            // E1, P1
        sz(32, 1), sz(32, 2),            // E2, P2
        sz(64, 2), sz(64, 3),            // E3, P3
        sz(128, 3), sz(128, 2),          // E4, U4
        sz(64, 2), sz(64, 2), sz(64, 1), // A5, D3, U3
        sz(32, 1), sz(32, 1), sz(32, 0), // A6, D2, U2
        sz(16, 0), sz(16, 0),            // A7, D1
        sz(1, 0)                         // Y
    };
    for (int t = 0; t < kTensors; ++t) {
        LT_TRY(act[t].alloc(std::max<long>(sizes[t], 1)));
        tsize[t] = sizes[t];
    }
    long maxsz = 0;
    for (int t = 0; t < kTensors; ++t) maxsz = std::max(maxsz, sizes[t]);
    LT_TRY(g0.alloc(maxsz));
    LT_TRY(g1.alloc(maxsz));
    LT_TRY(g2.alloc(maxsz));
    for (int t = 0; t < kTensors; ++t) LT_TRY(gact[t].alloc(std::max<long>(sizes[t], 1)));
    // weight-gradient partials: largest layer
    long maxpart = 1;
    for (int l = 0; l < kLayers; ++l) {
        const long P = (long)cap * (hw0 >> (2 * kLevel[l]));
        const long ns = (P + kDwChunk - 1) / kDwChunk;
        maxpart = std::max(maxpart, ns * kCout[l] * (kCin[l] * 9 + 1));
    }
    LT_TRY(part.alloc(maxpart));
    LT_TRY(stats.alloc(2L * cap));
    LT_TRY(gnorm.alloc(256));
    return LT_OK;
}

lt_status PcNet::init_random(unsigned seed) {
    // He initialisation for the ReLU layers, zero biases and a zero last
    // layer: an untrained network outputs 0 (the solver's correction starts
    // at none), and the last layer's gradient is still the full one
    std::vector<float> h(param_count(), 0.f);
    std::mt19937 gen(seed);
    for (int l = 0; l < kLayers - 1; ++l) {
        const int fan = kCin[l] * 9;
        const float sd = std::sqrt(2.f / fan);
        std::normal_distribution<float> nd(0.f, sd);
        for (long k = 0; k < (long)kCout[l] * fan; ++k) h[woff[l] + k] = nd(gen);
    }
    LT_TRY(params.upload(h.data(), h.size()));
    step = 0;
    return pack();
}

lt_status PcNet::pack() {
    for (int l = 0; l < kLayers; ++l) {
        const int K = kCin[l] * 9, Cp = round_up(kCout[l], 32), Kp = round_up(K, 16);
        hipLaunchKernelGGL(k_pack_w, dim3(blocks_for((long)Cp * Kp, 256)), dim3(256), 0, stream(), params.p + woff[l],
                           kCout[l], K, Cp, Kp, wbf[l].p);
        const int Cip = round_up(kCin[l], 32), Kp2 = round_up(kCout[l] * 9, 16);
        hipLaunchKernelGGL(k_pack_wflip, dim3(blocks_for((long)Cip * Kp2, 256)), dim3(256), 0, stream(),
                           params.p + woff[l], kCout[l], kCin[l], Cip, Kp2, wflip[l].p);
    }
    LT_CHECK_LAUNCH();
    return LT_OK;
}

lt_status PcNet::conv(const float *x, int ci, int co, int lev, int nb, const __bf16 *w, const float *bias, bool relu,
                      const float *skip, float *actout, float *y) {
    const int h = H >> lev, wd = W >> lev;
    ConvArgs a{x, w, bias, skip, actout, y, nb, ci, co, h, wd, ci * 9, round_up(ci * 9, 16), relu ? 1 : 0};
    const long P = (long)nb * h * wd;
    const unsigned nblk = (unsigned)((P + 127) / 128);
    const int mb = round_up(co, 32) / 32;
    switch (mb) {
        case 1: hipLaunchKernelGGL(k_conv<1>, dim3(nblk), dim3(kConvThreads), 0, stream(), a); break;
        case 2: hipLaunchKernelGGL(k_conv<2>, dim3(nblk), dim3(kConvThreads), 0, stream(), a); break;
        case 4: hipLaunchKernelGGL(k_conv<4>, dim3(nblk), dim3(kConvThreads), 0, stream(), a); break;
        default: set_error("pcnet: %d output channels", co); return LT_ERR_ARG;
    }
    LT_CHECK_LAUNCH();
    return LT_OK;
}

lt_status PcNet::conv_dw(const float *dz, const float *x, int l, int nb) {
    const int lev = kLevel[l], ci = kCin[l], co = kCout[l], K = ci * 9;
    const int h = H >> lev, wd = W >> lev;
    const long P = (long)nb * h * wd;
    const int nsplit = (int)((P + kDwChunk - 1) / kDwChunk);
    DwArgs a{dz, x, part.p, nb, ci, co, h, wd, K};
    const dim3 grid((unsigned)((K + 1 + 31) / 32), (unsigned)nsplit);
    switch (round_up(co, 32) / 32) {
        case 1: hipLaunchKernelGGL(k_conv_dw<1>, grid, dim3(64), 0, stream(), a); break;
        case 2: hipLaunchKernelGGL(k_conv_dw<2>, grid, dim3(64), 0, stream(), a); break;
        case 4: hipLaunchKernelGGL(k_conv_dw<4>, grid, dim3(64), 0, stream(), a); break;
        default: set_error("pcnet: %d output channels", co); return LT_ERR_ARG;
    }
    const int n = co * (K + 1);
    hipLaunchKernelGGL(k_dw_reduce, dim3(blocks_for(n, 256)), dim3(256), 0, stream(), part.p, nsplit, co, K,
                       grads.p + woff[l], grads.p + boff[l]);
    LT_CHECK_LAUNCH();
    return LT_OK;
}

lt_status PcNet::forward(int nb) {
    if (nb <= 0 || nb > nbcap) {
        set_error("pcnet: batch %d outside 1..%d", nb, nbcap);
        return LT_ERR_ARG;
    }
    const long hw = (long)H * W;
    auto launch_pool = [&](const float *x, float *y, int c, int lev) {
        const int h = H >> lev, w = W >> lev;
        const long n = (long)nb * c * (h / 2) * (w / 2);
        hipLaunchKernelGGL(k_pool, dim3(blocks_for(n, 256)), dim3(256), 0, stream(), (long)nb * c, h, w, x, y);
    };
    auto launch_up = [&](const float *x, float *y, int c, int lev) {  // lev of the coarse input
        const int h = H >> lev, w = W >> lev;
        const long n = (long)nb * c * 4 * h * w;
        hipLaunchKernelGGL(k_up, dim3(blocks_for(n, 256)), dim3(256), 0, stream(), (long)nb * c, h, w, x, y);
    };
    (void)hw;
    const __bf16 *wb[kLayers];
    const float *bb[kLayers];
    for (int l = 0; l < kLayers; ++l) {
        wb[l] = wbf[l].p;
        bb[l] = params.p + boff[l];
    }
    LT_TRY(conv(act[X0].p, 2, 16, 0, nb, wb[0], bb[0], true, nullptr, nullptr, act[E1].p));
    launch_pool(act[E1].p, act[P1].p, 16, 0);
    LT_TRY(conv(act[P1].p, 16, 32, 1, nb, wb[1], bb[1], true, nullptr, nullptr, act[E2].p));
    launch_pool(act[E2].p, act[P2].p, 32, 1);
    LT_TRY(conv(act[P2].p, 32, 64, 2, nb, wb[2], bb[2], true, nullptr, nullptr, act[E3].p));
    launch_pool(act[E3].p, act[P3].p, 64, 2);
    LT_TRY(conv(act[P3].p, 64, 128, 3, nb, wb[3], bb[3], true, nullptr, nullptr, act[E4].p));
    launch_up(act[E4].p, act[U4].p, 128, 3);
    LT_TRY(conv(act[U4].p, 128, 64, 2, nb, wb[4], bb[4], true, act[E3].p, act[A5].p, act[D3].p));
    launch_up(act[D3].p, act[U3].p, 64, 2);
    LT_TRY(conv(act[U3].p, 64, 32, 1, nb, wb[5], bb[5], true, act[E2].p, act[A6].p, act[D2].p));
    launch_up(act[D2].p, act[U2].p, 32, 1);
    LT_TRY(conv(act[U2].p, 32, 16, 0, nb, wb[6], bb[6], true, act[E1].p, act[A7].p, act[D1].p));
    LT_TRY(conv(act[D1].p, 16, 1, 0, nb, wb[7], bb[7], false, nullptr, nullptr, act[Y].p));
    LT_CHECK_LAUNCH();
    return LT_OK;
}

// gradient of the loss w.r.t. every parameter, given dL/dy in gact[Y]
lt_status PcNet::backward(int nb) {
    auto nel = [&](int c, int lev) { return (long)nb * c * ((long)(H >> lev) * (W >> lev)); };
    auto relu_bwd = [&](const float *g, const float *a, float *dz, long n) {
        hipLaunchKernelGGL(k_relu_bwd, dim3(blocks_for(n, 256)), dim3(256), 0, stream(), n, g, a, dz);
    };
    auto up_bwd = [&](const float *dy, float *dx, int c, int lev) {  // lev of the coarse tensor
        const long n = nel(c, lev);
        hipLaunchKernelGGL(k_up_bwd, dim3(blocks_for(n, 256)), dim3(256), 0, stream(), (long)nb * c, H >> lev,
                           W >> lev, dy, dx);
    };
    auto pool_bwd = [&](const float *dy, float *dx, int c, int lev) {  // lev of the fine tensor
        const long n = nel(c, lev);
        hipLaunchKernelGGL(k_pool_bwd, dim3(blocks_for(n, 256)), dim3(256), 0, stream(), (long)nb * c, H >> lev,
                           W >> lev, dy, dx);
    };
    auto dconv = [&](const float *dz, int l, float *dx) {  // data gradient of layer l
        return conv(dz, kCout[l], kCin[l], kLevel[l], nb, wflip[l].p, nullptr, false, nullptr, nullptr, dx);
    };
    float *G = g0.p, *G2 = g1.p;
    // y = conv8(d1)
    LT_TRY(conv_dw(gact[Y].p, act[D1].p, 7, nb));
    LT_TRY(dconv(gact[Y].p, 7, gact[D1].p));  // dL/dd1
    // d1 = a7 + e1; a7 = relu(conv7(u2))
    LT_HIP(hipMemcpyAsync(gact[E1].p, gact[D1].p, nel(16, 0) * sizeof(float), hipMemcpyDeviceToDevice, stream()));
    relu_bwd(gact[D1].p, act[A7].p, G, nel(16, 0));
    LT_TRY(conv_dw(G, act[U2].p, 6, nb));
    LT_TRY(dconv(G, 6, G2));              // dL/du2 (32 ch, level 0)
    up_bwd(G2, gact[D2].p, 32, 1);        // dL/dd2
    // d2 = a6 + e2
    LT_HIP(hipMemcpyAsync(gact[E2].p, gact[D2].p, nel(32, 1) * sizeof(float), hipMemcpyDeviceToDevice, stream()));
    relu_bwd(gact[D2].p, act[A6].p, G, nel(32, 1));
    LT_TRY(conv_dw(G, act[U3].p, 5, nb));
    LT_TRY(dconv(G, 5, G2));              // dL/du3 (64 ch, level 1)
    up_bwd(G2, gact[D3].p, 64, 2);
    // d3 = a5 + e3
    LT_HIP(hipMemcpyAsync(gact[E3].p, gact[D3].p, nel(64, 2) * sizeof(float), hipMemcpyDeviceToDevice, stream()));
    relu_bwd(gact[D3].p, act[A5].p, G, nel(64, 2));
    LT_TRY(conv_dw(G, act[U4].p, 4, nb));
    LT_TRY(dconv(G, 4, G2));              // dL/du4 (128 ch, level 2)
    up_bwd(G2, gact[E4].p, 128, 3);       // dL/de4
    // e4 = relu(conv4(p3))
    relu_bwd(gact[E4].p, act[E4].p, G, nel(128, 3));
    LT_TRY(conv_dw(G, act[P3].p, 3, nb));
    LT_TRY(dconv(G, 3, G2));              // dL/dp3 (64 ch, level 3)
    pool_bwd(G2, gact[E3].p, 64, 2);      // += into dL/de3
    relu_bwd(gact[E3].p, act[E3].p, G, nel(64, 2));
    LT_TRY(conv_dw(G, act[P2].p, 2, nb));
    LT_TRY(dconv(G, 2, G2));              // dL/dp2 (32 ch, level 2)
    pool_bwd(G2, gact[E2].p, 32, 1);
    relu_bwd(gact[E2].p, act[E2].p, G, nel(32, 1));
    LT_TRY(conv_dw(G, act[P1].p, 1, nb));
    LT_TRY(dconv(G, 1, G2));              // dL/dp1 (16 ch, level 1)
    pool_bwd(G2, gact[E1].p, 16, 0);
    relu_bwd(gact[E1].p, act[E1].p, G, nel(16, 0));
    LT_TRY(conv_dw(G, act[X0].p, 0, nb));
    LT_CHECK_LAUNCH();
    return LT_OK;
}

lt_status PcNet::loss(int nb, const float *target, double *loss_out) {
    const int hw = H * W;
    hipLaunchKernelGGL(k_loss_stats, dim3(nb), dim3(256), 0, stream(), hw, act[Y].p, target, stats.p);
    hipLaunchKernelGGL(k_loss_grad, dim3(blocks_for((long)nb * hw, 256)), dim3(256), 0, stream(), nb, hw, act[Y].p,
                       target, stats.p, gact[Y].p);
    LT_CHECK_LAUNCH();
    if (loss_out) {
        std::vector<double> h(2 * nb);
        LT_TRY(stats.download(h.data(), 2 * nb));
        LT_TRY(sync());
        double s = 0.0;
        for (int k = 0; k < nb; ++k) s += h[2 * k + 1] > 0 ? h[2 * k] / h[2 * k + 1] : 0.0;
        *loss_out = s / nb;
    }
    return LT_OK;
}

lt_status PcNet::adam_step(float lr, float max_norm) {
    const long np = param_count();
    ++step;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
    float clip = 1.f;
    if (max_norm > 0.f) {
        hipLaunchKernelGGL(k_sumsq_f32, dim3(256), dim3(256), 0, stream(), np, grads.p, gnorm.p);
        std::vector<double> h(256);
        LT_TRY(gnorm.download(h.data(), 256));
        LT_TRY(sync());
        double s = 0.0;
        for (double v : h) s += v;
        const double nrm = std::sqrt(s);
        if (nrm > max_norm) clip = (float)(max_norm / nrm);
    }
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(np, 256)), dim3(256), 0, stream(), np, params.p, grads.p, adam_m.p,
                       adam_v.p, lr, b1, b2, 1e-8f, c1, c2, clip);
    LT_CHECK_LAUNCH();
    return pack();
}

}  // namespace lt

// ---------------------------------------------------------------------------
// C ABI

using namespace lt;

struct lt_pcnet_s {
    PcNet net;
};

namespace lt {
PcNet *pcnet_of(lt_pcnet h) { return h ? &h->net : nullptr; }
}  // namespace lt

extern "C" {

lt_status lt_pcnet_create(int H, int W, int batch, const float *params, unsigned seed, lt_pcnet *out) {
    LT_TRY(ensure_init());
    if (!out) return LT_ERR_ARG;
    auto h = std::make_unique<lt_pcnet_s>();
    LT_TRY(h->net.create(H, W, batch));
    if (params) {
        LT_TRY(h->net.params.upload(params, PcNet::param_count()));
        LT_TRY(h->net.pack());
    } else {
        LT_TRY(h->net.init_random(seed));
    }
    LT_TRY(sync());
    *out = h.release();
    return LT_OK;
}

long lt_pcnet_param_count(void) { return PcNet::param_count(); }

lt_status lt_pcnet_get_params(lt_pcnet h, float *out) {
    if (!h || !out) return LT_ERR_ARG;
    LT_TRY(h->net.params.download(out, PcNet::param_count()));
    return sync();
}

lt_status lt_pcnet_set_params(lt_pcnet h, const float *in) {
    if (!h || !in) return LT_ERR_ARG;
    LT_TRY(h->net.params.upload(in, PcNet::param_count()));
    LT_TRY(h->net.pack());
    return sync();
}

lt_status lt_pcnet_forward_host(lt_pcnet h, const float *x0, int nb, float *y) {
    if (!h || !x0 || !y || nb <= 0 || nb > h->net.nbcap) {
        set_error("lt_pcnet_forward_host: bad arguments");
        return LT_ERR_ARG;
    }
    PcNet &N = h->net;
    const long hw = (long)N.H * N.W;
    LT_TRY(N.act[PcNet::X0].upload(x0, 2L * nb * hw));
    LT_TRY(N.forward(nb));
    LT_TRY(N.act[PcNet::Y].download(y, (long)nb * hw));
    return sync();
}

lt_status lt_pcnet_grad_host(lt_pcnet h, const float *x0, const float *dy, int nb, float *y, float *grad) {
    if (!h || !x0 || !dy || !grad || nb <= 0 || nb > h->net.nbcap) {
        set_error("lt_pcnet_grad_host: bad arguments");
        return LT_ERR_ARG;
    }
    PcNet &N = h->net;
    const long hw = (long)N.H * N.W;
    LT_TRY(N.act[PcNet::X0].upload(x0, 2L * nb * hw));
    LT_TRY(N.forward(nb));
    LT_TRY(N.gact[PcNet::Y].upload(dy, (long)nb * hw));
    LT_TRY(N.backward(nb));
    if (y) LT_TRY(N.act[PcNet::Y].download(y, (long)nb * hw));
    LT_TRY(N.grads.download(grad, PcNet::param_count()));
    return sync();
}

lt_status lt_pcnet_train(lt_pcnet h, const float *inputs, const float *targets, long n, int epochs, int batch,
                         float lr, float lr_final, unsigned seed, double *loss_hist) {
    if (!h || !inputs || !targets || n <= 0 || epochs <= 0 || batch <= 0 || batch > h->net.nbcap) {
        set_error("lt_pcnet_train: bad arguments");
        return LT_ERR_ARG;
    }
    PcNet &N = h->net;
    const long hw = (long)N.H * N.W;
    DevBuf<float> din, dtg;
    LT_TRY(din.upload(inputs, 2 * n * hw));
    LT_TRY(dtg.upload(targets, n * hw));
    DevBuf<float> tb;
    LT_TRY(tb.alloc((long)batch * hw));
    std::mt19937 gen(seed);
    std::vector<long> order(n);
    for (long k = 0; k < n; ++k) order[k] = k;
    const long nsteps_ep = n / batch;
    const long total = nsteps_ep * epochs;
    long it = 0;
    for (int ep = 0; ep < epochs; ++ep) {
        std::shuffle(order.begin(), order.end(), gen);
        double acc = 0.0;
        for (long st = 0; st < nsteps_ep; ++st, ++it) {
            for (int b = 0; b < batch; ++b) {
                const long k = order[st * batch + b];
                LT_HIP(hipMemcpyAsync(N.act[PcNet::X0].p + 2L * b * hw, din.p + 2L * k * hw, 2 * hw * sizeof(float),
                                      hipMemcpyDeviceToDevice, stream()));
                LT_HIP(hipMemcpyAsync(tb.p + (long)b * hw, dtg.p + k * hw, hw * sizeof(float),
                                      hipMemcpyDeviceToDevice, stream()));
            }
            LT_TRY(N.forward(batch));
            double l = 0.0;
            LT_TRY(N.loss(batch, tb.p, (st % 16 == 0 || st == nsteps_ep - 1) ? &l : nullptr));
            acc += l;
            LT_TRY(N.backward(batch));
            // cosine decay from lr to lr_final
            const float fr = total > 1 ? (float)it / (float)(total - 1) : 1.f;
            const float lrt = lr_final + 0.5f * (lr - lr_final) * (1.f + std::cos((float)M_PI * fr));
            LT_TRY(N.adam_step(lrt, 1.0f));
        }
        if (loss_hist) {
            const long nsamp = (nsteps_ep + 15) / 16 + ((nsteps_ep - 1) % 16 != 0 ? 1 : 0);
            loss_hist[ep] = acc / std::max<long>(nsamp, 1);
        }
    }
    return sync();
}

lt_status lt_pcnet_eval(lt_pcnet h, const float *inputs, const float *targets, long n, double *loss) {
    if (!h || !inputs || !targets || n <= 0 || !loss) return LT_ERR_ARG;
    PcNet &N = h->net;
    const long hw = (long)N.H * N.W;
    DevBuf<float> tb;
    LT_TRY(tb.alloc((long)N.nbcap * hw));
    double s = 0.0;
    for (long k0 = 0; k0 < n; k0 += N.nbcap) {
        const int nb = (int)std::min<long>(N.nbcap, n - k0);
        LT_TRY(N.act[PcNet::X0].upload(inputs + 2 * k0 * hw, 2L * nb * hw));
        LT_TRY(tb.upload(targets + k0 * hw, (long)nb * hw));
        LT_TRY(N.forward(nb));
        double l = 0.0;
        LT_TRY(N.loss(nb, tb.p, &l));
        s += l * nb;
    }
    *loss = s / n;
    return LT_OK;
}

lt_status lt_pcnet_destroy(lt_pcnet h) {
    delete h;
    return LT_OK;
}

}  // extern "C"
