// Kaiser-Bessel NUFFT plan: uv set binned into uv tiles, grid geometry,
// apodization, rocFFT plans and scratch, all device resident.
#pragma once

#include <rocfft/rocfft.h>

#include <memory>
#include <vector>

#include "lt_internal.h"

namespace lt {

template <class T>
struct Cplx {
    T re, im;
};

// Taylor coefficients of I0(beta*sqrt(s)) in s (all positive), truncated
// where the tail is below the working precision relative to I0(beta).
constexpr int kMaxPoly = 64;
struct KbPoly {
    double c[kMaxPoly];
    int deg;
};

// Per-tap polynomials of the kernel in the fractional offset f = u/du -
// round(u/du) in [-1/2, 1/2]: tap k (0..2W) sits at distance x = f + W - k
// and its value I0(beta sqrt(1 - (x/W)^2)) is an entire function of x, so a
// Chebyshev fit of modest degree on f in [-1/2, 1/2] is accurate to the
// working precision.  The edge taps are zero where |x| > W (f > 0 for k = 0,
// f < 0 for k = 2W).  Coefficients are monomials in f, [k][m], m = 0..D.
// degree per precision and support: FP32 keeps the fit error below ~3e-7
// of the kernel peak (narrow kernels vary faster per unit offset)
constexpr int kTapDeg64 = 16;
__host__ __device__ constexpr int tap_degree(bool fp64, int W) { return fp64 ? kTapDeg64 : (W <= 2 ? 9 : 7); }
template <class T, int W>
struct TapDeg {
    static constexpr int D = tap_degree(sizeof(T) == 8, W);
};

struct NufftPlan;

// per-tap monomial coefficients [2W+1][D+1] of the kernel (see TapDeg);
// returns the max fit error relative to the kernel peak
double kb_taps_fit(double beta, int W, int D, std::vector<double> &mono);
// the same for the 2W taps at distances s + W - 1/2 - j (j = 0..2W-1) from
// the visibility, s in [-1/2, 1/2]: [2W][D+1]
double kb_taps_fit_half(double beta, int W, int D, std::vector<double> &mono);
double kb_entire(double beta, int W, double x);

// D^T C D restricted to the uv grid, precomputed once per (uv set, weights):
// with the phase cancelling and a real kernel, degrid-weight-grid is a real,
// symmetric, translation-variant stencil on raw (pre-fold) grid cells,
//   R[r'] = sum_o coef[o][r'] S[r' + o],   o in [-2W, 2W]^2,
//   coef[o][r'] = 0.5 / (4W^2)^2 sum_v w_v k_v(r') k_v(r' + o),
// which replaces the per-visibility work inside every solver iteration by
// a stencil over the ~(active cells) x (4W+1)^2 coefficients.  The rows are
// folded onto the half-plane of the r2c spectrum at build time (nrow below).
struct NormalStencil {
    int W = 0, R = 0, NO = 0;  // half-width, stencil radius 2W, (4W+1)^2 offsets
    long ncell = 0;            // grid cells holding at least one visibility
    // stencil rows: the half-plane cells (a, b >= 0) of the final spectrum
    // where a raw cell within W of a visibility cell lies at (a, b) or at
    // its mirror (-a, -b); each row's coefficients are those of the raw
    // cell plus the mirror's reflected (the Hermitian fold done once, at
    // build time)
    long nrow = 0;
    long nrow_raw = 0;         // raw cells within W of the visibility cells (both half-planes)
    int ST = 16;               // stencil tile edge
    long ntile = 0;
    DevBuf<int32_t> rowmap;    // centred raw cell -> row or -1 (nu * nv)
    DevBuf<int32_t> row_cell;  // row -> centred raw cell index
    DevBuf<int32_t> row_dst;   // row -> its cell in a [v][u] band (v * nu + u)
    DevBuf<int32_t> tiles;     // stencil tiles holding rows
    DevBuf<double> coef;       // [NO][nrow]
    DevBuf<float> coef32;      // FP32 copy / coef32_scale for the precision form (made on first use)
    double coef32_scale = 1.0;
    int nbv = 0;               // spectral columns 0..nbv-1 hold every value the stencil reads
    bool use_band = true;      // band-pruned separable transform instead of the full 2-D one
    int nb_sel = 4;            // samples per stencil workgroup (stencil_autotune); < 0: the matrix-core
                               // form with -nb_sel groups of 8 samples per workgroup
    DevBuf<uint8_t> rbuf;      // per-sample stencil output of the full-spectrum form, [batch][nrow] complex T
    // band spectra of the row transforms (band_in) and of the stencil
    // (band_out, written at the rows' cells), per precision: written only
    // there and on the image rows, zero elsewhere from allocation on
    DevBuf<uint8_t> band_in[2], band_out[2];

    // matrix-core form of the FP64 band stencil (radius 6): 8 x 8 blocks of
    // rows, each as four 4 x 4 tiles whose 16 x 256 coefficient matrices are
    // packed in the MFMA operand layout (mfma_build, made on first use)
    long mf_nblock = -1;       // -1: not built
    int mf_ntv = 0;            // blocks along v
    DevBuf<int32_t> mf_blocks, mf_rows;
    DevBuf<double> mf_A;
    DevBuf<float> mf_A32;      // the same matrices of the FP32 coefficients (mfma_build32)
    int nb_sel32 = 0;          // the FP32 stencil's form (stencil_autotune32); 0: not tuned (vector form, 4)

    lt_status build(NufftPlan &P, const double *w_sorted);
    lt_status mfma_build(NufftPlan &P);
    lt_status mfma_build32(NufftPlan &P);
    lt_status apply(NufftPlan &P, const void *spec_in, long sstride, void *spec_out, const int *active, int nactive);
    lt_status apply32(NufftPlan &P, const void *spec_in, long sstride, void *spec_out, const int *active,
                      int nactive);  // FP32 spectra
};

struct NufftPlan {
    int prec = LT_F64;
    int W = 4;
    double beta = 0.0;
    // image side (only when built from an image space)
    bool has_image = false;
    int nx = 0, ny = 0, padx = 0, pady = 0, nxp = 0, nyp = 0;
    // Fourier grid (fourier_space of the padded image)
    int nu = 0, nv = 0, half_nv = 0;
    double du = 0, dv = 0, gcx = 0, gcy = 0;
    // visibilities (channel-expanded: index = nrows * ch + row)
    int nrows = 0, nch = 0;
    long nvis = 0;
    std::vector<double> uvw_host, chan_host;  // kept for derived plans (dirty beam)
    bool phase = false;
    // uv tiles of TU x TV grid cells in centred raw coordinates
    int TU = 32, TV = 32, ntu = 0, ntv = 0, cu = 0, cv = 0, WU = 0, WV = 0;
    int chunk = 8192;
    long nchunks = 0;
    KbPoly poly{};
    DevBuf<float> taps32;   // (2W+1) x (tap_degree(false, W)+1)
    DevBuf<double> taps64;  // (2W+1) x (kTapDeg64+1)
    double taps_err = 0.0;  // max fit error relative to the kernel peak
    // the 2W taps of the window that starts W - 1 cells below floor(u/du),
    // as polynomials of s = u/du - floor(u/du) - 1/2 (kb_taps_fit_half);
    // empty when their fit misses the precision's tolerance
    DevBuf<float> taps32h;   // 2W x (tap_degree(false, W)+1)
    DevBuf<double> taps64h;  // 2W x (kTapDeg64+1)
    DevBuf<int32_t> perm;        // sorted position -> channel-expanded index
    DevBuf<uint32_t> key_sorted;  // tile-major cell key of each sorted visibility
    DevBuf<int32_t> ngp;   // packed (ngp_u + 32768) | (ngp_v + 32768) << 16
    DevBuf<double> frac64, ph64;
    DevBuf<float> frac32, ph32;
    DevBuf<int32_t> ck_tile, ck_beg, ck_end, tile_ck0, tile_ck1;
    DevBuf<uint8_t> partial;  // nchunks x WU x WV complex T
    DevBuf<double> apod;      // nx*ny
    DevBuf<uint8_t> realbuf;  // nxp*nyp T
    DevBuf<uint8_t> spec;     // nu*half_nv complex T
    DevBuf<uint8_t> visbuf;   // nvis complex T (scratch for host-pointer applies)
    DevBuf<double> imgbuf;    // nx*ny
    rocfft_plan fwd = nullptr, inv = nullptr;
    rocfft_execution_info fwd_info = nullptr, inv_info = nullptr;
    DevBuf<uint8_t> fwd_work, inv_work;

    // batched images (normal operator D^T C D inside the solver): buffers
    // for `batch` images; rocFFT plans per transform count, so that a call
    // transforms only its active samples (rounded up to kBatchStep)
    static constexpr int kBatchStep = 8;
    struct BatchFFT {
        int n = 0, prec = LT_F64;
        rocfft_plan fwd = nullptr, inv = nullptr;
        rocfft_execution_info fi = nullptr, ii = nullptr;
        DevBuf<uint8_t> fw, iw;
        ~BatchFFT();
    };
    int batch = 0;
    std::vector<std::unique_ptr<BatchFFT>> bfft;
    // band-pruned transforms of the stencil path: rows along y, nbv columns
    // along u (see normal_band_impl)
    struct BandFFT {
        int n = 0, prec = LT_F64, nbv = 0;
        rocfft_plan plan[4] = {nullptr, nullptr, nullptr, nullptr};
        rocfft_execution_info info[4] = {nullptr, nullptr, nullptr, nullptr};
        DevBuf<uint8_t> work[4];
        ~BandFFT();
    };
    std::vector<std::unique_ptr<BandFFT>> bandfft;
    DevBuf<uint8_t> bandR, bandC, bandB, bandF;  // bandB / bandF: outputs of the column transforms
    // fused row transforms (power-of-two nyp, 64 .. 2048): twiddle tables
    // exp(-2 pi i e / nyp), e < nyp, complex FP64 and FP32
    DevBuf<double> twd64;
    DevBuf<float> twd32;
    bool fused_rows = true;  // use them where nyp allows
    // nyp = 1024: rows by the register-fed radix-4 transform (k_rows_fwd_r /
    // k_rows_inv_r, the same bits as the LDS form); false: the 256-thread
    // LDS radix-4 form
    bool wave_rows = true;
    // nu = nyp = 1024: the band's column transforms by k_cols_fwd_r /
    // k_cols_inv_r (the same passes); false: rocFFT
    bool own_cols = true;
    int force_nb = 0;  // stencil samples per workgroup when > 0, the matrix-core form with -force_nb
                       // sample groups per workgroup when < 0 (else the autotuned choice)
    bool degrid_half = true;  // degridder fast path with the window taps (taps32h / taps64h)
    lt_status band_fft(int n, int precision, int nbv, BandFFT *&out);
    DevBuf<uint8_t> realB, specB, specB2, partialB;
    lt_status batch_fft(int n, BatchFFT *&out, int precision = -1);

    ~NufftPlan();
    size_t tsize() const { return prec == LT_F64 ? sizeof(double) : sizeof(float); }

    lt_status create(const visibility_space &vs, const fourier_space &fs, int wsup, double kb_beta, int precision);
    lt_status attach_image(const image_space &im, int pad_factor);

    // device pointers; vis are complex T in sorted order
    lt_status degrid(const void *grid, void *vis, const double *weights, bool accumulate);
    lt_status grid(const void *vis, const double *weights, void *grid_out);
    lt_status forward(const double *img, void *vis, const double *weights);
    lt_status adjoint(const void *vis, const double *weights, double *img, double alpha, bool accumulate);
    // out_s = alpha * D^T diag(w) D img_s (+ out_s) for the samples in
    // `active` (device list of sample ids); images are nx*ny float64 each
    lt_status ensure_batch(int nb);
    // rowmask (band stencil form only; nullable): per (sample, image row x
    // index) 0 when that row of the input image is zero and its output is
    // not needed -- the fused row transforms then neither read nor write it
    lt_status normal_batched(const double *imgs, double *out, const double *w, const int *active, int nactive,
                             double alpha, bool accumulate, NormalStencil *stencil = nullptr,
                             const uint8_t *rowmask = nullptr);
    // out_s = crop(irfft2(spec_mul * rfft2(zpad(img_s)))) (BCB), spec_mul is
    // nu x half_nv complex128
    lt_status convolve_batched(const double *imgs, double *out, const double *spec_mul, const int *active,
                               int nactive, double alpha, bool accumulate);
    // FP32 form of the stencil normal operator (precision form of the
    // solver): FP32 images, spectra, rocFFT and stencil coefficients; the
    // result is D^T C D img / stencil.coef32_scale
    lt_status normal_batched32(const float *imgs, float *out, const int *active, int nactive, float alpha,
                               bool accumulate, NormalStencil &stencil, const uint8_t *rowmask = nullptr);
    // permute between caller order (complex128) and sorted complex T
    lt_status to_sorted(const double *vis_c128, void *sorted);
    lt_status from_sorted(const void *sorted, double *vis_c128, bool accumulate);

    // resident per-visibility weights (plan order) and, for FP64 plans with
    // W <= 4, the normal stencil D^T diag(w) D built from them once; solvers
    // created without weights share both
    DevBuf<double> wplan;
    std::unique_ptr<NormalStencil> wstencil;
    int wstencil_tuned_batch = -1;  // batch size the stencil was last autotuned for
    lt_status set_weights(const double *weights_caller_order);
};

// pick the stencil's samples per workgroup by timing the normal operator
lt_status stencil_autotune(NufftPlan &P, NormalStencil &St, const double *imgs, double *out, const int *all, int nb);
lt_status stencil_autotune32(NufftPlan &P, NormalStencil &St, const float *imgs, float *out, const int *all, int nb);

}  // namespace lt
