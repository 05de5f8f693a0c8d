/*
 * lentils_hip.h -- C ABI of liblentils_hip.so, the MI355X (gfx950) forward
 * model + source inversion path of Lentils.
 *
 * Two layers:
 *
 *  1. Drop-in symbols.  Same names, argument meaning and by-value struct
 *     layouts as the reference C backends that lentils/backend/load_backend.py
 *     binds through ctypes; host pointers, caller-owned zero-filled outputs,
 *     void return.  They copy in, run the HIP kernels, copy out.
 *
 *  2. lt_* entry points.  Device-resident handles (context, lens matrices,
 *     NUFFT plan, solver) so that the hot loop never crosses PCIe; every
 *     lt_* function returns an lt_status (0 = success).
 *
 * No torch or HIP types appear in any signature.
 */
#ifndef LENTILS_HIP_H
#define LENTILS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* By-value structs.  Field order and types are the ABI of                   */
/* lentils/backend/common.h:13-61 and lentils/backend/deflect.h:8-26.        */
/* ------------------------------------------------------------------------ */

typedef struct {
    int nx, ny, nchannels, nstokes;
    double xmin, xmax, ymin, ymax;
    double dx, dy;
    double *channels;
    char *mask; /* nx*ny bytes, x-major, nonzero = pixel used */
} image_space;

typedef struct {
    int nu, nv, half_nv, nchannels, nstokes;
    double du, dv, gcx, gcy;
    double *channels;
} fourier_space;

typedef struct {
    int nrows, nchannels, nstokes;
    double *uv;       /* nrows x 3 (u, v, w); w is ignored */
    double *channels; /* per-channel scale applied to u and v */
} visibility_space;

typedef struct {
    unsigned int type; /* 0 = PEMD (2F1 series), 1 = external shear */
    unsigned int _pad;
    double z_l, z_s, d_l, d_s, d_ls, sigma_c, beta;
    double fpars[24];
    int ipars[8];
    unsigned char flags[32];
} generic_mass_model; /* 320 bytes */

typedef struct {
    int num_lenses;
    int num_zplanes;
    generic_mass_model *lenses;
} global_lens_model;

/* lentils/backend/common.h:25-36: a triangulated (Delaunay) source plane */
typedef struct {
    int num_points, num_tris;
    int nchannels, nstokes;
    double *channels;
    double *points;    /* num_points x 2 */
    int *triangles;    /* num_tris x 3 vertex indices */
} delaunay_space;

/* ------------------------------------------------------------------------ */
/* 1. Drop-in symbols                                                         */
/* ------------------------------------------------------------------------ */

/* Replaces deflect_points, lentils/backend/deflect.c:60-66 (bound at
 * load_backend.py:59-61).  p_out = p_in - sum of lens deflections, FP64.
 * want_deriv != 0: as the reference, every point's PEMD parameter
 * derivatives d alpha / d(b_eff, qh, f, x, y, th) (deflect.c:304-338) are
 * added into the same deriv[0..11] as (x, y) pairs, points and components
 * in order; deriv must hold 12 doubles (the reference's caller passes a
 * larger zeroed array). */
void deflect_points(global_lens_model glm, double *p_in, int npoints, double *p_out, int want_deriv,
                    double *deriv);

/* Replaces manifold_lens_matrix_csr, lentils/backend/raster.c:272-395
 * (bound at load_backend.py:122-126).  Writes row_inds[nx*ny+1], cols[nnz],
 * vals[nnz]; the caller's cols/vals must hold nnz entries (query with
 * lt_lens_matrix_nnz), as in the reference. */
void manifold_lens_matrix_csr(image_space imspace, image_space srcspace, global_lens_model lensmodel,
                              int *row_inds, int *cols, double *vals);

/* Replaces grid_cpu, lentils/backend/nufft.c:42-122 (bound at
 * load_backend.py:95-97).  direction 0: vis += degrid(grid);
 * direction 1: grid += grid(vis).  vis and grid are complex128 arrays.
 * Taps beyond the band wrap periodically as in the reference (nufft.c:83-86);
 * such uv sets take a direct per-visibility pass (FP64 atomics when
 * gridding) instead of the tiled, atomic-free gridder. */
void grid_cpu(visibility_space visspace, double *vis, fourier_space gspace, double *grid, int wsup,
              double kb_beta, int direction);

/* Replaces dft_matrix_csr, lentils/backend/nufft.c:128-161
 * (load_backend.py:99-101). */
void dft_matrix_csr(visibility_space visspace, image_space imspace, int *row_inds, int *cols, double *vals);

/* Replaces curv_matrix_findiff_csr, lentils/backend/nufft.c:197-229
 * (load_backend.py:111-113). */
void curv_matrix_findiff_csr(image_space imspace, int *row_inds, int *cols, double *vals);

/* Replaces grad_matrix_findiff_csr, lentils/backend/nufft.c:163-194
 * (load_backend.py:107-109). */
void grad_matrix_findiff_csr(image_space imspace, int *row_inds, int *cols, double *vals);

/* Replaces triangle_gradient_csr, lentils/backend/triangles.c:66-103 (bound
 * at load_backend.py:69-71): two rows (x, y) of three entries per triangle,
 * the gradient of the linear interpolant times sqrt|area| (negated for a
 * clockwise triangle); row_inds[2 num_tris + 1], cols / vals[6 num_tris]. */
void triangle_gradient_csr(delaunay_space trispace, int *row_inds, int *cols, double *vals);

/* Replaces delaunay_lens_matrix_csr, triangles.c:106-145 (load_backend.py:72-74):
 * per image pixel (x-major) nothing when masked out, 1.0 on the next casted
 * vertex for i % ncast == j % ncast == 0, else the barycentric weights of the
 * pixel's deflected position (uncasted_points, in pixel order) in triangle
 * uncasted_tri_inds[k] (nothing when < 0). */
void delaunay_lens_matrix_csr(image_space imspace, delaunay_space trispace, int ncast, double *uncasted_points,
                              int *uncasted_tri_inds, int *row_inds, int *cols, double *vals);

/* Replaces triangle_geometry_from_inds, triangles.c:35-62 (triangles.h:4):
 * barycentric weights (3 per point) of s in the triangles tri_inds of pos. */
void triangle_geometry_from_inds(double *s, int *tri_inds, double *weights_out, int num_points, double *pos);

/* ------------------------------------------------------------------------ */
/* 2. lt_* entry points                                                       */
/* ------------------------------------------------------------------------ */

typedef int lt_status;
#define LT_OK 0
#define LT_ERR_ARG 1        /* invalid argument / shape */
#define LT_ERR_HIP 2        /* HIP runtime error */
#define LT_ERR_NYQUIST 3    /* a kernel tap falls outside the Nyquist band */
#define LT_ERR_UNSUPPORTED 4
#define LT_ERR_CAPACITY 5   /* caller buffer too small */
#define LT_ERR_FFT 6        /* rocFFT error */
#define LT_ERR_NOT_CONVERGED 7

#define LT_F64 0
#define LT_F32 1

/* raster status bits, per lens matrix (the reference only printf's these) */
#define LT_RASTER_STACK_OVERFLOW 1 /* raster.c:262-266 */
#define LT_RASTER_FRAG_OVERFLOW 2  /* raster.c:208-212 */
#define LT_RASTER_POP_LIMIT 4      /* a degenerate split that would never end */
#define LT_RASTER_BAD_LENS 8       /* unsupported lens type */

/* No reference counterpart (the reference is CPU only): library set-up. */
lt_status lt_init(int device);
const char *lt_last_error_message(void);
const char *lt_version(void);

/* Host-pointer deflection with a status code; same arithmetic as
 * deflect_points. */
lt_status lt_deflect_points(const global_lens_model *glm, const double *p_in, long npoints, double *p_out);

/* No reference counterpart (the reference only sums derivatives over all
 * points, deflect.c:60-66): per-point PEMD parameter derivatives.  dalpha
 * receives, per point p and per PEMD component c of the model (in order),
 * the 12 values d alpha / d(b_eff, qh, f, x, y, th) as (x, y) pairs at
 * dalpha[(p * n_pemd + c) * 12 + k], deflect.c:243-338 (b_eff the internal
 * radius of deflect.c:284, held fixed for the other five; th in degrees);
 * p_out as lt_deflect_points. */
lt_status lt_deflect_points_deriv(const global_lens_model *glm, const double *p_in, long npoints, double *p_out,
                                  double *dalpha);

/* Capacity query for manifold_lens_matrix_csr (replaces the 50-per-row guess
 * of lens_operators.py:30-34).  raster_status receives LT_RASTER_* bits. */
lt_status lt_lens_matrix_nnz(image_space imspace, image_space srcspace, global_lens_model lensmodel, long *nnz,
                             int *raster_status);

/* Checked variant of manifold_lens_matrix_csr. */
lt_status lt_lens_matrix_csr(image_space imspace, image_space srcspace, global_lens_model lensmodel, int *row_inds,
                             int *cols, double *vals, long capacity, long *nnz, int *raster_status);

/* Deflected stencil lattice ((2nx+1) x (2ny+1) x 2, x-major) of an image
 * space: the 9-point stencils of raster.c:319-351 share their points through
 * it.  p_out is a host pointer. */
lt_status lt_deflect_lattice(image_space imspace, global_lens_model lensmodel, double *p_out);

/* Lens matrix rasterized from a caller-supplied deflected lattice (host),
 * e.g. the oracle's deflections; used to separate raster parity from
 * deflection rounding. */
lt_status lt_lens_matrix_from_lattice(image_space imspace, image_space srcspace, const double *lattice,
                                      int *row_inds, int *cols, double *vals, long capacity, long *nnz,
                                      int *raster_status);

/* Checked variants of the Delaunay symbols: status codes, triangle and
 * vertex indices validated, lens-matrix capacity (entries; < 0 unchecked)
 * and the entry count returned. */
lt_status lt_triangle_gradient_csr(delaunay_space trispace, int *row_inds, int *cols, double *vals);
lt_status lt_delaunay_lens_matrix_csr(image_space imspace, delaunay_space trispace, int ncast,
                                      const double *uncasted_points, const int *uncasted_tri_inds, int *row_inds,
                                      int *cols, double *vals, long capacity, long *nnz);

/* Stand-alone pieces of the operator surface, host pointers.
 * rfft2 / irfft2 of `batch` real nx x ny arrays (x-major) with the
 * FFTOperator conventions, fourier_operators.py:12-24: forward unnormalised,
 * inverse unnormalised (norm='forward'); the inverse treats columns 0 and
 * ny/2 as Hermitian like numpy/scipy irfft2.  Complex data are interleaved
 * complex128 of shape (nx, ny/2+1). */
lt_status lt_fft_r2c_host(int nx, int ny, int batch, const double *in, double *out);
lt_status lt_fft_c2r_host(int nx, int ny, int batch, const double *in, double *out);

/* y = A x for a CSR matrix given by host arrays (Operator.apply on an
 * explicit matrix, operator_base.py:96-98). */
lt_status lt_csr_spmv_host(long nrows, long ncols, const int *row_ptr, const int *cols, const double *vals,
                           const double *x, double *y);
/* Same with the device kernel chosen: 0 the default (as lt_csr_spmv_host:
 * rows above 2048 entries in 2048-entry chunks); for the others, rows of up
 * to 64 entries by: 1..64 (powers of two) the lane-group kernel with that
 * many lanes per row, 128 the streaming (LDS-staged) kernel, 256 one thread
 * per row over a sliced copy (64-row slices, entries interleaved), -1..-64
 * the solver's L x: rows of 1..4 entries from a padded 4-wide table, rows of
 * 5..64 by that many lanes per row, rows without entries 0; longer rows as
 * in the solver's products: 65..512 entries one wave per row, longer ones in
 * 2048-entry chunks.  Every choice sums each row in a fixed order. */
lt_status lt_csr_spmv_host_kernel(long nrows, long ncols, const int *row_ptr, const int *cols, const double *vals,
                                  const double *x, double *y, int kernel);

/* No reference counterpart: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  y = A x with the CSR arrays (row_ptr nrows+1, cols and vals
 * nnz) and x, y already in device memory: the same kernels and summation
 * order as lt_csr_spmv_host (lane groups from the mean row length, rows of
 * more than 2048 entries in 2048-entry chunks added in order), so the same
 * bits.  The long-row plan and chunk sums live in a fixed 1.4 MiB device
 * array per GPU, set up by lt_init: at most 65536 chunks, which refuses
 * nnz above about 67M with LT_ERR_CAPACITY before touching the device.
 * No hipMalloc, hipFree or allocating library call; waits for the result.
 * Calls that overlap on one GPU race on the shared array. */
lt_status lt_csr_spmv_device_noalloc(long nrows, long ncols, long nnz, const int *d_row_ptr, const int *d_cols,
                                     const double *d_vals, const double *d_x, double *d_y);

/* No reference counterpart: a plain streaming copy of nbytes (multiple of
 * 32) between two device buffers, reps times; mean time per launch.  The
 * achievable-bandwidth ceiling and the calibration of HBM traffic counters. */
lt_status lt_stream_copy(long nbytes, int reps, double *ms_per_launch);

/* --- device-resident handles ---------------------------------------------- */

typedef struct lt_lensbatch_s *lt_lensbatch;
typedef struct lt_nufft_s *lt_nufft;
typedef struct lt_solver_s *lt_solver;

/* A batch of lens matrices L_s (image -> source s), one per lens-parameter
 * sample, built on the GPU in one pass (deflect lattice, rasterize, scan,
 * fill) and kept resident together with L_s^T.  All samples share imspace;
 * srcspaces[s] may differ in size.  lens_counts[s] components of `lenses`
 * (concatenated) belong to sample s. */
lt_status lt_lensbatch_create(image_space imspace, const image_space *srcspaces, int nsamples,
                              const generic_mass_model *lenses, const int *lens_counts, lt_lensbatch *out);
lt_status lt_lensbatch_info(lt_lensbatch lb, int sample, long *nnz, int *raster_status);
lt_status lt_lensbatch_download(lt_lensbatch lb, int sample, int *row_inds, int *cols, double *vals);
lt_status lt_lensbatch_destroy(lt_lensbatch lb);

/* NUFFT plan: gridder * rfft2 * zero-pad * apodization, as NUFFTOperator
 * (fourier_operators.py:96-126).  kb_beta <= 0 selects the reference default
 * pi*sqrt((2W/a)^2 (a-1/2)^2 - 0.8).  precision LT_F64 or LT_F32 for the
 * gridding arithmetic and grid storage.  The uv set is binned into uv tiles
 * once here. */
lt_status lt_nufft_create(visibility_space visspace, image_space imspace, int pad_factor, int wsup,
                          double kb_beta, int precision, lt_nufft *out);
/* host-pointer applies: img is nx*ny float64, vis is nchannels*nrows complex128 */
lt_status lt_nufft_forward_host(lt_nufft h, const double *img, double *vis);
lt_status lt_nufft_adjoint_host(lt_nufft h, const double *vis, double *img);
/* grid / degrid only (the GriddingOperator), host pointers, complex128
 * grid of nu x half_nv; accumulates into the output like grid_cpu */
lt_status lt_nufft_degrid_host(lt_nufft h, const double *grid, double *vis);
lt_status lt_nufft_grid_host(lt_nufft h, const double *vis, double *grid);
lt_status lt_nufft_info(lt_nufft h, int *nu, int *nv, int *half_nv, double *kb_beta, long *nvis);
/* Times `reps` device-resident degrid (direction 0) or grid (1) launches
 * with HIP events on the plan's stream; returns the mean per-launch time. */
lt_status lt_nufft_time(lt_nufft h, int direction, int reps, double *ms_per_launch);
lt_status lt_nufft_destroy(lt_nufft h);
/* No reference counterpart (the reference passes the data object's sigma to
 * every operator): keeps the per-visibility weights (nchannels*nrows
 * float64, caller order) resident in the plan, and for FP64 plans with
 * kernel_support <= 4 builds the normal stencil D^T diag(w) D from them
 * once.  Solvers created with weights == NULL share both. */
lt_status lt_nufft_set_weights(lt_nufft h, const double *weights);
/* No reference counterpart: forces the normal stencil's form for this plan
 * instead of the autotuned choice (0): 1, 2, 4, 8 = the vector form with that
 * many samples per workgroup (all four give the same bits); -1, -2, -4, -8 =
 * the matrix-core form (radius-6 stencils; v_mfma_f64_16x16x4f64 for the FP64
 * operator, v_mfma_f32_16x16x4f32 for the precision form's FP32 operator)
 * with that many groups of 8 samples per workgroup (the same bits among
 * themselves, the vector form's to rounding). */
lt_status lt_nufft_set_stencil_shape(lt_nufft h, int samples_per_group);
/* No reference counterpart: the degridder's tap evaluation for this plan:
 * 1 (default) = the 2W taps of the window below and above the visibility as
 * polynomials of its offset from the cell centre (nothing to select);
 * 0 = the 2W+1 per-offset taps around the nearest grid point, the window's
 * 2W picked by the offset's sign.  Both agree to the precision's rounding;
 * visibilities exactly on a grid line take the 2W+1 taps either way. */
lt_status lt_nufft_set_degrid_form(lt_nufft h, int form);
/* No reference counterpart: the band row transforms of padded rows of 1024
 * for this plan: 1 (default) = a radix-4 transform whose first and last
 * passes run in registers, twiddles from a quarter-period table in LDS;
 * 0 = the 256-thread radix-4 transform staged in LDS.  Both give the same
 * bits, in FP64 and in FP32. */
lt_status lt_nufft_set_row_form(lt_nufft h, int form);
/* No reference counterpart: the band's column transforms along u for this
 * plan when the padded grid is 1024 x 1024: 1 (default) = the row
 * transforms' register-fed radix-4 passes, reading only the image rows of
 * the forward input and writing only the image rows of the inverse output;
 * 0 = rocFFT over whole columns.  The two agree to rounding. */
lt_status lt_nufft_set_col_form(lt_nufft h, int form);
/* No reference counterpart: times the resident normal stencil's kernel as
 * the solver launches it for nb samples (band form, tuned shape) with HIP
 * events on the library stream; mean per-launch time, algorithmic bytes
 * (8 B per stencil coefficient, 16 B per sample per row written and per
 * distinct band value read) and flops (4 per row, offset and sample) per
 * launch, the stencil's rows and distinct window cells (any may be NULL). */
lt_status lt_nufft_stencil_time(lt_nufft h, int nb, int reps, double *ms_per_launch, double *bytes, double *flops,
                                long *rows, long *win_cells);

/* No reference counterpart: the resident stencil's form in use (the forced
 * one, else the autotuned one; the values of lt_nufft_set_stencil_shape). */
lt_status lt_nufft_stencil_form(lt_nufft h, int *form);

/* Device memory for resident inputs (plain hipMalloc'd bytes). */
lt_status lt_device_alloc(long nbytes, void **dptr);
lt_status lt_device_upload(void *dptr, const void *host, long nbytes);
lt_status lt_device_download(const void *dptr, void *host, long nbytes);
lt_status lt_device_free(void *dptr);

/* Replaces RadioDataset.dirty_beam (lentils/common/data_base.py:163-172,
 * beam space data_base.py:148-160): beam = N_b^T C 1 on the 2nx x 2ny beam
 * space with bounds +-dx*n (the x pixel size on both axes), N_b the NUFFT
 * with pad 2 and, for wsup <= 0 / kb_beta <= 0, the reference's defaults
 * (kernel support 4, Beatty's beta, fourier_operators.py:98,108).  weights:
 * nchannels*nrows float64 (sigma^-2 on unmasked visibilities, 0 elsewhere);
 * beam: 4*nx*ny float64 (x-major); peak (may be NULL): its maximum, which
 * equals the weight sum. */
lt_status lt_dirty_beam_host(visibility_space visspace, image_space imspace, const double *weights, int wsup,
                             double kb_beta, double *beam, double *peak);
/* Replaces RadioDataset.blurred_covariance_operator applied to images
 * (data_base.py:115-129): out_b = zpad^T irfft2(rfft2(roll(beam, (nx, ny)))
 * / (4 nx ny) * rfft2(zpad img_b)) for nimg images of nx x ny (even sizes),
 * beam from lt_dirty_beam_host.  Host pointers. */
lt_status lt_bcb_apply_host(image_space imspace, const double *beam, int nimg, const double *img, double *out);

/* No reference counterpart: the per-tap kernel polynomials the gridder
 * evaluates instead of kb_g (nufft.c:11-21).  Tap k of 2W+1 at fractional
 * offset f in [-1/2, 1/2] (distance f + W - k) has value
 * sum_m coeffs[k*(D+1) + m] f^m, D = 16 (LT_F64), 9 (LT_F32, W <= 2) or 7 (LT_F32), times
 * 1/(4W^2) per axis pair; edge taps are zero where |f + W - k| > W.
 * max_err: fit error relative to I0(kb_beta).  Host only. */
lt_status lt_kb_taps(double kb_beta, int wsup, int precision, double *coeffs, double *max_err);
/* No reference counterpart: the degridder's window taps (the default form of
 * lt_nufft_set_degrid_form).  Tap j of the 2W grid cells floor(x) - W + 1 ..
 * floor(x) + W, x = u/du, at s = x - floor(x) - 1/2 in [-1/2, 1/2] (distance
 * s + W - 1/2 - j, always inside the support) has value
 * sum_m coeffs[j*(D+1) + m] s^m, D as for lt_kb_taps.  The kernels evaluate
 * taps 0..W-1 and take tap 2W-1-j at s as tap j at -s.  Host only. */
lt_status lt_kb_taps_window(double kb_beta, int wsup, int precision, double *coeffs, double *max_err);

/* Source-inversion solver for every sample of a lens batch:
 *   (L^T D^T C D L + lambda_curv Cc^T Cc + lambda_tikh I) x = L^T D^T C d
 * with D the NUFFT, C = diag(weights) (sigma^-2 on unmasked visibilities,
 * data_base.py:36-44) and Cc the curvature matrix of each source grid.
 * op_mode LT_OP_NUFFT applies D^T C D exactly: for FP64 plans with
 * kernel_support <= 4 through a grid-space stencil precomputed here from the
 * visibilities and weights, otherwise through a fused per-visibility
 * degrid-weight-grid pass (LT_OP_NUFFT_EXPLICIT forces the latter);
 * LT_OP_BCB uses the dirty-beam convolution (data_base.py:115-129) with the
 * beam of lt_dirty_beam_host at the reference's default NUFFT (kernel
 * support 4, Beatty's beta), whatever the plan's own support. */
#define LT_OP_NUFFT 0          /* exact D^T C D, precomputed grid stencil when possible */
#define LT_OP_BCB 1            /* dirty-beam convolution */
#define LT_OP_NUFFT_EXPLICIT 2 /* exact D^T C D, per-visibility degrid-weight-grid each apply */
#define LT_PC_NONE 0
#define LT_PC_JACOBI 1
#define LT_PC_LEARNED 2 /* lens-structured M (as LT_PC_LENS) plus the network's correction (lt_solver_set_pcnet) */
#define LT_PC_LENS 3    /* M = lambda_c Cc^T Cc + c L_s^T L_s (the reference's R + w L^T L preconditioner,
                           test_all.py:63-65; L_s: L without its longest rows), built explicitly and
                           applied by a few Jacobi-PCG steps on M inside flexible CG
                           (lt_solver_set_lens_pc) */
lt_status lt_solver_create(lt_lensbatch lb, lt_nufft nf, const double *weights, const double *data, double lambda_curv,
                           double lambda_tikh, int op_mode, lt_solver *out);
/* As lt_solver_create, with the data given through their dirty images
 * D^T C d (nx*ny float64 each, e.g. lt_nufft_adjoint_host of weights * data):
 * ndirty = 1 shares one data set between all samples, ndirty = nsamples gives
 * every sample its own (one observation per lens model, same uv coverage). */
lt_status lt_solver_create_dirty(lt_lensbatch lb, lt_nufft nf, const double *weights, const double *dirty,
                                 int ndirty, double lambda_curv, double lambda_tikh, int op_mode, lt_solver *out);
/* weights == NULL in lt_solver_create / lt_solver_create_dirty: the plan's
 * resident weights and stencil (lt_nufft_set_weights; not for LT_OP_BCB).
 * lt_solver_create_resident: as lt_solver_create_dirty with the plan's
 * weights and the dirty images already in device memory (d_dirty: ndirty
 * x nx*ny float64, e.g. from lt_device_alloc), so that nothing crosses
 * PCIe but the lens parameters. */
lt_status lt_solver_create_resident(lt_lensbatch lb, lt_nufft nf, const double *d_dirty, int ndirty,
                                   double lambda_curv, double lambda_tikh, int op_mode, lt_solver *out);
/* CG to ||b - A x|| / ||b|| <= rtol (FP64 true residual).  x: host array
 * of all samples' sources concatenated (may be NULL); iters/resid: per
 * sample (may be NULL). */
lt_status lt_solver_solve(lt_solver s, double rtol, int maxiter, int precond, double *x, int *iters, double *resid);
lt_status lt_solver_rhs(lt_solver s, double *rhs);
/* Precision form of lt_solver_solve: LT_F32 applies A with FP32 values,
 * spectra, rocFFT and stencil coefficients (the SpMV index arrays are
 * shared) inside the FP64 CG (FP64 vectors and inner products), replacing
 * the recurrence residual by the FP64 true residual b - A x every
 * refine_every iterations; a sample whose final FP64 true residual misses
 * rtol restarts from its x.  Needs the stencil operator (LT_OP_NUFFT with
 * an FP64 plan and kernel_support <= 4).  LT_F64 (default): all FP64. */
lt_status lt_solver_set_precision(lt_solver s, int precision, int refine_every);
/* LT_PC_LENS settings: inner Jacobi-PCG steps on M per application (default
 * 3), the row-length limit of L_s (default 64) and the weight c of
 * L_s^T L_s (<= 0: the data operator's response to a unit impulse at the
 * image centre, the default). */
lt_status lt_solver_set_lens_pc(lt_solver s, int inner_steps, int max_row, double c);
/* The LT_PC_LENS matrix M = c L_s^T L_s + lambda Cc^T Cc (L_s: the rows of
 * L with at most max_row nonzero entries), built on first use, as block-diagonal CSR
 * over all samples' source pixels: nnz and the c used; row_ptr (ntot + 1),
 * cols and vals (nnz) are written when non-NULL. */
lt_status lt_solver_lens_pc_matrix(lt_solver s, long *nnz, double *c_used, int *row_ptr, int *cols, double *vals);
/* No reference counterpart: on (default), the FP64 CG (plain, Jacobi and
 * lens-preconditioned; not the precision form or the learned preconditioner)
 * runs its blocks of iterations between convergence checks as captured HIP
 * graphs, one per launch width (active samples rounded up to 8); off: the
 * same kernels launched one by one.  Results are identical. */
lt_status lt_solver_set_graphs(lt_solver s, int on);
/* replace the right-hand sides (all samples concatenated, host) */
/* Flexible CG (LT_PC_LENS, LT_PC_LEARNED) keeps each new direction
 * A-orthogonal to the last `depth` directions (1..32; 1, the default, is the
 * Polak-Ribiere form with the last direction only) and takes
 * alpha = p.r / p.A p; for a preconditioner that changes between
 * iterations this restores the conjugacy that one direction loses
 * (Notay's truncated FCG).  depth > 1 runs the iterations uncaptured. */
lt_status lt_solver_set_flexible_depth(lt_solver s, int depth);
/* LT_PC_LEARNED: the network corrects M for the first n iterations of a
 * solve only, M alone afterwards (0, the default: no limit) */
lt_status lt_solver_set_learned_iterations(lt_solver s, int n);
lt_status lt_solver_set_rhs(lt_solver s, const double *rhs);
/* y = A x for all samples, in the precision set by lt_solver_set_precision */
lt_status lt_solver_apply(lt_solver s, const double *x, double *y);
/* No reference counterpart: times the solver's two lens-matrix products
 * (L x over the nonempty rows, and L^T y, all samples) with HIP events on the
 * library stream, reps launches each after one untimed; mean ms per
 * product and each matrix's (rows, cols, nonzeros) (any may be NULL). */
lt_status lt_solver_time_products(lt_solver s, int reps, double *ms_L, double *ms_LT, long *shape_L,
                                  long *shape_LT);
lt_status lt_solver_destroy(lt_solver s);

/* --- learned preconditioner ------------------------------------------------
 * No reference counterpart (the reference preconditions its CG with a sparse
 * LU of R + w L^T L, test_all.py:63-72).  A small convolutional network on
 * the H x W source grid (H, W multiples of 8): 2 -> 16 -> 32 -> 64 -> 128
 * channels down by 2x2 average pooling, back up by nearest upsampling with
 * skip connections, 3x3 kernels, ReLU; BF16 operands on MFMA, FP32
 * accumulate, FP32 master weights.  Input per sample: the residual over its
 * RMS and diag(A) over its mean; output y, preconditioned direction
 * z = y * rms(r) / mean(diag(A)).  lt_pcnet_param_count() floats of
 * parameters: per layer, weights [co][ci][3][3] then biases [co]. */
typedef struct lt_pcnet_s *lt_pcnet;
long lt_pcnet_param_count(void);
/* params NULL: He-initialised from seed.  batch: largest batch a forward /
 * training step / solve may use. */
lt_status lt_pcnet_create(int H, int W, int batch, const float *params, unsigned seed, lt_pcnet *out);
lt_status lt_pcnet_get_params(lt_pcnet net, float *params);
lt_status lt_pcnet_set_params(lt_pcnet net, const float *params);
/* x0: nb x 2 x H x W, y: nb x H x W (host, FP32) */
lt_status lt_pcnet_forward_host(lt_pcnet net, const float *x0, int nb, float *y);
/* parameter gradient of sum(dy * y(x0)) (and y), for gradient checks */
lt_status lt_pcnet_grad_host(lt_pcnet net, const float *x0, const float *dy, int nb, float *y, float *grad);
/* Adam on mean_s |y_s - t_s|^2 / |t_s|^2 over n pairs (inputs n x 2 x H x W,
 * targets n x H x W), shuffled minibatches, cosine learning-rate decay from
 * lr to lr_final; loss_hist[epoch] (may be NULL). */
lt_status lt_pcnet_train(lt_pcnet net, const float *inputs, const float *targets, long n, int epochs, int batch,
                         float lr, float lr_final, unsigned seed, double *loss_hist);
lt_status lt_pcnet_eval(lt_pcnet net, const float *inputs, const float *targets, long n, double *loss);
lt_status lt_pcnet_destroy(lt_pcnet net);
/* use `net` for LT_PC_LEARNED solves (every sample's source grid must be the
 * network's H x W and the batch within its capacity) */
lt_status lt_solver_set_pcnet(lt_solver s, lt_pcnet net);
/* Train the network (epochs > 0) or evaluate it (epochs = 0) with this
 * solver's operators A_s.  The preconditioner it trains is
 *   z = m + y(r / rms(r), m / rms(m)) rms(m),  m = M^-1 r
 * with M the lens-structured preconditioner of LT_PC_LENS, so an untrained
 * (zero-output) network is exactly LT_PC_LENS.  residuals holds nres vectors
 * of all samples' sources concatenated (host, FP64), e.g. right-hand sides
 * A x of known sources x and the residuals of truncated solves (each such
 * r is the right-hand side of the pair (r, A^-1 r)); each step takes one
 * vector per sample and descends the energy loss
 *   mean_s (z^T A z - 2 z^T r) / (r^T m),
 * the A-norm error ||z - A^-1 r||_A^2 up to a constant.  loss_hist[epoch]:
 * mean loss (for z = m it is (m^T A m - 2 r^T m) / r^T m; its floor, at
 * z = A^-1 r, is -r^T A^-1 r / r^T m). */
lt_status lt_solver_train_pcnet(lt_solver s, const double *residuals, int nres, int epochs, float lr, float lr_final,
                                unsigned seed, double *loss_hist);
/* the diagonal of A by 8 x 8 probing (all samples concatenated) */
lt_status lt_solver_diag(lt_solver s, double *diag);
/* per sample, for the last LT_PC_LEARNED solve: counts = iterations in which
 * the corrected direction was rejected (r.z <= 0 or not finite) and m = M^-1 r
 * taken; plain = 1 if the sample switched its network off for the rest of
 * the solve (its residual norm fell less than 10x in 100 iterations) and
 * continued with M alone (may be NULL) */
lt_status lt_solver_pc_fallbacks(lt_solver s, int *counts, int *plain);

#ifdef __cplusplus
}
#endif

#endif /* LENTILS_HIP_H */
