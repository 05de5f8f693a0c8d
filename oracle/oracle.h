/*
 * oracle.h -- FP64 CPU restatement of the Lentils C backend, used ONLY as a
 * checker: by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg.  Nothing in the product library (lentils_amd/) links,
 * loads or calls this code.
 *
 * Struct layouts are the ABI the reference passes by value through ctypes
 * (lentils/backend/common.h:13-61, lentils/backend/deflect.h:8-26); they
 * must be byte-identical, so they are restated field for field.
 */
#ifndef LENTILS_ORACLE_H
#define LENTILS_ORACLE_H

#include <stdint.h>

typedef struct {
    int nx, ny, nchannels, nstokes;
    double xmin, xmax, ymin, ymax, dx, dy;
    double *channels;
    char *mask;
} or_image_space;

typedef struct {
    int nu, nv, half_nv, nchannels, nstokes;
    double du, dv, gcx, gcy;
    double *channels;
} or_fourier_space;

typedef struct {
    int nrows, nchannels, nstokes;
    double *uv;        /* (nrows, 3) u, v, w; w unused */
    double *channels;  /* frequency scale factors */
} or_visibility_space;

typedef struct {
    uint32_t type, pad;
    double z_l, z_s, d_l, d_s, d_ls, sigma_c, beta;
    double fpars[24];
    int32_t ipars[8];
    uint8_t flags[32];
} or_mass_model;        /* 320 bytes */

typedef struct {
    int num_lenses, num_zplanes;
    or_mass_model *lenses;
} or_lens_model;

/* status bits reported by the raster oracle (the reference only printf's) */
#define OR_RASTER_STACK_OVERFLOW 1
#define OR_RASTER_FRAG_OVERFLOW 2
#define OR_RASTER_POP_LIMIT 4
#define OR_RASTER_BAD_LENS 8

#endif
