// Device-resident batch of lens matrices (one per lens-parameter sample).
#pragma once

#include "lt_internal.h"
#include "raster_core.h"

namespace lt {

struct LensBatch {
    image_space im{};
    int nsamples = 0;
    long npix = 0;
    int nlong_rows = 0;
    std::vector<SrcGrid> src_host;
    std::vector<int64_t> srcoff_host;  // cumulative source sizes (ns + 1)
    std::vector<int64_t> nnzoff_host;  // cumulative nonzeros (ns + 1)
    std::vector<int> status_host;      // LT_RASTER_* bits per sample
    DevBuf<uint8_t> mask;
    DevBuf<SrcGrid> src;
    DevBuf<int64_t> srcoff;
    DevBuf<double> lattice;  // deflected stencil lattices, ns x (2nx+1)(2ny+1) x 2
    DevCSR L;                // block diagonal: rows ns*npix, cols sum(nsrc)
    DevCSR Lp;               // L without its exact-zero entries (the solver's operator), built on demand
    DevCSR LT;               // transpose of Lp, built on demand

    // lattice_host != nullptr: rasterize given lattices instead of deflecting
    lt_status build(const image_space &im, const image_space *srcs, int ns, const generic_mass_model *lenses,
                    const int *lens_counts, const double *lattice_host);
    lt_status download(int s, int *row_inds, int *cols, double *vals) const;
    // Lp and its transpose.  Rows without fragments hold one explicit zero
    // (the stale column, raster.c:388-391) and a dropped first fragment can
    // leave zero-valued entries; they add nothing to L x but make the stale
    // columns very long rows of the transpose.  L itself stays as built.
    lt_status ensure_transpose() {
        if (LT.nrows == 0 && L.nrows > 0) {
            LT_TRY(csr_drop_zeros(L, Lp));
            Lp.long_row = 64;  // the few lens-centre rows beyond this: one workgroup per 2048-entry chunk
            LT_TRY(csr_prepare(Lp));
            LT_TRY(csr_nonempty_rows(Lp));
            LT_TRY(csr_transpose(Lp, LT));
            // rows up to 256 entries summed where they are (the streaming
            // kernel stages 256 nonzeros past its range), 257..512 one wave
            // each, chunks beyond
            LT.long_row = 256;
            LT_TRY(csr_prepare(LT));
        }
        return LT_OK;
    }
};

}  // namespace lt
