// ccd.cpp — CCD++ epoch driver (kernels in kernels_ccd.hip).
//
// Canonical CCD++ (SG/ccd.cu:532-554, NC/ccd.cu:28-180) with NUM_INNER inner
// iterations: for f in 0..R-1: residual += component f (old values); for each
// inner iteration and mode m: column f of mode m <- sum(r h) / (lambda +
// sum h^2) over each row; residual -= component f (new values).  The
// subtraction of column f is deferred into the pass of column f+1 (or the next
// epoch's column 0) so each (column, mode) costs one pass over the tensor.
// At the start of column f every mode's per-row records are packed from the
// column-major mirror: (pending column's new value, f saved, f current).
//
// ccd_epoch_jacobi is the reference's own CCD form, run on request only.
#include <algorithm>
#include <vector>

#include "ctc_internal.h"
#include "kernels_ccd.h"
#include "kernels_launch.h"
#include "model_util.h"

namespace ctc {

ctc_status ccd_epoch_impl(ctc_model *m, const ctc_options *o) {
  const uint32_t N = m->nmodes, R = m->rank;
  const uint32_t S = span_tiles_of(o);
  const uint32_t inner = o->inner_iters ? o->inner_iters : 1;
  hipStream_t st = m->ctx->stream;
  size_t max_pieces = 0;
  for (uint32_t md = 0; md < N; ++md) {
    CTC_TRY(ensure_ciss(m, md));
    CTC_TRY(ensure_pieces(m, md, S));
    max_pieces = std::max<size_t>(max_pieces, m->ciss[md].npieces);
    if (!m->FT[md]) CTC_HIP(hipMalloc(&m->FT[md], m->dims[md] * R * sizeof(float)));
    if (!m->ccdx[md]) CTC_HIP(hipMalloc(&m->ccdx[md], std::max<uint64_t>(1, m->dims[md]) * 4 * sizeof(float)));
    if (!m->resid[md]) {
      CTC_HIP(hipMalloc(&m->resid[md], m->ciss[md].nnz * sizeof(float)));
      m->resid_valid = false;
    }
  }
  CTC_TRY(ensure_scratch(m, max_pieces * 2 * sizeof(double)));
  double *pieces = (double *)m->scratch;

  if (!m->resid_valid) {
    for (uint32_t md = 0; md < N; ++md)
      CTC_HIP(launch_ccd_resid(st, ciss_args(m->ciss[md]), rel_factors(m->F, N, md), (int)N, (int)R,
                               m->resid[md]));
    m->ccd_pending = -1;
    m->resid_valid = true;
  }
  for (uint32_t md = 0; md < N; ++md)
    CTC_HIP(launch_transpose(st, m->F[md], m->FT[md], m->dims[md], (int)R, true));

  const unsigned block = block_of(o);
  for (uint32_t f = 0; f < R; ++f) {
    for (uint32_t md = 0; md < N; ++md)
      CTC_HIP(launch_ccd_pack(st, m->FT[md], m->dims[md], (int)f, m->ccd_pending, (float4 *)m->ccdx[md]));
    for (uint32_t it = 0; it < inner; ++it) {
      for (uint32_t md = 0; md < N; ++md) {
        const DevCiss &d = m->ciss[md];
        CcdColArgs a{};
        a.c = ciss_args(d);
        for (uint32_t s = 0; s < N; ++s) a.x[s] = (const float4 *)m->ccdx[(md + s) % N];
        a.sub_col = it == 0 ? m->ccd_pending : -1;
        a.add_back = it == 0 ? 1 : 0;
        a.resid = m->resid[md];
        a.span_tiles = S;
        a.nspans = d.nspans;
        a.span_off = d.span_off;
        a.pieces = pieces;
        a.lds_s = -1;  // a small other mode's records are read from LDS (3-way tensors)
        if (N == 3)
          for (int s = 0; s < 2; ++s)
            if (m->dims[(md + 1 + s) % N] <= kCcdLdsRows) {
              a.lds_s = s;
              a.lds_rows = (uint32_t)m->dims[(md + 1 + s) % N];
              break;
            }
        float *col = m->FT[md] + (size_t)f * m->dims[md];
        {
          Prof p(m->ctx, "ccd_col");
          CTC_HIP(launch_ccd_col(st, a, (int)N, block));
        }
        CTC_HIP(launch_fill(st, col, m->dims[md], 0.f));  // rows with no nonzero -> 0
        {
          Prof p(m->ctx, "ccd_finish");
          CTC_HIP(launch_ccd_finish(st, pieces, d.slot_pb, d.directory, d.dlength, d.npieces, o->lambda, col,
                                    (float4 *)m->ccdx[md]));
        }
      }
    }
    m->ccd_pending = (int)f;
  }
  for (uint32_t md = 0; md < N; ++md)
    CTC_HIP(launch_transpose(st, m->FT[md], m->F[md], m->dims[md], (int)R, false));
  return CTC_OK;
}

// The reference's CCD (tc_ccd, SG/ccd.cu:504-560), for ref_quirks: one mode
// after another, num / den zeroed (:548-549), the Jacobi walk sums e h and
// h * h per row and column from the factors as they are (update_frac_gpu_as,
// with the tile defects as the weights of the quirk CISS), then every row of
// the mode's directory takes num / (lambda + den) (update_ccd_gpu).  A row
// whose every term was dropped becomes 0; rows without nonzeros are kept.
ctc_status ccd_epoch_jacobi(ctc_model *m, const ctc_options *o, uint32_t start) {
  const uint32_t N = m->nmodes, R = m->rank;
  const uint32_t wflags =
      o->ref_quirks & (CTC_QUIRK_TILE_HEAD | CTC_QUIRK_LAST_LANE | CTC_QUIRK_TRUNC_VALUES | CTC_QUIRK_CCD_GRID);
  hipStream_t st = m->ctx->stream;
  uint64_t maxdim = 0;
  for (uint32_t md = 0; md < N; ++md) maxdim = std::max<uint64_t>(maxdim, m->dims[md]);
  CTC_TRY(ensure_scratch(m, 2 * maxdim * R * sizeof(float)));
  float *num = (float *)m->scratch, *den = num + maxdim * R;
  for (uint32_t k = 0; k < N; ++k) {
    const uint32_t md = (start + k) % N;
    CTC_TRY(ensure_ciss(m, md));
    if (wflags) CTC_TRY(ensure_quirk_ciss(m, md, wflags));
    const DevCiss &walk = wflags ? m->qciss[md] : m->ciss[md];
    const DevCiss &d = m->ciss[md];
    CTC_HIP(launch_fill(st, num, m->dims[md] * R, 0.f));
    CTC_HIP(launch_fill(st, den, m->dims[md] * R, 0.f));
    {
      Prof p(m->ctx, "ccd_jacobi");
      CTC_HIP(launch_ccd_jacobi(st, ciss_args(walk), rel_factors(m->F, N, md), (int)N, (int)R, 256, num, den, 256));
    }
    CTC_HIP(launch_ccd_jacobi_rows(st, d.directory, d.dlength, (int)R, num, den, o->lambda, m->F[md]));
  }
  m->resid_valid = false;
  return CTC_OK;
}

}  // namespace ctc
