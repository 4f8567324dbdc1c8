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
        float *col = m->FT[md] + (size_t)f * m->dims[md];
        {
          Prof p(m->ctx, "ccd_col");
          CTC_HIP(launch_ccd_col(st, a, (int)N, block));
        }
        CTC_HIP(launch_fill(st, col, m->dims[md], 0.f));  // rows with no nonzero -> 0
        {
          Prof p(m->ctx, "ccd_finish");
          CTC_HIP(launch_ccd_finish(st, pieces, d.slot_pb, d.directory, d.dlength, o->lambda, col,
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

}  // namespace ctc
