// kernels_ccd.h — CCD++ kernel arguments and launchers (kernels_ccd.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "kernels_common.h"

namespace ctc {

struct CcdColArgs {
  CissArgs c;                 // CISS of the mode being updated
  const float *ft[4];         // column-major factors, CISS-relative mode order
  uint64_t dims[4];           // rows of each, CISS-relative
  const float *saved[4];      // pre-update column `col` of each mode, CISS-relative
  int col;                    // column being updated
  int sub_col;                // column whose new component is subtracted first (-1: none)
  int add_back;               // add column `col` back with its saved values
  float *resid;               // this mode's residual copy (CISS order)
  uint32_t span_tiles;
  uint64_t nspans;
  const uint32_t *span_off;
  double *pieces;             // (num, den) per piece
};

hipError_t launch_ccd_col(hipStream_t st, const CcdColArgs &a, int nmodes, unsigned block);
hipError_t launch_ccd_finish(hipStream_t st, const double *pieces, const uint32_t *slot_pb,
                             const uint32_t *directory, uint64_t dlength, double lam, float *col);
hipError_t launch_ccd_resid(hipStream_t st, const CissArgs &c, const FactorArgs &f, int nmodes, int R,
                            float *resid);
hipError_t launch_transpose(hipStream_t st, const float *in, float *out, uint64_t rows, int R,
                            bool to_colmajor);

}  // namespace ctc
