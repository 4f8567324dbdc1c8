// kernels_ccd.h — CCD++ kernel arguments and launchers (kernels_ccd.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "kernels_common.h"

namespace ctc {

struct CcdColArgs {
  CissArgs c;                 // CISS of the mode being updated
  const float4 *x[4];         // per-row records (pending, saved, current, -), CISS-relative
  int sub_col;                // a pending column's new component is subtracted first (-1: none)
  int add_back;               // add the column back with its saved values
  float *resid;               // this mode's residual copy (CISS order)
  uint32_t span_tiles;
  uint64_t nspans;
  const uint32_t *span_off;
  double *pieces;             // (num, den) per piece
  int lds_s;                  // other mode (CISS-relative 1 + lds_s) read from LDS, or -1
  uint32_t lds_rows;          // its rows (at most kCcdLdsRows)
};
constexpr uint32_t kCcdLdsRows = 2304;  // 36 KB of 16-byte records

hipError_t launch_ccd_col(hipStream_t st, const CcdColArgs &a, int nmodes, unsigned block);
// npieces: the mode's piece count (picks one lane or one wave per slot)
hipError_t launch_ccd_finish(hipStream_t st, const double *pieces, const uint32_t *slot_pb,
                             const uint32_t *directory, uint64_t dlength, uint64_t npieces, double lam,
                             float *col, float4 *x);
hipError_t launch_ccd_pack(hipStream_t st, const float *ft, uint64_t rows, int col, int sub_col, float4 *x);
hipError_t launch_ccd_resid(hipStream_t st, const CissArgs &c, const FactorArgs &f, int nmodes, int R,
                            float *resid);
// F(row, :) <- num(row, :) / (lambda + den(row, :)) for the rows of a
// directory (SG/ccd.cu:318-348)
hipError_t launch_ccd_jacobi_rows(hipStream_t st, const uint32_t *directory, uint64_t dlength, int R,
                                  const float *num, const float *den, double lam, float *F);
hipError_t launch_transpose(hipStream_t st, const float *in, float *out, uint64_t rows, int R,
                            bool to_colmajor);

}  // namespace ctc
