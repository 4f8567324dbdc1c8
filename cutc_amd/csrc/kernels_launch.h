// kernels_launch.h — host-side launchers of the HIP kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "kernels_common.h"

namespace ctc {

// code = PA * 16 + PB * 4 + PC with P in {0 racy, 1 atomic, 2 run, 3 hot-row bitmap} (see k_sgd)
hipError_t launch_sgd(hipStream_t st, const CissArgs &c, const FactorArgs &f, const HotRows &hot, int nmodes, int R,
                      uint32_t span, uint64_t stride, int code, float step, float lam, unsigned block);
// one sub-step s of the bulk-synchronous SGD schedule (k_sgd_sync); then
// launch_sgd_settle merges the rows' handed-in deltas (k_sgd_settle, and
// k_sgd_fold for the rows of more than one piece)
hipError_t launch_sgd_sync(hipStream_t st, const CissArgs &c, const FactorArgs &f, const SyncArgs &sy, int nmodes,
                           int R, uint32_t span, uint32_t sub, uint32_t s, uint32_t long_min, float step, float lam,
                           unsigned block, bool mixed);
// nb_merge is set here; pb / fb must be prefix sums with pb[0..1] = fb[0..1] = 0
hipError_t launch_sgd_settle(hipStream_t st, SettleArgs a, int R, bool mixed);
hipError_t launch_mttkrp(hipStream_t st, const CissArgs &c, const FactorArgs &f, int nmodes, int R,
                         uint32_t span, bool adjoint, float *out, double *loss_part,
                         unsigned *grid_out, unsigned block);
uint64_t mttkrp_grid(uint64_t nnz, int R, uint32_t span, unsigned block);
// the training loss as a walk of the CISS c (rows kept along their runs):
// part[b] = the squared residuals of block b, for the mttkrp_grid(c.nnz, R,
// span, block) blocks
hipError_t launch_loss_walk(hipStream_t st, const CissArgs &c, const FactorArgs &f, int nmodes, int R,
                            uint32_t span, double *part, unsigned block);
// walkers of k_mttkrp per workgroup of `block` threads (R / 4 lanes each)
uint64_t mttkrp_walkers_per_block(int R, unsigned block);
// the reference's CCD sums of one mode (SG/ccd.cu:157-195) over the CISS c:
// num += sum of e h, den += sum of h * h per row and column, e the residual
// of the factors as they are; num / den are zeroed dims x R arrays
hipError_t launch_ccd_jacobi(hipStream_t st, const CissArgs &c, const FactorArgs &f, int nmodes, int R,
                             uint32_t span, float *num, float *den, unsigned block);
hipError_t launch_als_gram(hipStream_t st, const CissArgs &c, const FactorArgs &f, int nmodes, int R,
                           uint32_t span_tiles, uint64_t nspans, const uint32_t *span_off,
                           double *pieces, unsigned block);
hipError_t launch_piece_reduce(hipStream_t st, const double *pieces, const uint32_t *grp_piece,
                               const uint32_t *grp_n, uint64_t ngroups, int R, double *gsum);
hipError_t launch_als_solve(hipStream_t st, const double *pieces, const uint32_t *slot_pb, const double *gsum,
                            const uint32_t *slot_gb, const uint32_t *directory, uint64_t dlength, int R,
                            double lam, bool refine, bool light, float *A, uint32_t *flag);
hipError_t launch_eval_coo(hipStream_t st, const CooArgs &c, const FactorArgs &f, int nmodes, int R,
                           unsigned nblocks, double *part);
// per-segment squared error: exactly one of c (mode-0 CISS order) / co (COO) is non-null
hipError_t launch_eval_seg(hipStream_t st, const CissArgs *c, const CooArgs *co, const FactorArgs &f,
                           int nmodes, int R, const uint64_t *seg, uint32_t nseg, double *out);
hipError_t launch_sumsq(hipStream_t st, const float *x, uint64_t n, unsigned nblocks, double *part);
hipError_t launch_reduce_parts(hipStream_t st, const double *part, uint32_t nb, uint32_t groups,
                               double *out);
hipError_t launch_fill(hipStream_t st, float *x, uint64_t n, float v);
hipError_t launch_copy(hipStream_t st, const void *src, void *dst, uint64_t bytes);  // bytes % 16 == 0
hipError_t launch_reg_rows(hipStream_t st, const uint32_t *directory, const uint32_t *rowptr,
                           uint64_t nrows, int R, float lam, int per_occ, const float *F, float *G,
                           unsigned nblocks, double *part);

}  // namespace ctc
