// kernels_mixed.h — row-fused ALS mode update (ranks 32, 64, 128): BF16-operand
// (mixed precision) or FP64-accumulated (FP32 factors) Gram, FP64 solve.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ctc {

// Work items of one mode: {slot, first nonzero, end nonzero, part}.  part ==
// kLightRow: the whole row, Gram and solve in one workgroup; otherwise the
// chunk writes its partial Gram and right-hand side to partial `part`.
constexpr uint32_t kLightRow = 0xFFFFFFFFu;
// rows with at most this many nonzeros are solved in the dual (n x n) form
constexpr int kThinRow = 32;
// rows with kThinRow < n <= kDualRow nonzeros at R = 128 (dual form, 48 or 64 rows)
constexpr int kDualRow = 64;
constexpr uint32_t kGroup = 32, kGroupMin = 64;

// Gram precision: BF16 operands with FP32 accumulation, or FP32 factors with
// h formed and accumulated in FP64
constexpr int kPrecBF16 = 0, kPrecF64 = 1;

struct MixedArgs {
  int prec;                 // kPrecBF16 / kPrecF64
  const uint32_t *oidx[3];  // other-mode ids in CISS order (CISS-relative modes 1..N-1)
  const float *vals;
  const uint32_t *directory;
  const uint16_t *fb[3];  // BF16 mirrors of the other modes' factors (CISS-relative 1..N-1)
  const float *fr[3];     // the other modes' FP32 factors (kPrecF64)
  float *out;             // factor of the mode being updated, row-major dims x R
  const uint4 *items;
  uint64_t nitems;
  const uint4 *thin;  // thin rows {slot, first, end, -}: n <= min(kThinRow, R)
  uint64_t nthin;
  uint64_t nthin16;   // the first nthin16 of them have n <= 16
  const uint4 *dual;  // rows {slot, first, end, -} with kThinRow < n <= kDualRow (R = 128)
  uint64_t ndual;
  uint64_t ndual48;   // the first ndual48 of them have n <= 48
  void *part;                  // partials of chunks, mixed_partial_bytes(R, prec) each
  const uint32_t *heavy_slot;  // heavy rows: directory slot
  const uint32_t *heavy_pb;    // heavy rows: partials [heavy_pb[h], heavy_pb[h + 1])
  uint64_t nheavy;
  // heavy rows with more than kGroupMin partials are first reduced in groups
  // of kGroup consecutive partials: group g sums {grp[g].x, grp[g].x + grp[g].y)
  // into gpart[g]; heavy row h then sums groups [heavy_gb[h], heavy_gb[h + 1])
  void *gpart;
  const uint2 *grp;
  uint64_t ngroups;
  const uint32_t *heavy_gb;
  double lam;
  int refine;  // iterative-refinement steps after the FP64 Cholesky solve
  uint32_t *flag;
};

// items pass (light rows solved, heavy chunks to partials), then heavy rows
hipError_t launch_mixed_items(hipStream_t st, const MixedArgs &a, int nmodes, int R);
// chunk items only (it.w != kLightRow): Gram into partials, no solve
hipError_t launch_mixed_chunks(hipStream_t st, const MixedArgs &a, int nmodes, int R);
hipError_t launch_mixed_heavy(hipStream_t st, const MixedArgs &a, int nmodes, int R);
hipError_t launch_mixed_groups(hipStream_t st, const MixedArgs &a, int R);
hipError_t launch_rows_thin(hipStream_t st, const MixedArgs &a, int nmodes, int R);
hipError_t launch_rows_dual(hipStream_t st, const MixedArgs &a, int nmodes, int R);
size_t mixed_partial_bytes(int R, int prec);
// FP32 -> BF16 (round to nearest even), n elements
hipError_t launch_to_bf16(hipStream_t st, const float *x, uint16_t *y, uint64_t n);

}  // namespace ctc
