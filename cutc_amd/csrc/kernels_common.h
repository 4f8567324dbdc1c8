// kernels_common.h — device helpers and kernel argument blocks shared by the
// HIP kernels of libctc (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ctc {

constexpr int kTile = 64;  // nonzeros per CISS tile (one wave load)

// CISS of one mode as the kernels see it.  oidx[s] holds the indices of mode
// (mode + 1 + s) % N for every nonzero in CISS order.
struct CissArgs {
  const uint32_t *oidx[3];
  const float *vals;
  const uint32_t *tile_slot;
  const uint64_t *tile_mask;
  const uint32_t *directory;
  const uint32_t *rowptr;
  uint64_t nnz;
  uint64_t ntiles;
  uint64_t dlength;
};

// COO tensor (validation / test) on the device.
struct CooArgs {
  const uint32_t *ind[4];
  const float *vals;
  uint64_t nnz;
};

// factor pointers in CISS-relative order: f[0] = the CISS mode's own factor,
// f[1 + s] = factor of mode (mode + 1 + s) % N.
struct FactorArgs {
  float *f[4];
};

// Hot-row bitmaps, CISS-relative modes: bit (i & 31) of bits[s][i >> 5] set
// when row i of that mode is written racily by the SGD sweep.
struct HotRows {
  const uint32_t *bits[4];
};

// Bulk-synchronous SGD (k_sgd_sync): per CISS-relative mode, the cold rows'
// delta sums in 2^-32 fixed point (dims x R).  A hot row's delta sum and its
// count of runs are spread over kSyncReplicas slots at
// different addresses (walker w adds into slot w mod kSyncReplicas), so that
// the walkers sharing it do not all queue on one memory channel; hidx[row] is
// the row's index among its mode's hot rows (0xFFFFFFFF: cold).
constexpr int kSyncReplicas = 32;
// Modes 1.. (CISS-relative) hand in each run's delta to a staging slot: the
// runs of walker w's chunk in sub-step s take slots sbase[k][s * nspans + w],
// +1, ... in walk order, and k_sgd_settle sums (cold row) or averages
// (hot row) every row's slots in a fixed order.  Mode 0's long rows add
// fixed-point sums (cold) or replicated sums (hot) through D / Dh.
struct SyncArgs {
  long long *D[4];
  const uint32_t *hidx[4];
  long long *Dh[4];   // hot rows x kSyncReplicas x R
  uint32_t *ch[4];    // hot rows x kSyncReplicas
  float *stage[4];    // modes 1..: runs of one sub-step x R (BF16 entries on the mixed path)
  const uint32_t *sbase[4];  // modes 1..: first slot of each (sub-step, walker) chunk
  const uint32_t *dest[4];   // modes 1.., mixed path: the sub-step's staging position of each slot (padded by one)
};

// Staged merge of one sub-step (k_sgd_settle + k_sgd_fold).  Per staged
// mode k, the sub-step's runs are numbered in walk order (slots) and listed
// by row, each row's in walk order: position q of the list is slot
// slots[q].  FP32 deltas are staged at their slot (a walker writes one
// contiguous range; a 128-byte delta is a whole cache line wherever the merge
// reads it); BF16 deltas at their position in the list (dest[slot] = q), so
// that the merge reads a row's runs as one contiguous range and no 128-byte
// line is shared by the runs of two rows.  Every row's list is cut into
// pieces of at most kStagePiece runs: piece p covers positions
// piece_start[p] .. piece_start[p + 1] (numbered over the epoch; the
// sub-step's first is slot0).  A row that fits one piece is settled by that
// piece (piece_row[p] = the row); the pieces of a longer row (piece_row[p] =
// 0xFFFFFFFF) leave FP64 partial sums in part[p - piece0], and k_sgd_fold
// adds them in order for each such row (fold[g]).  A cold row adds the sum of
// its runs' deltas, a hot row their mean.
constexpr uint32_t kStagePiece = 1024;
// rows of at most this many runs in a sub-step are settled by one walker each
// (most rows of a skewed mode); longer ones by a workgroup per piece
constexpr uint32_t kSettleLight = 32;
struct FoldSeg {
  uint32_t row, piece, npiece, runs;  // piece: first piece, sub-step relative
};
// Mode 0's merge: segment g covers merge items [off[g], off[g + 1]) of mode
// mode[g].  A cold segment's item k is row rows[g][k] (rows[g] null: row k),
// whose sum is in D; a hot segment's item k is hot row k of the mode, whose
// sums are spread over the replicas of Dh / ch.
struct MergeArgs {
  float *F[4];
  long long *D[4];
  long long *Dh[4];
  uint32_t *ch[4];
  const uint32_t *hrow[4];  // hot row k of each mode
  const uint32_t *rows[8];
  int mode[8];
  int is_hot[8];
  uint64_t off[9];
  int nseg;
};

struct SettleArgs {
  MergeArgs mg;        // mode 0's long and hot rows
  uint64_t nb_merge;   // k_sgd_fold's first blocks, which take mg's items
  float *F[4];
  const uint32_t *hot[4];  // hot-row bitmaps
  const float *stage[4];
  const uint32_t *piece_start[4];
  const uint32_t *piece_row[4];
  const uint32_t *slots[4];  // FP32: the (sub-step relative) slot at each position
  uint64_t slot0[4];       // the sub-step's first position
  double *part[4];
  const FoldSeg *fold[4];  // the sub-step's multi-piece rows
  uint64_t piece0[4];      // the sub-step's first piece
  const uint32_t *heavy[4];  // the sub-step's pieces settled by a workgroup each (epoch-numbered)
  const uint32_t *light[4];  // the sub-step's rows of at most kSettleLight runs, a walker each
  uint64_t pb[5];          // mode k's heavy pieces: k_sgd_settle's blocks pb[k] .. pb[k + 1]
  uint64_t lb[5];          // mode k's light rows: walkers lb[k] .. lb[k + 1] of the blocks after pb[4]
  uint64_t fb[5];          // mode k's multi-piece rows: k_sgd_fold's walkers fb[k] .. fb[k + 1] after the merge
};

__device__ inline int lane_id() { return (int)(threadIdx.x & 63); }

// FP32 -> BF16, round to nearest even (finite inputs); BF16 halves of a word
__device__ inline uint16_t f2bf(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ inline float bf_lo(uint32_t u) { return __uint_as_float(u << 16); }
__device__ inline float bf_hi(uint32_t u) { return __uint_as_float(u & 0xFFFF0000u); }

// 1/x to full FP64 precision: hardware reciprocal estimate + two Newton steps
__device__ inline double rcp_f64(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
  return __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
}

// LDS ordering between lanes of one wave (no workgroup barrier)
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sum across groups of W consecutive lanes (W in {1,2,...,64}); every lane of
// the group receives the total.  DPP quad permutes for distances 1 and 2,
// ds_swizzle (xor mode, within 32 lanes) for 4..16, ds_bpermute across halves.
template <int W>
__device__ inline float group_sum(float v) {
  if constexpr (W >= 2)
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));
  if constexpr (W >= 4)
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, false));
  if constexpr (W >= 8) v += __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x101F));
  if constexpr (W >= 16) v += __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x201F));
  if constexpr (W >= 32) v += __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x401F));
  if constexpr (W >= 64) v += __shfl_xor(v, 32, 64);
  return v;
}

__device__ inline double wave_sum_d(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Hardware float atomic add (global_atomic_add_f32) on device memory.
__device__ inline void atomic_add_f32(float *p, float v) { unsafeAtomicAdd(p, v); }

// slot of nonzero e (0..63) of a tile whose first nonzero has slot `first`
__device__ inline uint32_t tile_lane_slot(uint32_t first, uint64_t mask, int e) {
  uint64_t below = (e == 63) ? ~0ull : ((2ull << e) - 1ull);
  return first + (uint32_t)__popcll(mask & below & ~1ull);
}

}  // namespace ctc
