// ctc_internal.h — shared host-side declarations of libctc (not installed).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/ctc.h"

namespace ctc {
struct FoldSeg;  // kernels_common.h
}

namespace ctc {

// ---------------------------------------------------------------- errors --
void set_error(const char *fmt, ...);
ctc_status fail(ctc_status s, const char *fmt, ...);

#define CTC_HIP(call)                                                                 \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      return ::ctc::fail(CTC_ERR_HIP, "%s failed: %s (%s:%d)", #call,                 \
                         hipGetErrorString(e_), __FILE__, __LINE__);                  \
  } while (0)

#define CTC_TRY(call)                                                                 \
  do {                                                                                \
    ctc_status s_ = (call);                                                           \
    if (s_ != CTC_OK) return s_;                                                      \
  } while (0)

// ------------------------------------------------------------- host utils --
unsigned host_threads();

// Stable counting sort of `order` by key[order[p]] in [0, dim).
void counting_sort(const uint32_t *key, uint64_t dim, std::vector<uint32_t> &order,
                   std::vector<uint32_t> &tmp);

// ------------------------------------------------------------ host CISS --
struct HostCiss {
  uint32_t nmodes = 0, mode = 0;
  uint64_t nnz = 0, ntiles = 0, dlength = 0;
  uint64_t dims[CTC_MAX_MODES] = {0, 0, 0, 0};
  std::vector<uint32_t> directory, rowptr, perm, tile_slot;
  std::vector<uint32_t> oidx[CTC_MAX_MODES - 1];
  std::vector<float> vals;
  std::vector<uint64_t> tile_mask;
};
ctc_status build_host_ciss(const ctc_coo *t, uint32_t mode, HostCiss &out);
// the nonzeros of t repeated by their ref_quirks weights for `mode`
// (quirks.cpp); *out = nullptr when every weight is 0
ctc_status quirk_multiset(const ctc_coo *t, uint32_t mode, uint32_t flags, ctc_coo **out);

// ----------------------------------------------------------- device CISS --
// Device copy of one mode's CISS plus the per-span piece tables used by the
// deterministic segmented reductions (see kernels.hip).
struct DevCiss {
  uint32_t nmodes = 0, mode = 0;
  uint64_t nnz = 0, ntiles = 0, dlength = 0;
  uint32_t *oidx[CTC_MAX_MODES - 1] = {nullptr, nullptr, nullptr};
  float *vals = nullptr;
  uint32_t *tile_slot = nullptr;
  uint64_t *tile_mask = nullptr;
  uint32_t *directory = nullptr;
  uint32_t *rowptr = nullptr;
  // piece tables for span_tiles tiles per span
  uint32_t span_tiles = 0;
  uint64_t nspans = 0, npieces = 0;
  uint32_t *span_off = nullptr;  // nspans: piece id = slot + span_off[span]
  uint32_t *slot_pb = nullptr;   // dlength + 1: pieces of slot s
  // heavy slots (more than kHeavyPieces pieces): their pieces are summed in
  // groups of kGroupPieces consecutive pieces before the ALS solve
  uint64_t ngroups = 0;
  uint32_t *grp_piece = nullptr;  // ngroups: first piece of group g
  uint32_t *grp_n = nullptr;      // ngroups: pieces in group g
  uint32_t *slot_gb = nullptr;    // dlength + 1: groups of slot s (empty for light slots)
  // BF16-operand ALS work items (see kernels_mixed.h) for rows cut at mix_chunk nonzeros
  uint32_t mix_chunk = 0;
  uint64_t mix_nitems = 0, mix_nheavy = 0, mix_nparts = 0, mix_nthin = 0, mix_nthin16 = 0, mix_ndual = 0, mix_ndual48 = 0,
           mix_ngroups = 0;
  uint4 *mix_items = nullptr;
  uint32_t *mix_heavy_slot = nullptr, *mix_heavy_pb = nullptr, *mix_heavy_gb = nullptr;
  uint2 *mix_grp = nullptr;  // long rows' partial groups {first partial, count}
  // host copies needed by the piece-table builder
  std::vector<uint32_t> h_tile_slot;
  std::vector<uint64_t> h_tile_mask;
  std::vector<uint32_t> h_rowptr;
};

// Device COO used for evaluation (validation / test).
struct DevCoo {
  uint32_t nmodes = 0;
  uint64_t nnz = 0;
  uint32_t *ind[CTC_MAX_MODES] = {nullptr, nullptr, nullptr, nullptr};
  float *vals = nullptr;
};

}  // namespace ctc

// ------------------------------------------------------- opaque handles --
struct ctc_ciss {
  ctc::HostCiss h;
};

struct KernelStat {
  uint64_t launches = 0;
  double total_ms = 0.0;
};

struct ctc_context {
  // holders: the handle returned by ctc_context_create plus every model and
  // batch created on it; the context is freed when the last one lets go, so
  // handles may be destroyed in any order
  int refs = 1;
  int device = 0;
  hipStream_t stream = nullptr;
  bool profiling = false;
  uint32_t prof_every = 1;                  // time every n-th launch of each name
  std::map<std::string, uint64_t> prof_seen;  // launches of each name since profiling was set
  std::map<std::string, KernelStat> stats;
  // pending profiled launches: name, start, stop
  struct Pending {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> event_pool;
  hipDeviceProp_t props{};
};

void context_retain(ctc_context *c);
void context_release(ctc_context *c);

struct ctc_model {
  ctc_context *ctx = nullptr;
  uint32_t nmodes = 0, rank = 0;
  uint64_t dims[CTC_MAX_MODES] = {0, 0, 0, 0};
  uint64_t train_nnz = 0;
  // host copy of the training tensor (CISS copies are built from it on demand)
  std::unique_ptr<ctc_coo, void (*)(ctc_coo *)> train{nullptr, ctc_coo_free};
  ctc::DevCiss ciss[CTC_MAX_MODES];
  bool ciss_ready[CTC_MAX_MODES] = {false, false, false, false};
  // ref_quirks: each mode's CISS of the nonzeros repeated by their weights,
  // and the quirk bits it was built for (0: not built)
  ctc::DevCiss qciss[CTC_MAX_MODES];
  uint32_t qciss_flags[CTC_MAX_MODES] = {0, 0, 0, 0};
  ctc::DevCoo valid, test;
  float *F[CTC_MAX_MODES] = {nullptr, nullptr, nullptr, nullptr};
  float *best[CTC_MAX_MODES] = {nullptr, nullptr, nullptr, nullptr};
  // scratch
  void *scratch = nullptr;
  size_t scratch_bytes = 0;
  double *dstat = nullptr;  // small device buffer for reductions (4096 doubles)
  double *eval_part = nullptr;  // per-block squared residuals of the training-loss walk
  uint64_t eval_part_n = 0;
  double *hstat = nullptr;  // pinned host mirror
  uint32_t *dflag = nullptr;  // device status flags (not-SPD etc.)
  // CCD++ residuals per mode (in that mode's CISS order)
  float *resid[CTC_MAX_MODES] = {nullptr, nullptr, nullptr, nullptr};
  bool resid_valid = false;
  // CCD++ per-row records of the column being updated, 4 floats per row:
  // (pending column's new value, saved value, current value, unused)
  float *ccdx[CTC_MAX_MODES] = {nullptr, nullptr, nullptr, nullptr};
  float *FT[CTC_MAX_MODES] = {nullptr, nullptr, nullptr, nullptr};  // column-major mirrors
  int ccd_pending = -1;  // column whose new component the residual copies still hold
  // SGD hot-row bitmaps (rows with more nonzeros than hot_threshold)
  uint32_t *hotbits[CTC_MAX_MODES] = {nullptr, nullptr, nullptr, nullptr};
  // BF16 mirrors of the factors (mixed-precision ALS)
  uint16_t *Fbf[CTC_MAX_MODES] = {nullptr, nullptr, nullptr, nullptr};
  uint64_t hot_threshold = 0;
  // bulk-synchronous SGD: per mode, the cold rows' fixed-point delta sums
  // (zero between sub-steps), and the mode-0 rows long enough to meet two
  // walkers in one sub-step (for the span / sub-step they were built for)
  long long *sgd_D[CTC_MAX_MODES] = {nullptr, nullptr, nullptr, nullptr};
  // hot rows: index map (dims, 0xFFFFFFFF cold), row list, replicated sums
  uint32_t *sgd_hidx[CTC_MAX_MODES] = {nullptr, nullptr, nullptr, nullptr};
  uint32_t *sgd_hrow[CTC_MAX_MODES] = {nullptr, nullptr, nullptr, nullptr};
  long long *sgd_Dh[CTC_MAX_MODES] = {nullptr, nullptr, nullptr, nullptr};
  uint32_t *sgd_ch[CTC_MAX_MODES] = {nullptr, nullptr, nullptr, nullptr};
  uint64_t sgd_nhot[CTC_MAX_MODES] = {0, 0, 0, 0};
  // modes 1..N-1: staging slots (runs of the largest sub-step x R), each
  // chunk's first slot (sub-steps x walkers), every row's slots grouped per
  // sub-step and cut into pieces (first slot and whole row of each piece),
  // the pieces' partial sums, the rows of several pieces, and per sub-step
  // the first piece and first multi-piece row (host)
  float *sgd_stage[CTC_MAX_MODES] = {nullptr, nullptr, nullptr, nullptr};
  uint32_t *sgd_sbase[CTC_MAX_MODES] = {nullptr, nullptr, nullptr, nullptr};
  uint32_t *sgd_slots[CTC_MAX_MODES] = {nullptr, nullptr, nullptr, nullptr};
  uint32_t *sgd_dest[CTC_MAX_MODES] = {nullptr, nullptr, nullptr, nullptr};
  uint32_t *sgd_piece_start[CTC_MAX_MODES] = {nullptr, nullptr, nullptr, nullptr};
  uint32_t *sgd_piece_row[CTC_MAX_MODES] = {nullptr, nullptr, nullptr, nullptr};
  double *sgd_part[CTC_MAX_MODES] = {nullptr, nullptr, nullptr, nullptr};
  ctc::FoldSeg *sgd_fold[CTC_MAX_MODES] = {nullptr, nullptr, nullptr, nullptr};
  // the pieces settled by a workgroup each (heavy) and the rows of at most
  // kSettleLight runs settled by one walker each (light), per sub-step
  uint32_t *sgd_heavy[CTC_MAX_MODES] = {nullptr, nullptr, nullptr, nullptr};
  uint32_t *sgd_light[CTC_MAX_MODES] = {nullptr, nullptr, nullptr, nullptr};
  std::vector<uint64_t> sgd_heavy_off[CTC_MAX_MODES];
  std::vector<uint64_t> sgd_light_off[CTC_MAX_MODES];
  std::vector<uint64_t> sgd_slot_off[CTC_MAX_MODES];  // first staging position of each sub-step
  std::vector<uint64_t> sgd_piece_off[CTC_MAX_MODES];
  std::vector<uint64_t> sgd_fold_off[CTC_MAX_MODES];
  uint32_t *sgd_long = nullptr;
  uint64_t sgd_nlong = 0;
  uint32_t sgd_long_span = 0, sgd_long_sub = 0;
  uint64_t sgd_long_thr = 0;  // hot threshold the staged groups were built for

};

namespace ctc {
ctc_status ensure_scratch(ctc_model *m, size_t bytes);
ctc_status ensure_ciss(ctc_model *m, uint32_t mode);
ctc_status ensure_pieces(ctc_model *m, uint32_t mode, uint32_t span_tiles);
// every ctc_ref_quirk bit
constexpr uint32_t kQuirkAll = CTC_QUIRK_TILE_HEAD | CTC_QUIRK_LAST_LANE | CTC_QUIRK_TRUNC_VALUES |
                               CTC_QUIRK_NO_ADDBACK | CTC_QUIRK_CCD_GRID;
// the device CISS of `mode`'s nonzeros repeated by their quirk weights under
// `flags` (model.cpp), kept in m->qciss[mode] until other flags are asked for;
// it has nnz = 0 (and no arrays) when every weight is 0
ctc_status ensure_quirk_ciss(ctc_model *m, uint32_t mode, uint32_t flags);
// one CCD++ epoch on the device (ccd.cpp)
ctc_status ccd_epoch_impl(ctc_model *m, const ctc_options *o);
// one epoch of the reference's CCD form (ref_quirks with CTC_QUIRK_NO_ADDBACK;
// ccd.cpp), modes start, start + 1, ... in turn
ctc_status ccd_epoch_jacobi(ctc_model *m, const ctc_options *o, uint32_t start);
// one row-fused ALS mode update (ranks 32, 64, 128), enqueued on the model's
// stream: BF16-operand Gram (bf16) or FP64-accumulated Gram from FP32 factors
ctc_status als_mode_rows(ctc_model *m, uint32_t mode, const ctc_options *o, bool bf16);
// everything als_mode_rows allocates or uploads, for all modes
ctc_status als_rows_prepare(ctc_model *m, const ctc_options *o, bool bf16);
// which FP32 ALS kernels a model of rank R uses under these options
inline bool als_uses_rows(const ctc_options *o, uint32_t R) {
  return o->precision == CTC_PREC_MIXED || o->als_kernel == CTC_ALS_ROWS || (o->als_kernel == CTC_ALS_AUTO && R >= 64);
}
// one ALS mode update, enqueued on the model's stream (no synchronisation, so
// it can be captured into a graph once the model is prepared)
ctc_status als_mode_impl(ctc_model *m, uint32_t mode, const ctc_options *o);
// reads and clears the model's not-positive-definite flag (synchronises)
ctc_status check_not_spd(ctc_model *m);
// profiled launch bracketing
void prof_begin(ctc_context *c, const char *name, hipEvent_t *a, hipEvent_t *b, hipStream_t s);
void prof_end(ctc_context *c, const char *name, hipEvent_t a, hipEvent_t b, hipStream_t s);
ctc_status prof_flush(ctc_context *c);
// RAII bracket of one profiled launch (no-op unless profiling is enabled)
struct Prof {
  ctc_context *c;
  const char *name;
  hipEvent_t a, b;
  hipStream_t s;  // the stream the timed launches go to (null: the context's)
  Prof(ctc_context *c_, const char *n, hipStream_t s_ = nullptr) : c(c_), name(n), s(s_) {
    prof_begin(c, n, &a, &b, s);
  }
  ~Prof() { prof_end(c, name, a, b, s); }
};
}  // namespace ctc
