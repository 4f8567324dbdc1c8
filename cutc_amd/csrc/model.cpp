// model.cpp — device context, device-resident model and the solver loops of
// libctc (C ABI in include/ctc.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "ctc_internal.h"
#include "host_parallel.h"
#include "kernels_launch.h"
#include "model_util.h"

using namespace ctc;

namespace {

constexpr unsigned kEvalBlocks = 512;  // fixed grid for deterministic reductions
// SGD hot rows (updated with plain stores): more than kSgdHotRows training
// nonzeros and more than kSgdHotRel times the mode's mean.  Measured on the
// 100M-nonzero tensor (480189 x 17770 x 2182, R = 32), k_sgd ms per epoch:
// threshold 1048576 -> 27.7, 262144 -> 19.8, 65536 -> 18.3 (same test-RMSE
// trajectory); at 32768 and below mode 2's rows (~45K nonzeros each) turn
// hot and the sweep stalls at test RMSE 2.02, which the relative bound
// rules out whatever the threshold.
constexpr uint64_t kSgdHotRows = 65536;
constexpr double kSgdHotRel = 4.0;
// bulk-synchronous SGD work units: walker span and sub-step (nonzeros)
constexpr uint32_t kSgdSpan = 2048;
constexpr uint32_t kSgdSubstep = 32;

template <class T>
ctc_status dev_alloc(T **p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  CTC_HIP(hipMalloc((void **)p, n * sizeof(T)));
  return CTC_OK;
}

template <class T>
ctc_status upload(T **p, const T *src, size_t n, hipStream_t st) {
  CTC_TRY(dev_alloc(p, n));
  if (n) CTC_HIP(hipMemcpyAsync(*p, src, n * sizeof(T), hipMemcpyHostToDevice, st));
  return CTC_OK;
}

void dev_free(void *p) {
  if (p) (void)hipFree(p);
}

bool rank_supported(uint32_t R) { return R == 8 || R == 16 || R == 32 || R == 64 || R == 128; }

double median(std::vector<double> v) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  size_t n = v.size();
  return n % 2 ? v[n / 2] : 0.5 * (v[n / 2 - 1] + v[n / 2]);
}

}  // namespace

namespace ctc {

// ------------------------------------------------------------- profiling --
void prof_begin(ctc_context *c, const char *name, hipEvent_t *a, hipEvent_t *b, hipStream_t s) {
  *a = *b = nullptr;
  if (!c->profiling) return;
  if (c->prof_every > 1 && c->prof_seen[name]++ % c->prof_every != 0) return;  // not sampled
  if (c->event_pool.size() < 2) {
    // timing only: no system-scope fence (cache write-back and invalidation)
    // between the launches they bracket
    hipEvent_t e1, e2;
    if (hipEventCreateWithFlags(&e1, hipEventDisableSystemFence) != hipSuccess ||
        hipEventCreateWithFlags(&e2, hipEventDisableSystemFence) != hipSuccess)
      return;
    c->event_pool.push_back(e1);
    c->event_pool.push_back(e2);
  }
  *a = c->event_pool.back();
  c->event_pool.pop_back();
  *b = c->event_pool.back();
  c->event_pool.pop_back();
  (void)hipEventRecord(*a, s ? s : c->stream);
}

void prof_end(ctc_context *c, const char *name, hipEvent_t a, hipEvent_t b, hipStream_t s) {
  if (!c->profiling || !a) return;
  (void)hipEventRecord(b, s ? s : c->stream);
  c->pending.push_back({name, a, b});
}

ctc_status prof_flush(ctc_context *c) {
  if (c->pending.empty()) return CTC_OK;
  for (auto &p : c->pending) {
    CTC_HIP(hipEventSynchronize(p.b));
    float ms = 0.f;
    CTC_HIP(hipEventElapsedTime(&ms, p.a, p.b));
    auto &s = c->stats[p.name];
    s.launches++;
    s.total_ms += ms;
    c->event_pool.push_back(p.a);
    c->event_pool.push_back(p.b);
  }
  c->pending.clear();
  return CTC_OK;
}

// ------------------------------------------------------------- scratch --
ctc_status ensure_scratch(ctc_model *m, size_t bytes) {
  if (bytes <= m->scratch_bytes) return CTC_OK;
  if (m->scratch) {
    CTC_HIP(hipStreamSynchronize(m->ctx->stream));
    dev_free(m->scratch);
    m->scratch = nullptr;
    m->scratch_bytes = 0;
  }
  CTC_HIP(hipMalloc(&m->scratch, bytes));
  m->scratch_bytes = bytes;
  return CTC_OK;
}

// ------------------------------------------------------------ CISS upload --
static ctc_status upload_ciss(ctc_model *m, HostCiss &h, DevCiss &d, uint32_t mode) {
  hipStream_t st = m->ctx->stream;
  d.nmodes = h.nmodes;
  d.mode = mode;
  d.nnz = h.nnz;
  d.ntiles = h.ntiles;
  d.dlength = h.dlength;
  for (uint32_t s = 0; s + 1 < h.nmodes; ++s) CTC_TRY(upload(&d.oidx[s], h.oidx[s].data(), h.nnz, st));
  CTC_TRY(upload(&d.vals, h.vals.data(), h.nnz, st));
  CTC_TRY(upload(&d.tile_slot, h.tile_slot.data(), h.ntiles, st));
  CTC_TRY(upload(&d.tile_mask, h.tile_mask.data(), h.ntiles, st));
  CTC_TRY(upload(&d.directory, h.directory.data(), h.dlength, st));
  CTC_TRY(upload(&d.rowptr, h.rowptr.data(), h.dlength + 1, st));
  CTC_HIP(hipStreamSynchronize(st));
  d.h_tile_slot.swap(h.tile_slot);
  d.h_tile_mask.swap(h.tile_mask);
  d.h_rowptr.swap(h.rowptr);
  d.span_tiles = 0;
  return CTC_OK;
}

ctc_status ensure_ciss(ctc_model *m, uint32_t mode) {
  if (mode >= m->nmodes) return fail(CTC_ERR_INVALID_ARG, "mode %u out of range", mode);
  if (m->ciss_ready[mode]) return CTC_OK;
  HostCiss h;
  CTC_TRY(build_host_ciss(m->train.get(), mode, h));
  CTC_TRY(upload_ciss(m, h, m->ciss[mode], mode));
  m->ciss_ready[mode] = true;
  return CTC_OK;
}

static void free_dev_ciss(DevCiss &d) {
  for (int s = 0; s < 3; ++s) dev_free(d.oidx[s]);
  for (void *p : {(void *)d.vals, (void *)d.tile_slot, (void *)d.tile_mask, (void *)d.directory, (void *)d.rowptr,
                  (void *)d.span_off, (void *)d.slot_pb, (void *)d.grp_piece, (void *)d.grp_n, (void *)d.slot_gb,
                  (void *)d.mix_items, (void *)d.mix_heavy_slot, (void *)d.mix_heavy_pb, (void *)d.mix_grp,
                  (void *)d.mix_heavy_gb})
    dev_free(p);
  d = DevCiss();
}

// ref_quirks: the CISS of the mode's nonzeros repeated by their weights
// (quirks.cpp), built once per set of quirk bits
ctc_status ensure_quirk_ciss(ctc_model *m, uint32_t mode, uint32_t flags) {
  if (m->qciss_flags[mode] == flags) return CTC_OK;
  CTC_HIP(hipStreamSynchronize(m->ctx->stream));
  free_dev_ciss(m->qciss[mode]);
  m->qciss_flags[mode] = 0;
  ctc_coo *q = nullptr;
  CTC_TRY(quirk_multiset(m->train.get(), mode, flags, &q));
  if (!q) {  // every term dropped
    m->qciss_flags[mode] = flags;
    return CTC_OK;
  }
  HostCiss h;
  const ctc_status s = build_host_ciss(q, mode, h);
  ctc_coo_free(q);
  CTC_TRY(s);
  CTC_TRY(upload_ciss(m, h, m->qciss[mode], mode));
  m->qciss_flags[mode] = flags;
  return CTC_OK;
}

// Piece tables for spans of `span_tiles` tiles (see k_als_gram).  Piece ids
// run over (span, slot) pairs in order, so piece(span, slot) = slot +
// span_off[span] with span_off[span] = pieces before the span - its first slot.
ctc_status ensure_pieces(ctc_model *m, uint32_t mode, uint32_t span_tiles) {
  DevCiss &d = m->ciss[mode];
  if (d.span_tiles == span_tiles && d.span_off) return CTC_OK;
  const uint64_t nspans = (d.ntiles + span_tiles - 1) / span_tiles;
  std::vector<uint32_t> off(nspans), pb(d.dlength + 1);
  uint64_t base = 0;
  for (uint64_t w = 0; w < nspans; ++w) {
    const uint64_t t0 = w * span_tiles;
    const uint64_t last = std::min<uint64_t>((t0 + span_tiles) * 64, d.nnz) - 1;
    const uint32_t first = d.h_tile_slot[t0];
    const uint64_t tl = last >> 6;
    const int e = (int)(last & 63);
    const uint64_t below = e == 63 ? ~0ull : ((2ull << e) - 1ull);
    const uint32_t lastslot = d.h_tile_slot[tl] + (uint32_t)__builtin_popcountll(d.h_tile_mask[tl] & below & ~1ull);
    off[w] = (uint32_t)(base - first);
    base += (uint64_t)(lastslot - first) + 1;
  }
  if (base > 0xFFFFFFFFull) return fail(CTC_ERR_UNSUPPORTED, "too many pieces");
  for (uint64_t s = 0; s < d.dlength; ++s) {
    const uint64_t w = ((uint64_t)d.h_rowptr[s] >> 6) / span_tiles;
    pb[s] = (uint32_t)(s + off[w]);
  }
  pb[d.dlength] = (uint32_t)base;
  // group tables for heavy slots
  constexpr uint32_t kHeavyPieces = 64, kGroupPieces = 32;
  std::vector<uint32_t> gpiece, gn, gb(d.dlength + 1);
  for (uint64_t s = 0; s < d.dlength; ++s) {
    gb[s] = (uint32_t)gpiece.size();
    const uint32_t n = pb[s + 1] - pb[s];
    if (n <= kHeavyPieces) continue;
    for (uint32_t q = 0; q < n; q += kGroupPieces) {
      gpiece.push_back(pb[s] + q);
      gn.push_back(std::min(kGroupPieces, n - q));
    }
  }
  gb[d.dlength] = (uint32_t)gpiece.size();
  dev_free(d.span_off);
  dev_free(d.slot_pb);
  dev_free(d.grp_piece);
  dev_free(d.grp_n);
  dev_free(d.slot_gb);
  d.span_off = d.slot_pb = d.grp_piece = d.grp_n = d.slot_gb = nullptr;
  hipStream_t st = m->ctx->stream;
  CTC_TRY(upload(&d.span_off, off.data(), nspans, st));
  CTC_TRY(upload(&d.slot_pb, pb.data(), d.dlength + 1, st));
  d.ngroups = gpiece.size();
  if (d.ngroups) {
    CTC_TRY(upload(&d.grp_piece, gpiece.data(), d.ngroups, st));
    CTC_TRY(upload(&d.grp_n, gn.data(), d.ngroups, st));
    CTC_TRY(upload(&d.slot_gb, gb.data(), d.dlength + 1, st));
  }
  CTC_HIP(hipStreamSynchronize(st));
  d.span_tiles = span_tiles;
  d.nspans = nspans;
  d.npieces = base;
  return CTC_OK;
}

}  // namespace ctc

// =========================================================================
// context
// =========================================================================
extern "C" ctc_status ctc_context_create(int device, ctc_context **out) {
  if (!out) return fail(CTC_ERR_INVALID_ARG, "ctc_context_create: null out");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(CTC_ERR_NO_DEVICE, "no HIP device");
  if (device < 0 || device >= n) return fail(CTC_ERR_INVALID_ARG, "device %d of %d", device, n);
  CTC_HIP(hipSetDevice(device));
  ctc_context *c = new ctc_context();
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return fail(CTC_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
  }
  (void)hipGetDeviceProperties(&c->props, device);
  *out = c;
  return CTC_OK;
}

void context_retain(ctc_context *c) { ++c->refs; }

void context_release(ctc_context *c) {
  if (!c || --c->refs > 0) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (auto &p : c->pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : c->event_pool) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(c->stream);
  delete c;
}

extern "C" void ctc_context_destroy(ctc_context *c) { context_release(c); }

extern "C" ctc_status ctc_context_stream(ctc_context *c, void **stream) {
  if (!c || !stream) return fail(CTC_ERR_INVALID_ARG, "ctc_context_stream");
  *stream = (void *)c->stream;
  return CTC_OK;
}

extern "C" ctc_status ctc_context_synchronize(ctc_context *c) {
  if (!c) return fail(CTC_ERR_INVALID_ARG, "ctc_context_synchronize");
  CTC_HIP(hipSetDevice(c->device));
  CTC_HIP(hipStreamSynchronize(c->stream));
  return CTC_OK;
}

extern "C" ctc_status ctc_context_set_profiling(ctc_context *c, int enabled) {
  if (!c || enabled < 0) return fail(CTC_ERR_INVALID_ARG, "ctc_context_set_profiling");
  CTC_TRY(prof_flush(c));
  c->profiling = enabled != 0;
  c->prof_every = enabled > 1 ? (uint32_t)enabled : 1u;
  c->prof_seen.clear();
  return CTC_OK;
}

extern "C" ctc_status ctc_context_kernel_stats(ctc_context *c, const char *kernel, uint64_t *launches,
                                               double *total_ms) {
  if (!c || !kernel) return fail(CTC_ERR_INVALID_ARG, "ctc_context_kernel_stats");
  CTC_TRY(prof_flush(c));
  auto it = c->stats.find(kernel);
  if (launches) *launches = it == c->stats.end() ? 0 : it->second.launches;
  if (total_ms) *total_ms = it == c->stats.end() ? 0.0 : it->second.total_ms;
  return CTC_OK;
}

extern "C" ctc_status ctc_context_reset_stats(ctc_context *c) {
  if (!c) return fail(CTC_ERR_INVALID_ARG, "null context");
  CTC_TRY(prof_flush(c));
  c->stats.clear();
  return CTC_OK;
}

extern "C" ctc_status ctc_context_copy_bandwidth(ctc_context *c, uint64_t bytes, double *gbps) {
  if (!c || !gbps || bytes < 1024) return fail(CTC_ERR_INVALID_ARG, "ctc_context_copy_bandwidth");
  CTC_HIP(hipSetDevice(c->device));
  void *a = nullptr, *b = nullptr;
  CTC_HIP(hipMalloc(&a, bytes));
  if (hipMalloc(&b, bytes) != hipSuccess) {
    dev_free(a);
    return fail(CTC_ERR_NOMEM, "copy bandwidth buffers");
  }
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  bytes &= ~(uint64_t)15;
  (void)launch_copy(c->stream, a, b, bytes);
  (void)hipEventRecord(e0, c->stream);
  const int reps = 10;
  for (int r = 0; r < reps; ++r) (void)launch_copy(c->stream, a, b, bytes);
  (void)hipEventRecord(e1, c->stream);
  (void)hipEventSynchronize(e1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  dev_free(a);
  dev_free(b);
  *gbps = 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;  // read + write
  return CTC_OK;
}

// =========================================================================
// options
// =========================================================================
extern "C" ctc_status ctc_options_default(int32_t algorithm, ctc_options *o) {
  if (!o || algorithm < 0 || algorithm > 2) return fail(CTC_ERR_INVALID_ARG, "ctc_options_default");
  memset(o, 0, sizeof(*o));
  o->algorithm = algorithm;
  o->rank = 16;                                                        // SG/base.h:21
  o->lambda = algorithm == CTC_ALS ? 10.0 : algorithm == CTC_SGD ? 1e-3 : 100.0;  // :39-41
  o->learning_rate = 1e-3;                                             // :58
  o->lr_schedule = CTC_LR_BOLD_DRIVER;                                 // SG/sgd.cu:314-323
  o->lr_decay = 0.95;
  o->sgd_scale = 1e-5;                                                 // :42-44
  o->max_epochs = algorithm == CTC_SGD ? 999 : 1000;                   // :22
  o->tolerance = 1e-4;                                                 // :36
  o->max_bad_epochs = 3;                                               // :32
  // work-unit shapes measured on the 100M-nonzero benchmark tensor (the
  // reference's GA searched the same two knobs, GA/fitness.c:67-79)
  // measured on the 100M-nonzero tensor at R = 32: SGD 2048, CCD++ 1024; ALS
  // leaves it 0, so that each of its kernel families takes its own measured
  // default (span and FP32 row-fused kernels 1024, BF16 row-fused 4096 or 8192)
  o->tile_len = algorithm == CTC_SGD ? kSgdSpan : algorithm == CTC_ALS ? 0 : 1024;
  o->block = algorithm == CTC_SGD ? 64 : 256;
  o->seed = 0;
  o->mode_order = CTC_ORDER_FIXED;
  o->precision = CTC_PREC_FP32;
  o->inner_iters = 1;                                                  // :30
  o->refine = 0;
  o->als_kernel = CTC_ALS_AUTO;
  o->eval_every_epoch = 1;
  o->verbose = 0;
  o->sgd_update = CTC_SGD_SYNC;
  o->spread = 0;
  o->sgd_hot_rows = kSgdHotRows;
  o->sgd_substep = kSgdSubstep;
  return CTC_OK;
}

// =========================================================================
// model
// =========================================================================
static ctc_status copy_coo(const ctc_coo *src, ctc_coo **dst) {
  CTC_TRY(ctc_coo_alloc(src->nmodes, src->nnz, src->dims, dst));
  for (uint32_t m = 0; m < src->nmodes; ++m) {
    memcpy((*dst)->ind[m], src->ind[m], src->nnz * sizeof(uint32_t));
    (*dst)->base[m] = src->base[m];
  }
  memcpy((*dst)->vals, src->vals, src->nnz * sizeof(double));
  return CTC_OK;
}

static ctc_status upload_coo(const ctc_coo *t, DevCoo &d, hipStream_t st) {
  d.nmodes = t->nmodes;
  d.nnz = t->nnz;
  for (uint32_t m = 0; m < t->nmodes; ++m) CTC_TRY(upload(&d.ind[m], t->ind[m], t->nnz, st));
  std::vector<float> v(t->nnz);
  for (uint64_t p = 0; p < t->nnz; ++p) v[p] = (float)t->vals[p];
  CTC_TRY(upload(&d.vals, v.data(), t->nnz, st));
  CTC_HIP(hipStreamSynchronize(st));
  return CTC_OK;
}

static ctc_status check_coo(const ctc_coo *t, const ctc_coo *train, const char *what) {
  if (!t) return CTC_OK;
  if (t->nmodes != train->nmodes) return fail(CTC_ERR_INVALID_ARG, "%s: mode count differs", what);
  for (uint32_t m = 0; m < t->nmodes; ++m) {
    const uint64_t dim = train->dims[m];
    for (uint64_t p = 0; p < t->nnz; ++p)
      if (t->ind[m][p] >= dim)
        return fail(CTC_ERR_INVALID_ARG, "%s: index %u of mode %u beyond training dim %llu", what,
                    t->ind[m][p], m, (unsigned long long)dim);
  }
  return CTC_OK;
}

extern "C" void ctc_model_destroy(ctc_model *m) {
  if (!m) return;
  (void)hipSetDevice(m->ctx->device);
  (void)hipStreamSynchronize(m->ctx->stream);
  for (uint32_t md = 0; md < CTC_MAX_MODES; ++md) {
    free_dev_ciss(m->ciss[md]);
    free_dev_ciss(m->qciss[md]);
    dev_free(m->Fbf[md]);
    dev_free(m->F[md]);
    dev_free(m->best[md]);
    dev_free(m->resid[md]);
    dev_free(m->FT[md]);
    dev_free(m->ccdx[md]);
    dev_free(m->hotbits[md]);
    dev_free(m->sgd_D[md]);
    dev_free(m->sgd_stage[md]);
    dev_free(m->sgd_slots[md]);
    dev_free(m->sgd_dest[md]);
    dev_free(m->sgd_sbase[md]);
    dev_free(m->sgd_piece_start[md]);
    dev_free(m->sgd_piece_row[md]);
    dev_free(m->sgd_part[md]);
    dev_free(m->sgd_fold[md]);
    dev_free(m->sgd_heavy[md]);
    dev_free(m->sgd_light[md]);
    dev_free(m->sgd_hidx[md]);
    dev_free(m->sgd_hrow[md]);
    dev_free(m->sgd_Dh[md]);
    dev_free(m->sgd_ch[md]);
  }
  dev_free(m->sgd_long);
  for (DevCoo *c : {&m->valid, &m->test}) {
    for (int s = 0; s < 4; ++s) dev_free(c->ind[s]);
    dev_free(c->vals);
  }
  dev_free(m->scratch);
  dev_free(m->eval_part);
  dev_free(m->dstat);
  dev_free(m->dflag);
  if (m->hstat) (void)hipHostFree(m->hstat);
  ctc_context *ctx = m->ctx;
  delete m;
  context_release(ctx);
}

extern "C" ctc_status ctc_model_create(ctc_context *ctx, const ctc_coo *train, const ctc_coo *validation,
                                       const ctc_coo *test, uint32_t rank, ctc_model **out) {
  if (!ctx || !train || !out) return fail(CTC_ERR_INVALID_ARG, "ctc_model_create: null argument");
  if (!rank_supported(rank)) return fail(CTC_ERR_UNSUPPORTED, "rank %u: supported ranks are 8,16,32,64,128", rank);
  if (train->nmodes < 3 || train->nmodes > 4)
    return fail(CTC_ERR_UNSUPPORTED, "%u-way tensors: supported are 3- and 4-way", train->nmodes);
  if (train->nnz == 0) return fail(CTC_ERR_INVALID_ARG, "empty training tensor");
  for (uint32_t md = 0; md < train->nmodes; ++md)
    for (uint64_t p = 0; p < train->nnz; ++p)
      if (train->ind[md][p] >= train->dims[md])
        return fail(CTC_ERR_INVALID_ARG, "train index beyond dims in mode %u", md);
  CTC_TRY(check_coo(validation, train, "validation"));
  CTC_TRY(check_coo(test, train, "test"));
  CTC_HIP(hipSetDevice(ctx->device));
  ctc_model *m = new ctc_model();
  m->ctx = ctx;
  context_retain(ctx);
  m->nmodes = train->nmodes;
  m->rank = rank;
  m->train_nnz = train->nnz;
  for (uint32_t md = 0; md < m->nmodes; ++md) m->dims[md] = train->dims[md];
  ctc_coo *tc = nullptr;
  ctc_status s = copy_coo(train, &tc);
  if (s != CTC_OK) {
    delete m;
    context_release(ctx);
    return s;
  }
  m->train.reset(tc);
  auto bail = [&](ctc_status st) {
    ctc_model_destroy(m);
    return st;
  };
  hipStream_t st = ctx->stream;
  if (validation && validation->nnz) {
    if ((s = upload_coo(validation, m->valid, st)) != CTC_OK) return bail(s);
  }
  if (test && test->nnz) {
    if ((s = upload_coo(test, m->test, st)) != CTC_OK) return bail(s);
  }
  for (uint32_t md = 0; md < m->nmodes; ++md) {
    if ((s = dev_alloc(&m->F[md], m->dims[md] * rank)) != CTC_OK) return bail(s);
    if ((s = dev_alloc(&m->best[md], m->dims[md] * rank)) != CTC_OK) return bail(s);
  }
  if ((s = dev_alloc(&m->dstat, 4096)) != CTC_OK) return bail(s);
  if ((s = dev_alloc(&m->dflag, 16)) != CTC_OK) return bail(s);
  if (hipMemsetAsync(m->dflag, 0, 16 * sizeof(uint32_t), st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
    return bail(fail(CTC_ERR_HIP, "status flag initialisation"));
  if (hipHostMalloc((void **)&m->hstat, 64 * sizeof(double), hipHostMallocDefault) != hipSuccess)
    return bail(fail(CTC_ERR_NOMEM, "pinned host buffer"));
  *out = m;
  return CTC_OK;
}

extern "C" ctc_status ctc_model_init_random(ctc_model *m, uint64_t seed) {
  if (!m) return fail(CTC_ERR_INVALID_ARG, "null model");
  for (uint32_t md = 0; md < m->nmodes; ++md) {
    std::vector<float> h(m->dims[md] * m->rank);
    // U(-1,1) entries (SG/matrixprocess.c:156-165) from a seeded generator
    std::mt19937_64 g(seed * 0x9E3779B97F4A7C15ull + md + 1);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    for (auto &x : h) x = (float)u(g);
    CTC_TRY(ctc_model_set_factor(m, md, h.data()));
  }
  return CTC_OK;
}

extern "C" ctc_status ctc_model_set_factor(ctc_model *m, uint32_t mode, const float *host) {
  if (!m || !host || mode >= m->nmodes) return fail(CTC_ERR_INVALID_ARG, "ctc_model_set_factor");
  CTC_HIP(hipSetDevice(m->ctx->device));
  const size_t n = m->dims[mode] * m->rank;
  CTC_HIP(hipMemcpyAsync(m->F[mode], host, n * sizeof(float), hipMemcpyHostToDevice, m->ctx->stream));
  CTC_HIP(hipMemcpyAsync(m->best[mode], m->F[mode], n * sizeof(float), hipMemcpyDeviceToDevice,
                         m->ctx->stream));
  CTC_HIP(hipStreamSynchronize(m->ctx->stream));
  m->resid_valid = false;
  return CTC_OK;
}

extern "C" ctc_status ctc_model_get_factor(ctc_model *m, uint32_t mode, int best, float *host) {
  if (!m || !host || mode >= m->nmodes) return fail(CTC_ERR_INVALID_ARG, "ctc_model_get_factor");
  CTC_HIP(hipSetDevice(m->ctx->device));
  const size_t n = m->dims[mode] * m->rank;
  CTC_HIP(hipMemcpyAsync(host, best ? m->best[mode] : m->F[mode], n * sizeof(float),
                         hipMemcpyDeviceToHost, m->ctx->stream));
  CTC_HIP(hipStreamSynchronize(m->ctx->stream));
  return CTC_OK;
}

extern "C" ctc_status ctc_model_factor_device(ctc_model *m, uint32_t mode, float **dev) {
  if (!m || !dev || mode >= m->nmodes) return fail(CTC_ERR_INVALID_ARG, "ctc_model_factor_device");
  *dev = m->F[mode];
  return CTC_OK;
}

extern "C" ctc_status ctc_model_prepare(ctc_model *m, const ctc_options *o) {
  if (!m || !o) return fail(CTC_ERR_INVALID_ARG, "ctc_model_prepare");
  CTC_HIP(hipSetDevice(m->ctx->device));
  if (o->algorithm == CTC_SGD) return ensure_ciss(m, 0);
  if (o->algorithm == CTC_ALS && als_uses_rows(o, m->rank))
    return als_rows_prepare(m, o, o->precision == CTC_PREC_MIXED);
  for (uint32_t md = 0; md < m->nmodes; ++md) {
    CTC_TRY(ensure_ciss(m, md));
    if (o->algorithm == CTC_ALS) CTC_TRY(ensure_pieces(m, md, span_tiles_of(o)));
  }
  return CTC_OK;
}

// =========================================================================
// single steps
// =========================================================================
static ctc_status als_mode_plain(ctc_model *m, uint32_t mode, const ctc_options *o) {
  const uint32_t R = m->rank;
  if (als_uses_rows(o, R)) return als_mode_rows(m, mode, o, o->precision == CTC_PREC_MIXED);
  if (R > 64) return fail(CTC_ERR_UNSUPPORTED, "the span ALS kernels support ranks up to 64 (got %u)", R);
  const uint32_t S = span_tiles_of(o);
  CTC_TRY(ensure_ciss(m, mode));
  CTC_TRY(ensure_pieces(m, mode, S));
  DevCiss &d = m->ciss[mode];
  const size_t psz = (size_t)R * R + R;
  CTC_TRY(ensure_scratch(m, (d.npieces + d.ngroups) * psz * sizeof(double)));
  hipStream_t st = m->ctx->stream;
  double *pieces = (double *)m->scratch;
  double *gsum = pieces + d.npieces * psz;
  CissArgs ca = ciss_args(d);
  FactorArgs fa = rel_factors(m->F, m->nmodes, mode);
  {
    Prof p(m->ctx, "als_gram");
    CTC_HIP(launch_als_gram(st, ca, fa, (int)m->nmodes, (int)R, S, d.nspans, d.span_off, pieces,
                            block_of(o)));
  }
  // rows absent from the directory become 0 (SG/als.cu:699 zeroes the factor)
  CTC_HIP(launch_fill(st, m->F[mode], m->dims[mode] * R, 0.f));
  if (d.ngroups) {
    Prof p(m->ctx, "als_reduce");
    CTC_HIP(launch_piece_reduce(st, pieces, d.grp_piece, d.grp_n, d.ngroups, (int)R, gsum));
  }
  {
    Prof p(m->ctx, "als_solve");
    CTC_HIP(launch_als_solve(st, pieces, d.slot_pb, d.ngroups ? gsum : nullptr, d.slot_gb, d.directory,
                             d.dlength, (int)R, o->lambda, o->refine != 0, d.npieces <= 2 * d.dlength, m->F[mode],
                             m->dflag));
  }
  m->resid_valid = false;
  return CTC_OK;
}

// With ref_quirks the update runs on the mode's quirk CISS (the nonzeros
// repeated by their weights), swapped in for the update and out after it.
// The CCD bits (Q3, Q4) do not concern ALS and are ignored here.
static ctc_status als_mode(ctc_model *m, uint32_t mode, const ctc_options *o) {
  if (o->ref_quirks & ~kQuirkAll) return fail(CTC_ERR_INVALID_ARG, "ref_quirks: unknown bits 0x%x", o->ref_quirks & ~kQuirkAll);
  const uint32_t q = o->ref_quirks & (CTC_QUIRK_TILE_HEAD | CTC_QUIRK_LAST_LANE | CTC_QUIRK_TRUNC_VALUES);
  if (!q) return als_mode_plain(m, mode, o);
  if (mode >= m->nmodes) return fail(CTC_ERR_INVALID_ARG, "mode %u out of range", mode);
  CTC_TRY(ensure_ciss(m, mode));  // the plain CISS stays resident for the other algorithms
  CTC_TRY(ensure_quirk_ciss(m, mode, q));
  if (m->qciss[mode].nnz == 0) {  // the reference zeroes the factor and solves lambda I a = 0
    CTC_HIP(launch_fill(m->ctx->stream, m->F[mode], m->dims[mode] * m->rank, 0.f));
    m->resid_valid = false;
    return CTC_OK;
  }
  std::swap(m->ciss[mode], m->qciss[mode]);
  const ctc_status s = als_mode_plain(m, mode, o);
  std::swap(m->ciss[mode], m->qciss[mode]);
  return s;
}

static ctc_status check_flag(ctc_model *m) {
  uint32_t f = 0;
  CTC_HIP(hipMemcpyAsync(&f, m->dflag, sizeof(uint32_t), hipMemcpyDeviceToHost, m->ctx->stream));
  CTC_HIP(hipStreamSynchronize(m->ctx->stream));
  if (f) {
    CTC_HIP(hipMemsetAsync(m->dflag, 0, sizeof(uint32_t), m->ctx->stream));
    CTC_HIP(hipStreamSynchronize(m->ctx->stream));
    return fail(CTC_ERR_NOT_SPD, "an ALS normal system (G + lambda I) was not positive definite");
  }
  return CTC_OK;
}

namespace ctc {
ctc_status als_mode_impl(ctc_model *m, uint32_t mode, const ctc_options *o) { return als_mode(m, mode, o); }
ctc_status check_not_spd(ctc_model *m) { return check_flag(m); }
}  // namespace ctc

// stride coprime with n near n / golden ratio (1 when spreading is off)
static uint64_t spread_stride(uint64_t n, bool on) {
  if (!on || n < 3) return 1;
  uint64_t s = (uint64_t)((double)n * 0.6180339887498949) | 1ull;
  auto gcd = [](uint64_t x, uint64_t y) {
    while (y) {
      uint64_t t = x % y;
      x = y;
      y = t;
    }
    return x;
  };
  while (gcd(s, n) != 1) s += 2;
  return s % n ? s % n : 1;
}

// Hot-row bitmaps of every mode: rows with more training nonzeros than the
// threshold and more than kSgdHotRel times the mode's mean.  The relative
// bound keeps a mode whose rows are all alike (every row shared by many
// walkers, none far hotter than the rest) out of the hot set.
// The bulk-synchronous schedule also keeps, per mode, each row's index among
// the hot rows, the hot rows' list and their replicated (zeroed) delta sums.
static ctc_status ensure_hotbits(ctc_model *m, uint64_t thr) {
  if (thr == m->hot_threshold && m->hotbits[0]) return CTC_OK;
  const ctc_coo *t = m->train.get();
  hipStream_t st = m->ctx->stream;
  CTC_HIP(hipStreamSynchronize(st));
  for (uint32_t md = 0; md < m->nmodes; ++md) {
    std::vector<uint32_t> cnt(m->dims[md], 0), bits((m->dims[md] + 31) / 32, 0);
    std::vector<uint32_t> hidx(m->dims[md], 0xFFFFFFFFu), hrow;
    for (uint64_t p = 0; p < t->nnz; ++p) cnt[t->ind[md][p]]++;
    const double rel = kSgdHotRel * (double)t->nnz / (double)std::max<uint64_t>(1, m->dims[md]);
    for (uint64_t i = 0; i < m->dims[md]; ++i)
      if (cnt[i] > thr && (double)cnt[i] > rel) {
        bits[i >> 5] |= 1u << (i & 31);
        hidx[i] = (uint32_t)hrow.size();
        hrow.push_back((uint32_t)i);
      }
    if (!m->hotbits[md]) CTC_HIP(hipMalloc(&m->hotbits[md], std::max<size_t>(1, bits.size()) * sizeof(uint32_t)));
    if (!bits.empty())
      CTC_HIP(hipMemcpy(m->hotbits[md], bits.data(), bits.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    for (void *p : {(void *)m->sgd_hidx[md], (void *)m->sgd_hrow[md], (void *)m->sgd_Dh[md], (void *)m->sgd_ch[md]})
      dev_free(p);
    m->sgd_hidx[md] = m->sgd_hrow[md] = m->sgd_ch[md] = nullptr;
    m->sgd_Dh[md] = nullptr;
    CTC_TRY(upload(&m->sgd_hidx[md], hidx.data(), hidx.size(), st));
    CTC_TRY(upload(&m->sgd_hrow[md], hrow.data(), hrow.size(), st));
    const size_t nrep = hrow.size() * kSyncReplicas;
    CTC_TRY(dev_alloc(&m->sgd_Dh[md], nrep * m->rank));
    CTC_TRY(dev_alloc(&m->sgd_ch[md], nrep));
    if (nrep) {
      CTC_HIP(hipMemsetAsync(m->sgd_Dh[md], 0, nrep * m->rank * sizeof(long long), st));
      CTC_HIP(hipMemsetAsync(m->sgd_ch[md], 0, nrep * sizeof(uint32_t), st));
    }
    m->sgd_nhot[md] = hrow.size();
    CTC_HIP(hipStreamSynchronize(st));
  }
  m->hot_threshold = thr;
  return CTC_OK;
}

// nonzeros per walker and per sub-step of the bulk-synchronous schedule
static ctc_status sync_shape(const ctc_options *o, uint32_t *span, uint32_t *sub) {
  *span = o->tile_len ? o->tile_len : kSgdSpan;
  *sub = o->sgd_substep ? o->sgd_substep : kSgdSubstep;
  if (*sub > *span) return fail(CTC_ERR_INVALID_ARG, "sgd_substep %u exceeds tile_len %u", *sub, *span);
  return CTC_OK;
}

// Buffers of the bulk-synchronous schedule: zeroed delta sums per mode (the
// merge leaves them zero), and the list of mode-0 rows that
// two walkers can meet in one sub-step (length >= span - sub + 2).
static ctc_status ensure_sync(ctc_model *m, uint32_t span, uint32_t sub) {
  hipStream_t st = m->ctx->stream;
  if (!m->sgd_D[0]) {  // mode 0's long rows: fixed-point delta sums
    CTC_TRY(dev_alloc(&m->sgd_D[0], (size_t)m->dims[0] * m->rank));
    CTC_HIP(hipMemsetAsync(m->sgd_D[0], 0, (size_t)m->dims[0] * m->rank * sizeof(long long), st));
  }
  if (m->sgd_long && m->sgd_long_span == span && m->sgd_long_sub == sub && m->sgd_long_thr == m->hot_threshold)
    return CTC_OK;
  const DevCiss &d = m->ciss[0];
  // Modes 1.. hand in each run's delta to a staging slot.  The runs of a
  // walker's chunk (walker w, sub-step s) are the maximal runs of equal index
  // in it, numbered in walk order (slots) from base[s][w] (the runs of the
  // walkers before w in the same sub-step); per sub-step the slots are listed
  // by row in slot (= walk) order for the merge: position q holds slot
  // slots[q], and dest[slot] = q (the mixed path stages at q).
  const uint64_t nnz = d.nnz, nspans = (nnz + span - 1) / span, nsub = (span + sub - 1) / sub;
  for (uint32_t k = 1; k < m->nmodes; ++k) {
    for (void *p : {(void *)m->sgd_stage[k], (void *)m->sgd_slots[k], (void *)m->sgd_dest[k], (void *)m->sgd_sbase[k],
                    (void *)m->sgd_piece_start[k], (void *)m->sgd_piece_row[k], (void *)m->sgd_part[k],
                    (void *)m->sgd_fold[k], (void *)m->sgd_heavy[k], (void *)m->sgd_light[k]})
      dev_free(p);
    m->sgd_heavy[k] = m->sgd_light[k] = nullptr;
    m->sgd_stage[k] = nullptr;
    m->sgd_part[k] = nullptr;
    m->sgd_fold[k] = nullptr;
    m->sgd_slots[k] = m->sgd_dest[k] = m->sgd_sbase[k] = m->sgd_piece_start[k] = m->sgd_piece_row[k] = nullptr;
    std::vector<uint32_t> idx(nnz);
    CTC_HIP(hipMemcpy(idx.data(), d.oidx[k - 1], nnz * sizeof(uint32_t), hipMemcpyDeviceToHost));
    auto starts_run = [&](uint64_t p) { return (p % span) % sub == 0 || idx[p] != idx[p - 1]; };
    // runs per (sub-step, walker) -> bases, and per sub-step
    std::vector<uint32_t> base(nsub * nspans, 0);
    std::vector<uint64_t> boff(nsub + 1, 0);
    for (uint64_t p = 0; p < nnz; ++p)
      if (starts_run(p)) {
        const uint64_t w = p / span, s2 = (p - w * span) / sub;
        base[s2 * nspans + w]++;
      }
    uint64_t max_runs = 0;
    for (uint64_t s2 = 0; s2 < nsub; ++s2) {
      uint64_t acc = 0;
      for (uint64_t w = 0; w < nspans; ++w) {
        const uint32_t c = base[s2 * nspans + w];
        base[s2 * nspans + w] = (uint32_t)acc;
        acc += c;
      }
      boff[s2 + 1] = boff[s2] + acc;
      max_runs = std::max(max_runs, acc);
    }
    // (row, slot) keys of every run, bucketed by sub-step and sorted by row
    std::vector<uint64_t> key(boff[nsub]);
    {
      uint64_t w_prev = ~0ull, s_prev = ~0ull, run = 0;
      for (uint64_t p = 0; p < nnz; ++p) {
        if (!starts_run(p)) continue;
        const uint64_t w = p / span, s2 = (p - w * span) / sub;
        if (w != w_prev || s2 != s_prev) run = 0, w_prev = w, s_prev = s2;
        const uint64_t slotr = (uint64_t)base[s2 * nspans + w] + run++;
        key[boff[s2] + slotr] = ((uint64_t)idx[p] << 32) | slotr;
      }
    }
    parallel_for(nsub, host_threads(), [&](uint64_t b, uint64_t e, unsigned) {
      for (uint64_t s2 = b; s2 < e; ++s2) std::sort(key.begin() + boff[s2], key.begin() + boff[s2 + 1]);
    });
    // every row's slots of a sub-step cut into pieces of at most kStagePiece
    // slots; a row of one piece is settled by it, the others are folded
    std::vector<uint32_t> piece_start, piece_row, slots(key.size());
    std::vector<uint32_t> dest(key.size() + 1, 0);  // + 1: the walk reads one ahead
    std::vector<FoldSeg> fold;
    std::vector<uint64_t> &poff = m->sgd_piece_off[k], &foff = m->sgd_fold_off[k];
    poff.assign(nsub + 1, 0);
    foff.assign(nsub + 1, 0);
    uint64_t max_pieces = 0;
    for (uint64_t s2 = 0; s2 < nsub; ++s2) {
      poff[s2] = piece_start.size();
      foff[s2] = fold.size();
      for (uint64_t q = boff[s2]; q < boff[s2 + 1];) {
        const uint32_t row = (uint32_t)(key[q] >> 32);
        uint64_t e = q + 1;
        while (e < boff[s2 + 1] && (uint32_t)(key[e] >> 32) == row) ++e;
        const uint64_t np = (e - q + kStagePiece - 1) / kStagePiece;
        if (np > 1)
          fold.push_back({row, (uint32_t)(piece_start.size() - poff[s2]), (uint32_t)np, (uint32_t)(e - q)});
        for (uint64_t t = 0; t < np; ++t) {
          piece_start.push_back((uint32_t)(q + t * kStagePiece));
          piece_row.push_back(np == 1 ? row : 0xFFFFFFFFu);
        }
        for (uint64_t t = q; t < e; ++t) {
          slots[t] = (uint32_t)key[t];
          dest[boff[s2] + slots[t]] = (uint32_t)(t - boff[s2]);
        }
        q = e;
      }
      max_pieces = std::max<uint64_t>(max_pieces, piece_start.size() - poff[s2]);
    }
    poff[nsub] = piece_start.size();
    foff[nsub] = fold.size();
    piece_start.push_back((uint32_t)key.size());
    // per sub-step: rows of at most kSettleLight runs (one piece each) go to
    // the light list, every other piece to the heavy one
    std::vector<uint32_t> heavy, light;
    std::vector<uint64_t> &hoff = m->sgd_heavy_off[k], &loff = m->sgd_light_off[k];
    hoff.assign(nsub + 1, 0);
    loff.assign(nsub + 1, 0);
    for (uint64_t s2 = 0; s2 < nsub; ++s2) {
      hoff[s2] = heavy.size();
      loff[s2] = light.size();
      for (uint64_t gp = poff[s2]; gp < poff[s2 + 1]; ++gp) {
        const bool lt = piece_row[gp] != 0xFFFFFFFFu && piece_start[gp + 1] - piece_start[gp] <= ctc::kSettleLight;
        (lt ? light : heavy).push_back((uint32_t)gp);
      }
    }
    hoff[nsub] = heavy.size();
    loff[nsub] = light.size();
    CTC_TRY(upload(&m->sgd_heavy[k], heavy.data(), heavy.size(), st));
    CTC_TRY(upload(&m->sgd_light[k], light.data(), light.size(), st));
    CTC_TRY(upload(&m->sgd_piece_start[k], piece_start.data(), piece_start.size(), st));
    CTC_TRY(upload(&m->sgd_piece_row[k], piece_row.data(), piece_row.size(), st));
    CTC_TRY(upload(&m->sgd_fold[k], fold.data(), fold.size(), st));
    CTC_TRY(upload(&m->sgd_slots[k], slots.data(), slots.size(), st));
    CTC_TRY(upload(&m->sgd_dest[k], dest.data(), dest.size(), st));
    m->sgd_slot_off[k] = boff;
    CTC_TRY(upload(&m->sgd_sbase[k], base.data(), base.size(), st));
    CTC_TRY(dev_alloc(&m->sgd_stage[k], (size_t)std::max<uint64_t>(1, max_runs) * m->rank));
    CTC_TRY(dev_alloc(&m->sgd_part[k], (size_t)std::max<uint64_t>(1, max_pieces) * m->rank));
    CTC_HIP(hipStreamSynchronize(st));
  }
  const uint64_t long_min = (uint64_t)span - sub + 2;
  std::vector<uint32_t> dir(d.dlength), rows;
  CTC_HIP(hipMemcpy(dir.data(), d.directory, d.dlength * sizeof(uint32_t), hipMemcpyDeviceToHost));
  for (uint64_t sl = 0; sl < d.dlength; ++sl)
    if ((uint64_t)d.h_rowptr[sl + 1] - d.h_rowptr[sl] >= long_min) rows.push_back(dir[sl]);
  CTC_HIP(hipStreamSynchronize(st));
  dev_free(m->sgd_long);
  m->sgd_long = nullptr;
  CTC_TRY(upload(&m->sgd_long, rows.data(), rows.size(), st));
  CTC_HIP(hipStreamSynchronize(st));
  m->sgd_nlong = rows.size();
  m->sgd_long_span = span;
  m->sgd_long_sub = sub;
  m->sgd_long_thr = m->hot_threshold;
  return CTC_OK;
}

// One epoch of the bulk-synchronous schedule: per sub-step, k_sgd_sync over
// every walker, then k_sgd_settle over the pieces of the other modes' rows and
// k_sgd_fold over mode 0's long rows and the other modes' rows of several
// pieces.
// Enqueued only (no synchronisation).
static ctc_status sgd_epoch_sync(ctc_model *m, const ctc_options *o, double lr) {
  uint32_t span = 0, sub = 0;
  CTC_TRY(sync_shape(o, &span, &sub));
  CTC_TRY(ensure_hotbits(m, o->sgd_hot_rows ? o->sgd_hot_rows : kSgdHotRows));
  CTC_TRY(ensure_sync(m, span, sub));
  const uint32_t N = m->nmodes, R = m->rank;
  const bool mixed = o->precision == CTC_PREC_MIXED;  // run deltas of modes 1.. staged in BF16
  hipStream_t st = m->ctx->stream;
  SyncArgs sy{};
  SettleArgs sa{};
  MergeArgs &mg = sa.mg;
  mg.nseg = 0;
  mg.off[0] = 0;
  for (uint32_t k = 0; k < N; ++k) {
    sy.D[k] = m->sgd_D[k];  // CISS mode 0: CISS-relative order is the identity
    sy.hidx[k] = m->sgd_hidx[k];
    sy.Dh[k] = m->sgd_Dh[k];
    sy.ch[k] = m->sgd_ch[k];
    sy.stage[k] = m->sgd_stage[k];
    sy.sbase[k] = m->sgd_sbase[k];
    mg.F[k] = m->F[k];
    mg.D[k] = m->sgd_D[k];
    mg.Dh[k] = m->sgd_Dh[k];
    mg.ch[k] = m->sgd_ch[k];
    mg.hrow[k] = m->sgd_hrow[k];
    sa.F[k] = m->F[k];
    sa.hot[k] = m->hotbits[k];
    sa.stage[k] = m->sgd_stage[k];
    sa.piece_start[k] = m->sgd_piece_start[k];
    sa.piece_row[k] = m->sgd_piece_row[k];
    sa.slots[k] = m->sgd_slots[k];
    sa.part[k] = m->sgd_part[k];
    if (k > 0) continue;  // modes 1.. are merged from their staging slots
    // mode 0: its long rows' sums, then its hot rows' replicated sums
    mg.mode[mg.nseg] = 0;
    mg.is_hot[mg.nseg] = 0;
    mg.rows[mg.nseg] = m->sgd_long;
    mg.off[mg.nseg + 1] = mg.off[mg.nseg] + m->sgd_nlong;
    ++mg.nseg;
    if (m->sgd_nhot[k]) {  // hot segment: the mode's hot rows
      mg.mode[mg.nseg] = (int)k;
      mg.is_hot[mg.nseg] = 1;
      mg.rows[mg.nseg] = nullptr;
      mg.off[mg.nseg + 1] = mg.off[mg.nseg] + m->sgd_nhot[k];
      ++mg.nseg;
    }
  }
  const float step = (float)(lr * o->sgd_scale);
  const CissArgs c = ciss_args(m->ciss[0]);
  const FactorArgs f = rel_factors(m->F, N, 0);
  const uint32_t nsub = (span + sub - 1) / sub;
  Prof pe(m->ctx, "sgd");
  for (uint32_t s = 0; s < nsub; ++s) {
    for (uint32_t k = 1; k < N; ++k) sy.dest[k] = m->sgd_dest[k] + m->sgd_slot_off[k][s];
    {
      Prof p(m->ctx, "sgd_walk");
      CTC_HIP(launch_sgd_sync(st, c, f, sy, (int)N, (int)R, span, sub, s, span - sub + 2, step, (float)o->lambda,
                              block_of(o), mixed));
    }
    {
      Prof p(m->ctx, "sgd_merge");
      for (uint32_t k = 0; k < 4; ++k) {
        const bool staged = k >= 1 && k < N;
        sa.piece0[k] = staged ? m->sgd_piece_off[k][s] : 0;
        sa.slot0[k] = staged ? m->sgd_slot_off[k][s] : 0;
        sa.fold[k] = staged ? m->sgd_fold[k] + m->sgd_fold_off[k][s] : nullptr;
        sa.heavy[k] = staged ? m->sgd_heavy[k] + m->sgd_heavy_off[k][s] : nullptr;
        sa.light[k] = staged ? m->sgd_light[k] + m->sgd_light_off[k][s] : nullptr;
        sa.pb[k + 1] = sa.pb[k] + (staged ? m->sgd_heavy_off[k][s + 1] - m->sgd_heavy_off[k][s] : 0);
        sa.lb[k + 1] = sa.lb[k] + (staged ? m->sgd_light_off[k][s + 1] - m->sgd_light_off[k][s] : 0);
        sa.fb[k + 1] = sa.fb[k] + (staged ? m->sgd_fold_off[k][s + 1] - m->sgd_fold_off[k][s] : 0);
      }
      CTC_HIP(launch_sgd_settle(st, sa, (int)R, mixed));
    }
  }
  m->resid_valid = false;
  return CTC_OK;
}

static ctc_status sgd_epoch(ctc_model *m, const ctc_options *o, double lr) {
  if (o->ref_quirks) return fail(CTC_ERR_UNSUPPORTED, "ref_quirks reproduces the reference's ALS and CCD kernels only");
  CTC_TRY(ensure_ciss(m, 0));
  if (o->sgd_update == CTC_SGD_SYNC) return sgd_epoch_sync(m, o, lr);
  if (o->precision == CTC_PREC_MIXED)
    return fail(CTC_ERR_INVALID_ARG, "mixed-precision SGD runs the bulk-synchronous schedule (sgd_update %d)",
                (int)CTC_SGD_SYNC);
  // policy codes of k_sgd: PA * 64 + PB * 8 + PC with 0 racy, 1 atomic, 3 hot
  // bitmap (racy / atomic), 4 hot bitmap (racy / run-accumulated atomic)
  static const int code_of[5] = {0 * 64 + 4 * 8 + 4, 0 * 64 + 1 * 8 + 1, 0, 0 * 64 + 0 * 8 + 1, 0 * 64 + 3 * 8 + 3};
  if (o->sgd_update < 0 || o->sgd_update > 4) return fail(CTC_ERR_INVALID_ARG, "sgd_update %d", o->sgd_update);
  const int code = code_of[o->sgd_update];
  HotRows hot{};
  if (o->sgd_update == CTC_SGD_AUTO || o->sgd_update == CTC_SGD_HOT_ATOMIC) {
    CTC_TRY(ensure_hotbits(m, o->sgd_hot_rows ? o->sgd_hot_rows : kSgdHotRows));
    for (uint32_t s2 = 0; s2 < m->nmodes; ++s2) hot.bits[s2] = m->hotbits[s2];  // CISS mode 0: identity order
  }
  const float step = (float)(lr * o->sgd_scale);
  const uint32_t span = span_tiles_of(o) * 64;
  const uint64_t nspans = (m->ciss[0].nnz + span - 1) / span;
  {
    Prof p(m->ctx, "sgd");
    CTC_HIP(launch_sgd(m->ctx->stream, ciss_args(m->ciss[0]), rel_factors(m->F, m->nmodes, 0), hot, (int)m->nmodes,
                       (int)m->rank, span, spread_stride(nspans, o->spread != 0), code, step,
                       (float)o->lambda, block_of(o)));
  }
  m->resid_valid = false;
  return CTC_OK;
}

// `start`: the first mode of the reference's CCD form (ref_quirks), which
// updates whole modes in turn; CCD++ visits every mode per column.
static ctc_status ccd_epoch(ctc_model *m, const ctc_options *o, uint32_t start = 0) {
  if (!o->ref_quirks) return ctc::ccd_epoch_impl(m, o);
  if (o->ref_quirks & ~kQuirkAll) return fail(CTC_ERR_INVALID_ARG, "ref_quirks: unknown bits 0x%x", o->ref_quirks & ~kQuirkAll);
  if (!(o->ref_quirks & CTC_QUIRK_NO_ADDBACK))
    return fail(CTC_ERR_INVALID_ARG, "ref_quirks for CCD need CTC_QUIRK_NO_ADDBACK: the reference's CCD has that form only");
  return ctc::ccd_epoch_jacobi(m, o, start % m->nmodes);
}

extern "C" ctc_status ctc_als_mode_update(ctc_model *m, uint32_t mode, const ctc_options *o) {
  if (!m || !o || mode >= m->nmodes) return fail(CTC_ERR_INVALID_ARG, "ctc_als_mode_update");
  CTC_HIP(hipSetDevice(m->ctx->device));
  CTC_HIP(hipMemsetAsync(m->dflag, 0, sizeof(uint32_t), m->ctx->stream));
  CTC_TRY(als_mode(m, mode, o));
  CTC_TRY(check_flag(m));
  return prof_flush(m->ctx);
}

extern "C" ctc_status ctc_sgd_epoch(ctc_model *m, const ctc_options *o, double lr) {
  if (!m || !o) return fail(CTC_ERR_INVALID_ARG, "ctc_sgd_epoch");
  CTC_HIP(hipSetDevice(m->ctx->device));
  CTC_TRY(sgd_epoch(m, o, lr));
  CTC_HIP(hipStreamSynchronize(m->ctx->stream));
  return prof_flush(m->ctx);
}

extern "C" ctc_status ctc_mttkrp(ctc_model *m, uint32_t mode, float *host_out) {
  if (!m || !host_out || mode >= m->nmodes) return fail(CTC_ERR_INVALID_ARG, "ctc_mttkrp");
  CTC_HIP(hipSetDevice(m->ctx->device));
  CTC_TRY(ensure_ciss(m, mode));
  const uint32_t R = m->rank;
  const size_t n = m->dims[mode] * R;
  CTC_TRY(ensure_scratch(m, n * sizeof(float)));
  float *out = (float *)m->scratch;
  hipStream_t st = m->ctx->stream;
  CTC_HIP(launch_fill(st, out, n, 0.f));
  unsigned grid = 0;
  {
    Prof p(m->ctx, "mttkrp");
    CTC_HIP(launch_mttkrp(st, ciss_args(m->ciss[mode]), rel_factors(m->F, m->nmodes, mode), (int)m->nmodes,
                          (int)R, 256, false, out, nullptr, &grid, 256));
  }
  CTC_HIP(hipMemcpyAsync(host_out, out, n * sizeof(float), hipMemcpyDeviceToHost, st));
  CTC_HIP(hipStreamSynchronize(st));
  return prof_flush(m->ctx);
}

// =========================================================================
// evaluation (replaces tc_loss_sq / tc_rmse / tc_frob_sq,
// SG/completion.c:123-214, with FP64 device reductions)
// =========================================================================
extern "C" ctc_status ctc_evaluate(ctc_model *m, double lambda, double *loss, double *frob,
                                   double *rmse_tr, double *rmse_vl, double *rmse_te) {
  if (!m) return fail(CTC_ERR_INVALID_ARG, "ctc_evaluate: null model");
  CTC_HIP(hipSetDevice(m->ctx->device));
  CTC_TRY(ensure_ciss(m, 0));
  hipStream_t st = m->ctx->stream;
  const uint32_t N = m->nmodes, R = m->rank;
  const uint32_t groups = 3 + N;
  double *part = m->dstat + 64;  // [groups][kEvalBlocks] after the 64 result slots
  if ((size_t)groups * kEvalBlocks + 64 > 4096) return fail(CTC_ERR_INVALID_ARG, "eval buffer");
  FactorArgs f0 = rel_factors(m->F, N, 0);
  FactorArgs fa{};
  for (uint32_t md = 0; md < N; ++md) fa.f[md] = m->F[md];
  // the training loss as a walk of the mode-0 CISS, each walker keeping its
  // rows along their runs (kWalkLoss); its per-block sums are reduced in order
  // into group 0's first slot, the group's other slots zeroed
  const uint32_t span = 256;
  const uint64_t nb = mttkrp_grid(m->ciss[0].nnz, (int)R, span, 256);
  if (nb > m->eval_part_n) {
    CTC_HIP(hipStreamSynchronize(st));
    dev_free(m->eval_part);
    m->eval_part = nullptr;
    m->eval_part_n = 0;
    CTC_TRY(dev_alloc(&m->eval_part, nb));
    m->eval_part_n = nb;
  }
  {
    Prof p(m->ctx, "eval");
    CTC_HIP(launch_loss_walk(st, ciss_args(m->ciss[0]), f0, (int)N, (int)R, span, m->eval_part, 256));
    CTC_HIP(launch_fill(st, (float *)(part + 1), 2 * (kEvalBlocks - 1), 0.f));  // 0.0 doubles
    CTC_HIP(launch_reduce_parts(st, m->eval_part, (uint32_t)nb, 1, part));
    CooArgs cv{}, ct{};
    for (uint32_t md = 0; md < N; ++md) {
      cv.ind[md] = m->valid.ind[md];
      ct.ind[md] = m->test.ind[md];
    }
    cv.vals = m->valid.vals;
    cv.nnz = m->valid.nnz;
    ct.vals = m->test.vals;
    ct.nnz = m->test.nnz;
    CTC_HIP(launch_eval_coo(st, cv, fa, (int)N, (int)R, kEvalBlocks, part + kEvalBlocks));
    CTC_HIP(launch_eval_coo(st, ct, fa, (int)N, (int)R, kEvalBlocks, part + 2 * kEvalBlocks));
    for (uint32_t md = 0; md < N; ++md)
      CTC_HIP(launch_sumsq(st, m->F[md], m->dims[md] * R, kEvalBlocks, part + (3 + md) * kEvalBlocks));
    CTC_HIP(launch_reduce_parts(st, part, kEvalBlocks, groups, m->dstat));
  }
  CTC_HIP(hipMemcpyAsync(m->hstat, m->dstat, groups * sizeof(double), hipMemcpyDeviceToHost, st));
  CTC_HIP(hipStreamSynchronize(st));
  const double ltr = m->hstat[0], lvl = m->hstat[1], lte = m->hstat[2];
  double fr = 0.0;
  for (uint32_t md = 0; md < N; ++md) fr += lambda * m->hstat[3 + md];
  if (loss) *loss = ltr;
  if (frob) *frob = fr;
  if (rmse_tr) *rmse_tr = std::sqrt(ltr / (double)m->train_nnz);
  if (rmse_vl) *rmse_vl = m->valid.nnz ? std::sqrt(lvl / (double)m->valid.nnz) : 0.0;
  if (rmse_te) *rmse_te = m->test.nnz ? std::sqrt(lte / (double)m->test.nnz) : 0.0;
  return prof_flush(m->ctx);
}

// =========================================================================
// gradient (adjoint)
// =========================================================================
extern "C" ctc_status ctc_grad(ctc_model *m, const float *const *factors, float *const *grads,
                               double lambda, int reg_per_occurrence, double *loss, void *stream) {
  if (!m || !factors || !grads) return fail(CTC_ERR_INVALID_ARG, "ctc_grad: null argument");
  CTC_HIP(hipSetDevice(m->ctx->device));
  const uint32_t N = m->nmodes, R = m->rank;
  hipStream_t st = stream ? (hipStream_t)stream : m->ctx->stream;
  float *F[CTC_MAX_MODES] = {nullptr, nullptr, nullptr, nullptr};
  for (uint32_t md = 0; md < N; ++md) {
    if (!factors[md] || !grads[md]) return fail(CTC_ERR_INVALID_ARG, "ctc_grad: null factor/grad");
    F[md] = (float *)factors[md];
    CTC_TRY(ensure_ciss(m, md));
  }
  // partial sums: N walks + N regularisers, each up to `nb` blocks
  const uint32_t span = 256;
  uint64_t nb = 0;
  for (uint32_t md = 0; md < N; ++md) nb = std::max<uint64_t>(nb, mttkrp_grid(m->train_nnz, (int)R, span, 256));
  nb = std::max<uint64_t>(nb, kEvalBlocks);
  CTC_TRY(ensure_scratch(m, (size_t)2 * N * nb * sizeof(double)));
  double *part = (double *)m->scratch;
  CTC_HIP(hipMemsetAsync(part, 0, (size_t)2 * N * nb * sizeof(double), st));
  for (uint32_t md = 0; md < N; ++md) {
    CTC_HIP(launch_fill(st, grads[md], m->dims[md] * R, 0.f));
    unsigned grid = 0;
    {
      Prof p(m->ctx, "adjoint", st);
      CTC_HIP(launch_mttkrp(st, ciss_args(m->ciss[md]), rel_factors(F, N, md), (int)N, (int)R, span, true,
                            grads[md], part + (size_t)md * nb, &grid, 256));
    }
    const DevCiss &d = m->ciss[md];
    const uint64_t nrows = reg_per_occurrence ? d.dlength : m->dims[md];
    CTC_HIP(launch_reg_rows(st, d.directory, d.rowptr, nrows, (int)R, (float)lambda, reg_per_occurrence, F[md],
                            grads[md], kEvalBlocks, part + (size_t)(N + md) * nb));
  }
  CTC_HIP(launch_reduce_parts(st, part, (uint32_t)nb, 2 * N, m->dstat));
  CTC_HIP(hipMemcpyAsync(m->hstat, m->dstat, 2 * N * sizeof(double), hipMemcpyDeviceToHost, st));
  CTC_HIP(hipStreamSynchronize(st));
  if (loss) {
    double l = 0.5 * m->hstat[0];  // squared residuals from the mode-0 walk
    for (uint32_t md = 0; md < N; ++md) l += m->hstat[N + md];
    *loss = l;
  }
  return prof_flush(m->ctx);
}

// =========================================================================
// solver loop
// =========================================================================
extern "C" ctc_status ctc_ccd_epoch(ctc_model *m, const ctc_options *o) {
  if (!m || !o) return fail(CTC_ERR_INVALID_ARG, "ctc_ccd_epoch");
  CTC_HIP(hipSetDevice(m->ctx->device));
  CTC_TRY(ccd_epoch(m, o));
  CTC_HIP(hipStreamSynchronize(m->ctx->stream));
  return prof_flush(m->ctx);
}

static void print_progress(uint32_t epoch, const ctc_epoch_stats &s) {
  // the reference's progress line (SG/completion.c:111-113) plus RMSE-te
  printf("epoch:%u   loss: %0.5e   RMSE-tr: %0.5e   RMSE-vl: %0.5e   RMSE-te: %0.5e   ms: %0.3f\n", epoch,
         s.loss, s.rmse_tr, s.rmse_vl, s.rmse_te, s.epoch_ms);
  fflush(stdout);
}

extern "C" ctc_status ctc_solve(ctc_model *m, const ctc_options *o, ctc_result *res) {
  if (!m || !o || !res) return fail(CTC_ERR_INVALID_ARG, "ctc_solve: null argument");
  if (o->rank != 0 && o->rank != m->rank)
    return fail(CTC_ERR_INVALID_ARG, "options rank %u differs from model rank %u", o->rank, m->rank);
  if (o->algorithm < 0 || o->algorithm > 2) return fail(CTC_ERR_INVALID_ARG, "unknown algorithm");
  CTC_HIP(hipSetDevice(m->ctx->device));
  CTC_TRY(ctc_model_prepare(m, o));
  hipStream_t st = m->ctx->stream;
  const uint32_t N = m->nmodes;
  res->epochs = 0;
  res->best_epoch = 0;
  res->converged = 0;
  res->train_nnz = m->train_nnz;
  hipEvent_t e0, e1;
  CTC_HIP(hipEventCreate(&e0));
  CTC_HIP(hipEventCreate(&e1));
  struct EvGuard {
    hipEvent_t a, b;
    ~EvGuard() {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  } guard{e0, e1};

  // convergence state (SG/completion.c:454-458 initial values)
  double best_rmse = 1e7;
  uint32_t nbad = 0;
  auto record = [&](uint32_t e, const ctc_epoch_stats &s) {
    if (res->history && e < res->history_cap) res->history[e] = s;
  };
  auto evaluate = [&](ctc_epoch_stats &s) -> ctc_status {
    return ctc_evaluate(m, o->lambda, &s.loss, &s.frob, &s.rmse_tr, &s.rmse_vl, &s.rmse_te);
  };
  // returns 1 when converged (SG/completion.c:296-362)
  auto converge = [&](uint32_t e, const ctc_epoch_stats &s) -> ctc_status {
    const double r = m->valid.nnz ? s.rmse_vl : s.rmse_tr;
    if (r - best_rmse < -o->tolerance) {
      nbad = 0;
      best_rmse = r;
      res->best_epoch = e;
      for (uint32_t md = 0; md < N; ++md)
        CTC_HIP(hipMemcpyAsync(m->best[md], m->F[md], m->dims[md] * m->rank * sizeof(float),
                               hipMemcpyDeviceToDevice, st));
    } else {
      ++nbad;
      if (nbad == o->max_bad_epochs) res->converged = 1;
    }
    return CTC_OK;
  };

  ctc_epoch_stats s0{};
  CTC_TRY(evaluate(s0));
  s0.lr = o->learning_rate;
  record(0, s0);
  if (o->verbose) print_progress(0, s0);
  CTC_TRY(converge(0, s0));
  double lr = o->learning_rate;
  double prev_obj = s0.loss + s0.frob;
  std::mt19937_64 rot(o->seed ^ 0xA15ull);
  std::vector<double> times;
  CTC_HIP(hipMemsetAsync(m->dflag, 0, sizeof(uint32_t), st));
  for (uint32_t e = 1; e <= o->max_epochs; ++e) {
    CTC_HIP(hipEventRecord(e0, st));
    const uint32_t start = o->mode_order == CTC_ORDER_ROTATE ? (uint32_t)(rot() % N) : 0;
    if (o->algorithm == CTC_ALS) {
      for (uint32_t k = 0; k < N; ++k) CTC_TRY(als_mode(m, (start + k) % N, o));
    } else if (o->algorithm == CTC_SGD) {
      CTC_TRY(sgd_epoch(m, o, lr));
    } else {
      CTC_TRY(ccd_epoch(m, o, start));
    }
    CTC_HIP(hipEventRecord(e1, st));
    CTC_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    CTC_HIP(hipEventElapsedTime(&ms, e0, e1));
    times.push_back(ms);
    if (o->algorithm == CTC_ALS) CTC_TRY(check_flag(m));
    res->epochs = e;
    ctc_epoch_stats s{};
    s.epoch_ms = ms;
    s.lr = lr;
    if (o->eval_every_epoch || e == o->max_epochs) {
      CTC_TRY(evaluate(s));
      record(e, s);
      if (o->verbose) print_progress(e, s);
      CTC_TRY(converge(e, s));
    } else {
      record(e, s);
    }
    if (o->algorithm == CTC_SGD) {
      if (o->lr_schedule == CTC_LR_BOLD_DRIVER && o->eval_every_epoch) {
        const double obj = s.loss + s.frob;  // SG/sgd.cu:309, 314-323
        if (e > 1) lr *= obj < prev_obj ? 1.05 : 0.50;
        prev_obj = obj;
      } else if (o->lr_schedule == CTC_LR_EXP_DECAY) {
        lr *= o->lr_decay;
      }
    }
    if (res->converged) break;
  }
  CTC_TRY(prof_flush(m->ctx));
  std::vector<double> tail(times.size() > 1 ? times.begin() + 1 : times.begin(), times.end());
  res->ms_per_epoch = median(tail);
  res->nnz_per_s = res->ms_per_epoch > 0 ? (double)m->train_nnz / (res->ms_per_epoch * 1e-3) : 0.0;
  res->best_rmse = best_rmse;
  return CTC_OK;
}

// =========================================================================
// reference-shaped entry points
// =========================================================================
namespace {

ctc_status run_reference_shaped(const ctc_coo *traina, const ctc_coo *validation, const ctc_coo *test,
                                ctc_matrix **mats, ctc_matrix **best_mats, int alg, double lambda, double lr,
                                double *best_rmse, double *tolerance, uint64_t *nbad, uint64_t *bestep,
                                uint64_t *maxbad) {
  if (!traina || !mats) return fail(CTC_ERR_INVALID_ARG, "reference entry: null argument");
  const uint32_t R = (uint32_t)mats[0]->J;
  ctc_context *ctx = nullptr;
  CTC_TRY(ctc_context_create(0, &ctx));
  ctc_model *m = nullptr;
  ctc_status s = ctc_model_create(ctx, traina, validation, test, R, &m);
  if (s != CTC_OK) {
    ctc_context_destroy(ctx);
    return s;
  }
  for (uint32_t md = 0; md < traina->nmodes && s == CTC_OK; ++md) {
    std::vector<float> h(mats[md]->I * R);
    for (size_t x = 0; x < h.size(); ++x) h[x] = (float)mats[md]->values[x];
    s = ctc_model_set_factor(m, md, h.data());
  }
  ctc_options o;
  ctc_options_default(alg, &o);
  o.rank = R;
  o.lambda = lambda;
  o.learning_rate = lr;
  if (tolerance) o.tolerance = *tolerance;
  if (maxbad) o.max_bad_epochs = (uint32_t)*maxbad;
  o.verbose = 1;
  ctc_result r{};
  if (s == CTC_OK) s = ctc_solve(m, &o, &r);
  for (uint32_t md = 0; md < traina->nmodes && s == CTC_OK; ++md) {
    std::vector<float> h(mats[md]->I * R);
    s = ctc_model_get_factor(m, md, 0, h.data());
    for (size_t x = 0; x < h.size() && s == CTC_OK; ++x) mats[md]->values[x] = h[x];
    if (best_mats && best_mats[md] && s == CTC_OK) {
      s = ctc_model_get_factor(m, md, 1, h.data());
      for (size_t x = 0; x < h.size() && s == CTC_OK; ++x) best_mats[md]->values[x] = h[x];
    }
  }
  if (best_rmse) *best_rmse = r.best_rmse;
  if (bestep) *bestep = r.best_epoch;
  if (nbad) *nbad = r.converged ? o.max_bad_epochs : 0;
  ctc_model_destroy(m);
  ctc_context_destroy(ctx);
  return s;
}

}  // namespace

extern "C" ctc_status ctc_tc_als(const ctc_coo *traina, const ctc_coo *, const ctc_coo *,
                                 const ctc_coo *validation, const ctc_coo *test, ctc_matrix **mats,
                                 ctc_matrix **best_mats, int, double regularization_index, double *best_rmse,
                                 double *tolerance, uint64_t *nbadepochs, uint64_t *bestepochs,
                                 uint64_t *max_badepochs) {
  return run_reference_shaped(traina, validation, test, mats, best_mats, CTC_ALS, regularization_index, 1e-3,
                              best_rmse, tolerance, nbadepochs, bestepochs, max_badepochs);
}

extern "C" ctc_status ctc_tc_sgd(const ctc_coo *traina, const ctc_coo *, const ctc_coo *,
                                 const ctc_coo *validation, const ctc_coo *test, ctc_matrix **mats,
                                 ctc_matrix **best_mats, int, double regularization_index, double learning_rate,
                                 double *best_rmse, double *tolerance, uint64_t *nbadepochs,
                                 uint64_t *bestepochs, uint64_t *max_badepochs) {
  return run_reference_shaped(traina, validation, test, mats, best_mats, CTC_SGD, regularization_index,
                              learning_rate, best_rmse, tolerance, nbadepochs, bestepochs, max_badepochs);
}

extern "C" ctc_status ctc_tc_ccd(const ctc_coo *traina, const ctc_coo *, const ctc_coo *,
                                 const ctc_coo *validation, const ctc_coo *test, ctc_matrix **mats,
                                 ctc_matrix **best_mats, int, double regularization_index, double *best_rmse,
                                 double *tolerance, uint64_t *nbadepochs, uint64_t *bestepochs,
                                 uint64_t *max_badepochs) {
  return run_reference_shaped(traina, validation, test, mats, best_mats, CTC_CCD, regularization_index, 1e-3,
                              best_rmse, tolerance, nbadepochs, bestepochs, max_badepochs);
}

extern "C" ctc_status ctc_tc_main_ciss(const ctc_coo *traina, const ctc_coo *, const ctc_coo *,
                                       const ctc_coo *validation, const ctc_coo *test, int algorithm_index,
                                       const ctc_options *opt) {
  if (!traina) return fail(CTC_ERR_INVALID_ARG, "ctc_tc_main_ciss: null train");
  ctc_options o;
  if (opt)
    o = *opt;
  else
    CTC_TRY(ctc_options_default(algorithm_index, &o));
  o.algorithm = algorithm_index;
  o.verbose = 1;
  ctc_context *ctx = nullptr;
  CTC_TRY(ctc_context_create(0, &ctx));
  ctc_model *m = nullptr;
  ctc_status s = ctc_model_create(ctx, traina, validation, test, o.rank, &m);
  if (s == CTC_OK) s = ctc_model_init_random(m, o.seed);
  ctc_result r{};
  if (s == CTC_OK) s = ctc_solve(m, &o, &r);
  if (m) ctc_model_destroy(m);
  ctc_context_destroy(ctx);
  return s;
}

// =========================================================================
// autotuning (replaces the GA over blocksize x tile length, GA/fitness.c:31-100)
// =========================================================================
extern "C" ctc_status ctc_autotune(ctc_model *m, const ctc_options *o, ctc_tune_result *out) {
  if (!m || !o || !out) return fail(CTC_ERR_INVALID_ARG, "ctc_autotune");
  CTC_HIP(hipSetDevice(m->ctx->device));
  const uint32_t N = m->nmodes;
  hipStream_t st = m->ctx->stream;
  // save the factors and the best-model copy: tuning sweeps modify the
  // factors, and the restore must leave the best copy as it was
  std::vector<std::vector<float>> saved(N), saved_best(N);
  for (uint32_t md = 0; md < N; ++md) {
    saved[md].resize(m->dims[md] * m->rank);
    saved_best[md].resize(m->dims[md] * m->rank);
    CTC_TRY(ctc_model_get_factor(m, md, 0, saved[md].data()));
    CTC_TRY(ctc_model_get_factor(m, md, 1, saved_best[md].data()));
  }
  // search space: (tile_len, block), the GA's two knobs (GA/fitness.c:67-79);
  // the bulk-synchronous SGD schedule also tunes its sub-step, and its walker
  // span and sub-step define the schedule, so the tuned shape is what a later
  // solve with these options computes
  const bool sync = o->algorithm == CTC_SGD && o->sgd_update == CTC_SGD_SYNC;
  std::vector<uint32_t> tiles = {64, 128, 256, 512, 1024, 2048, 4096};
  std::vector<uint32_t> subs = {o->sgd_substep};
  std::vector<unsigned> blocks = {64, 128, 256};
  if (sync) {
    tiles = {1024, 2048, 4096, 8192};
    subs = {16, 32, 64};
    blocks = {64, 256};
  }
  hipEvent_t e0, e1;
  CTC_HIP(hipEventCreate(&e0));
  CTC_HIP(hipEventCreate(&e1));
  double best = 1e300;
  ctc_tune_result r{o->tile_len, o->block, 0.0, 0, o->sgd_substep};
  ctc_status s = CTC_OK;
  for (uint32_t tl : tiles) {
    for (uint32_t sb : subs) {
      // the shape is part of the schedule: only shapes of at least the
      // default's span / sub-step = 64 sub-steps per epoch, so that no more
      // than 1/64 of the nonzeros is computed from one snapshot (fewer
      // sub-steps ran faster and diverged: 256 / 64 turned to NaN in six
      // epochs on 12.5M nonzeros of the benchmark generator)
      if (sync && tl < 64 * sb) continue;
      for (unsigned b : blocks) {
        ctc_options t = *o;
        t.tile_len = tl;
        t.block = b;
        t.sgd_substep = sb;
        double tot = 0.0;
        const int reps = 3;
        for (int rep = 0; rep <= reps && s == CTC_OK; ++rep) {  // rep 0: warm-up (builds tables)
          (void)hipEventRecord(e0, st);
          if (o->algorithm == CTC_SGD)
            s = sgd_epoch(m, &t, 0.0);  // lr 0: same memory traffic, factors unchanged
          else if (o->algorithm == CTC_ALS)
            for (uint32_t md = 0; md < N && s == CTC_OK; ++md) s = als_mode(m, md, &t);
          else
            s = ccd_epoch(m, &t);
          (void)hipEventRecord(e1, st);
          (void)hipEventSynchronize(e1);
          float ms = 0.f;
          (void)hipEventElapsedTime(&ms, e0, e1);
          if (rep > 0) tot += ms;
        }
        if (s != CTC_OK) break;
        r.trials++;
        const double avg = tot / reps;
        if (avg < best) {
          best = avg;
          r.tile_len = tl;
          r.block = b;
          r.sgd_substep = sb;
        }
      }
      if (s != CTC_OK) break;
    }
    if (s != CTC_OK) break;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  for (uint32_t md = 0; md < N; ++md) {
    const size_t bytes = saved[md].size() * sizeof(float);
    hipError_t e = hipMemcpyAsync(m->F[md], saved[md].data(), bytes, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(m->best[md], saved_best[md].data(), bytes, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess && s == CTC_OK) s = fail(CTC_ERR_HIP, "autotune restore: %s", hipGetErrorString(e));
  }
  m->resid_valid = false;
  // the trial sweeps' status flags do not describe the restored model
  (void)hipMemsetAsync(m->dflag, 0, sizeof(uint32_t), st);
  (void)hipStreamSynchronize(st);
  r.ms = best;
  *out = r;
  return s;
}
