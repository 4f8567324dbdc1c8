// batch.cpp — many independent ALS problems solved together.
//
// Problems are grouped by rank.  The problems of one rank class are embedded
// as one block-diagonal tensor: problem p's rows of mode m are offset past the
// rows of the problems before it, so its nonzeros only ever meet its own
// factor rows.  ALS on that tensor is ALS on every problem at once (the Gram,
// MTTKRP and solve of a row only involve the row's own nonzeros), and each
// phase of a sweep is one kernel launch over all problems of the class, with
// the row-segmented reductions keeping light and heavy rows of different
// problems in the same launch.  Each class runs on its own stream, forked
// from and joined to the caller's stream; a run of `iters` iterations can be
// captured once into a hipGraph with one branch per class and replayed.
//
// Because mode-0 global ids preserve problem order, each problem's training
// nonzeros form one contiguous range of the class's mode-0 CISS and its test
// nonzeros one contiguous range of the class's test COO; per-problem RMSE is
// one block per range.
#include <algorithm>
#include <cmath>
#include <vector>

#include "ctc_internal.h"
#include "kernels_launch.h"
#include "model_util.h"

namespace {

struct RankClass {
  uint32_t rank = 0;
  std::vector<uint32_t> problems;  // batch problem ids, in order
  ctc_context *ctx = nullptr;      // own stream
  ctc_model *model = nullptr;
  uint64_t *seg_train = nullptr, *seg_test = nullptr;  // device, nprob + 1 each
  double *dsum = nullptr;                              // device, 2 * nprob
  bool tuned = false;
  uint32_t tile_len = 0, block = 0;
  hipEvent_t join = nullptr;
};

struct Placement {
  uint32_t cls = 0, slot = 0;  // class and position inside the class
  uint64_t row_off[CTC_MAX_MODES] = {0, 0, 0, 0};
  uint64_t dims[CTC_MAX_MODES] = {0, 0, 0, 0};
  uint64_t ntrain = 0, ntest = 0;
};

}  // namespace

struct ctc_batch {
  ctc_context *ctx = nullptr;
  uint32_t nmodes = 0, nproblems = 0;
  uint64_t total_train = 0;
  std::vector<RankClass> cls;
  std::vector<Placement> place;
  hipEvent_t fork = nullptr, t0 = nullptr, t1 = nullptr;
  // captured run
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint32_t graph_iters = 0;
  std::vector<uintptr_t> graph_key;  // options and buffer addresses the graph baked in
};

using namespace ctc;

static void batch_release_graph(ctc_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->graph_key.clear();
}

extern "C" void ctc_batch_destroy(ctc_batch *b) {
  if (!b) return;
  (void)hipSetDevice(b->ctx->device);
  (void)hipStreamSynchronize(b->ctx->stream);
  batch_release_graph(b);
  for (auto &c : b->cls) {
    if (c.model) ctc_model_destroy(c.model);
    if (c.seg_train) (void)hipFree(c.seg_train);
    if (c.seg_test) (void)hipFree(c.seg_test);
    if (c.dsum) (void)hipFree(c.dsum);
    if (c.join) (void)hipEventDestroy(c.join);
    if (c.ctx) ctc_context_destroy(c.ctx);
  }
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->t0) (void)hipEventDestroy(b->t0);
  if (b->t1) (void)hipEventDestroy(b->t1);
  ctc_context *ctx = b->ctx;
  delete b;
  context_release(ctx);
}

// concatenates the given problems' tensors with per-mode row offsets
static ctc_status concat(uint32_t nmodes, const std::vector<const ctc_coo *> &parts,
                         const std::vector<const uint64_t *> &offs, const uint64_t *dims, ctc_coo **out) {
  uint64_t nnz = 0;
  for (auto *t : parts) nnz += t ? t->nnz : 0;
  CTC_TRY(ctc_coo_alloc(nmodes, nnz, dims, out));
  uint64_t p = 0;
  for (size_t k = 0; k < parts.size(); ++k) {
    const ctc_coo *t = parts[k];
    if (!t) continue;
    for (uint32_t m = 0; m < nmodes; ++m) {
      const uint32_t o = (uint32_t)offs[k][m];
      for (uint64_t q = 0; q < t->nnz; ++q) (*out)->ind[m][p + q] = t->ind[m][q] + o;
    }
    std::copy(t->vals, t->vals + t->nnz, (*out)->vals + p);
    p += t->nnz;
  }
  return CTC_OK;
}

extern "C" ctc_status ctc_batch_create(ctc_context *ctx, uint32_t nproblems, const ctc_coo *const *train,
                                       const ctc_coo *const *test, const uint32_t *ranks, ctc_batch **out) {
  if (!ctx || !train || !ranks || !out || nproblems == 0) return fail(CTC_ERR_INVALID_ARG, "ctc_batch_create");
  *out = nullptr;
  const uint32_t N = train[0] ? train[0]->nmodes : 0;
  if (N < 3 || N > CTC_MAX_MODES) return fail(CTC_ERR_INVALID_ARG, "ctc_batch_create: problem 0 has %u modes", N);
  for (uint32_t p = 0; p < nproblems; ++p) {
    if (!train[p] || train[p]->nmodes != N)
      return fail(CTC_ERR_INVALID_ARG, "ctc_batch_create: problem %u: missing or with a different mode count", p);
    if (test && test[p] && test[p]->nmodes != N)
      return fail(CTC_ERR_INVALID_ARG, "ctc_batch_create: test tensor of problem %u", p);
    const uint32_t R = ranks[p];
    if (R != 8 && R != 16 && R != 32 && R != 64)
      return fail(CTC_ERR_UNSUPPORTED, "ctc_batch_create: problem %u has rank %u (8, 16, 32 or 64)", p, R);
  }
  CTC_HIP(hipSetDevice(ctx->device));
  ctc_batch *b = new ctc_batch();
  b->ctx = ctx;
  context_retain(ctx);
  b->nmodes = N;
  b->nproblems = nproblems;
  b->place.resize(nproblems);
  ctc_status s = CTC_OK;
  const uint32_t class_ranks[4] = {8, 16, 32, 64};
  for (uint32_t R : class_ranks) {
    RankClass c;
    c.rank = R;
    for (uint32_t p = 0; p < nproblems; ++p)
      if (ranks[p] == R) c.problems.push_back(p);
    if (!c.problems.empty()) b->cls.push_back(std::move(c));
  }
  for (uint32_t ci = 0; ci < b->cls.size() && s == CTC_OK; ++ci) {
    RankClass &c = b->cls[ci];
    const size_t np = c.problems.size();
    uint64_t dims[CTC_MAX_MODES] = {0, 0, 0, 0};
    std::vector<const ctc_coo *> tr(np), te(np);
    std::vector<const uint64_t *> offs(np);
    std::vector<uint64_t> seg_tr(np + 1, 0), seg_te(np + 1, 0);
    for (size_t k = 0; k < np; ++k) {
      const uint32_t p = c.problems[k];
      Placement &pl = b->place[p];
      pl.cls = ci;
      pl.slot = (uint32_t)k;
      for (uint32_t m = 0; m < N; ++m) {
        pl.row_off[m] = dims[m];
        pl.dims[m] = train[p]->dims[m];
        if (test && test[p]) pl.dims[m] = std::max(pl.dims[m], test[p]->dims[m]);
        dims[m] += pl.dims[m];
      }
      pl.ntrain = train[p]->nnz;
      pl.ntest = test && test[p] ? test[p]->nnz : 0;
      tr[k] = train[p];
      te[k] = test ? test[p] : nullptr;
      offs[k] = pl.row_off;
      seg_tr[k + 1] = seg_tr[k] + pl.ntrain;
      seg_te[k + 1] = seg_te[k] + pl.ntest;
    }
    for (uint32_t m = 0; m < N && s == CTC_OK; ++m)
      if (dims[m] > 0xFFFFFFFFull) s = fail(CTC_ERR_UNSUPPORTED, "rank class %u: more than 2^32 rows", c.rank);
    ctc_coo *ctr = nullptr, *cte = nullptr;
    if (s == CTC_OK) s = concat(N, tr, offs, dims, &ctr);
    if (s == CTC_OK) s = concat(N, te, offs, dims, &cte);
    if (s == CTC_OK) s = ctc_context_create(ctx->device, &c.ctx);
    if (s == CTC_OK) s = ctc_model_create(c.ctx, ctr, nullptr, cte, c.rank, &c.model);
    ctc_coo_free(ctr);
    ctc_coo_free(cte);
    if (s != CTC_OK) break;
    b->total_train += seg_tr[np];
    if (hipMalloc(&c.seg_train, (np + 1) * sizeof(uint64_t)) != hipSuccess ||
        hipMalloc(&c.seg_test, (np + 1) * sizeof(uint64_t)) != hipSuccess ||
        hipMalloc(&c.dsum, 2 * np * sizeof(double)) != hipSuccess) {
      s = fail(CTC_ERR_NOMEM, "batch segment tables");
      break;
    }
    if (hipMemcpy(c.seg_train, seg_tr.data(), (np + 1) * sizeof(uint64_t), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c.seg_test, seg_te.data(), (np + 1) * sizeof(uint64_t), hipMemcpyHostToDevice) != hipSuccess ||
        hipEventCreateWithFlags(&c.join, hipEventDisableTiming) != hipSuccess) {
      s = fail(CTC_ERR_HIP, "batch segment upload");
      break;
    }
  }
  if (s == CTC_OK &&
      (hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
       hipEventCreate(&b->t0) != hipSuccess || hipEventCreate(&b->t1) != hipSuccess))
    s = fail(CTC_ERR_HIP, "batch events");
  if (s != CTC_OK) {
    ctc_batch_destroy(b);
    return s;
  }
  *out = b;
  return CTC_OK;
}

extern "C" ctc_status ctc_batch_info(const ctc_batch *b, uint32_t *nproblems, uint32_t *nclasses,
                                     uint64_t *train_nnz) {
  if (!b) return fail(CTC_ERR_INVALID_ARG, "ctc_batch_info");
  if (nproblems) *nproblems = b->nproblems;
  if (nclasses) *nclasses = (uint32_t)b->cls.size();
  if (train_nnz) *train_nnz = b->total_train;
  return CTC_OK;
}

static ctc_status factor_io(ctc_batch *b, uint32_t p, uint32_t mode, float *host, bool get) {
  if (!b || !host || p >= b->nproblems || mode >= b->nmodes)
    return fail(CTC_ERR_INVALID_ARG, "ctc_batch_%s_factor", get ? "get" : "set");
  CTC_HIP(hipSetDevice(b->ctx->device));
  const Placement &pl = b->place[p];
  RankClass &c = b->cls[pl.cls];
  const size_t R = c.rank;
  float *dev = c.model->F[mode] + pl.row_off[mode] * R;
  const size_t bytes = pl.dims[mode] * R * sizeof(float);
  hipStream_t st = c.ctx->stream;
  if (get)
    CTC_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, st));
  else
    CTC_HIP(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, st));
  CTC_HIP(hipStreamSynchronize(st));
  return CTC_OK;
}

extern "C" ctc_status ctc_batch_set_factor(ctc_batch *b, uint32_t problem, uint32_t mode, const float *host) {
  return factor_io(b, problem, mode, const_cast<float *>(host), false);
}

extern "C" ctc_status ctc_batch_get_factor(ctc_batch *b, uint32_t problem, uint32_t mode, float *host) {
  return factor_io(b, problem, mode, host, true);
}

// the precision a class runs in: the BF16-operand path exists for the
// row-fused kernels' ranks, of which a batch has 32 and 64; ranks 8 and 16
// keep the FP32 kernels
static int32_t class_precision(uint32_t rank, int32_t asked) {
  return asked == CTC_PREC_MIXED && (rank == 32 || rank == 64) ? CTC_PREC_MIXED : CTC_PREC_FP32;
}

static ctc_options class_options(const RankClass &c, const ctc_options *o) {
  ctc_options t = *o;
  t.algorithm = CTC_ALS;
  t.precision = class_precision(c.rank, o->precision);
  if (c.tuned) {
    t.tile_len = c.tile_len;
    t.block = c.block;
  }
  return t;
}

// everything an ALS sweep allocates or uploads, done before any enqueue so
// that the sweep itself only launches kernels
static ctc_status prepare_class(RankClass &c, const ctc_options *t) {
  ctc_model *m = c.model;
  if (als_uses_rows(t, c.rank)) return als_rows_prepare(m, t, t->precision == CTC_PREC_MIXED);
  const uint32_t S = span_tiles_of(t);
  size_t need = 0;
  for (uint32_t md = 0; md < m->nmodes; ++md) {
    CTC_TRY(ensure_ciss(m, md));
    CTC_TRY(ensure_pieces(m, md, S));
    const DevCiss &d = m->ciss[md];
    need = std::max(need, (size_t)(d.npieces + d.ngroups) * ((size_t)c.rank * c.rank + c.rank) * sizeof(double));
  }
  return ensure_scratch(m, need);
}

extern "C" ctc_status ctc_batch_autotune(ctc_batch *b, const ctc_options *o, ctc_tune_result *results,
                                         uint32_t *nclasses) {
  if (!b || !o) return fail(CTC_ERR_INVALID_ARG, "ctc_batch_autotune");
  CTC_HIP(hipSetDevice(b->ctx->device));
  batch_release_graph(b);
  for (size_t ci = 0; ci < b->cls.size(); ++ci) {
    RankClass &c = b->cls[ci];
    ctc_options t = *o;
    t.algorithm = CTC_ALS;
    t.precision = class_precision(c.rank, o->precision);
    ctc_tune_result r{};
    CTC_TRY(ctc_autotune(c.model, &t, &r));
    c.tuned = true;
    c.tile_len = r.tile_len;
    c.block = r.block;
    if (results) results[ci] = r;
  }
  if (nclasses) *nclasses = (uint32_t)b->cls.size();
  return CTC_OK;
}

extern "C" ctc_status ctc_batch_class_rank(const ctc_batch *b, uint32_t cls, uint32_t *rank, uint32_t *nproblems) {
  if (!b || cls >= b->cls.size()) return fail(CTC_ERR_INVALID_ARG, "ctc_batch_class_rank");
  if (rank) *rank = b->cls[cls].rank;
  if (nproblems) *nproblems = (uint32_t)b->cls[cls].problems.size();
  return CTC_OK;
}

// enqueue `iters` sweeps of every class: fork from the caller's stream into
// one branch per class stream, join back.  Under capture the same calls
// record a graph whose class branches are concurrent.
static ctc_status enqueue_run(ctc_batch *b, const ctc_options *o, uint32_t iters) {
  hipStream_t main = b->ctx->stream;
  CTC_HIP(hipEventRecord(b->fork, main));
  for (auto &c : b->cls) {
    const ctc_options t = class_options(c, o);
    hipStream_t st = c.ctx->stream;
    CTC_HIP(hipStreamWaitEvent(st, b->fork, 0));
    for (uint32_t it = 0; it < iters; ++it)
      for (uint32_t md = 0; md < b->nmodes; ++md) CTC_TRY(als_mode_impl(c.model, md, &t));
    CTC_HIP(hipEventRecord(c.join, st));
    CTC_HIP(hipStreamWaitEvent(main, c.join, 0));
  }
  return CTC_OK;
}

static std::vector<uintptr_t> graph_key_of(ctc_batch *b, const ctc_options *o, uint32_t iters) {
  std::vector<uintptr_t> k{iters, (uintptr_t)std::llround(o->lambda * 1e12), (uintptr_t)o->refine,
                           (uintptr_t)o->precision};
  for (auto &c : b->cls) {
    const ctc_options t = class_options(c, o);
    k.push_back(t.tile_len);
    k.push_back(t.block);
    k.push_back((uintptr_t)t.als_kernel);
    k.push_back((uintptr_t)c.model->scratch);
    for (uint32_t md = 0; md < b->nmodes; ++md) {
      const DevCiss &d = c.model->ciss[md];
      k.push_back((uintptr_t)d.slot_pb);
      k.push_back((uintptr_t)d.span_off);
      k.push_back((uintptr_t)d.slot_gb);
      k.push_back((uintptr_t)d.mix_items);
      k.push_back((uintptr_t)d.mix_chunk);
      k.push_back((uintptr_t)c.model->F[md]);
      k.push_back((uintptr_t)c.model->Fbf[md]);
    }
  }
  return k;
}

extern "C" ctc_status ctc_batch_als(ctc_batch *b, const ctc_options *o, uint32_t iters, int32_t use_graph,
                                    double *ms) {
  if (!b || !o) return fail(CTC_ERR_INVALID_ARG, "ctc_batch_als");
  if (o->precision != CTC_PREC_FP32 && o->precision != CTC_PREC_MIXED)
    return fail(CTC_ERR_INVALID_ARG, "ctc_batch_als: unknown precision %d", (int)o->precision);
  if (o->ref_quirks) return fail(CTC_ERR_UNSUPPORTED, "ctc_batch_als: ref_quirks applies to single models");
  CTC_HIP(hipSetDevice(b->ctx->device));
  for (auto &c : b->cls) {
    const ctc_options t = class_options(c, o);
    CTC_TRY(prepare_class(c, &t));
    // the not-positive-definite flag starts clear (outside any capture)
    CTC_HIP(hipMemsetAsync(c.model->dflag, 0, sizeof(uint32_t), c.ctx->stream));
    CTC_HIP(hipStreamSynchronize(c.ctx->stream));
  }
  hipStream_t main = b->ctx->stream;
  if (use_graph) {
    std::vector<uintptr_t> key = graph_key_of(b, o, iters);
    if (!b->exec || key != b->graph_key) {
      batch_release_graph(b);
      CTC_HIP(hipStreamBeginCapture(main, hipStreamCaptureModeRelaxed));
      ctc_status s = enqueue_run(b, o, iters);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(main, &g);
      if (s != CTC_OK) {
        if (g) (void)hipGraphDestroy(g);
        return s;
      }
      if (e != hipSuccess) return fail(CTC_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
      b->graph = g;
      CTC_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
      b->graph_key = key;
      b->graph_iters = iters;
    }
    CTC_HIP(hipEventRecord(b->t0, main));
    CTC_HIP(hipGraphLaunch(b->exec, main));
    CTC_HIP(hipEventRecord(b->t1, main));
  } else {
    CTC_HIP(hipEventRecord(b->t0, main));
    CTC_TRY(enqueue_run(b, o, iters));
    CTC_HIP(hipEventRecord(b->t1, main));
  }
  CTC_HIP(hipEventSynchronize(b->t1));
  float t = 0.f;
  CTC_HIP(hipEventElapsedTime(&t, b->t0, b->t1));
  if (ms) *ms = t;
  ctc_status s = CTC_OK;
  for (auto &c : b->cls) {
    if (check_not_spd(c.model) != CTC_OK && s == CTC_OK)
      s = fail(CTC_ERR_NOT_SPD, "rank-%u class: a normal system (G + lambda I) was not positive definite", c.rank);
  }
  return s;
}

extern "C" ctc_status ctc_batch_rmse(ctc_batch *b, double *rmse_tr, double *rmse_te) {
  if (!b) return fail(CTC_ERR_INVALID_ARG, "ctc_batch_rmse");
  CTC_HIP(hipSetDevice(b->ctx->device));
  for (auto &c : b->cls) {
    ctc_model *m = c.model;
    const uint32_t np = (uint32_t)c.problems.size();
    hipStream_t st = c.ctx->stream;
    CTC_TRY(ensure_ciss(m, 0));
    const CissArgs ca = ciss_args(m->ciss[0]);
    CooArgs ct{};
    for (uint32_t md = 0; md < m->nmodes; ++md) ct.ind[md] = m->test.ind[md];
    ct.vals = m->test.vals;
    ct.nnz = m->test.nnz;
    FactorArgs fa{};
    for (uint32_t md = 0; md < m->nmodes; ++md) fa.f[md] = m->F[md];
    CTC_HIP(launch_eval_seg(st, &ca, nullptr, fa, (int)m->nmodes, (int)c.rank, c.seg_train, np, c.dsum));
    if (m->test.nnz)
      CTC_HIP(launch_eval_seg(st, nullptr, &ct, fa, (int)m->nmodes, (int)c.rank, c.seg_test, np, c.dsum + np));
    std::vector<double> h(2 * (size_t)np, 0.0);
    CTC_HIP(hipMemcpyAsync(h.data(), c.dsum, (m->test.nnz ? 2 : 1) * np * sizeof(double), hipMemcpyDeviceToHost,
                           st));
    CTC_HIP(hipStreamSynchronize(st));
    for (uint32_t k = 0; k < np; ++k) {
      const Placement &pl = b->place[c.problems[k]];
      if (rmse_tr) rmse_tr[c.problems[k]] = pl.ntrain ? std::sqrt(h[k] / (double)pl.ntrain) : 0.0;
      if (rmse_te) rmse_te[c.problems[k]] = pl.ntest ? std::sqrt(h[np + k] / (double)pl.ntest) : 0.0;
    }
  }
  return CTC_OK;
}
