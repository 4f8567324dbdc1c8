// mixed.cpp — host driver of the row-fused ALS mode update (kernels in
// kernels_mixed.hip), in its BF16-operand (mixed precision) and FP64-Gram
// (FP32 factors) forms.
//
// Per mode update: for BF16 the other modes' factors are rounded to mirrors, the
// factor being updated is zeroed (rows without nonzeros stay 0, as
// SG/als.cu:699 does), the items pass solves every row of at most
// `chunk` nonzeros in one workgroup and writes partial systems for the chunks
// of longer rows, the heavy pass sums each long row's partials and solves it,
// the thin pass solves the rows of at most min(32, R) nonzeros in their dual
// n x n form, one wave each, and at R = 128 the dual pass does the same for
// rows of 33..64 nonzeros (systems padded to 48 or 64).
#include <algorithm>
#include <vector>

#include "ctc_internal.h"
#include "kernels_launch.h"
#include "kernels_mixed.h"
#include "model_util.h"

namespace ctc {

// nonzeros per work item (rows longer than this are cut into chunks): the
// option's tile length, a multiple of 32.  Defaults (tile_len 0): 1024 for
// the FP64-Gram path (measured on the 100M-nonzero tensor at R = 32); for the
// BF16 path, whose per-row Gram is cheaper, 8192 at rank 128 (450M-nonzero
// rank-128 tensor, s per sweep: 1024 -> 0.308, 2048 -> 0.297, 4096 -> 0.291,
// 8192 -> 0.288) and 4096 below it (what the autotuner picks for the rank-32
// and rank-64 classes of the batched workload).
static uint32_t mix_chunk_of(const ctc_options *o, uint32_t R, bool bf16) {
  uint32_t c = o && o->tile_len ? o->tile_len : (!bf16 ? 1024 : R >= 128 ? 8192 : 4096);
  return std::max<uint32_t>(32, (c + 31) / 32 * 32);
}

static ctc_status ensure_mix_items(ctc_model *m, uint32_t mode, uint32_t chunk) {
  DevCiss &d = m->ciss[mode];
  if (d.mix_items && d.mix_chunk == chunk) return CTC_OK;
  const uint32_t thin = std::min<uint32_t>((uint32_t)kThinRow, m->rank);
  std::vector<uint4> chunks, light, thins, thins32, duals, duals64;
  // dual 64 x 64 systems for rows of 33..64 nonzeros pay at R = 128 (at R = 64
  // the primal 4 x 4-tile solve measured faster)
  const uint32_t dual = m->rank >= 2 * (uint32_t)kDualRow ? (uint32_t)kDualRow : 0u;
  std::vector<uint32_t> hslot, hpb{0};
  uint32_t nparts = 0;
  for (uint64_t s = 0; s < d.dlength; ++s) {
    const uint32_t p0 = d.h_rowptr[s], p1 = d.h_rowptr[s + 1];
    if (p1 - p0 <= thin) {  // 16-row systems first, then 32-row ones
      (p1 - p0 <= 16 ? thins : thins32).push_back(make_uint4((uint32_t)s, p0, p1, kLightRow));
      continue;
    }
    if (p1 - p0 <= dual) {  // 48-row systems first, then 64-row ones
      (p1 - p0 <= 48 ? duals : duals64).push_back(make_uint4((uint32_t)s, p0, p1, kLightRow));
      continue;
    }
    if (p1 - p0 <= chunk) {
      light.push_back(make_uint4((uint32_t)s, p0, p1, kLightRow));
      continue;
    }
    for (uint32_t p = p0; p < p1; p += chunk) chunks.push_back(make_uint4((uint32_t)s, p, std::min(p1, p + chunk), nparts++));
    hslot.push_back((uint32_t)s);
    hpb.push_back(nparts);
  }
  // the long rows' chunks first, so the longest work starts earliest; the
  // thin rows (dual-form kernel) last
  const size_t nthin16 = thins.size();
  thins.insert(thins.end(), thins32.begin(), thins32.end());
  const size_t ndual48 = duals.size();
  duals.insert(duals.end(), duals64.begin(), duals64.end());
  std::vector<uint4> items;
  items.reserve(chunks.size() + light.size() + thins.size());
  items.insert(items.end(), chunks.begin(), chunks.end());
  items.insert(items.end(), light.begin(), light.end());
  items.insert(items.end(), thins.begin(), thins.end());
  items.insert(items.end(), duals.begin(), duals.end());
  // long rows' partials in groups of kGroup (rows with more than kGroupMin)
  std::vector<uint2> grp;
  std::vector<uint32_t> hgb{0};
  for (size_t h = 0; h + 1 < hpb.size(); ++h) {
    if (hpb[h + 1] - hpb[h] > kGroupMin)
      for (uint32_t q = hpb[h]; q < hpb[h + 1]; q += kGroup)
        grp.push_back(make_uint2(q, std::min(kGroup, hpb[h + 1] - q)));
    hgb.push_back((uint32_t)grp.size());
  }
  for (void *p : {(void *)d.mix_items, (void *)d.mix_heavy_slot, (void *)d.mix_heavy_pb, (void *)d.mix_grp,
                  (void *)d.mix_heavy_gb})
    if (p) CTC_HIP(hipFree(p));
  d.mix_items = nullptr;
  d.mix_heavy_slot = d.mix_heavy_pb = d.mix_heavy_gb = nullptr;
  d.mix_grp = nullptr;
  CTC_HIP(hipMalloc(&d.mix_grp, std::max<size_t>(1, grp.size()) * sizeof(uint2)));
  CTC_HIP(hipMalloc(&d.mix_heavy_gb, hgb.size() * sizeof(uint32_t)));
  if (!grp.empty())
    CTC_HIP(hipMemcpyAsync(d.mix_grp, grp.data(), grp.size() * sizeof(uint2), hipMemcpyHostToDevice, m->ctx->stream));
  CTC_HIP(hipMemcpyAsync(d.mix_heavy_gb, hgb.data(), hgb.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                         m->ctx->stream));
  hipStream_t st = m->ctx->stream;
  CTC_HIP(hipMalloc(&d.mix_items, std::max<size_t>(1, items.size()) * sizeof(uint4)));
  CTC_HIP(hipMalloc(&d.mix_heavy_slot, std::max<size_t>(1, hslot.size()) * sizeof(uint32_t)));
  CTC_HIP(hipMalloc(&d.mix_heavy_pb, hpb.size() * sizeof(uint32_t)));
  if (!items.empty())
    CTC_HIP(hipMemcpyAsync(d.mix_items, items.data(), items.size() * sizeof(uint4), hipMemcpyHostToDevice, st));
  if (!hslot.empty())
    CTC_HIP(hipMemcpyAsync(d.mix_heavy_slot, hslot.data(), hslot.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                           st));
  CTC_HIP(hipMemcpyAsync(d.mix_heavy_pb, hpb.data(), hpb.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st));
  CTC_HIP(hipStreamSynchronize(st));
  d.mix_chunk = chunk;
  d.mix_nitems = items.size() - thins.size() - duals.size();
  d.mix_nthin = thins.size();
  d.mix_nthin16 = nthin16;
  d.mix_ndual = duals.size();
  d.mix_ndual48 = ndual48;
  d.mix_nheavy = hslot.size();
  d.mix_nparts = nparts;
  d.mix_ngroups = grp.size();
  return CTC_OK;
}

ctc_status als_rows_prepare(ctc_model *m, const ctc_options *o, bool bf16) {
  const uint32_t R = m->rank;
  size_t need = 1;
  for (uint32_t md = 0; md < m->nmodes; ++md) {
    CTC_TRY(ensure_ciss(m, md));
    CTC_TRY(ensure_mix_items(m, md, mix_chunk_of(o, R, bf16)));
    need = std::max(need, std::max<size_t>(1, m->ciss[md].mix_nparts + m->ciss[md].mix_ngroups) *
                              mixed_partial_bytes((int)R, bf16 ? kPrecBF16 : kPrecF64));
    if (bf16 && !m->Fbf[md]) CTC_HIP(hipMalloc(&m->Fbf[md], m->dims[md] * R * sizeof(uint16_t)));
  }
  return ensure_scratch(m, need);
}

ctc_status als_mode_rows(ctc_model *m, uint32_t mode, const ctc_options *o, bool bf16) {
  const uint32_t R = m->rank, N = m->nmodes;
  if (R != 32 && R != 64 && R != 128)
    return fail(CTC_ERR_UNSUPPORTED, "%s ALS supports rank 32, 64 or 128 (got %u)",
                bf16 ? "mixed-precision" : "row-fused FP32", R);
  const uint32_t chunk = mix_chunk_of(o, R, bf16);
  const int prec = bf16 ? kPrecBF16 : kPrecF64;
  CTC_TRY(ensure_ciss(m, mode));
  CTC_TRY(ensure_mix_items(m, mode, chunk));
  DevCiss &d = m->ciss[mode];
  const size_t pbytes = mixed_partial_bytes((int)R, prec);
  CTC_TRY(ensure_scratch(m, std::max<size_t>(1, d.mix_nparts + d.mix_ngroups) * pbytes));
  hipStream_t st = m->ctx->stream;
  MixedArgs a{};
  a.prec = prec;
  for (uint32_t s = 0; s + 1 < N; ++s) {
    const uint32_t md = (mode + 1 + s) % N;
    a.oidx[s] = d.oidx[s];
    a.fr[s] = m->F[md];
    if (bf16) {
      if (!m->Fbf[md]) CTC_HIP(hipMalloc(&m->Fbf[md], m->dims[md] * R * sizeof(uint16_t)));
      CTC_HIP(launch_to_bf16(st, m->F[md], m->Fbf[md], m->dims[md] * R));
      a.fb[s] = m->Fbf[md];
    }
  }
  a.vals = d.vals;
  a.directory = d.directory;
  a.out = m->F[mode];
  a.items = d.mix_items;
  a.nitems = d.mix_nitems;
  a.thin = d.mix_items + d.mix_nitems;
  a.nthin = d.mix_nthin;
  a.nthin16 = d.mix_nthin16;
  a.dual = a.thin + a.nthin;
  a.ndual = d.mix_ndual;
  a.ndual48 = d.mix_ndual48;
  a.part = m->scratch;
  a.heavy_slot = d.mix_heavy_slot;
  a.heavy_pb = d.mix_heavy_pb;
  a.nheavy = d.mix_nheavy;
  a.gpart = (char *)m->scratch + (size_t)d.mix_nparts * pbytes;
  a.grp = d.mix_grp;
  a.ngroups = d.mix_ngroups;
  a.heavy_gb = d.mix_heavy_gb;
  a.lam = o->lambda;
  // the BF16 path always refines once against its FP32 accumulators; the FP32
  // path's Gram is accumulated in FP64, so it refines only when asked (as the
  // span kernels do)
  a.refine = bf16 ? (o->refine > 0 ? o->refine : 1) : (o->refine > 0 ? o->refine : 0);
  a.flag = m->dflag;
  CTC_HIP(launch_fill(st, m->F[mode], m->dims[mode] * R, 0.f));
  // chunks of long rows and whole rows are the same kernel, launched apart so
  // that their times are separate
  const uint64_t nchunk = d.mix_nparts, nall = a.nitems;
  if (nchunk) {
    Prof p(m->ctx, bf16 ? "als_mixed_chunks" : "als_rows_chunks");
    a.nitems = nchunk;
    CTC_HIP(launch_mixed_chunks(st, a, (int)N, (int)R));
  }
  if (nall > nchunk) {
    Prof p(m->ctx, bf16 ? "als_mixed_items" : "als_rows_items");
    a.items = d.mix_items + nchunk;
    a.nitems = nall - nchunk;
    CTC_HIP(launch_mixed_items(st, a, (int)N, (int)R));
  }
  if (a.ngroups) {
    Prof p(m->ctx, bf16 ? "als_mixed_groups" : "als_rows_groups");
    CTC_HIP(launch_mixed_groups(st, a, (int)R));
  }
  if (a.nheavy) {
    Prof p(m->ctx, bf16 ? "als_mixed_heavy" : "als_rows_heavy");
    CTC_HIP(launch_mixed_heavy(st, a, (int)N, (int)R));
  }
  if (a.nthin) {
    Prof p(m->ctx, bf16 ? "als_mixed_thin" : "als_rows_thin");
    CTC_HIP(launch_rows_thin(st, a, (int)N, (int)R));
  }
  if (a.ndual) {
    Prof p(m->ctx, bf16 ? "als_mixed_dual" : "als_rows_dual");
    CTC_HIP(launch_rows_dual(st, a, (int)N, (int)R));
  }
  m->resid_valid = false;
  return CTC_OK;
}

}  // namespace ctc
