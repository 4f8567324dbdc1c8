// host_ciss.cpp — CISS (TB-COO) construction on the host.
//
// Logical content follows the reference's ciss_alloc (SG/ciss.c:68-195):
// a directory of the distinct mode indices, per-nonzero records of the other
// modes' indices and the value, and per-tile headers that locate each
// nonzero's directory slot.  The layout is MI355X-native: structure-of-arrays
// u32 index streams and an f32 value stream (coalesced wave64 loads), tiles
// of 64 nonzeros (one wave load), and per-tile {first slot u32, row-start mask
// u64} headers in place of the reference's in-band [-slot, -slot, bitmap]
// double records (decoded with s_ff1/popcount instead of brev + shifts).
#include <algorithm>
#include <cstring>

#include "ctc_internal.h"
#include "host_parallel.h"

namespace ctc {

// Stable LSD counting sort: least significant mode first, the CISS mode last,
// giving lexicographic order over (mode, mode+1, ..., mode+N-1).  The
// reference leaves sorting to its caller (SG/ciss.c:74, quirk Q12).
void mode_order_perm(const ctc_coo *t, uint32_t mode, std::vector<uint32_t> &order) {
  const uint32_t N = t->nmodes;
  const uint64_t nnz = t->nnz;
  std::vector<uint32_t> tmp;
  order.resize(nnz);
  unsigned T = host_threads();
  parallel_for(nnz, T, [&](uint64_t lo, uint64_t hi, unsigned) {
    for (uint64_t p = lo; p < hi; ++p) order[p] = (uint32_t)p;
  });
  for (int s = (int)N - 1; s >= 0; --s) {
    uint32_t m = (mode + (uint32_t)s) % N;
    counting_sort(t->ind[m], t->dims[m], order, tmp);
  }
}

ctc_status build_host_ciss(const ctc_coo *t, uint32_t mode, HostCiss &c) {
  if (!t || mode >= t->nmodes) return fail(CTC_ERR_INVALID_ARG, "ciss: bad mode %u", mode);
  if (t->nnz == 0) return fail(CTC_ERR_INVALID_ARG, "ciss: empty tensor");
  const uint32_t N = t->nmodes;
  const uint64_t nnz = t->nnz;
  c.nmodes = N;
  c.mode = mode;
  c.nnz = nnz;
  for (uint32_t m = 0; m < N; ++m) {
    c.dims[m] = t->dims[m];
    if (t->dims[m] == 0 || t->dims[m] > 0xFFFFFFFFull)
      return fail(CTC_ERR_INVALID_ARG, "ciss: dims[%u] out of range", m);
  }
  std::vector<uint32_t> order;
  mode_order_perm(t, mode, order);
  c.perm.swap(order);
  unsigned T = host_threads();

  // streams
  for (uint32_t s = 0; s + 1 < N; ++s) c.oidx[s].resize(nnz);
  c.vals.resize(nnz);
  std::vector<uint32_t> rows(nnz);
  parallel_for(nnz, T, [&](uint64_t lo, uint64_t hi, unsigned) {
    for (uint64_t p = lo; p < hi; ++p) {
      const uint32_t q = c.perm[p];
      rows[p] = t->ind[mode][q];
      for (uint32_t s = 0; s + 1 < N; ++s) c.oidx[s][p] = t->ind[(mode + 1 + s) % N][q];
      c.vals[p] = (float)t->vals[q];
    }
  });

  // directory and row pointers (SG/ciss.c:89-115; rowptr replaces dcounter)
  c.directory.clear();
  c.rowptr.clear();
  c.directory.reserve(std::min<uint64_t>(nnz, t->dims[mode]));
  for (uint64_t p = 0; p < nnz; ++p) {
    if (p == 0 || rows[p] != rows[p - 1]) {
      c.directory.push_back(rows[p]);
      c.rowptr.push_back((uint32_t)p);
    }
  }
  c.rowptr.push_back((uint32_t)nnz);
  c.dlength = c.directory.size();

  // tile headers: row-start mask per 64 nonzeros and the first slot
  const uint64_t ntiles = (nnz + CTC_TILE_LEN - 1) / CTC_TILE_LEN;
  c.ntiles = ntiles;
  c.tile_mask.assign(ntiles, 0);
  c.tile_slot.assign(ntiles, 0);
  std::vector<uint32_t> pc(ntiles);
  parallel_for(ntiles, T, [&](uint64_t lo, uint64_t hi, unsigned) {
    for (uint64_t tl = lo; tl < hi; ++tl) {
      uint64_t mask = 0;
      uint64_t b = tl * CTC_TILE_LEN, e = std::min<uint64_t>(nnz, b + CTC_TILE_LEN);
      for (uint64_t p = b; p < e; ++p)
        if (p == 0 || rows[p] != rows[p - 1]) mask |= 1ull << (p - b);
      c.tile_mask[tl] = mask;
      pc[tl] = (uint32_t)__builtin_popcountll(mask);
    }
  });
  uint64_t cum = 0;  // heads before the tile
  for (uint64_t tl = 0; tl < ntiles; ++tl) {
    // slot of the tile's first nonzero = heads in [0, 64 tl] - 1
    c.tile_slot[tl] = (uint32_t)(cum + (c.tile_mask[tl] & 1ull) - 1);
    cum += pc[tl];
  }
  return CTC_OK;
}

}  // namespace ctc

using namespace ctc;

extern "C" ctc_status ctc_ciss_build(const ctc_coo *t, uint32_t mode, ctc_ciss **out) {
  if (!out) return fail(CTC_ERR_INVALID_ARG, "ctc_ciss_build: null out");
  ctc_ciss *c = new (std::nothrow) ctc_ciss();
  if (!c) return fail(CTC_ERR_NOMEM, "ctc_ciss_build");
  ctc_status s = build_host_ciss(t, mode, c->h);
  if (s != CTC_OK) {
    delete c;
    return s;
  }
  *out = c;
  return CTC_OK;
}

extern "C" ctc_status ctc_ciss_get_view(const ctc_ciss *c, ctc_ciss_view *v) {
  if (!c || !v) return fail(CTC_ERR_INVALID_ARG, "ctc_ciss_get_view: null argument");
  memset(v, 0, sizeof(*v));
  const HostCiss &h = c->h;
  v->nmodes = h.nmodes;
  v->mode = h.mode;
  v->nnz = h.nnz;
  v->ntiles = h.ntiles;
  v->dlength = h.dlength;
  for (uint32_t m = 0; m < h.nmodes; ++m) v->dims[m] = h.dims[m];
  v->directory = h.directory.data();
  v->rowptr = h.rowptr.data();
  for (uint32_t s = 0; s + 1 < h.nmodes; ++s) v->oidx[s] = h.oidx[s].data();
  v->vals = h.vals.data();
  v->tile_slot = h.tile_slot.data();
  v->tile_mask = h.tile_mask.data();
  v->perm = h.perm.data();
  return CTC_OK;
}

extern "C" void ctc_ciss_free(ctc_ciss *c) { delete c; }

// Reference array-of-doubles layout, for fixtures and interchange.  Field by
// field as the reference fills it (SG/ciss.c:121-188): header record
// [-(slot in progress), -(last slot, only on a row change), bitmap with a
// leading sentinel bit], then (next idx, next-next idx, value) records with
// 1-based indices, and a (-1,-1,-1) record after the last nonzero.
extern "C" ctc_status ctc_ciss_reference_layout(const ctc_ciss *c, const ctc_coo *src,
                                                uint32_t L, int truncate_vals,
                                                uint64_t *directory, uint64_t *dcounter,
                                                double *entries, uint64_t *dlength,
                                                uint64_t *size) {
  if (!c || !src || !directory || !dcounter || !entries || !dlength || !size)
    return fail(CTC_ERR_INVALID_ARG, "ctc_ciss_reference_layout: null argument");
  const HostCiss &h = c->h;
  if (h.nmodes != 3) return fail(CTC_ERR_UNSUPPORTED, "reference layout is 3-way only");
  if (L < 1 || L > 52)
    return fail(CTC_ERR_INVALID_ARG, "reference tile length must be 1..52 (exact bitmap)");
  if (src->nnz != h.nnz) return fail(CTC_ERR_INVALID_ARG, "src is not the CISS source");
  const uint64_t nnz = h.nnz;
  for (uint64_t s = 0; s < h.dlength; ++s) directory[s] = (uint64_t)h.directory[s] + 1;
  dcounter[0] = 0;
  for (uint64_t s = 1; s < h.dlength; ++s) dcounter[s] = h.rowptr[s] - 1;  // last nnz of slot s-1
  dcounter[h.dlength] = nnz - 1;
  *dlength = h.dlength;
  const uint64_t ntile = nnz / L + 1;
  *size = ntile + nnz + 1;
  uint64_t slot = 0;  // slot of the nonzero processed last
  uint64_t p = 0;
  for (uint64_t tl = 0; tl < ntile; ++tl) {
    double *rec = entries + tl * (uint64_t)(L + 1) * 3;
    rec[0] = -1.0 * (double)slot;
    uint64_t bitmap = 1;
    for (uint32_t e = 0; e < L; ++e, ++p) {
      double *r = rec + (uint64_t)(e + 1) * 3;
      if (p == nnz) {
        r[0] = r[1] = r[2] = -1.0;
        ++p;
        break;
      }
      const bool starts = p > 0 && h.rowptr[slot + 1] == p;  // row changes here
      if (starts) {
        ++slot;
        rec[1] = -1.0 * (double)slot;
        bitmap <<= 1;
      } else {
        bitmap = (bitmap << 1) | 1u;
      }
      r[0] = (double)h.oidx[0][p] + 1.0;
      r[1] = (double)h.oidx[1][p] + 1.0;
      double v = src->vals[h.perm[p]];
      if (truncate_vals) v = (double)(uint64_t)v;
      r[2] = v;
      rec[2] = (double)bitmap;
    }
  }
  return CTC_OK;
}
