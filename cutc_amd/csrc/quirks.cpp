// quirks.cpp — the reference's tile-reduction defects, reproduced for
// diagnosis (ctc_options.ref_quirks).
//
// The reference reduces each tile of 32 consecutive nonzeros of a mode's order
// with one warp, lane per nonzero (SG/als.cu:226-312).  Two defects of that
// reduction make a nonzero's term reach its row's Gram and MTTKRP sums a fixed
// number of times, which depends only on where rows start and end in the
// tile:
//   Q1: lane 0's mybit is forced to 1 and only lanes with mybit = 0 write
//       (SG/als.cu:245-249,285-288,307-310): the tile's first row segment is
//       never written (weight 0);
//   Q2: __shfl_down_sync from lane 31 returns lane 31's own value
//       (SG/als.cu:279-284): in a full tile whose lane 31 continues a row
//       segment that starts at lane h, the head receives lane 31's term
//       itercounter - (31 - h) extra times, itercounter being the number of
//       continuation lanes 1..31 (SG/als.cu:233).
// The last lane of a partial tile reads an inactive lane (undefined in the
// reference) and gets no extra.  The CCD kernel reduces its sums the same way
// (SG/ccd.cu:120-195), and its launch covers only 8 * (T / 256 + 1) of the
// T = nnz / 32 + 1 tiles (Q3, SG/ccd.cu:520): the others have weight 0.  An
// ALS or CCD update with these defects is then the plain update on a multiset
// of the nonzeros, each repeated by its weight; Q5 stores the values through
// an integer (SG/ciss.c:123,157,176).
#include <cmath>
#include <vector>

#include "ctc_internal.h"

namespace ctc {

void mode_order_perm(const ctc_coo *t, uint32_t mode, std::vector<uint32_t> &order);

static constexpr uint32_t kRefTile = 32;    // DEFAULT_T_TILE_LENGTH = ALS_WARPSIZE (SG/base.h:47,62)
static constexpr uint32_t kRefBlock = 256;  // DEFAULT_BLOCKSIZE (SG/base.h:61)

// w[p] for nonzero p of t (its position in t), given the mode's order
static void quirk_weights(const ctc_coo *t, uint32_t mode, uint32_t flags, const std::vector<uint32_t> &order,
                          uint32_t *w) {
  const uint64_t n = t->nnz;
  const uint32_t *row = t->ind[mode];
  for (uint64_t b = 0; b < n; b += kRefTile) {
    const uint32_t L = (uint32_t)std::min<uint64_t>(kRefTile, n - b);
    bool cont[kRefTile];
    uint32_t iters = 0;
    for (uint32_t l = 0; l < L; ++l) {
      const uint64_t p = b + l;
      cont[l] = p > 0 && row[order[p]] == row[order[p - 1]];
      if (l > 0) iters += cont[l] ? 1u : 0u;
      w[order[p]] = 1;
    }
    if (flags & CTC_QUIRK_TILE_HEAD)
      for (uint32_t l = 0; l < L && (l == 0 || cont[l]); ++l) w[order[b + l]] = 0;
    if ((flags & CTC_QUIRK_LAST_LANE) && L == kRefTile && cont[L - 1]) {
      uint32_t h = L - 1;
      while (h > 0 && cont[h]) --h;
      if (!((flags & CTC_QUIRK_TILE_HEAD) && h == 0)) w[order[b + L - 1]] += iters - (L - 1 - h);
    }
  }
  if (flags & CTC_QUIRK_CCD_GRID) {
    // blocknum_m = tilenum / DEFAULT_BLOCKSIZE + 1 blocks of 256 / 32 warps, one tile each
    const uint64_t tiles = n / kRefTile + 1;
    const uint64_t covered = (tiles / kRefBlock + 1) * (kRefBlock / kRefTile);
    for (uint64_t p = covered * kRefTile; p < n; ++p) w[order[p]] = 0;
  }
}

ctc_status quirk_multiset(const ctc_coo *t, uint32_t mode, uint32_t flags, ctc_coo **out) {
  std::vector<uint32_t> order, w(t->nnz);
  mode_order_perm(t, mode, order);
  quirk_weights(t, mode, flags, order, w.data());
  uint64_t n = 0;
  for (uint64_t p = 0; p < t->nnz; ++p) n += w[p];
  *out = nullptr;
  if (n == 0) return CTC_OK;  // every term dropped: the reference's rows of the mode all become 0
  CTC_TRY(ctc_coo_alloc(t->nmodes, n, t->dims, out));
  uint64_t q = 0;
  for (uint64_t i = 0; i < t->nnz; ++i) {  // in mode order: repeated copies stay together
    const uint32_t p = order[i];
    for (uint32_t k = 0; k < w[p]; ++k, ++q) {
      for (uint32_t m = 0; m < t->nmodes; ++m) (*out)->ind[m][q] = t->ind[m][p];
      (*out)->vals[q] = (flags & CTC_QUIRK_TRUNC_VALUES) ? std::trunc(t->vals[p]) : t->vals[p];
    }
  }
  return CTC_OK;
}

}  // namespace ctc

using namespace ctc;

extern "C" ctc_status ctc_ref_quirk_weights(const ctc_coo *t, uint32_t mode, uint32_t flags, uint32_t *weights) {
  if (!t || !weights || mode >= t->nmodes) return fail(CTC_ERR_INVALID_ARG, "ctc_ref_quirk_weights");
  std::vector<uint32_t> order;
  mode_order_perm(t, mode, order);
  quirk_weights(t, mode, flags, order, weights);
  return CTC_OK;
}
