// kernels.hip — gfx950 kernels of libctc: CISS tile walks for SGD, ALS
// (MFMA Gram + MTTKRP, FP64 Cholesky solve), MTTKRP, the loss adjoint and the
// RMSE reductions.
//
// Execution model shared by the walks ("walkers"):
//   * The nonzeros of a mode's CISS are cut into spans of whole 64-nonzero
//     tiles.  A walker owns one span and visits its nonzeros in order.
//   * A walker is W = min(R, 64) lanes; lane w holds columns w, w+W, ... of
//     every factor row it touches, so each row access is one contiguous
//     W*4-byte load or atomic (coalesced), and a dot product is a W-lane DPP /
//     swizzle reduction.  64/W walkers share a wave.
//   * A wave loads W nonzeros at a time (indices, value, row-start bit from the
//     tile mask) with one coalesced load per stream, then hands nonzero q to
//     its walker lanes with ds_bpermute; the row starts come from one ballot.
//   * Row (directory-slot) changes are found from the tile mask; the slot of a
//     walker's first nonzero is the tile header's first slot.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels_common.h"
#include "kernels_launch.h"

namespace ctc {

// ==========================================================================
// SGD epoch (reference rule SG/sgd.cu:75-99: every row of a nonzero is
// updated from its pre-update values; step = lr * scale).
//
// A walker keeps the rows of its current nonzero in registers.  How a row it
// shares with concurrent walkers is read and written is chosen per mode
// (PA for the CISS mode, PB for the next mode, PC for the others):
//   P_RACY     read at the start of a run of equal indices, apply own updates
//              in registers, store the row after every nonzero (concurrent
//              writers overwrite each other).
//   P_ATOMIC   read every nonzero, add every update with a device-scope float
//              atomic (the reference's atomicAdd, SG/sgd.cu:96-98).
//   P_RUN      read at the start of a run, apply own updates in registers, add
//              the run's accumulated delta with atomics when the run ends.
//   P_HOT      per row: P_RACY if the row's bit is set in the mode's hot-row
//              bitmap (rows with more updates per sweep than a threshold,
//              whose atomics would serialise on one L2 channel), else P_ATOMIC.
//   P_HOTRUN   per row: P_RACY for a hot row, else P_RUN with runs cut
//              after kRunCap nonzeros (the row is then added and re-read, so
//              a walker's copy is never more than kRunCap own updates away
//              from the shared row: uncut runs of thousands of nonzeros on a
//              row many walkers share add up deltas taken from one stale copy
//              and diverge).  In mode-0 order the nonzeros of one mode-0 row
//              come sorted by the next mode, so runs of one next-mode row are
//              common and each run costs one atomic row update instead of one
//              per nonzero.
// The float atomics are the kernel's limit (the L2 completes about one
// 4-byte atomic per channel per clock), so fewer atomic rows is what pays.
// Walker w processes span (w * stride) mod nspans (stride 1 = natural order).
// ==========================================================================
enum { P_RACY = 0, P_ATOMIC = 1, P_RUN = 2, P_HOT = 3, P_HOTRUN = 4 };
constexpr int kRunCap = 8;

template <int R, int N, int PA, int PB, int PC>
__global__ void __launch_bounds__(256) k_sgd(CissArgs c, FactorArgs fa, HotRows hot, uint32_t span,
                                             uint64_t nspans, uint64_t stride, float step, float lam) {
  constexpr int W = R < 64 ? R : 64;
  constexpr int CPL = R / W;
  constexpr int WPW = 64 / W;
  constexpr uint32_t NONE = 0xFFFFFFFFu;
  const int lane = lane_id();
  const int sub = lane / W;
  const int wl = lane % W;
  const uint64_t wave = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const uint64_t walker = wave * WPW + sub;
  if (wave * WPW >= nspans) return;  // whole wave idle (uniform)
  const bool live = walker < nspans;
  const uint64_t sp = live ? (walker * stride) % nspans : 0;
  const uint64_t s0 = sp * span;
  const uint64_t s1 = !live ? 0 : (s0 + span < c.nnz ? s0 + span : c.nnz);

  // row slots: 0 = CISS mode (A), 1.. = modes mode+1, ... (B, C, D)
  float *Fm[N];
#pragma unroll
  for (int s = 0; s < N; ++s) Fm[s] = fa.f[s];
  float v[N][CPL], del[N][CPL];
  uint32_t cur[N];
#pragma unroll
  for (int s = 0; s < N; ++s) {
    cur[s] = NONE;
#pragma unroll
    for (int q = 0; q < CPL; ++q) v[s][q] = 0.f, del[s][q] = 0.f;
  }
  uint32_t slot = 0;

  auto pol = [](int s) { return s == 0 ? PA : (s == 1 ? PB : PC); };
  auto bitmapped = [&](int s) { return pol(s) == P_HOT || pol(s) == P_HOTRUN; };
  auto flush_run = [&](int s) {
#pragma unroll
    for (int q2 = 0; q2 < CPL; ++q2) atomic_add_f32(Fm[s] + (size_t)cur[s] * R + wl + q2 * W, del[s][q2]);
  };
  // the cached row of slot s is a P_RUN row whose delta is not yet added, and
  // the nonzeros its run has taken so far
  bool pend[N];
  int runlen[N];
#pragma unroll
  for (int s = 0; s < N; ++s) pend[s] = false, runlen[s] = 0;

  for (uint32_t ch = 0; ch < span; ch += W) {
    const uint64_t el = s0 + ch + wl;
    const bool vl = el < s1;
    uint32_t il[N - 1];
    float xl = 0.f;
    bool hl = false;
    if (vl) {
#pragma unroll
      for (int s = 0; s < N - 1; ++s) il[s] = c.oidx[s][el];
      xl = c.vals[el];
      hl = (c.tile_mask[el >> 6] >> (el & 63)) & 1ull;
    } else {
#pragma unroll
      for (int s = 0; s < N - 1; ++s) il[s] = 0;
    }
    const uint64_t hm = __ballot(vl && hl);
    // hot-row bits of this lane's nonzero, one bit per slot, looked up once
    // per chunk rather than on every nonzero's dependency chain
    int hotl = 0;
#pragma unroll
    for (int s = 1; s < N; ++s)
      if (bitmapped(s) && vl) hotl |= (int)((hot.bits[s][il[s - 1] >> 5] >> (il[s - 1] & 31)) & 1u) << s;
    for (int q = 0; q < W; ++q) {
      const int src = sub * W + q;
      uint32_t idx[N];
#pragma unroll
      for (int s = 0; s < N - 1; ++s) idx[1 + s] = (uint32_t)__shfl((int)il[s], src, 64);
      const float x = __shfl(xl, src, 64);
      const int hotq = __shfl(hotl, src, 64);
      const uint64_t e = s0 + ch + q;
      if (e < s1) {
        const bool head = (hm >> src) & 1ull;
        if (cur[0] == NONE || head) {
          slot = cur[0] == NONE ? c.tile_slot[s0 >> 6] : slot + 1;
          idx[0] = c.directory[slot];
        } else {
          idx[0] = cur[0];
        }
        int pr[N];  // per-row policy of this nonzero
#pragma unroll
        for (int s = 0; s < N; ++s) {
          int P = pol(s);
          if (P == P_HOT) P = ((hotq >> s) & 1) ? P_RACY : P_ATOMIC;
          if (P == P_HOTRUN) P = ((hotq >> s) & 1) ? P_RACY : P_RUN;
          pr[s] = P;
        }
#pragma unroll
        for (int s = 0; s < N; ++s) {
          const int P = pr[s];
          const bool capped = pol(s) == P_HOTRUN && pend[s] && runlen[s] >= kRunCap;
          if (P == P_ATOMIC || idx[s] != cur[s] || capped) {
            if (pend[s]) flush_run(s);  // the cached row's run ends here
            pend[s] = P == P_RUN;
            runlen[s] = 0;
            // an atomic row is re-read at every nonzero: forget it as the cached row
            cur[s] = P == P_ATOMIC ? NONE : idx[s];
#pragma unroll
            for (int q2 = 0; q2 < CPL; ++q2) {
              v[s][q2] = Fm[s][(size_t)idx[s] * R + wl + q2 * W];
              del[s][q2] = 0.f;
            }
          }
        }
        // prediction and error
        float part = 0.f;
#pragma unroll
        for (int q2 = 0; q2 < CPL; ++q2) {
          float pv = v[0][q2];
#pragma unroll
          for (int s = 1; s < N; ++s) pv *= v[s][q2];
          part += pv;
        }
        const float err = x - group_sum<W>(part);
#pragma unroll
        for (int q2 = 0; q2 < CPL; ++q2) {
          // products of all rows but one, from pre-update values
          float oth[N];
#pragma unroll
          for (int s = 0; s < N; ++s) {
            float o = 1.f;
#pragma unroll
            for (int s2 = 0; s2 < N; ++s2)
              if (s2 != s) o *= v[s2][q2];
            oth[s] = o;
          }
          const int col = wl + q2 * W;
#pragma unroll
          for (int s = 0; s < N; ++s) {
            const int P = pr[s];
            const float d = step * (err * oth[s] - lam * v[s][q2]);
            float *dst = Fm[s] + (size_t)idx[s] * R + col;
            if (P == P_ATOMIC) {
              atomic_add_f32(dst, d);
            } else {
              v[s][q2] += d;
              del[s][q2] += d;
              if (P == P_RACY) *dst = v[s][q2];
            }
          }
        }
#pragma unroll
        for (int s = 0; s < N; ++s) runlen[s] += 1;
      }
    }
  }
#pragma unroll
  for (int s = 0; s < N; ++s)
    if (pend[s]) flush_run(s);
}

// ==========================================================================
// Bulk-synchronous SGD (the default schedule).  One launch of k_sgd_sync is
// one sub-step: walker w runs nonzeros [w*span + s*sub, +sub) of its span
// from the factors as they stood when the sub-step began; k_sgd_settle and
// k_sgd_fold then add every row's handed-in deltas to the factors.
//   * Along a run of consecutive nonzeros with the same index in a mode the
//     walker keeps its copy of that row in registers (read where the run
//     starts) and applies the reference step to it (SG/sgd.cu:86-99).
//   * Modes 1..: when a run ends its delta is written to the run's staging
//     slot (a static table numbers every chunk's runs in walk order), and
//     k_sgd_settle adds, per row, the sum of its slots (a cold row: the
//     reference's atomicAdd, SG/sgd.cu:96-98) or their mean (a hot row:
//     replica averaging, NC/sgd.cu:229-262 with the sum of :523-525), summed
//     in a fixed order.  No atomics: the sweep is bitwise reproducible.
//   * Mode 0: a row of at most long_min - 1 nonzeros cannot meet a second
//     walker within one sub-step (two walkers' nonzeros of one sub-step are at
//     least span - sub + 1 positions apart), so its walker stores its copy in
//     place.  A longer row's run deltas are added as 64-bit fixed-point
//     integers (2^-32 units) into its delta sum D (integer addition is
//     associative, so the sum does not depend on the order the walkers finish
//     in), a hot row's into kSyncReplicas replicas with a count of runs, and
//     k_sgd_fold adds the sum (cold) or the mean (hot) and clears them.
// ==========================================================================
constexpr double kFixScale = 4294967296.0;        // 2^32
constexpr double kFixInv = 1.0 / 4294967296.0;    // 2^-32

// A walker is W = R / 4 lanes of a wave; lane l holds the row elements
// 4l .. 4l + 3, so a row moves with one 16-byte access per lane and a wave
// carries 256 / R walkers (eight at R = 32).  Each walker is a chain of
// dependent steps whose row reads miss the L2 (the launch starts with it
// invalidated), so the more walkers a wave carries, the more of those reads
// are in flight at once.
template <int R>
struct SyncShape {
  static constexpr int CPL = 4;
  static constexpr int W = R / CPL;
  static constexpr int WPW = 64 / W;
};

template <int CPL>
__device__ inline void load_row4(const float *p, float (&v)[CPL]) {
#pragma unroll
  for (int j = 0; j < CPL; j += 4) {
    const float4 t = *reinterpret_cast<const float4 *>(p + j);
    v[j] = t.x, v[j + 1] = t.y, v[j + 2] = t.z, v[j + 3] = t.w;
  }
}

template <int CPL>
__device__ inline void store_row4(float *p, const float (&v)[CPL]) {
#pragma unroll
  for (int j = 0; j < CPL; j += 4) *reinterpret_cast<float4 *>(p + j) = make_float4(v[j], v[j + 1], v[j + 2], v[j + 3]);
}

// Staged deltas are written once and read once by the next launch: streamed
// past the caches.
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <int CPL>
__device__ inline void store_row4_nt(float *p, const float (&v)[CPL]) {
#pragma unroll
  for (int j = 0; j < CPL; j += 4) {
    f32x4 t = {v[j], v[j + 1], v[j + 2], v[j + 3]};
    __builtin_nontemporal_store(t, reinterpret_cast<f32x4 *>(p + j));
  }
}

template <int CPL>
__device__ inline void load_row4_nt(const float *p, float (&v)[CPL]) {
#pragma unroll
  for (int j = 0; j < CPL; j += 4) {
    const f32x4 t = __builtin_nontemporal_load(reinterpret_cast<const f32x4 *>(p + j));
    v[j] = t.x, v[j + 1] = t.y, v[j + 2] = t.z, v[j + 3] = t.w;
  }
}

// Four BF16 entries per lane (8 bytes): the mixed path's staged deltas.
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

template <int CPL>
__device__ inline void load_bf4_nt(const uint16_t *p, float (&v)[CPL]) {
#pragma unroll
  for (int j = 0; j < CPL; j += 4) {
    const u32x2 t = __builtin_nontemporal_load(reinterpret_cast<const u32x2 *>(p + j));
    v[j] = bf_lo(t.x), v[j + 1] = bf_hi(t.x), v[j + 2] = bf_lo(t.y), v[j + 3] = bf_hi(t.y);
  }
}

template <int CPL>
__device__ inline void store_bf4_nt(uint16_t *p, const float (&v)[CPL]) {
#pragma unroll
  for (int j = 0; j < CPL; j += 4) {
    const u32x2 t = {(uint32_t)f2bf(v[j]) | ((uint32_t)f2bf(v[j + 1]) << 16),
                     (uint32_t)f2bf(v[j + 2]) | ((uint32_t)f2bf(v[j + 3]) << 16)};
    __builtin_nontemporal_store(t, reinterpret_cast<u32x2 *>(p + j));
  }
}

// MIX (the mixed-precision path): the run deltas of modes 1.. are staged in
// BF16 (round to nearest even), halving the staging traffic; the factors, the
// walker's arithmetic and the sums of the deltas stay FP32 / FP64.
template <int R, int N, bool MIX>
__global__ void __launch_bounds__(256) k_sgd_sync(CissArgs c, FactorArgs fa, SyncArgs sy, uint32_t span,
                                                  uint32_t sub, uint32_t s, uint64_t nspans, uint32_t long_min,
                                                  float step, float lam) {
  constexpr int W = SyncShape<R>::W;
  constexpr int CPL = SyncShape<R>::CPL;
  constexpr int WPW = SyncShape<R>::WPW;
  constexpr uint32_t NONE = 0xFFFFFFFFu;
  const int lane = lane_id();
  const int sub_w = lane / W;
  const int wl = lane % W;
  const uint64_t wave = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const uint64_t walker = wave * WPW + sub_w;
  if (wave * WPW >= nspans) return;  // whole wave idle (uniform)
  uint64_t e0 = 0, e1 = 0;
  if (walker < nspans) {
    const uint64_t send = (walker + 1) * span < c.nnz ? (walker + 1) * span : c.nnz;
    e0 = walker * span + (uint64_t)s * sub;
    e1 = e0 + sub < send ? e0 + sub : send;
    if (e0 > e1) e0 = e1;
  }
  float *Fm[N];
#pragma unroll
  for (int q = 0; q < N; ++q) Fm[q] = fa.f[q];
  float v[N][CPL], del[N][CPL];
  uint32_t cur[N];
#pragma unroll
  for (int q = 0; q < N; ++q) {
    cur[q] = NONE;
#pragma unroll
    for (int k = 0; k < CPL; ++k) v[q][k] = 0.f, del[q][k] = 0.f;
  }
  uint32_t slot = 0;
  bool own0 = false;  // the current mode-0 row is stored in place (no other walker meets it)
  // modes 1..: the next slot of this chunk's runs and, on the mixed path, its
  // staging position, read one run ahead (the table has one entry of padding)
  uint32_t rslot[N], rpos[N];
#pragma unroll
  for (int k = 0; k < N; ++k) {
    rslot[k] = (k >= 1 && walker < nspans) ? sy.sbase[k][(size_t)s * nspans + walker] : 0u;
    rpos[k] = (MIX && k >= 1) ? sy.dest[k][rslot[k]] : 0u;
  }
  auto hand_in = [&](int q) {
    if (q >= 1) {  // the run's delta goes to its staging slot / position
      if constexpr (MIX) {
        const uint32_t pos = rpos[q];
        rpos[q] = sy.dest[q][++rslot[q]];
        store_bf4_nt<CPL>(reinterpret_cast<uint16_t *>(sy.stage[q]) + (size_t)pos * R + wl * CPL, del[q]);
      } else {
        store_row4_nt<CPL>(sy.stage[q] + (size_t)rslot[q]++ * R + wl * CPL, del[q]);
      }
      return;
    }
    if (own0) {
      store_row4<CPL>(Fm[0] + (size_t)cur[0] * R + wl * CPL, v[0]);
      return;
    }
    // a long mode-0 row: fixed-point sums
    const uint32_t h = sy.hidx[0][cur[0]];
    unsigned long long *d;
    if (h == 0xFFFFFFFFu) {
      d = (unsigned long long *)sy.D[0] + (size_t)cur[0] * R;
    } else {
      // a hot row's mean needs its count of runs; a cold row adds its delta
      // sum, so the merge needs no count (an untouched row's sum is 0)
      const size_t r = (size_t)h * kSyncReplicas + walker % kSyncReplicas;
      d = (unsigned long long *)sy.Dh[0] + r * R;
      if (wl == 0) atomicAdd(sy.ch[0] + r, 1u);
    }
#pragma unroll
    for (int k = 0; k < CPL; ++k)
      atomicAdd(d + wl * CPL + k, (unsigned long long)__double2ll_rn((double)del[q][k] * kFixScale));
  };

  for (uint32_t ch = 0; ch < sub; ch += W) {
    const uint64_t el = e0 + ch + wl;
    const bool vl = el < e1;
    uint32_t il[N - 1];
    float xl = 0.f;
    bool hl = false;
    if (vl) {
#pragma unroll
      for (int q = 0; q < N - 1; ++q) il[q] = c.oidx[q][el];
      xl = c.vals[el];
      hl = (c.tile_mask[el >> 6] >> (el & 63)) & 1ull;
    } else {
#pragma unroll
      for (int q = 0; q < N - 1; ++q) il[q] = 0;
    }
    const uint64_t hm = __ballot(vl && hl);
#pragma unroll 1  // rolled: 76 VGPRs, six waves per SIMD
    for (int q = 0; q < W; ++q) {
      const int src = sub_w * W + q;
      uint32_t idx[N];
#pragma unroll
      for (int k = 0; k < N - 1; ++k) idx[1 + k] = (uint32_t)__shfl((int)il[k], src, 64);
      const float x = __shfl(xl, src, 64);
      const uint64_t e = e0 + ch + q;
      if (e < e1) {
        const bool head = (hm >> src) & 1ull;
        bool new0 = false;
        if (cur[0] == NONE) {
          slot = tile_lane_slot(c.tile_slot[e >> 6], c.tile_mask[e >> 6], (int)(e & 63));
          new0 = true;
        } else if (head) {
          slot = slot + 1;
          new0 = true;
        }
        if (new0) {
          idx[0] = c.directory[slot];
        } else {
          idx[0] = cur[0];
        }
#pragma unroll
        for (int k = 0; k < N; ++k) {
          if (idx[k] != cur[k]) {
            if (cur[k] != NONE) hand_in(k);
            if (k == 0) own0 = c.rowptr[slot + 1] - c.rowptr[slot] < long_min;
            cur[k] = idx[k];
            load_row4<CPL>(Fm[k] + (size_t)idx[k] * R + wl * CPL, v[k]);
#pragma unroll
            for (int j = 0; j < CPL; ++j) del[k][j] = 0.f;
          }
        }
        float part = 0.f;
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
          float pv = v[0][j];
#pragma unroll
          for (int k = 1; k < N; ++k) pv *= v[k][j];
          part += pv;
        }
        const float err = x - group_sum<W>(part);
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
          float d[N];
#pragma unroll
          for (int k = 0; k < N; ++k) {
            float o = 1.f;
#pragma unroll
            for (int k2 = 0; k2 < N; ++k2)
              if (k2 != k) o *= v[k2][j];
            d[k] = step * (err * o - lam * v[k][j]);
          }
#pragma unroll
          for (int k = 0; k < N; ++k) {
            v[k][j] += d[k];
            del[k][j] += d[k];
          }
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < N; ++k)
    if (cur[k] != NONE) hand_in(k);
}

// Mode 0's long rows (k_sgd_sync's fixed-point sums) as block `blk` of a
// 256-thread launch: a cold row adds its delta sum D, a hot row the mean of
// its replicas' sums Dh over their count of runs ch; both are cleared.
template <int R>
__device__ inline void sgd_merge_block(const MergeArgs &a, uint64_t blk) {
  constexpr int W = R < 64 ? R : 64;
  constexpr int CPL = R / W;
  constexpr int WPW = 64 / W;
  const int lane = lane_id();
  const int wl = lane % W;
  const uint64_t wave = blk * 4 + (threadIdx.x >> 6);
  const uint64_t i = wave * WPW + lane / W;
  if (i >= a.off[a.nseg]) return;
  int g = 0;
  while (g + 1 < a.nseg && i >= a.off[g + 1]) ++g;
  const int m = a.mode[g];
  const uint64_t k = i - a.off[g];
  if (a.is_hot[g]) {
    // hot row k of mode m: sum its replicas, add the mean of its runs' deltas
    uint32_t cnt = 0;
    for (int r = 0; r < kSyncReplicas; ++r) cnt += a.ch[m][k * kSyncReplicas + r];
    if (cnt == 0) return;
    const uint32_t row = a.hrow[m][k];
    float *gr = a.F[m] + (size_t)row * R;
    long long *d = a.Dh[m] + k * kSyncReplicas * R;
    const double sc = 1.0 / (double)cnt;
#pragma unroll
    for (int j = 0; j < CPL; ++j) {
      const int col = wl + j * W;
      long long sum = 0;
      for (int r = 0; r < kSyncReplicas; ++r) {
        sum += d[(size_t)r * R + col];
        d[(size_t)r * R + col] = 0;
      }
      gr[col] = (float)((double)gr[col] + (double)sum * kFixInv * sc);
    }
    if (wl < kSyncReplicas) a.ch[m][k * kSyncReplicas + wl] = 0u;
    if (W < kSyncReplicas)
      for (int r = wl + W; r < kSyncReplicas; r += W) a.ch[m][k * kSyncReplicas + r] = 0u;
    return;
  }
  // cold row: add its delta sum where it is nonzero
  const uint32_t row = a.rows[g] ? a.rows[g][k] : (uint32_t)k;
  float *gr = a.F[m] + (size_t)row * R;
  long long *d = a.D[m] + (size_t)row * R;
#pragma unroll
  for (int j = 0; j < CPL; ++j) {
    const int col = wl + j * W;
    const long long dv = d[col];
    if (dv != 0) {
      gr[col] = (float)((double)gr[col] + (double)dv * kFixInv);
      d[col] = 0;
    }
  }
}

// One launch per sub-step settles every piece of the staged modes: a
// workgroup per piece, whose walkers take the piece's positions in turn
// (FP64 partial sums) and add their partials in
// walker order.  A piece that holds a whole row applies it; the others leave
// their sums for k_sgd_fold.  The order is fixed, so the result is
// reproducible.
// A row of at most kSettleLight runs, settled by one walker (W lanes) of
// light block lblk: its runs' deltas summed in position order in FP64, the
// sum (cold row) or the mean (hot row) added to the row.
template <int R, bool MIX>
__device__ inline void sgd_settle_light(const SettleArgs &a, uint64_t lblk) {
  constexpr int CPL = SyncShape<R>::CPL;
  constexpr int W = SyncShape<R>::W;
  constexpr int NW = 256 / W;
  const uint64_t li = lblk * NW + threadIdx.x / W;
  const int wl = threadIdx.x % W;
  int k = 1;
  while (k < 3 && li >= a.lb[k + 1]) ++k;
  if (li >= a.lb[k + 1]) return;
  const uint32_t gp = a.light[k][li - a.lb[k]];
  const uint32_t b = a.piece_start[k][gp], e = a.piece_start[k][gp + 1];
  const uint32_t row = a.piece_row[k][gp];
  const float *stage = a.stage[k];
  const uint32_t s0 = (uint32_t)a.slot0[k];
  const uint32_t *slots = a.slots[k];
  auto at = [&](uint32_t q) -> size_t { return MIX ? (size_t)(q - s0) : (size_t)slots[q]; };
  float *fr = a.F[k] + (size_t)row * R + wl * CPL;
  float fv[CPL];
  load_row4<CPL>(fr, fv);
  const uint32_t hw = a.hot[k][row >> 5];
  double acc[CPL];
#pragma unroll
  for (int j = 0; j < CPL; ++j) acc[j] = 0.0;
  uint32_t p = b;
  for (; p + 3 < e; p += 4) {  // four slots in flight, added in position order
    float x[4][CPL];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if constexpr (MIX)
        load_bf4_nt<CPL>(reinterpret_cast<const uint16_t *>(stage) + at(p + u) * R + wl * CPL, x[u]);
      else
        load_row4_nt<CPL>(stage + at(p + u) * R + wl * CPL, x[u]);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int j = 0; j < CPL; ++j) acc[j] += (double)x[u][j];
  }
  for (; p < e; ++p) {
    float x[CPL];
    if constexpr (MIX)
      load_bf4_nt<CPL>(reinterpret_cast<const uint16_t *>(stage) + at(p) * R + wl * CPL, x);
    else
      load_row4_nt<CPL>(stage + at(p) * R + wl * CPL, x);
#pragma unroll
    for (int j = 0; j < CPL; ++j) acc[j] += (double)x[j];
  }
  const double sc = ((hw >> (row & 31)) & 1u) ? 1.0 / (double)(e - b) : 1.0;
  float out[CPL];
#pragma unroll
  for (int j = 0; j < CPL; ++j) out[j] = (float)((double)fv[j] + acc[j] * sc);
  store_row4<CPL>(fr, out);
}

// Blocks 0 .. pb[4] - 1: one heavy piece each (a workgroup sums its slots);
// the blocks after them: NW light rows each (sgd_settle_light).
template <int R, bool MIX>
__global__ void __launch_bounds__(256) k_sgd_settle(SettleArgs a) {
  constexpr int CPL = SyncShape<R>::CPL;  // a slot's row moves as 16-byte pieces
  constexpr int W = SyncShape<R>::W;
  constexpr int NW = 256 / W;  // walkers per workgroup
  constexpr int G = 256 / R;   // first-level groups of the reduction
  __shared__ double part[NW][R];
  __shared__ double red[G][R];
  const uint64_t pblk = blockIdx.x;
  if (pblk >= a.pb[4]) {  // uniform per block: the light path has no barrier
    sgd_settle_light<R, MIX>(a, pblk - a.pb[4]);
    return;
  }
  int k = 1;
  while (k < 3 && pblk >= a.pb[k + 1]) ++k;
  const uint64_t gp = a.heavy[k][pblk - a.pb[k]];
  const int wid = threadIdx.x / W;
  const int wl = threadIdx.x % W;
  const uint32_t b = a.piece_start[k][gp], e = a.piece_start[k][gp + 1];
  const float *stage = a.stage[k];
  const uint32_t s0 = (uint32_t)a.slot0[k];  // positions are numbered over the epoch
  const uint32_t *slots = a.slots[k];
  // the staged delta at position q: FP32 at its slot, BF16 at q itself
  auto at = [&](uint32_t q) -> size_t { return MIX ? (size_t)(q - s0) : (size_t)slots[q]; };
  // a whole row's factor entry and hot bit, read while the slots are summed
  const uint32_t row = a.piece_row[k][gp];
  float fv = 0.f;
  uint32_t hw = 0;
  if (row != 0xFFFFFFFFu && threadIdx.x < R) {
    fv = a.F[k][(size_t)row * R + threadIdx.x];
    hw = a.hot[k][row >> 5];
  }
  double acc[CPL];
#pragma unroll
  for (int j = 0; j < CPL; ++j) acc[j] = 0.0;
  uint32_t p = b + wid;
  for (; p + 3 * NW < e; p += 4 * NW) {  // four independent slots in flight
    float x[4][CPL];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if constexpr (MIX)
        load_bf4_nt<CPL>(reinterpret_cast<const uint16_t *>(stage) + at(p + u * NW) * R + wl * CPL, x[u]);
      else
        load_row4_nt<CPL>(stage + at(p + u * NW) * R + wl * CPL, x[u]);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int j = 0; j < CPL; ++j) acc[j] += (double)x[u][j];
  }
  for (; p < e; p += NW) {
    float x[CPL];
    if constexpr (MIX)
      load_bf4_nt<CPL>(reinterpret_cast<const uint16_t *>(stage) + at(p) * R + wl * CPL, x);
    else
      load_row4_nt<CPL>(stage + at(p) * R + wl * CPL, x);
#pragma unroll
    for (int j = 0; j < CPL; ++j) acc[j] += (double)x[j];
  }
#pragma unroll
  for (int j = 0; j < CPL; ++j) part[wid][wl * CPL + j] = acc[j];
  __syncthreads();
  // walker partials in a fixed order: group g sums walkers g, g + G, ... of
  // those that took a slot, then the groups in order
  const uint32_t used = e - b < (uint32_t)NW ? e - b : (uint32_t)NW;
  {
    const int col = threadIdx.x % R, g = threadIdx.x / R;
    double t = 0.0;
    for (uint32_t w = g; w < used; w += G) t += part[w][col];
    red[g][col] = t;
  }
  __syncthreads();
  if (threadIdx.x < R) {
    double t = 0.0;
#pragma unroll
    for (int g = 0; g < G; ++g) t += red[g][threadIdx.x];
    if (row == 0xFFFFFFFFu) {
      a.part[k][(gp - a.piece0[k]) * R + threadIdx.x] = t;
    } else {
      const double sc = ((hw >> (row & 31)) & 1u) ? 1.0 / (double)(e - b) : 1.0;
      a.F[k][(size_t)row * R + threadIdx.x] = (float)((double)fv + t * sc);
    }
  }
}

// Blocks 0 .. nb_merge - 1 merge mode 0's long rows (sgd_merge_block); after
// them one walker per staged row of more than one piece adds its pieces'
// partial sums in order, then the total (a cold row) or the total over its
// count of runs (a hot row) to the row.
template <int R>
__global__ void __launch_bounds__(256) k_sgd_fold(SettleArgs a) {
  constexpr int W = R < 64 ? R : 64;
  constexpr int CPL = R / W;
  constexpr int WPW = 64 / W;
  if (blockIdx.x < a.nb_merge) {
    sgd_merge_block<R>(a.mg, blockIdx.x);
    return;
  }
  const int wl = lane_id() % W;
  const uint64_t i = ((uint64_t)(blockIdx.x - a.nb_merge) * (blockDim.x >> 6) + (threadIdx.x >> 6)) * WPW + lane_id() / W;
  int k = 1;
  while (k < 3 && i >= a.fb[k + 1]) ++k;
  if (i >= a.fb[k + 1]) return;
  const FoldSeg sg = a.fold[k][i - a.fb[k]];
  const double *pt = a.part[k];
  double t[CPL];
#pragma unroll
  for (int j = 0; j < CPL; ++j) t[j] = 0.0;
  for (uint32_t q = 0; q < sg.npiece; ++q)
#pragma unroll
    for (int j = 0; j < CPL; ++j) t[j] += pt[(size_t)(sg.piece + q) * R + wl + j * W];
  const uint32_t *hot = a.hot[k];
  const double sc = ((hot[sg.row >> 5] >> (sg.row & 31)) & 1u) ? 1.0 / (double)sg.runs : 1.0;
  float *gr = a.F[k] + (size_t)sg.row * R;
#pragma unroll
  for (int j = 0; j < CPL; ++j) gr[wl + j * W] = (float)((double)gr[wl + j * W] + t[j] * sc);
}

// ==========================================================================
// MTTKRP / adjoint walk: per row, acc = sum over its nonzeros of w * h with
// h = hadamard of the other modes' rows and
//   w = x                       (kWalkMttkrp: SG/als.cu:295-312 intent), or
//   w = x - <F_mode(row), h>    (kWalkAdjoint, kWalkJacobi: the residual).
// Rows may span walkers, so each walker adds its partial row with atomics
// into a zeroed output.  For the adjoint, out receives -acc and loss_part
// the per-block sum of squared residuals.  kWalkJacobi is the reference's
// CCD sums (SG/ccd.cu:157-195): out receives acc = sum e h and den the
// sum of h * h, per column, all from the factors as they were at launch.
// kWalkLoss only sums the squared residuals into loss_part (the training
// loss of tc_loss_sq, SG/completion.c:166-189).
// ==========================================================================
enum : int { kWalkMttkrp = 0, kWalkAdjoint = 1, kWalkJacobi = 2, kWalkLoss = 3 };

template <int R, int N, int KIND>
__global__ void __launch_bounds__(256) k_mttkrp(CissArgs c, FactorArgs fa, uint32_t span,
                                                float *out, double *loss_part, float *den) {
  constexpr bool ADJ = KIND != kWalkMttkrp;  // w is the residual
  constexpr bool SQ = KIND == kWalkJacobi;
  // a walker is R / 4 lanes, each holding 4 consecutive entries of a row, so
  // a gathered row is one 16-byte load per lane and a wave carries 256 / R
  // walkers (the shape of the SGD walk)
  constexpr int W = SyncShape<R>::W;
  constexpr int CPL = SyncShape<R>::CPL;
  constexpr int WPW = SyncShape<R>::WPW;
  constexpr uint32_t NONE = 0xFFFFFFFFu;
  __shared__ double red[4];
  const int lane = lane_id();
  const int sub = lane / W;
  const int wl = lane % W;
  const uint64_t wave = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const uint64_t walker = wave * WPW + sub;
  const uint64_t s0 = walker * span;
  const uint64_t s1 = s0 + span < c.nnz ? s0 + span : c.nnz;
  const bool wave_idle = wave * WPW * (uint64_t)span >= c.nnz;

  float acc[CPL], own[CPL], sq[CPL];
  double lsum = 0.0;
  uint32_t arow = NONE, slot = 0;
#pragma unroll
  for (int q = 0; q < CPL; ++q) acc[q] = 0.f, own[q] = 0.f, sq[q] = 0.f;
  constexpr bool LOSS = KIND == kWalkAdjoint || KIND == kWalkLoss;  // the squared residuals are summed
  auto flush = [&]() {
    if (KIND == kWalkLoss) return;
    float *o = out + (size_t)arow * R + wl * CPL;
#pragma unroll
    for (int q2 = 0; q2 < CPL; ++q2) {
      atomic_add_f32(o + q2, KIND == kWalkAdjoint ? -acc[q2] : acc[q2]);
      if (SQ) atomic_add_f32(den + (size_t)arow * R + wl * CPL + q2, sq[q2]);
    }
  };

  if (!wave_idle) {
    for (uint32_t ch = 0; ch < span; ch += W) {
      const uint64_t el = s0 + ch + wl;
      const bool vl = el < s1;
      uint32_t il[N - 1];
      float xl = 0.f;
      bool hl = false;
      if (vl) {
#pragma unroll
        for (int s = 0; s < N - 1; ++s) il[s] = c.oidx[s][el];
        xl = c.vals[el];
        hl = (c.tile_mask[el >> 6] >> (el & 63)) & 1ull;
      } else {
#pragma unroll
        for (int s = 0; s < N - 1; ++s) il[s] = 0;
      }
      const uint64_t hm = __ballot(vl && hl);
#pragma unroll 1
      for (int q = 0; q < W; ++q) {
        const int src = sub * W + q;
        uint32_t idx[N - 1];
#pragma unroll
        for (int s = 0; s < N - 1; ++s) idx[s] = (uint32_t)__shfl((int)il[s], src, 64);
        const float x = __shfl(xl, src, 64);
        const uint64_t e = s0 + ch + q;
        if (e < s1) {
          const bool head = (hm >> src) & 1ull;
          if (arow == NONE || head) {
            if (arow != NONE) {
              flush();
              ++slot;
            } else {
              slot = c.tile_slot[s0 >> 6];
            }
            arow = c.directory[slot];
#pragma unroll
            for (int q2 = 0; q2 < CPL; ++q2) acc[q2] = 0.f, sq[q2] = 0.f;
            if (ADJ) load_row4<CPL>(fa.f[0] + (size_t)arow * R + wl * CPL, own);
          }
          float h[CPL];
          load_row4<CPL>(fa.f[1] + (size_t)idx[0] * R + wl * CPL, h);
#pragma unroll
          for (int s = 1; s < N - 1; ++s) {
            float g[CPL];
            load_row4<CPL>(fa.f[1 + s] + (size_t)idx[s] * R + wl * CPL, g);
#pragma unroll
            for (int q2 = 0; q2 < CPL; ++q2) h[q2] *= g[q2];
          }
          float w = x;
          if (ADJ) {
            float part = 0.f;
#pragma unroll
            for (int q2 = 0; q2 < CPL; ++q2) part += own[q2] * h[q2];
            w = x - group_sum<W>(part);
            if (LOSS && wl == 0) lsum += (double)w * (double)w;
          }
#pragma unroll
          for (int q2 = 0; q2 < CPL; ++q2) {
            acc[q2] += w * h[q2];
            if (SQ) sq[q2] += h[q2] * h[q2];
          }
        }
      }
    }
    if (arow != NONE) flush();
  }
  if (LOSS) {
    double t = wave_sum_d(lsum);
    if (lane == 0) red[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
      double s = 0.0;
      for (int w2 = 0; w2 < (int)(blockDim.x >> 6); ++w2) s += red[w2];
      loss_part[blockIdx.x] = s;
    }
  }
}

// ==========================================================================
// ALS, phase 1: Gram + MTTKRP walk on the FP64 matrix cores.
//
// One wave per span.  Per row segment the wave accumulates
//   G += h h^T  with v_mfma_f64_16x16x4_f64 (h = hadamard of the other
//               modes' rows, formed exactly in FP64 from the FP32 factors)
//   m += x h    in FP64 on the VALUs
// and writes (G, m) as an FP64 "piece" when the row ends or the span ends.
// Piece ids are slot + span_off[span], so the pieces of a slot are contiguous
// and written exactly once: no atomics, fixed summation order
// (deterministic).  Lane l holds column (l % 16) + 16 q of nonzero l / 16 of
// the current 4-nonzero step, which is exactly the A (16 x 4) and B (4 x 16)
// operand layout of the 16x16x4 shape, so both operands are the same h
// registers; block (bi, bj) of G (bi >= bj) takes A = h[bi], B = h[bj].
// FP64 accumulation keeps the normal equations within 1e-5 of an FP64
// reference even when (G + lambda I) is poorly conditioned.
// ==========================================================================
typedef double doublex4 __attribute__((ext_vector_type(4)));

template <int R, int N>
__device__ __forceinline__ void als_gram_walk(const CissArgs &c, const FactorArgs &fa, uint32_t span_tiles,
                                              uint64_t nspans, const uint32_t *span_off, double *pieces) {
  constexpr int CG = (R + 15) / 16;      // 16-column groups
  constexpr int NB = CG * (CG + 1) / 2;  // lower-triangle 16x16 blocks
  constexpr int NPS = 4;                 // nonzeros per MFMA step
  constexpr size_t PSZ = (size_t)R * R + R;
  const int lane = lane_id();
  const uint64_t w = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (w >= nspans) return;
  const uint64_t t0 = w * span_tiles;
  const uint64_t t1 = t0 + span_tiles < c.ntiles ? t0 + span_tiles : c.ntiles;
  const uint32_t poff = span_off[w];
  uint32_t slot = c.tile_slot[t0];
  const int col = lane & 15;
  const int nsub = lane >> 4;

  doublex4 acc[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) acc[b] = doublex4{0.0, 0.0, 0.0, 0.0};
  double macc[CG];
#pragma unroll
  for (int q = 0; q < CG; ++q) macc[q] = 0.0;

  auto flush = [&](uint32_t pid) {
    double *P = pieces + (size_t)pid * PSZ;
    int b = 0;
#pragma unroll
    for (int bi = 0; bi < CG; ++bi) {
#pragma unroll
      for (int bj = 0; bj <= bi; ++bj, ++b) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          // f64 16x16x4 output: accumulator v of lane l is row 4 v + l / 16
          const int row = 16 * bi + 4 * v + nsub, cc = 16 * bj + col;
          if (row < R && cc < R) P[row * R + cc] = acc[b][v];
          acc[b][v] = 0.0;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < CG; ++q) {
      double mt = macc[q];
      mt += __shfl_xor(mt, 16, 64);
      mt += __shfl_xor(mt, 32, 64);
      if (lane < 16 && 16 * q + lane < R) P[R * R + 16 * q + lane] = mt;
      macc[q] = 0.0;
    }
  };

  for (uint64_t t = t0; t < t1; ++t) {
    const uint64_t base = t * kTile;
    const uint32_t cnt = (uint32_t)((c.nnz - base) < (uint64_t)kTile ? (c.nnz - base) : kTile);
    uint32_t jl = 0, kl = 0, ll = 0;
    float xl = 0.f;
    if ((uint32_t)lane < cnt) {
      jl = c.oidx[0][base + lane];
      kl = c.oidx[1][base + lane];
      if (N == 4) ll = c.oidx[2][base + lane];
      xl = c.vals[base + lane];
    }
    uint64_t mask = c.tile_mask[t];
    if (cnt < 64) mask &= (1ull << cnt) - 1ull;
    uint32_t pos = 0;
    while (pos < cnt) {
      if ((mask >> pos) & 1ull) {
        if (!(t == t0 && pos == 0)) {
          flush(slot + poff);
          ++slot;
        }
      }
      const uint64_t rest = (pos >= 63) ? 0ull : ((mask >> (pos + 1)) << (pos + 1));
      uint32_t nxt = rest ? (uint32_t)__ffsll((unsigned long long)rest) - 1u : cnt;
      if (nxt > cnt) nxt = cnt;
      // the factor values of step e are gathered one step ahead (inside the
      // row segment), so a wave has the next step's loads in flight under
      // the current step's MFMAs
      float f1[CG], f2[CG], f3[CG];
      auto fetch = [&](uint32_t e, float (&g1)[CG], float (&g2)[CG], float (&g3)[CG]) {
        const uint32_t my = e + (uint32_t)nsub;
        const int src = (int)(my < 64 ? my : 63);
        const uint32_t j = (uint32_t)__shfl((int)jl, src, 64);
        const uint32_t k = (uint32_t)__shfl((int)kl, src, 64);
        const uint32_t l = N == 4 ? (uint32_t)__shfl((int)ll, src, 64) : 0u;
#pragma unroll
        for (int q = 0; q < CG; ++q) {
          const int cq = col + 16 * q;
          g1[q] = g2[q] = g3[q] = 0.f;
          if (my < nxt && cq < R) {
            g1[q] = fa.f[1][(size_t)j * R + cq];
            g2[q] = fa.f[2][(size_t)k * R + cq];
            if (N == 4) g3[q] = fa.f[3][(size_t)l * R + cq];
          }
        }
      };
      fetch(pos, f1, f2, f3);
      for (uint32_t e = pos; e < nxt; e += NPS) {
        const uint32_t my = e + (uint32_t)nsub;
        const bool valid = my < nxt;
        const int src = (int)(my < 64 ? my : 63);
        const double x = (double)__shfl(xl, src, 64);
        float n1[CG], n2[CG], n3[CG];
        if (e + NPS < nxt) fetch(e + NPS, n1, n2, n3);
        double h[CG];
#pragma unroll
        for (int q = 0; q < CG; ++q) {
          double v = 0.0;
          if (valid && col + 16 * q < R) {
            v = (double)f1[q] * (double)f2[q];
            if (N == 4) v *= (double)f3[q];
          }
          h[q] = v;
          macc[q] += x * v;
        }
        int b = 0;
#pragma unroll
        for (int bi = 0; bi < CG; ++bi)
#pragma unroll
          for (int bj = 0; bj <= bi; ++bj, ++b)
            acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(h[bi], h[bj], acc[b], 0, 0, 0);
#pragma unroll
        for (int q = 0; q < CG; ++q) f1[q] = n1[q], f2[q] = n2[q], f3[q] = n3[q];
      }
      pos = nxt;
    }
  }
  flush(slot + poff);
}

// Occupancy: at R <= 32 the walk (with its one-step gather lookahead) is held
// to 6 waves per SIMD, about 80 registers; at R = 64, and at R = 32 with four
// modes (which would spill at 80), it keeps the compiler's default.  Measured on the 100M-nonzero tensor at R = 32: 4.56 ms per mode
// without lookahead at the default occupancy, 4.27 ms without lookahead at 7
// waves, 4.00 ms with lookahead at 6.
template <int R, int N>
__global__ void __launch_bounds__(256, 6) k_als_gram(CissArgs c, FactorArgs fa, uint32_t span_tiles,
                                                     uint64_t nspans, const uint32_t *span_off, double *pieces) {
  als_gram_walk<R, N>(c, fa, span_tiles, nspans, span_off, pieces);
}
template <int R, int N>
__global__ void __launch_bounds__(256) k_als_gram_wide(CissArgs c, FactorArgs fa, uint32_t span_tiles,
                                                       uint64_t nspans, const uint32_t *span_off, double *pieces) {
  als_gram_walk<R, N>(c, fa, span_tiles, nspans, span_off, pieces);
}

// ==========================================================================
// ALS, phase 2: per directory slot, sum its pieces in FP64, add lambda to the
// diagonal (p_update_matrice, SG/als.cu:484-487), factor the R x R system by
// Cholesky (right-looking, Golub & Van Loan 4.2) in FP64 in LDS, solve, and
// optionally refine once: x += (G+lambda I)^-1 (m - (G+lambda I) x).
// One 64-thread workgroup per slot.  Replaces cusolverDnDpotrfBatched +
// DpotrsBatched (SG/als.cu:715-716).
// ==========================================================================
__device__ inline int tri_row(int t) {
  int a = (int)((sqrtf(8.f * (float)t + 1.f) - 1.f) * 0.5f);
  if ((a + 1) * (a + 2) / 2 <= t) ++a;
  if (a * (a + 1) / 2 > t) --a;
  return a;
}

// Heavy slots (more pieces than fit one workgroup's serial sum) are first
// reduced in fixed groups of consecutive pieces, one workgroup per group, so
// the solve sums a handful of group sums instead of thousands of pieces.
template <int R>
__global__ void __launch_bounds__(64) k_piece_reduce(const double *pieces, const uint32_t *grp_piece,
                                                     const uint32_t *grp_n, uint64_t ngroups, double *gsum) {
  constexpr size_t PSZ = (size_t)R * R + R;
  const uint64_t g = blockIdx.x;
  if (g >= ngroups) return;
  const double *P = pieces + (size_t)grp_piece[g] * PSZ;
  const uint32_t n = grp_n[g];
  double *out = gsum + g * PSZ;
  for (size_t e = threadIdx.x; e < PSZ; e += 64) {
    double acc = 0.0;
    for (uint32_t q = 0; q < n; ++q) acc += P[(size_t)q * PSZ + e];
    out[e] = acc;
  }
}

template <int R, bool REFINE>
__global__ void __launch_bounds__(64) k_als_solve(const double *pieces, const uint32_t *slot_pb,
                                                  const double *gsum, const uint32_t *slot_gb,
                                                  const uint32_t *directory, uint64_t dlength,
                                                  double lam, float *A, uint32_t *flag) {
  constexpr int LD = R + 1;
  constexpr size_t PSZ = (size_t)R * R + R;
  constexpr int PER = (R * R + 63) / 64;
  __shared__ double Gs[R * LD];
  __shared__ double G0[REFINE ? R * LD : 1];
  __shared__ double ys[R], rs[R];
  const int lane = threadIdx.x;
  const uint64_t s = blockIdx.x;
  if (s >= dlength) return;
  uint32_t p0 = slot_pb[s], p1 = slot_pb[s + 1];
  const double *src = pieces;
  if (slot_gb && slot_gb[s + 1] > slot_gb[s]) {  // heavy slot: read its group sums
    p0 = slot_gb[s];
    p1 = slot_gb[s + 1];
    src = gsum;
  }

  double g[PER];
#pragma unroll
  for (int t = 0; t < PER; ++t) g[t] = 0.0;
  double mv = 0.0;
  for (uint32_t p = p0; p < p1; ++p) {
    const double *P = src + (size_t)p * PSZ;
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      const int idx = lane + 64 * t;
      if (idx < R * R) {
        const int i = idx / R, j = idx % R;
        const int hi = i > j ? i : j, lo = i > j ? j : i;
        g[t] += P[hi * R + lo];
      }
    }
    if (lane < R) mv += P[R * R + lane];
  }
#pragma unroll
  for (int t = 0; t < PER; ++t) {
    const int idx = lane + 64 * t;
    if (idx < R * R) {
      const int i = idx / R, j = idx % R;
      const double v = g[t] + (i == j ? lam : 0.0);
      Gs[i * LD + j] = v;
      if (REFINE) G0[i * LD + j] = v;
    }
  }
  if (lane < R) {
    ys[lane] = mv;
    rs[lane] = mv;
  }
  __syncthreads();

  bool bad = false;
  for (int k = 0; k < R; ++k) {
    double d = Gs[k * LD + k];
    if (!(d > 0.0)) {
      bad = true;
      d = 1.0;
    }
    const double lkk = sqrt(d), inv = 1.0 / lkk;
    __syncthreads();
    if (lane == 0) Gs[k * LD + k] = lkk;
    for (int i = k + 1 + lane; i < R; i += 64) Gs[i * LD + k] *= inv;
    __syncthreads();
    const int n = R - k - 1;
    const int T = n * (n + 1) / 2;
    for (int t = lane; t < T; t += 64) {
      const int a = tri_row(t), b = t - a * (a + 1) / 2;
      const int i = k + 1 + a, j = k + 1 + b;
      Gs[i * LD + j] -= Gs[i * LD + k] * Gs[j * LD + k];
    }
    __syncthreads();
  }

  auto solve = [&](double *y) __attribute__((always_inline)) {
    for (int k = 0; k < R; ++k) {  // L y = b
      const double yk = y[k] / Gs[k * LD + k];
      __syncthreads();
      if (lane == 0) y[k] = yk;
      for (int i = k + 1 + lane; i < R; i += 64) y[i] -= Gs[i * LD + k] * yk;
      __syncthreads();
    }
    for (int k = R - 1; k >= 0; --k) {  // L^T x = y
      const double xk = y[k] / Gs[k * LD + k];
      __syncthreads();
      if (lane == 0) y[k] = xk;
      for (int i = lane; i < k; i += 64) y[i] -= Gs[k * LD + i] * xk;
      __syncthreads();
    }
  };
  solve(ys);
  if (REFINE) {
    if (lane < R) {
      double r = rs[lane];
      for (int j = 0; j < R; ++j) {
        const int hi = lane > j ? lane : j, lo = lane > j ? j : lane;
        r -= G0[hi * LD + lo] * ys[j];
      }
      rs[lane] = r;
    }
    __syncthreads();
    solve(rs);
    if (lane < R) ys[lane] += rs[lane];
    __syncthreads();
  }
  const uint32_t row = directory[s];
  if (lane < R) A[(size_t)row * R + lane] = bad ? 0.f : (float)ys[lane];
  if (bad && lane == 0) atomicOr(flag, 1u);
}

// Same solve without refinement for R <= 32, one half-wave (32 lanes) per
// slot, eight slots per 256-thread workgroup, no workgroup barriers.  Lane c
// keeps column c of G in registers (rows i >= c; the upper part of the
// register column holds don't-care values).  Root-free right-looking LDL^T:
// at step k lane k publishes its column and right-hand-side entry to a
// per-slot LDS buffer; d_k = buf[k] + lambda (lambda only ever sits on the
// diagonal, so it is added when the pivot is read); every lane c > k
// subtracts buf[i] * buf[c] / d_k from its column and l_ck z_k from its
// right-hand-side entry, so the forward solve L z = m rides along.  Masking is
// done by zero coefficients, not branches.  Then x = L^-T (D^-1 z), with
// l_kc = g[k] / d_c read from lane c's own registers.  Instruction count per
// system is about R^2 FMAs plus R reciprocals; the kernel is issue-bound, so
// that count is its time.
template <int R>
__global__ void __launch_bounds__(256) k_als_solve_h(const double *pieces, const uint32_t *slot_pb,
                                                     const double *gsum, const uint32_t *slot_gb,
                                                     const uint32_t *directory, uint64_t dlength, double lam,
                                                     float *A, uint32_t *flag) {
  static_assert(R <= 32, "one half-wave per slot");
  constexpr size_t PSZ = (size_t)R * R + R;
  constexpr int CB = R + 2;  // column k, then z_k / x_k; even, so every half's buffer is 16-byte aligned
  __shared__ __align__(16) double cbuf[8][CB];
  const int lane = lane_id();
  const int c = lane & 31;
  const int cc = c < R ? c : R - 1;  // clamped column for lanes past the rank
  const int hw = (int)(threadIdx.x >> 5);
  const uint64_t s = (uint64_t)blockIdx.x * 8 + (uint64_t)hw;
  const bool live = s < dlength;
  double *buf = cbuf[hw];

  double g[R];
#pragma unroll
  for (int i = 0; i < R; ++i) g[i] = 0.0;
  double b = 0.0;
  if (live && c < R) {
    uint32_t p0 = slot_pb[s], p1 = slot_pb[s + 1];
    const double *src = pieces;
    if (slot_gb && slot_gb[s + 1] > slot_gb[s]) {  // heavy slot: its group sums
      p0 = slot_gb[s];
      p1 = slot_gb[s + 1];
      src = gsum;
    }
    for (uint32_t p = p0; p < p1; ++p) {
      const double *P = src + (size_t)p * PSZ;
      // rows i < c are above the diagonal: don't-care (inside the piece, never used)
#pragma unroll
      for (int i = 0; i < R; ++i) g[i] += P[i * R + c];
      b += P[R * R + c];
    }
  }

  bool bad = false;
  double rd = 1.0;  // lane c: 1 / d_c
#pragma unroll
  for (int k = 0; k < R; ++k) {
    if (c == k) {
#pragma unroll
      for (int i = k; i < R; ++i) buf[i] = g[i];
      buf[R] = b;
    }
    wave_sync();
    double d = buf[k] + lam;
    if (!(d > 0.0)) {
      bad = true;
      d = 1.0;
    }
    const double rk = rcp_f64(d);
    const double zk = buf[R];
    const double lck = c > k ? buf[cc] * rk : 0.0;  // l_ck for the lanes still to factor, else 0
#pragma unroll
    for (int i = k + 1; i < R; ++i) g[i] = __builtin_fma(-buf[i], lck, g[i]);
    b = __builtin_fma(-lck, zk, b);
    if (c == k) rd = rk;
    wave_sync();
  }
  // w = D^-1 z, then L^T x = w
  b *= rd;
#pragma unroll
  for (int k = R - 1; k >= 0; --k) {
    if (c == k) buf[R] = b;
    wave_sync();
    const double xk = buf[R];
    const double lkc = c < k ? g[k] * rd : 0.0;  // l_kc (unit lower factor) for lanes before k
    b = __builtin_fma(-lkc, xk, b);
    wave_sync();
  }
  if (live && c < R) A[(size_t)directory[s] * R + c] = bad ? 0.f : (float)b;
  if (live && bad && c == 0) atomicOr(flag, 1u);
}

// ==========================================================================
// Evaluation: sum of squared errors over a COO (validation / test) or per
// problem of a batch, FP64 accumulation, one partial per block (fixed grid,
// fixed order -> deterministic).  The training loss is a walk of the mode-0
// CISS (k_mttkrp, kWalkLoss).
// ==========================================================================
template <int N>
__device__ inline float predict_row(const float *const *F, const uint32_t *ix, int R) {
  const float *r0 = F[0] + (size_t)ix[0] * R;
  const float *r1 = F[1] + (size_t)ix[1] * R;
  const float *r2 = N > 2 ? F[2] + (size_t)ix[2] * R : nullptr;
  const float *r3 = N > 3 ? F[3] + (size_t)ix[3] * R : nullptr;
  float acc = 0.f;
  for (int f = 0; f < R; f += 4) {
    float4 a = *(const float4 *)(r0 + f);
    float4 b = *(const float4 *)(r1 + f);
    float4 v = make_float4(a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w);
    if (N > 2) {
      float4 cc = *(const float4 *)(r2 + f);
      v = make_float4(v.x * cc.x, v.y * cc.y, v.z * cc.z, v.w * cc.w);
    }
    if (N > 3) {
      float4 d = *(const float4 *)(r3 + f);
      v = make_float4(v.x * d.x, v.y * d.y, v.z * d.z, v.w * d.w);
    }
    acc += (v.x + v.y) + (v.z + v.w);
  }
  return acc;
}

__device__ inline void block_store_sum(double v, double *out) {
  __shared__ double red[8];
  v = wave_sum_d(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
    *out = s;
  }
}

template <int N>
__global__ void __launch_bounds__(256) k_eval_coo(CooArgs c, FactorArgs fa, int R, double *part) {
  double acc = 0.0;
  const float *F[4] = {fa.f[0], fa.f[1], fa.f[2], fa.f[3]};
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < c.nnz;
       p += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t ix[4];
#pragma unroll
    for (int m = 0; m < N; ++m) ix[m] = c.ind[m][p];
    const double e = (double)c.vals[p] - (double)predict_row<N>(F, ix, R);
    acc += e * e;
  }
  block_store_sum(acc, part + blockIdx.x);
}

// squared error per segment [seg[b], seg[b+1]) of the nonzeros, one block per
// segment (a segment = one problem of a batch); CISS=true reads mode-0 CISS
// order, false a COO
template <int N, bool CISS>
__global__ void __launch_bounds__(256) k_eval_seg(CissArgs c, CooArgs co, FactorArgs fa, int R,
                                                  const uint64_t *seg, double *out) {
  double acc = 0.0;
  const float *F[4] = {fa.f[0], fa.f[1], fa.f[2], fa.f[3]};
  const uint64_t p1 = seg[blockIdx.x + 1];
  for (uint64_t p = seg[blockIdx.x] + threadIdx.x; p < p1; p += blockDim.x) {
    uint32_t ix[4];
    float x;
    if (CISS) {
      const uint64_t t = p >> 6;
      ix[0] = c.directory[tile_lane_slot(c.tile_slot[t], c.tile_mask[t], (int)(p & 63))];
#pragma unroll
      for (int s = 0; s < N - 1; ++s) ix[1 + s] = c.oidx[s][p];
      x = c.vals[p];
    } else {
#pragma unroll
      for (int m = 0; m < N; ++m) ix[m] = co.ind[m][p];
      x = co.vals[p];
    }
    const double e = (double)x - (double)predict_row<N>(F, ix, R);
    acc += e * e;
  }
  block_store_sum(acc, out + blockIdx.x);
}

__global__ void __launch_bounds__(256) k_sumsq(const float *x, uint64_t n, double *part) {
  double acc = 0.0;
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (uint64_t)gridDim.x * blockDim.x) {
    const double v = x[p];
    acc += v * v;
  }
  block_store_sum(acc, part + blockIdx.x);
}

// out[k] = sum of part[k * nb .. k * nb + nb) in a fixed order
__global__ void __launch_bounds__(64) k_reduce_parts(const double *part, uint32_t nb, double *out) {
  const double *p = part + (size_t)blockIdx.x * nb;
  double acc = 0.0;
  for (uint32_t i = threadIdx.x; i < nb; i += 64) acc += p[i];
  acc = wave_sum_d(acc);
  if (threadIdx.x == 0) out[blockIdx.x] = acc;
}

// ==========================================================================
// elementwise helpers
// ==========================================================================
__global__ void __launch_bounds__(256) k_fill(float *x, uint64_t n, float v) {
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (uint64_t)gridDim.x * blockDim.x)
    x[p] = v;
}

// plain streaming copy of n 16-byte words: the bandwidth yardstick and the
// known-size kernel that calibrates HBM traffic counters.  (Four loads in
// flight per lane before the stores measured the same rate, 4.59 TB/s.)
__global__ void __launch_bounds__(256) k_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, uint64_t n) {
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (uint64_t)gridDim.x * blockDim.x)
    dst[p] = src[p];
}

// grad += lambda * c_i * F  for the rows of a CISS directory (c_i = row nnz)
// or, when per_occurrence == 0, for every row (c_i = 1).
__global__ void __launch_bounds__(256) k_reg_rows(const uint32_t *directory, const uint32_t *rowptr,
                                                  uint64_t nrows, int R, float lam, int per_occ,
                                                  const float *F, float *G, double *part) {
  double acc = 0.0;
  const uint64_t total = nrows * (uint64_t)R;
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < total;
       p += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t s = p / R;
    const int f = (int)(p % R);
    uint64_t row = s;
    float cnt = 1.f;
    if (per_occ) {
      row = directory[s];
      cnt = (float)(rowptr[s + 1] - rowptr[s]);
    }
    const float a = F[row * R + f];
    G[row * R + f] += lam * cnt * a;
    acc += 0.5 * (double)lam * (double)cnt * (double)a * (double)a;
  }
  block_store_sum(acc, part + blockIdx.x);
}

// ==========================================================================
// launchers
// ==========================================================================
#define CTC_R_SWITCH(R, BODY)                                       \
  switch (R) {                                                      \
    case 8: { constexpr int RR = 8; BODY; } break;                  \
    case 16: { constexpr int RR = 16; BODY; } break;                \
    case 32: { constexpr int RR = 32; BODY; } break;                \
    case 64: { constexpr int RR = 64; BODY; } break;                \
    case 128: { constexpr int RR = 128; BODY; } break;              \
    default: return hipErrorInvalidValue;                           \
  }

static inline unsigned grid_for(uint64_t waves, unsigned block) {
  const uint64_t wpb = block / 64;
  return (unsigned)((waves + wpb - 1) / wpb);
}

// policy code = PA * 64 + PB * 8 + PC
template <int N>
static hipError_t launch_sgd_n(hipStream_t st, const CissArgs &c, const FactorArgs &f, const HotRows &hot, int R,
                               uint32_t span,
                               uint64_t nspans, uint64_t stride, int code, float step, float lam, unsigned grid,
                               unsigned block) {
#define CTC_SGD_P(PA, PB, PC) \
  case PA * 64 + PB * 8 + PC:  \
    CTC_R_SWITCH(R, hipLaunchKernelGGL((k_sgd<RR, N, PA, PB, PC>), dim3(grid), dim3(block), 0, st, c, f, hot, span, nspans, stride, step, lam)); \
    break;
  switch (code) {
    CTC_SGD_P(P_RACY, P_HOT, P_HOT)
    CTC_SGD_P(P_RACY, P_HOTRUN, P_HOTRUN)
    CTC_SGD_P(P_RACY, P_ATOMIC, P_ATOMIC)
    CTC_SGD_P(P_RACY, P_RACY, P_RACY)
    CTC_SGD_P(P_RACY, P_RACY, P_ATOMIC)
    default:
      return hipErrorInvalidValue;
  }
#undef CTC_SGD_P
  return hipGetLastError();
}

hipError_t launch_sgd(hipStream_t st, const CissArgs &c, const FactorArgs &f, const HotRows &hot, int nmodes, int R,
                      uint32_t span, uint64_t stride, int code, float step, float lam, unsigned block) {
  if (c.nnz == 0) return hipSuccess;
  const uint64_t nspans = (c.nnz + span - 1) / span;
  const int W = R < 64 ? R : 64;
  const uint64_t waves = (nspans + (64 / W) - 1) / (64 / W);
  const unsigned grid = grid_for(waves, block);
  if (nmodes == 3) return launch_sgd_n<3>(st, c, f, hot, R, span, nspans, stride, code, step, lam, grid, block);
  if (nmodes == 4) return launch_sgd_n<4>(st, c, f, hot, R, span, nspans, stride, code, step, lam, grid, block);
  return hipErrorInvalidValue;
}

hipError_t launch_sgd_sync(hipStream_t st, const CissArgs &c, const FactorArgs &f, const SyncArgs &sy, int nmodes,
                           int R, uint32_t span, uint32_t sub, uint32_t s, uint32_t long_min, float step, float lam,
                           unsigned block, bool mixed) {
  if (c.nnz == 0) return hipSuccess;
  const uint64_t nspans = (c.nnz + span - 1) / span;
  const int wpw = 256 / R;  // walkers per wave (SyncShape)
  const uint64_t waves = (nspans + wpw - 1) / wpw;
  const unsigned grid = grid_for(waves, block);
#define CTC_SYNC_LAUNCH(NN, MM) \
  CTC_R_SWITCH(R, hipLaunchKernelGGL((k_sgd_sync<RR, NN, MM>), dim3(grid), dim3(block), 0, st, c, f, sy, span, sub, s, nspans, long_min, step, lam))
  if (nmodes == 3) {
    if (mixed) {
      CTC_SYNC_LAUNCH(3, true);
    } else {
      CTC_SYNC_LAUNCH(3, false);
    }
  } else if (nmodes == 4) {
    if (mixed) {
      CTC_SYNC_LAUNCH(4, true);
    } else {
      CTC_SYNC_LAUNCH(4, false);
    }
  } else {
    return hipErrorInvalidValue;
  }
#undef CTC_SYNC_LAUNCH
  return hipGetLastError();
}

hipError_t launch_sgd_settle(hipStream_t st, SettleArgs a, int R, bool mixed) {
  const int W = R < 64 ? R : 64;
  const uint64_t items = a.mg.off[a.mg.nseg];
  a.nb_merge = grid_for((items + (64 / W) - 1) / (64 / W), 256);
  const uint64_t nw = 1024 / (uint64_t)R;  // light rows per block: 256 / (R / 4) walkers
  const uint64_t sb = a.pb[4] + (a.lb[4] + nw - 1) / nw;
  if (sb > 0x7FFFFFFFull) return hipErrorInvalidValue;
  if (sb > 0) {
    if (mixed) {
      CTC_R_SWITCH(R, hipLaunchKernelGGL((k_sgd_settle<RR, true>), dim3((unsigned)sb), dim3(256), 0, st, a));
    } else {
      CTC_R_SWITCH(R, hipLaunchKernelGGL((k_sgd_settle<RR, false>), dim3((unsigned)sb), dim3(256), 0, st, a));
    }
  }
  const uint64_t fg = a.nb_merge + grid_for((a.fb[4] + (64 / W) - 1) / (64 / W), 256);
  if (fg > 0) CTC_R_SWITCH(R, hipLaunchKernelGGL((k_sgd_fold<RR>), dim3((unsigned)fg), dim3(256), 0, st, a));
  return hipGetLastError();
}

hipError_t launch_mttkrp(hipStream_t st, const CissArgs &c, const FactorArgs &f, int nmodes, int R,
                         uint32_t span, bool adjoint, float *out, double *loss_part,
                         unsigned *grid_out, unsigned block) {
  const unsigned grid = (unsigned)mttkrp_grid(c.nnz, R, span, block);
  *grid_out = grid;
  if (c.nnz == 0) return hipSuccess;
  float *none = nullptr;
  if (nmodes == 3) {
    if (adjoint) {
      CTC_R_SWITCH(R, hipLaunchKernelGGL((k_mttkrp<RR, 3, kWalkAdjoint>), dim3(grid), dim3(block), 0, st, c, f, span, out, loss_part, none));
    } else {
      CTC_R_SWITCH(R, hipLaunchKernelGGL((k_mttkrp<RR, 3, kWalkMttkrp>), dim3(grid), dim3(block), 0, st, c, f, span, out, loss_part, none));
    }
  } else if (nmodes == 4) {
    if (adjoint) {
      CTC_R_SWITCH(R, hipLaunchKernelGGL((k_mttkrp<RR, 4, kWalkAdjoint>), dim3(grid), dim3(block), 0, st, c, f, span, out, loss_part, none));
    } else {
      CTC_R_SWITCH(R, hipLaunchKernelGGL((k_mttkrp<RR, 4, kWalkMttkrp>), dim3(grid), dim3(block), 0, st, c, f, span, out, loss_part, none));
    }
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_ccd_jacobi(hipStream_t st, const CissArgs &c, const FactorArgs &f, int nmodes, int R,
                             uint32_t span, float *num, float *den, unsigned block) {
  if (c.nnz == 0) return hipSuccess;
  const unsigned grid = (unsigned)mttkrp_grid(c.nnz, R, span, block);
  double *none = nullptr;
  if (nmodes == 3) {
    CTC_R_SWITCH(R, hipLaunchKernelGGL((k_mttkrp<RR, 3, kWalkJacobi>), dim3(grid), dim3(block), 0, st, c, f, span, num, none, den));
  } else if (nmodes == 4) {
    CTC_R_SWITCH(R, hipLaunchKernelGGL((k_mttkrp<RR, 4, kWalkJacobi>), dim3(grid), dim3(block), 0, st, c, f, span, num, none, den));
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_loss_walk(hipStream_t st, const CissArgs &c, const FactorArgs &f, int nmodes, int R,
                            uint32_t span, double *part, unsigned block) {
  if (c.nnz == 0) return hipSuccess;
  const unsigned grid = (unsigned)mttkrp_grid(c.nnz, R, span, block);
  float *none = nullptr;
  if (nmodes == 3) {
    CTC_R_SWITCH(R, hipLaunchKernelGGL((k_mttkrp<RR, 3, kWalkLoss>), dim3(grid), dim3(block), 0, st, c, f, span, none, part, none));
  } else if (nmodes == 4) {
    CTC_R_SWITCH(R, hipLaunchKernelGGL((k_mttkrp<RR, 4, kWalkLoss>), dim3(grid), dim3(block), 0, st, c, f, span, none, part, none));
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

uint64_t mttkrp_walkers_per_block(int R, unsigned block) { return (uint64_t)(block / 64) * (64 / (R / 4)); }

uint64_t mttkrp_grid(uint64_t nnz, int R, uint32_t span, unsigned block) {
  const uint64_t walkers = (nnz + span - 1) / span;
  const uint64_t wpw = 64 / (R / 4);  // walkers per wave (SyncShape<R>)
  const uint64_t waves = (walkers + wpw - 1) / wpw;
  return grid_for(waves, block);
}

hipError_t launch_als_gram(hipStream_t st, const CissArgs &c, const FactorArgs &f, int nmodes, int R,
                           uint32_t span_tiles, uint64_t nspans, const uint32_t *span_off,
                           double *pieces, unsigned block) {
  if (c.nnz == 0) return hipSuccess;
  const unsigned grid = grid_for(nspans, block);
#define CTC_GRAM(RR)                                                                                     \
  do {                                                                                                    \
    if (nmodes == 3)                                                                                      \
      hipLaunchKernelGGL((k_als_gram<RR, 3>), dim3(grid), dim3(block), 0, st, c, f, span_tiles, nspans, \
                         span_off, pieces);                                                               \
    else                                                                                                  \
      hipLaunchKernelGGL((k_als_gram<RR, 4>), dim3(grid), dim3(block), 0, st, c, f, span_tiles, nspans, \
                         span_off, pieces);                                                               \
  } while (0)
  switch (R) {
    case 8: CTC_GRAM(8); break;
    case 16: CTC_GRAM(16); break;
    case 32:
      if (nmodes == 3)
        CTC_GRAM(32);
      else  // four modes need more than the 80 registers of 6 waves per SIMD (it would spill)
        hipLaunchKernelGGL((k_als_gram_wide<32, 4>), dim3(grid), dim3(block), 0, st, c, f, span_tiles, nspans,
                           span_off, pieces);
      break;
    case 64:
      if (nmodes == 3)
        hipLaunchKernelGGL((k_als_gram_wide<64, 3>), dim3(grid), dim3(block), 0, st, c, f, span_tiles, nspans,
                           span_off, pieces);
      else
        hipLaunchKernelGGL((k_als_gram_wide<64, 4>), dim3(grid), dim3(block), 0, st, c, f, span_tiles, nspans,
                           span_off, pieces);
      break;
    default: return hipErrorInvalidValue;
  }
#undef CTC_GRAM
  return hipGetLastError();
}

hipError_t launch_piece_reduce(hipStream_t st, const double *pieces, const uint32_t *grp_piece,
                               const uint32_t *grp_n, uint64_t ngroups, int R, double *gsum) {
  if (ngroups == 0) return hipSuccess;
  const unsigned grid = (unsigned)ngroups;
  switch (R) {
    case 8: hipLaunchKernelGGL((k_piece_reduce<8>), dim3(grid), dim3(64), 0, st, pieces, grp_piece, grp_n, ngroups, gsum); break;
    case 16: hipLaunchKernelGGL((k_piece_reduce<16>), dim3(grid), dim3(64), 0, st, pieces, grp_piece, grp_n, ngroups, gsum); break;
    case 32: hipLaunchKernelGGL((k_piece_reduce<32>), dim3(grid), dim3(64), 0, st, pieces, grp_piece, grp_n, ngroups, gsum); break;
    case 64: hipLaunchKernelGGL((k_piece_reduce<64>), dim3(grid), dim3(64), 0, st, pieces, grp_piece, grp_n, ngroups, gsum); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_als_solve(hipStream_t st, const double *pieces, const uint32_t *slot_pb, const double *gsum,
                            const uint32_t *slot_gb, const uint32_t *directory, uint64_t dlength, int R,
                            double lam, bool refine, bool light, float *A, uint32_t *flag) {
  if (dlength == 0) return hipSuccess;
  const unsigned grid = (unsigned)dlength;
  // light: few pieces per slot (short rows), where the per-slot factorisation
  // dominates; slots that sum many pieces keep the 64-lane summation
  if (!refine && light && R <= 32) {
    const unsigned g8 = (unsigned)((dlength + 7) / 8);
    switch (R) {
      case 8: hipLaunchKernelGGL((k_als_solve_h<8>), dim3(g8), dim3(256), 0, st, pieces, slot_pb, gsum, slot_gb, directory, dlength, lam, A, flag); break;
      case 16: hipLaunchKernelGGL((k_als_solve_h<16>), dim3(g8), dim3(256), 0, st, pieces, slot_pb, gsum, slot_gb, directory, dlength, lam, A, flag); break;
      case 32: hipLaunchKernelGGL((k_als_solve_h<32>), dim3(g8), dim3(256), 0, st, pieces, slot_pb, gsum, slot_gb, directory, dlength, lam, A, flag); break;
      default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
#define CTC_SOLVE(RR)                                                                                      \
  do {                                                                                                      \
    if (refine)                                                                                             \
      hipLaunchKernelGGL((k_als_solve<RR, true>), dim3(grid), dim3(64), 0, st, pieces, slot_pb, gsum,      \
                         slot_gb, directory, dlength, lam, A, flag);                                        \
    else                                                                                                    \
      hipLaunchKernelGGL((k_als_solve<RR, false>), dim3(grid), dim3(64), 0, st, pieces, slot_pb, gsum,     \
                         slot_gb, directory, dlength, lam, A, flag);                                        \
  } while (0)
  switch (R) {
    case 8: CTC_SOLVE(8); break;
    case 16: CTC_SOLVE(16); break;
    case 32: CTC_SOLVE(32); break;
    case 64: CTC_SOLVE(64); break;
    default: return hipErrorInvalidValue;
  }
#undef CTC_SOLVE
  return hipGetLastError();
}

hipError_t launch_eval_coo(hipStream_t st, const CooArgs &c, const FactorArgs &f, int nmodes, int R,
                           unsigned nblocks, double *part) {
  if (nmodes == 3)
    hipLaunchKernelGGL((k_eval_coo<3>), dim3(nblocks), dim3(256), 0, st, c, f, R, part);
  else if (nmodes == 4)
    hipLaunchKernelGGL((k_eval_coo<4>), dim3(nblocks), dim3(256), 0, st, c, f, R, part);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_eval_seg(hipStream_t st, const CissArgs *c, const CooArgs *co, const FactorArgs &f,
                           int nmodes, int R, const uint64_t *seg, uint32_t nseg, double *out) {
  if (nseg == 0) return hipSuccess;
  const CissArgs ca = c ? *c : CissArgs{};
  const CooArgs cb = co ? *co : CooArgs{};
#define CTC_SEG(NN)                                                                                    \
  do {                                                                                                \
    if (c)                                                                                            \
      hipLaunchKernelGGL((k_eval_seg<NN, true>), dim3(nseg), dim3(256), 0, st, ca, cb, f, R, seg, out);  \
    else                                                                                              \
      hipLaunchKernelGGL((k_eval_seg<NN, false>), dim3(nseg), dim3(256), 0, st, ca, cb, f, R, seg, out); \
  } while (0)
  if (nmodes == 3)
    CTC_SEG(3);
  else if (nmodes == 4)
    CTC_SEG(4);
  else
    return hipErrorInvalidValue;
#undef CTC_SEG
  return hipGetLastError();
}

hipError_t launch_sumsq(hipStream_t st, const float *x, uint64_t n, unsigned nblocks, double *part) {
  hipLaunchKernelGGL(k_sumsq, dim3(nblocks), dim3(256), 0, st, x, n, part);
  return hipGetLastError();
}

hipError_t launch_reduce_parts(hipStream_t st, const double *part, uint32_t nb, uint32_t groups,
                               double *out) {
  hipLaunchKernelGGL(k_reduce_parts, dim3(groups), dim3(64), 0, st, part, nb, out);
  return hipGetLastError();
}

hipError_t launch_fill(hipStream_t st, float *x, uint64_t n, float v) {
  if (n == 0) return hipSuccess;
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(256), 0, st, x, n, v);
  return hipGetLastError();
}

hipError_t launch_copy(hipStream_t st, const void *src, void *dst, uint64_t bytes) {
  const uint64_t n = bytes / 16;
  if (n == 0) return hipSuccess;
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(256), 0, st, (const uint4 *)src, (uint4 *)dst, n);
  return hipGetLastError();
}

hipError_t launch_reg_rows(hipStream_t st, const uint32_t *directory, const uint32_t *rowptr,
                           uint64_t nrows, int R, float lam, int per_occ, const float *F, float *G,
                           unsigned nblocks, double *part) {
  hipLaunchKernelGGL(k_reg_rows, dim3(nblocks), dim3(256), 0, st, directory, rowptr, nrows, R, lam,
                     per_occ, F, G, part);
  return hipGetLastError();
}

}  // namespace ctc
