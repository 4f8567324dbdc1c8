// kernels_ccd.hip — CCD++ column kernels for gfx950 (canonical CCD++ as the
// reference states it in comments, SG/ccd.cu:532-554 and NC/ccd.cu:28-180:
// for each column f, add the rank-one component back to the residual, update
// column f of every mode as a 1-D ridge solve a_if = sum(r h) / (lambda +
// sum h^2), then subtract the new component).
//
// Layout.  Every mode's CISS carries its own copy of the residual in its own
// nonzero order, so each column pass streams it coalesced.  The factors are
// mirrored column-major (R x dims) during an epoch.  A copy stores "x minus
// every component except column `pending`"; the pass of column f over mode m
// first completes the previous column (subtracting its new component), then
// adds column f back with its pre-update values (saved before the column's
// first mode update), and sums num/den.  The three values a pass needs of a
// row (pending column's new value, column f saved, column f current) sit in
// one 16-byte record per row, built at the start of each column: a nonzero
// costs one gather per mode instead of three.  Divergent gathers are what
// the pass is limited by (the texture addresser is the busy unit), so the
// number of gather instructions, not their width, sets its time.
//
// Reductions.  One wave walks a span of tiles, one lane per nonzero.  Row
// sums inside a tile use a segmented inclusive scan (DPP row shifts and
// ds_bpermute for the cross-row steps) with the row starts of the tile
// mask as flags; the row still open at the end of a tile is carried in a
// wave-uniform register into the next tile.  Completed rows are written as
// (num, den) pieces with the same piece numbering as the ALS Gram pieces, so
// the per-row sums have a fixed order (deterministic, no atomics).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels_ccd.h"
#include "kernels_common.h"

namespace ctc {

// inclusive segmented scan over the wave: f = segment-start flag
__device__ inline void seg_scan(float &a, float &b, bool f) {
  const int lane = lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const float ua = __shfl_up(a, d, 64);
    const float ub = __shfl_up(b, d, 64);
    const int uf = __shfl_up((int)f, d, 64);
    if (lane >= d) {
      if (!f) {
        a += ua;
        b += ub;
      }
      f = f || uf;
    }
  }
}

// The pass of wave w over its span.  LS >= 0: the records of other mode LS
// (a mode of at most kCcdLdsRows rows) are read from the workgroup's copy xs
// in LDS instead of gathered from memory.
template <int N, int LS>
__device__ __forceinline__ void ccd_col_wave(const CcdColArgs &a, uint64_t w, const float4 *xs) {
  const int lane = lane_id();
  const CissArgs &c = a.c;
  const uint64_t t0 = w * a.span_tiles;
  const uint64_t t1 = t0 + a.span_tiles < c.ntiles ? t0 + a.span_tiles : c.ntiles;
  const uint32_t poff = a.span_off[w];
  const bool fix = a.sub_col >= 0 || a.add_back;  // this pass rewrites the residual
  double carry_n = 0.0, carry_d = 0.0;
  uint32_t carry_slot = 0;
  bool have_carry = false;

  for (uint64_t t = t0; t < t1; ++t) {
    const uint64_t base = t * kTile;
    const uint32_t cnt = (uint32_t)((c.nnz - base) < (uint64_t)kTile ? (c.nnz - base) : kTile);
    uint64_t mask = c.tile_mask[t];
    if (cnt < 64) mask &= (1ull << cnt) - 1ull;
    const uint32_t first_slot = c.tile_slot[t];
    const bool valid = (uint32_t)lane < cnt;
    float num = 0.f, den = 0.f;
    uint32_t myslot = first_slot;
    if (valid) {
      const uint64_t p = base + lane;
      myslot = tile_lane_slot(first_slot, mask, lane);
      float4 q[N];
#pragma unroll
      for (int s = 0; s < N - 1; ++s) q[1 + s] = s == LS ? xs[c.oidx[s][p]] : a.x[1 + s][c.oidx[s][p]];
      float e = a.resid[p];
      if (fix) {
        q[0] = a.x[0][c.directory[myslot]];  // the own mode's row: pending and saved values
        if (a.sub_col >= 0) {
          float v = q[0].x;
#pragma unroll
          for (int s = 1; s < N; ++s) v *= q[s].x;
          e -= v;
        }
        if (a.add_back) {
          float v = q[0].y;
#pragma unroll
          for (int s = 1; s < N; ++s) v *= q[s].y;
          e += v;
        }
        a.resid[p] = e;
      }
      float h = q[1].z;
#pragma unroll
      for (int s = 2; s < N; ++s) h *= q[s].z;
      num = e * h;
      den = h * h;
    }
    const bool head = valid && ((mask >> lane) & 1ull);
    seg_scan(num, den, lane == 0 || head);

    // first segment: lanes [0, fh); fh = first row start after lane 0
    const uint64_t later = mask & ~1ull;
    const uint32_t fh = later ? (uint32_t)__ffsll((unsigned long long)later) - 1u : cnt;
    const bool join = (t > t0) && !(mask & 1ull);  // lane 0 continues the carried row
    if (have_carry && !join) {
      if (lane == 0) {
        a.pieces[(size_t)(carry_slot + poff) * 2] = carry_n;
        a.pieces[(size_t)(carry_slot + poff) * 2 + 1] = carry_d;
      }
      have_carry = false;
      carry_n = carry_d = 0.0;
    }
    const double fn = (double)__shfl(num, (int)fh - 1, 64) + (join ? carry_n : 0.0);
    const double fd = (double)__shfl(den, (int)fh - 1, 64) + (join ? carry_d : 0.0);
    if (fh < cnt) {
      if (lane == 0) {
        a.pieces[(size_t)(first_slot + poff) * 2] = fn;
        a.pieces[(size_t)(first_slot + poff) * 2 + 1] = fd;
      }
      // complete middle segments end at lanes i in [fh, cnt-1) whose next lane
      // starts a row; the last segment (ending at cnt-1) becomes the carry
      const bool tail = valid && (uint32_t)lane >= fh && (uint32_t)lane + 1 < cnt &&
                        ((mask >> (lane + 1)) & 1ull);
      if (tail) {
        a.pieces[(size_t)(myslot + poff) * 2] = (double)num;
        a.pieces[(size_t)(myslot + poff) * 2 + 1] = (double)den;
      }
      carry_n = (double)__shfl(num, (int)cnt - 1, 64);
      carry_d = (double)__shfl(den, (int)cnt - 1, 64);
      carry_slot = (uint32_t)__shfl((int)myslot, (int)cnt - 1, 64);
    } else {
      carry_n = fn;
      carry_d = fd;
      carry_slot = first_slot;
    }
    have_carry = true;
  }
  if (have_carry && lane == 0) {
    a.pieces[(size_t)(carry_slot + poff) * 2] = carry_n;
    a.pieces[(size_t)(carry_slot + poff) * 2 + 1] = carry_d;
  }
}

template <int N>
__global__ void __launch_bounds__(256) k_ccd_col(CcdColArgs a) {
  const uint64_t w = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (w >= a.nspans) return;
  ccd_col_wave<N, -1>(a, w, nullptr);
}

// 1024-thread workgroups (16 spans) sharing one LDS copy of the records of a
// small other mode: its gathers no longer go through the texture addresser.
template <int N, int LS>
__global__ void __launch_bounds__(1024) k_ccd_col_lds(CcdColArgs a) {
  __shared__ float4 xs[kCcdLdsRows];
  const float4 *src = a.x[1 + LS];
  for (uint32_t i = threadIdx.x; i < a.lds_rows; i += blockDim.x) xs[i] = src[i];
  __syncthreads();
  const uint64_t w = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (w >= a.nspans) return;
  ccd_col_wave<N, LS>(a, w, xs);
}

// a_f(row) = num / (lambda + den) per directory slot (SG/ccd.cu:345 form),
// written to the factor mirror and to the row's record (the current value
// later modes' passes gather).  Each slot's pieces are summed in a fixed order
// (deterministic).
__device__ inline void ccd_finish_slot(uint64_t slot, double n, double d, const uint32_t *directory, double lam,
                                       float *col, float4 *x) {
  const uint32_t row = directory[slot];
  const float v = (float)(n / (lam + d));
  col[row] = v;
  x[row].z = v;
}

// one wave per slot: lane l sums pieces l, l + 64, ... into four partial sums
// (piece l + 64 i into partial i mod 4, so four loads are in flight), adds
// them in order, and a fixed shuffle tree adds the lanes (parallel over the
// thousands of pieces of a long row)
__device__ inline void ccd_finish_wave(uint64_t slot, const double *pieces, const uint32_t *slot_pb,
                                       const uint32_t *directory, double lam, float *col, float4 *x) {
  const int lane = lane_id();
  const double2 *pc = reinterpret_cast<const double2 *>(pieces);
  const uint32_t p1 = slot_pb[slot + 1];
  double n4[4] = {0.0, 0.0, 0.0, 0.0}, d4[4] = {0.0, 0.0, 0.0, 0.0};
  uint32_t p = slot_pb[slot] + lane;
  for (; p + 3 * 64 < p1; p += 4 * 64) {
    double2 q[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) q[u] = pc[p + u * 64];
#pragma unroll
    for (int u = 0; u < 4; ++u) n4[u] += q[u].x, d4[u] += q[u].y;
  }
  for (int u = 0; p < p1; p += 64, ++u) n4[u] += pc[p].x, d4[u] += pc[p].y;
  double n = (n4[0] + n4[1]) + (n4[2] + n4[3]), d = (d4[0] + d4[1]) + (d4[2] + d4[3]);
  for (int o = 32; o > 0; o >>= 1) {
    n += __shfl_xor(n, o, 64);
    d += __shfl_xor(d, o, 64);
  }
  if (lane == 0) ccd_finish_slot(slot, n, d, directory, lam, col, x);
}

__global__ void __launch_bounds__(256) k_ccd_finish(const double *pieces, const uint32_t *slot_pb,
                                                    const uint32_t *directory, uint64_t dlength,
                                                    double lam, float *col, float4 *x) {
  const uint64_t s = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (s < dlength) ccd_finish_wave(s, pieces, slot_pb, directory, lam, col, x);
}

// For modes whose slots hold about one piece (mode 0 of the benchmark tensor:
// 480K rows, each meeting a span boundary rarely): a block takes 256 slots;
// a slot of at most kFinishLight pieces is summed by its own lane in piece
// order, the heavier ones are listed in LDS and summed by a wave each as
// above.  Measured per pass on the benchmark tensor: mode 0 35.5 us (129 us
// with a wave for every slot, 119 us with 256 consecutive slots per block,
// where the hottest rows, numbered first, queued in the first blocks); the
// other modes, whose slots hold tens to thousands of pieces, 9 us with a wave
// per slot and longer with lanes.
constexpr uint32_t kFinishLight = 8;

__global__ void __launch_bounds__(256) k_ccd_finish_light(const double *pieces, const uint32_t *slot_pb,
                                                          const uint32_t *directory, uint64_t dlength,
                                                          double lam, float *col, float4 *x) {
  __shared__ uint32_t heavy[256];
  __shared__ uint32_t nheavy;
  // slots are dealt to the blocks in a stride (block b takes b, b + G, ...):
  // the long rows are the hottest, numbered first, and land in different
  // blocks instead of queueing in the first ones
  const uint64_t s = (uint64_t)blockIdx.x + (uint64_t)threadIdx.x * gridDim.x;
  if (threadIdx.x == 0) nheavy = 0;
  __syncthreads();
  if (s < dlength) {
    const uint32_t p0 = slot_pb[s], p1 = slot_pb[s + 1];
    if (p1 - p0 <= kFinishLight) {
      double n = 0.0, d = 0.0;
      for (uint32_t p = p0; p < p1; ++p) {
        const double2 q = reinterpret_cast<const double2 *>(pieces)[p];
        n += q.x;
        d += q.y;
      }
      ccd_finish_slot(s, n, d, directory, lam, col, x);
    } else {
      heavy[atomicAdd(&nheavy, 1u)] = threadIdx.x;
    }
  }
  __syncthreads();
  for (uint32_t h = threadIdx.x >> 6; h < nheavy; h += blockDim.x >> 6)
    ccd_finish_wave((uint64_t)blockIdx.x + (uint64_t)heavy[h] * gridDim.x, pieces, slot_pb, directory, lam, col, x);
}

// per-row records of column `col` at the start of its update
__global__ void __launch_bounds__(256) k_ccd_pack(const float *ft, uint64_t rows, int col, int sub_col, float4 *x) {
  const float *cur = ft + (size_t)col * rows;
  const float *sub = sub_col >= 0 ? ft + (size_t)sub_col * rows : nullptr;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < rows; i += (uint64_t)gridDim.x * blockDim.x) {
    const float v = cur[i];
    x[i] = make_float4(sub ? sub[i] : 0.f, v, v, 0.f);
  }
}

// residual copy of a mode: r = x - <rows>, row-major factors in CISS-relative
// order, the products and their sum in FP64 (the residual is updated from
// here on for a whole solve).  Walkers of R / 4 lanes with 4 entries of a row
// per lane (one 16-byte load per row), walker g of G taking nonzeros g, g + G,
// ...  (One lane per nonzero with R scalar loads per row took 15.1 ms per mode
// on the 100M-nonzero tensor; this shape 3.9 ms; four nonzeros per walker at a
// time 5.4 ms.)
template <int R, int N>
__global__ void __launch_bounds__(256) k_ccd_resid(CissArgs c, FactorArgs fa, float *resid) {
  constexpr int CPL = 4, W = R / CPL, WPW = 64 / W;
  const int lane = lane_id();
  const int sub = lane / W, wl = lane % W;
  const uint64_t G = (uint64_t)gridDim.x * (blockDim.x >> 6) * WPW;
  const uint64_t g = ((uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * WPW + sub;
  for (uint64_t p = g; p < c.nnz; p += G) {  // uniform over the walker's lanes
    const uint64_t t = p >> 6;
    const uint32_t slot = tile_lane_slot(c.tile_slot[t], c.tile_mask[t], (int)(p & 63));
    const float4 a = *reinterpret_cast<const float4 *>(fa.f[0] + (size_t)c.directory[slot] * R + wl * CPL);
    double v[CPL] = {a.x, a.y, a.z, a.w};
#pragma unroll
    for (int s = 0; s < N - 1; ++s) {
      const float4 o = *reinterpret_cast<const float4 *>(fa.f[1 + s] + (size_t)c.oidx[s][p] * R + wl * CPL);
      v[0] *= o.x, v[1] *= o.y, v[2] *= o.z, v[3] *= o.w;
    }
    double pr = (v[0] + v[1]) + (v[2] + v[3]);
#pragma unroll
    for (int o = 1; o < W; o <<= 1) pr += __shfl_xor(pr, o, 64);  // within the walker's lanes
    if (wl == 0) resid[p] = (float)((double)c.vals[p] - pr);
  }
}

// The reference's CCD update of a mode (SG/ccd.cu:318-348): every column of
// each directory row from the sums of the Jacobi walk, in place; rows outside
// the directory keep their values.
__global__ void __launch_bounds__(256) k_ccd_jacobi_rows(const uint32_t *directory, uint64_t dlength, int R,
                                                         const float *num, const float *den, double lam, float *F) {
  const uint64_t n = dlength * (uint64_t)R;
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (uint64_t)gridDim.x * blockDim.x) {
    const size_t e = (size_t)directory[p / R] * R + p % R;
    F[e] = (float)((double)num[e] / (lam + (double)den[e]));
  }
}

// out (R x rows, column-major) <-> in (rows x R, row-major)
__global__ void __launch_bounds__(256) k_transpose(const float *in, float *out, uint64_t rows, int R,
                                                   int to_colmajor) {
  const uint64_t n = rows * (uint64_t)R;
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (uint64_t)gridDim.x * blockDim.x) {
    if (to_colmajor) {
      const uint64_t f = p / rows, i = p % rows;  // write-coalesced
      out[p] = in[i * R + f];
    } else {
      const uint64_t i = p / R, f = p % R;
      out[p] = in[f * rows + i];
    }
  }
}

hipError_t launch_ccd_col(hipStream_t st, const CcdColArgs &a, int nmodes, unsigned block) {
  if (a.c.nnz == 0) return hipSuccess;
  if (a.lds_s >= 0 && a.lds_rows <= kCcdLdsRows) {
    const unsigned g = (unsigned)((a.nspans + 15) / 16);
    if (nmodes == 3 && a.lds_s == 0)
      hipLaunchKernelGGL((k_ccd_col_lds<3, 0>), dim3(g), dim3(1024), 0, st, a);
    else if (nmodes == 3 && a.lds_s == 1)
      hipLaunchKernelGGL((k_ccd_col_lds<3, 1>), dim3(g), dim3(1024), 0, st, a);
    else
      return hipErrorInvalidValue;
    return hipGetLastError();
  }
  const uint64_t wpb = block / 64;
  const unsigned grid = (unsigned)((a.nspans + wpb - 1) / wpb);
  if (nmodes == 3)
    hipLaunchKernelGGL((k_ccd_col<3>), dim3(grid), dim3(block), 0, st, a);
  else if (nmodes == 4)
    hipLaunchKernelGGL((k_ccd_col<4>), dim3(grid), dim3(block), 0, st, a);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_ccd_finish(hipStream_t st, const double *pieces, const uint32_t *slot_pb,
                             const uint32_t *directory, uint64_t dlength, uint64_t npieces, double lam,
                             float *col, float4 *x) {
  if (dlength == 0) return hipSuccess;
  if (npieces <= 2 * dlength)  // about one piece per slot
    hipLaunchKernelGGL(k_ccd_finish_light, dim3((unsigned)((dlength + 255) / 256)), dim3(256), 0, st, pieces,
                       slot_pb, directory, dlength, lam, col, x);
  else
    hipLaunchKernelGGL(k_ccd_finish, dim3((unsigned)((dlength + 3) / 4)), dim3(256), 0, st, pieces, slot_pb,
                       directory, dlength, lam, col, x);
  return hipGetLastError();
}

hipError_t launch_ccd_pack(hipStream_t st, const float *ft, uint64_t rows, int col, int sub_col, float4 *x) {
  if (rows == 0) return hipSuccess;
  uint64_t blocks = (rows + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_ccd_pack, dim3((unsigned)blocks), dim3(256), 0, st, ft, rows, col, sub_col, x);
  return hipGetLastError();
}

hipError_t launch_ccd_resid(hipStream_t st, const CissArgs &c, const FactorArgs &f, int nmodes, int R,
                            float *resid) {
  if (c.nnz == 0) return hipSuccess;
  const unsigned blocks = 4096;  // 4 waves each: 16 per CU
  switch (R * 8 + nmodes) {
#define CTC_RESID(RR, NN) \
  case RR * 8 + NN: hipLaunchKernelGGL((k_ccd_resid<RR, NN>), dim3(blocks), dim3(256), 0, st, c, f, resid); break;
    CTC_RESID(8, 3) CTC_RESID(16, 3) CTC_RESID(32, 3) CTC_RESID(64, 3) CTC_RESID(128, 3)
    CTC_RESID(8, 4) CTC_RESID(16, 4) CTC_RESID(32, 4) CTC_RESID(64, 4) CTC_RESID(128, 4)
#undef CTC_RESID
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_ccd_jacobi_rows(hipStream_t st, const uint32_t *directory, uint64_t dlength, int R,
                                  const float *num, const float *den, double lam, float *F) {
  const uint64_t n = dlength * (uint64_t)R;
  if (n == 0) return hipSuccess;
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_ccd_jacobi_rows, dim3((unsigned)blocks), dim3(256), 0, st, directory, dlength, R, num, den,
                     lam, F);
  return hipGetLastError();
}

hipError_t launch_transpose(hipStream_t st, const float *in, float *out, uint64_t rows, int R,
                            bool to_colmajor) {
  const uint64_t n = rows * (uint64_t)R;
  if (n == 0) return hipSuccess;
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_transpose, dim3((unsigned)blocks), dim3(256), 0, st, in, out, rows, R,
                     to_colmajor ? 1 : 0);
  return hipGetLastError();
}

}  // namespace ctc
