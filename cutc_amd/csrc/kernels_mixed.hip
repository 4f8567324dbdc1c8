// kernels_mixed.hip — BF16-operand ALS mode update for ranks 32, 64 and 128.
//
// The same rule as the FP32 path (SG/als.cu:186-317 Gram + MTTKRP, then
// SG/als.cu:467-491,715-716 (G + lambda I) a = m per row), in the precision
// split of the mixed path: h = product of the BF16 factor rows (exact in
// FP32), G = sum bf16(h) bf16(h)^T on BF16 MFMA with FP32 accumulation,
// m = sum x bf16(h) in FP32, and the R x R system factored and solved in FP64
// with iterative refinement against the FP32 Gram.
//
// One workgroup of 2R threads (R/32 waves) owns one row.  The row-major
// per-row Gram never touches HBM: wave w accumulates the 32-row stripe w of
// G (lower 32x32 blocks only) in MFMA accumulators, while the workgroup
// stages 32 nonzeros' h vectors at a time in LDS as BF16.  The system then
// moves into LDS as FP64 16x16 tiles (lower triangle, R/16 tile rows) and is
// factored by a blocked right-looking Cholesky: a wave factors and inverts
// the diagonal tile in registers, the panel below is multiplied by that
// inverse and the trailing tiles are updated, both on FP64 MFMA.  The forward
// substitution rides along the panel steps; the back substitution and the
// refinement solves walk the tile rows with the inverted diagonal tiles.
// Rows too long for one workgroup are cut into chunks whose partial Gram and
// right-hand side go to HBM and are summed by a second pass before the solve.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "kernels_common.h"
#include "kernels_mixed.h"

namespace ctc {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef double f64x4 __attribute__((ext_vector_type(4)));


// v_mfma_f32_32x32x16_bf16 accumulator v of lane l: row 8(v/4) + 4(l/32) + v%4, column l%32
__device__ inline int row32(int v, int l) { return 8 * (v >> 2) + 4 * (l >> 5) + (v & 3); }
// v_mfma_f64_16x16x4f64 accumulator v of lane l: row 4v + l/16, column l%16
__device__ inline int row16(int v, int l) { return 4 * v + (l >> 4); }

// 16x16 FP64 tile, element (r, c) with an XOR swizzle of the column
__device__ inline int tix(int r, int c) { return (r << 4) | (c ^ r); }
__device__ inline int tileno(int I, int J) { return I * (I + 1) / 2 + J; }

constexpr int kHTS = 20;  // dwords per staged feature row: 32 BF16 nonzeros + pad

template <int R>
struct MixedSmem {
  static constexpr int NW = R / 32, RT = R / 16, NTILE = RT * (RT + 1) / 2;
  static constexpr int TPW = (NTILE + NW - 1) / NW;  // FP64 path: lower tiles per wave
  static constexpr int HS = R + 2;                   // FP64 path: staged row stride (doubles)
  static constexpr int NT64 = NTILE * 256 > 32 * HS ? NTILE * 256 : 32 * HS;
  double tiles[NT64];  // FP64 lower tiles; the Gram phase stages h here
  double vm[R], vx[R], vy[R];
  double contrib[NW][R];
  int bad;
};

// lower tile u (row-major over I >= J) -> (I, J)
__device__ inline void tile_ij(int u, int &I, int &J) {
  I = 0;
  while ((I + 1) * (I + 2) / 2 <= u) ++I;
  J = u - I * (I + 1) / 2;
}

// ------------------------------------------------------------ Gram phase --
// acc[tt] (tt <= w) += stripe w x stripe tt of sum bf16(h) bf16(h)^T over
// nonzeros [p0, p1); returns this thread's FP32 partial of m for features
// 8 fg .. 8 fg + 7 (fg = thread % (R/8)) in mp.
template <int R, int N>
__device__ inline void gram_chunk(const MixedArgs &a, uint32_t p0, uint32_t p1, f32x16 *acc, float *mp,
                                  uint32_t *ht) {
  constexpr int NW = R / 32, FG8 = R / 8;
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
  const int fg = tid % FG8, qp = tid / FG8;  // features 8fg..8fg+7, nonzeros 2qp, 2qp+1
  // software pipeline: ids two steps ahead, factor rows one step ahead, so
  // the gathers overlap the staging and MFMA work of the current step
  uint32_t id_n[2][N - 1];
  float x_n[2], x_c[2];
  uint4 u_n[2][N - 1], u_c[2][N - 1];
  auto load_ids = [&](uint32_t base) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const uint32_t p = base + 2 * qp + h;
      const bool ok = p < p1;
#pragma unroll
      for (int s = 0; s < N - 1; ++s) id_n[h][s] = ok ? a.oidx[s][p] : 0u;
      x_n[h] = ok ? a.vals[p] : 0.f;
    }
  };
  auto load_rows = [&]() {
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int s = 0; s < N - 1; ++s) u_n[h][s] = *(const uint4 *)(a.fb[s] + (size_t)id_n[h][s] * R + 8 * fg);
  };
  load_ids(p0);
  load_rows();
  float x_r[2] = {x_n[0], x_n[1]};
  load_ids(p0 + 32);
  for (uint32_t base = p0; base < p1; base += 32) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      x_c[h] = x_r[h];
      x_r[h] = x_n[h];
#pragma unroll
      for (int s = 0; s < N - 1; ++s) u_c[h][s] = u_n[h][s];
    }
    load_rows();           // rows of step base + 32
    load_ids(base + 64);   // ids of step base + 64
    uint32_t hv[8][2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if (base + 2 * qp + h < p1) {
        float hf[8];
#pragma unroll
        for (int f = 0; f < 8; ++f) hf[f] = 1.f;
#pragma unroll
        for (int s = 0; s < N - 1; ++s) {
          const uint4 u = u_c[h][s];
          hf[0] *= bf_lo(u.x);
          hf[1] *= bf_hi(u.x);
          hf[2] *= bf_lo(u.y);
          hf[3] *= bf_hi(u.y);
          hf[4] *= bf_lo(u.z);
          hf[5] *= bf_hi(u.z);
          hf[6] *= bf_lo(u.w);
          hf[7] *= bf_hi(u.w);
        }
#pragma unroll
        for (int f = 0; f < 8; ++f) {
          hv[f][h] = f2bf(hf[f]);
          mp[f] += x_c[h] * __uint_as_float(hv[f][h] << 16);
        }
      } else {
#pragma unroll
        for (int f = 0; f < 8; ++f) hv[f][h] = 0u;
      }
    }
#pragma unroll
    for (int f = 0; f < 8; ++f) {
      const int r = 8 * fg + f;
      ht[r * kHTS + (qp ^ (4 * ((r >> 3) & 3)))] = hv[f][0] | (hv[f][1] << 16);
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int ra = 32 * w + (l & 31), ch = 2 * s + (l >> 5);
      const uint4 ua = *(const uint4 *)(ht + ra * kHTS + 4 * (ch ^ ((ra >> 3) & 3)));
      const bf16x8 av = __builtin_bit_cast(bf16x8, ua);
#pragma unroll
      for (int tt = 0; tt < NW; ++tt) {
        if (tt > w) continue;  // wave-uniform: lower stripes only
        const int rb = 32 * tt + (l & 31);
        const uint4 ub = *(const uint4 *)(ht + rb * kHTS + 4 * (ch ^ ((rb >> 3) & 3)));
        acc[tt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, __builtin_bit_cast(bf16x8, ub), acc[tt], 0, 0, 0);
      }
    }
    __syncthreads();
  }
}

// FP64 variant: h = product of the FP32 rows formed in FP64, staged in LDS
// as FP64 (32 nonzeros x R), and every lower 16x16 tile of G accumulated on
// v_mfma_f64_16x16x4f64, tiles dealt round-robin to the waves (wave w owns
// tiles w, w + NW, ...); m accumulated in FP64.
template <int R, int N>
__device__ inline void gram_f64(const MixedArgs &a, uint32_t p0, uint32_t p1, f64x4 *acc, double *mp,
                                double *hs) {
  constexpr int NW = R / 32, FG8 = R / 8, HS = MixedSmem<R>::HS;
  constexpr int NTILE = MixedSmem<R>::NTILE, TPW = MixedSmem<R>::TPW;
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
  const int fg = tid % FG8, qp = tid / FG8;
  // software pipeline as in gram_chunk: ids two steps ahead, rows one ahead
  uint32_t id_n[2][N - 1];
  float x_n[2], x_r[2];
  float4 u_n[2][N - 1][2], u_c[2][N - 1][2];
  auto load_ids = [&](uint32_t base) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const uint32_t p = base + 2 * qp + h;
      const bool ok = p < p1;
#pragma unroll
      for (int s = 0; s < N - 1; ++s) id_n[h][s] = ok ? a.oidx[s][p] : 0u;
      x_n[h] = ok ? a.vals[p] : 0.f;
    }
  };
  auto load_rows = [&]() {
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int s = 0; s < N - 1; ++s) {
        const float *row = a.fr[s] + (size_t)id_n[h][s] * R + 8 * fg;
        u_n[h][s][0] = *(const float4 *)row;
        u_n[h][s][1] = *(const float4 *)(row + 4);
      }
  };
  load_ids(p0);
  load_rows();
  x_r[0] = x_n[0];
  x_r[1] = x_n[1];
  load_ids(p0 + 32);
  for (uint32_t base = p0; base < p1; base += 32) {
    float x_c[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      x_c[h] = x_r[h];
      x_r[h] = x_n[h];
#pragma unroll
      for (int s = 0; s < N - 1; ++s) u_c[h][s][0] = u_n[h][s][0], u_c[h][s][1] = u_n[h][s][1];
    }
    load_rows();           // rows of step base + 32
    load_ids(base + 64);   // ids of step base + 64
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const uint32_t p = base + 2 * qp + h;
      double hd[8];
      if (p < p1) {
#pragma unroll
        for (int f = 0; f < 8; ++f) hd[f] = 1.0;
#pragma unroll
        for (int s = 0; s < N - 1; ++s) {
          const float4 u0 = u_c[h][s][0], u1 = u_c[h][s][1];
          hd[0] *= u0.x;
          hd[1] *= u0.y;
          hd[2] *= u0.z;
          hd[3] *= u0.w;
          hd[4] *= u1.x;
          hd[5] *= u1.y;
          hd[6] *= u1.z;
          hd[7] *= u1.w;
        }
        const double x = x_c[h];
#pragma unroll
        for (int f = 0; f < 8; ++f) mp[f] += x * hd[f];
      } else {
#pragma unroll
        for (int f = 0; f < 8; ++f) hd[f] = 0.0;
      }
      double *dst = hs + (2 * qp + h) * HS + 8 * fg;
#pragma unroll
      for (int f = 0; f < 8; f += 2) *(double2 *)(dst + f) = make_double2(hd[f], hd[f + 1]);
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const double *hq = hs + (4 * s + (l >> 4)) * HS + (l & 15);
#pragma unroll
      for (int j = 0; j < TPW; ++j) {
        const int u = w + NW * j;
        if (u >= NTILE) continue;  // wave-uniform
        int I, J;
        tile_ij(u, I, J);
        acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(hq[16 * I], hq[16 * J], acc[j], 0, 0, 0);
      }
    }
    __syncthreads();
  }
}

// sums the per-thread m partials of the workgroup into sm.vm (FP64); ends
// with a barrier
template <int R, class T, class SM>
__device__ inline void reduce_m(T *mp, SM &sm) {
  constexpr int NW = R / 32, FG8 = R / 8;
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, fg = tid % FG8;
#pragma unroll
  for (int o = FG8; o < 64; o <<= 1)
#pragma unroll
    for (int f = 0; f < 8; ++f) mp[f] += __shfl_xor(mp[f], o, 64);
  if (l < FG8)
#pragma unroll
    for (int f = 0; f < 8; ++f) sm.contrib[w][8 * fg + f] = (double)mp[f];
  __syncthreads();
  if (tid < R) {
    double s = 0.0;
#pragma unroll
    for (int ww = 0; ww < NW; ++ww) s += sm.contrib[ww][tid];
    sm.vm[tid] = s;
  }
  __syncthreads();
}

// ---------------------------------------------------------- solve phase --
// wave-level: factor tile (K, K) (Cholesky), replace it by the inverse of its
// lower factor; with fold, also y_K = inverse * vy_K in place.
// The factorisation runs without square roots (after step k, a_ij for i > j
// holds l_ij d_j and a_jj the pivot d_j of A = L D L^T) on the tile held in
// registers, lane (q, c) owning entries (4q .. 4q+3, c).  At step k the four
// lanes owning column k publish it to a 16-double buffer in the tile's own LDS
// slot (free while the tile is in registers; two buffers alternate, so one
// wave barrier per step suffices) and every lane updates its entries from it:
// each of the 15 dependent steps is one LDS round trip and one reciprocal.
// The same elimination builds M = L^-1 in the otherwise unused upper triangle,
// transposed (entry (j, r), j < r, holds m_rj): step k subtracts l_rk times
// row k of M (final by then, and published with column k) from every row
// r > k, as Gauss-Jordan does on an identity.  The inverse of the Cholesky
// factor C = L D^1/2 is then D^-1/2 M, one scaling pass.
template <int R>
__device__ __attribute__((always_inline)) inline void diag_step(MixedSmem<R> &sm, int K, bool fold) {
  double *T = sm.tiles + tileno(K, K) * 256;
  const int l = threadIdx.x & 63, c = l & 15, q = l >> 4;
  double a[4];  // entries (4q + e, c): lower part of A, upper part m_c,4q+e (starts at 0)
#pragma unroll
  for (int e = 0; e < 4; ++e) a[e] = 4 * q + e >= c ? T[tix(4 * q + e, c)] : 0.0;
  wave_sync();  // the column buffers overwrite the tile's first 32 entries
  bool ok = true;
  // kept rolled: unrolled, each step's constant addresses stay live and the
  // enclosing kernels lose a wave per SIMD to the extra registers
#pragma unroll 1
  for (int k = 0; k < 15; ++k) {
    double *B = T + 16 * (k & 1);
    if (c == k)
#pragma unroll
      for (int e = 0; e < 4; ++e) B[4 * q + e] = a[e];
    wave_sync();
    const double akk = B[k];
    ok = ok && (akk > 0.0);
    const double rk = rcp_f64(akk > 0.0 ? akk : 1.0);
    const double ack = B[c] * rk;  // l_ck (used for c > k)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int i = 4 * q + e;
      // column k: a_ik below the pivot, m_ki (row k of M, final) above it
      // lower: a_ic -= a_ik l_ck; upper: m_ci -= l_ck m_ki (m_kk = 1)
      const double coef = i == k ? 1.0 : B[i];
      if (c > k && (i >= c || i <= k)) a[e] = __builtin_fma(-coef, ack, a[e]);
    }
  }
  wave_sync();  // the last step's reads of the buffers are done
#pragma unroll
  for (int e = 0; e < 4; ++e) T[tix(4 * q + e, c)] = a[e];
  wave_sync();
  ok = ok && (T[tix(15, 15)] > 0.0);
  // X = C^-1 = D^-1/2 M: x_ic = m_ic / sqrt(d_i) below the diagonal, 1 / sqrt(d_i) on it
  double x[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int i = 4 * q + e;
    const double di = T[tix(i, i)];
    const double m = i > c ? T[tix(c, i)] : (i == c ? 1.0 : 0.0);
    x[e] = m * rcp_f64(sqrt(di > 0.0 ? di : 1.0));
  }
  wave_sync();
#pragma unroll
  for (int e = 0; e < 4; ++e) T[tix(4 * q + e, c)] = x[e];
  if (!ok && l == 0) sm.bad = 1;
  wave_sync();
  if (fold) {
    // y_K = X b_K: lane i < 16 takes row i of X
    double y = 0.0;
    if (l < 16) {
#pragma unroll
      for (int j = 0; j < 16; ++j) y = __builtin_fma(T[tix(l, j)], sm.vy[16 * K + j], y);
    }
    wave_sync();
    if (l < 16) sm.vy[16 * K + l] = y;
  }
  wave_sync();
}

// wave-level: L_IK = A_IK X_KK^T (panel), then vy_I -= L_IK vy_K
template <int R>
__device__ inline void panel_step(MixedSmem<R> &sm, int I, int K) {
  double *Tik = sm.tiles + tileno(I, K) * 256;
  const double *X = sm.tiles + tileno(K, K) * 256;
  const int l = threadIdx.x & 63;
  f64x4 d = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int k0 = 0; k0 < 16; k0 += 4) {
    const double av = Tik[tix(l & 15, k0 + (l >> 4))];
    const double bv = X[tix(l & 15, k0 + (l >> 4))];
    d = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, d, 0, 0, 0);
  }
  wave_sync();
  double pv[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    Tik[tix(row16(v, l), l & 15)] = d[v];
    pv[v] = d[v] * sm.vy[16 * K + (l & 15)];
  }
#pragma unroll
  for (int o = 1; o < 16; o <<= 1)
#pragma unroll
    for (int v = 0; v < 4; ++v) pv[v] += __shfl_xor(pv[v], o, 64);
  if ((l & 15) == 0)
#pragma unroll
    for (int v = 0; v < 4; ++v) sm.vy[16 * I + row16(v, l)] -= pv[v];
  wave_sync();
}

// wave-level: A_IJ -= L_IK L_JK^T
template <int R>
__device__ inline void trail_step(MixedSmem<R> &sm, int I, int J, int K) {
  double *Tij = sm.tiles + tileno(I, J) * 256;
  const double *Tik = sm.tiles + tileno(I, K) * 256;
  const double *Tjk = sm.tiles + tileno(J, K) * 256;
  const int l = threadIdx.x & 63;
  f64x4 d;
#pragma unroll
  for (int v = 0; v < 4; ++v) d[v] = Tij[tix(row16(v, l), l & 15)];
#pragma unroll
  for (int k0 = 0; k0 < 16; k0 += 4) {
    const double av = -Tik[tix(l & 15, k0 + (l >> 4))];
    const double bv = Tjk[tix(l & 15, k0 + (l >> 4))];
    d = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, d, 0, 0, 0);
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) Tij[tix(row16(v, l), l & 15)] = d[v];
}

// wave-level forward pieces: vy_I -= L_IK vy_K ; vy_K = X_KK vy_K
template <int R>
__device__ inline void fwd_update(MixedSmem<R> &sm, int I, int K) {
  const double *T = sm.tiles + tileno(I, K) * 256;
  const int l = threadIdx.x & 63, c = l & 15, q = l >> 4;
  double pv[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) pv[e] = T[tix(4 * q + e, c)] * sm.vy[16 * K + c];
#pragma unroll
  for (int o = 1; o < 16; o <<= 1)
#pragma unroll
    for (int e = 0; e < 4; ++e) pv[e] += __shfl_xor(pv[e], o, 64);
  if (c == 0)
#pragma unroll
    for (int e = 0; e < 4; ++e) sm.vy[16 * I + 4 * q + e] -= pv[e];
  wave_sync();
}

template <int R>
__device__ inline void fwd_diag(MixedSmem<R> &sm, int K) {
  const double *X = sm.tiles + tileno(K, K) * 256;
  const int l = threadIdx.x & 63, c = l & 15, q = l >> 4;
  double pv[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) pv[e] = X[tix(4 * q + e, c)] * sm.vy[16 * K + c];
#pragma unroll
  for (int o = 1; o < 16; o <<= 1)
#pragma unroll
    for (int e = 0; e < 4; ++e) pv[e] += __shfl_xor(pv[e], o, 64);
  wave_sync();
  if (c == 0)
#pragma unroll
    for (int e = 0; e < 4; ++e) sm.vy[16 * K + 4 * q + e] = pv[e];
  wave_sync();
}

// wave-level backward pieces: vy_J -= L_KJ^T dst_K ; dst_K = X_KK^T vy_K
template <int R>
__device__ inline void back_update(MixedSmem<R> &sm, const double *dst, int K, int J) {
  const double *T = sm.tiles + tileno(K, J) * 256;
  const int l = threadIdx.x & 63, c = l & 15, q = l >> 4;
  double s = 0.0;
#pragma unroll
  for (int e = 0; e < 4; ++e) s += T[tix(4 * q + e, c)] * dst[16 * K + 4 * q + e];
  s += __shfl_xor(s, 16, 64);
  s += __shfl_xor(s, 32, 64);
  if (q == 0) sm.vy[16 * J + c] -= s;
  wave_sync();
}

template <int R>
__device__ inline void back_diag(MixedSmem<R> &sm, double *dst, int K) {
  const double *X = sm.tiles + tileno(K, K) * 256;
  const int l = threadIdx.x & 63, c = l & 15, q = l >> 4;
  double s = 0.0;
#pragma unroll
  for (int e = 0; e < 4; ++e) s += X[tix(4 * q + e, c)] * sm.vy[16 * K + 4 * q + e];
  s += __shfl_xor(s, 16, 64);
  s += __shfl_xor(s, 32, 64);
  if (q == 0) dst[16 * K + c] = s;
  wave_sync();
}

// dst = L^-T vy (vy is consumed)
template <int R>
__device__ inline void back_solve(MixedSmem<R> &sm, double *dst) {
  constexpr int NW = R / 32, RT = R / 16;
  const int w = threadIdx.x >> 6;
  if (w == (RT - 1) % NW) back_diag(sm, dst, RT - 1);
  __syncthreads();
  for (int K = RT - 1; K >= 1; --K) {
    for (int J = 0; J < K; ++J)
      if (w == J % NW) {
        back_update(sm, dst, K, J);
        if (J == K - 1) back_diag(sm, dst, K - 1);
      }
    __syncthreads();
  }
}

// vy = L^-1 vy
template <int R>
__device__ inline void fwd_solve(MixedSmem<R> &sm) {
  constexpr int NW = R / 32, RT = R / 16;
  const int w = threadIdx.x >> 6;
  if (w == 0) fwd_diag(sm, 0);
  __syncthreads();
  for (int K = 0; K < RT - 1; ++K) {
    for (int I = K + 1; I < RT; ++I)
      if (w == I % NW) {
        fwd_update(sm, I, K);
        if (I == K + 1) fwd_diag(sm, K + 1);
      }
    __syncthreads();
  }
}

// A = G + lambda I into the lower FP64 tiles, from the BF16 path's stripe
// accumulators (wave w: 32x32 blocks (w, tt), tt <= w)
template <int R>
__device__ inline void tiles_from(MixedSmem<R> &sm, const f32x16 *acc, double lam) {
  constexpr int NW = R / 32;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int tt = 0; tt < NW; ++tt) {
    if (tt > w) continue;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int i = 32 * w + row32(v, l), j = 32 * tt + (l & 31);
      if ((i >> 4) >= (j >> 4))
        sm.tiles[tileno(i >> 4, j >> 4) * 256 + tix(i & 15, j & 15)] = (double)acc[tt][v] + (i == j ? lam : 0.0);
    }
  }
}

// ... from the FP64 path's tile accumulators
template <int R>
__device__ inline void tiles_from(MixedSmem<R> &sm, const f64x4 *acc, double lam) {
  constexpr int NW = R / 32, NTILE = MixedSmem<R>::NTILE, TPW = MixedSmem<R>::TPW;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int j = 0; j < TPW; ++j) {
    const int u = w + NW * j;
    if (u >= NTILE) continue;
    int I, J;
    tile_ij(u, I, J);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = row16(v, l), c = l & 15;
      sm.tiles[u * 256 + tix(r, c)] = acc[j][v] + (I == J && r == c ? lam : 0.0);
    }
  }
}

// contrib[w][*] = this wave's share of G x (BF16 path: stripe rows of every
// lower block plus, for off-diagonal blocks, the transposed block)
template <int R>
__device__ inline void gx_share(MixedSmem<R> &sm, const f32x16 *acc) {
  constexpr int NW = R / 32;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  double rs[16];
#pragma unroll
  for (int v = 0; v < 16; ++v) rs[v] = 0.0;
  for (int k = l; k < R; k += 64) sm.contrib[w][k] = 0.0;
  wave_sync();
#pragma unroll
  for (int tt = 0; tt < NW; ++tt) {
    if (tt > w) continue;
    const double xj = sm.vx[32 * tt + (l & 31)];
    double cs = 0.0;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      rs[v] += (double)acc[tt][v] * xj;
      cs += (double)acc[tt][v] * sm.vx[32 * w + row32(v, l)];
    }
    if (tt < w) {
      cs += __shfl_xor(cs, 32, 64);
      if (l < 32) sm.contrib[w][32 * tt + l] = cs;
    }
  }
#pragma unroll
  for (int o = 1; o < 32; o <<= 1)
#pragma unroll
    for (int v = 0; v < 16; ++v) rs[v] += __shfl_xor(rs[v], o, 64);
  if ((l & 31) == 0)
#pragma unroll
    for (int v = 0; v < 16; ++v) sm.contrib[w][32 * w + row32(v, l)] = rs[v];
}

// ... FP64 path: per owned tile (I, J), rows of I and (I != J) rows of J
template <int R>
__device__ inline void gx_share(MixedSmem<R> &sm, const f64x4 *acc) {
  constexpr int NW = R / 32, NTILE = MixedSmem<R>::NTILE, TPW = MixedSmem<R>::TPW;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  for (int k = l; k < R; k += 64) sm.contrib[w][k] = 0.0;
  wave_sync();
#pragma unroll
  for (int j = 0; j < TPW; ++j) {
    const int u = w + NW * j;
    if (u >= NTILE) continue;
    int I, J;
    tile_ij(u, I, J);
    double rv[4], cs = 0.0;
    const double xc = sm.vx[16 * J + (l & 15)];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      rv[v] = acc[j][v] * xc;
      cs += acc[j][v] * sm.vx[16 * I + row16(v, l)];
    }
#pragma unroll
    for (int o = 1; o < 16; o <<= 1)
#pragma unroll
      for (int v = 0; v < 4; ++v) rv[v] += __shfl_xor(rv[v], o, 64);
    cs += __shfl_xor(cs, 16, 64);
    cs += __shfl_xor(cs, 32, 64);
    if ((l & 15) == 0)
#pragma unroll
      for (int v = 0; v < 4; ++v) sm.contrib[w][16 * I + row16(v, l)] += rv[v];
    wave_sync();
    if (I != J && l < 16) sm.contrib[w][16 * J + l] += cs;
    wave_sync();
  }
}

// (G + lambda I) x = m with G in the accumulators and m in sm.vm; returns
// false if the system is not positive definite
template <int R, class Acc>
__device__ inline bool solve_row(MixedSmem<R> &sm, const Acc *acc, double lam, int refine) {
  constexpr int NW = R / 32, RT = R / 16;
  const int tid = threadIdx.x, w = tid >> 6;
  tiles_from<R>(sm, acc, lam);
  if (tid < R) sm.vy[tid] = sm.vm[tid];
  if (tid == 0) sm.bad = 0;
  __syncthreads();
  // blocked right-looking Cholesky with one step of lookahead; vy becomes
  // y = L^-1 m on the way.  After the panels of step K, the wave that owns
  // diagonal block K+1 updates that tile and factors it at once, while the
  // other waves apply the rest of step K's trailing updates: the serial
  // diagonal factorisations overlap the trailing work instead of following it.
  if (w == 0) diag_step(sm, 0, true);
  __syncthreads();
  for (int K = 0; K < RT; ++K) {
    for (int I = K + 1; I < RT; ++I)
      if (w == I % NW) panel_step(sm, I, K);
    __syncthreads();
    if (K + 1 < RT) {
      const int dw = (K + 1) % NW;  // owner of diagonal block K+1
      if (w == dw) {
        trail_step(sm, K + 1, K + 1, K);
        diag_step(sm, K + 1, true);
      }
      // the other tiles of step K, dealt round-robin to the other waves (all
      // waves when there is only one)
      const int nother = NW > 1 ? NW - 1 : 1;
      const int me = NW > 1 ? (w - dw - 1 + NW) % NW : 0;  // rank among the other waves
      int u = 0;
      for (int I = K + 1; I < RT; ++I)
        for (int J = K + 1; J <= I; ++J) {
          if (I == K + 1 && J == K + 1) continue;
          if ((NW == 1 || w != dw) && u % nother == me) trail_step(sm, I, J, K);
          ++u;
        }
      __syncthreads();
    }
  }
  if (sm.bad) return false;
  back_solve(sm, sm.vx);
  for (int it = 0; it < refine; ++it) {
    // r = m - lambda x - G x against the unfactored accumulators
    gx_share<R>(sm, acc);
    __syncthreads();
    if (tid < R) {
      double r = sm.vm[tid] - lam * sm.vx[tid];
#pragma unroll
      for (int ww = 0; ww < NW; ++ww) r -= sm.contrib[ww][tid];
      sm.vy[tid] = r;
    }
    __syncthreads();
    fwd_solve(sm);
    back_solve(sm, sm.contrib[0]);
    if (tid < R) sm.vx[tid] += sm.contrib[0][tid];
    __syncthreads();
  }
  return true;
}

template <int R>
__device__ inline void store_row(const MixedArgs &a, MixedSmem<R> &sm, uint32_t slot, bool ok) {
  const int tid = threadIdx.x;
  float *dst = a.out + (size_t)a.directory[slot] * R;
  if (tid < R) dst[tid] = ok ? (float)sm.vx[tid] : 0.f;
  if (!ok && tid == 0) atomicOr(a.flag, 1u);
}

// ------------------------------------------------------------- thin rows --
// A row with n <= kThinRow nonzeros is solved in its dual form,
//   a = H^T (H H^T + lambda I)^-1 x,
// which equals (H^T H + lambda I)^-1 H^T x exactly (H = the row's n x R
// matrix of h vectors, x its values): an n x n FP64 system in place of the
// R x R one.  One wave per row: h staged in LDS, K = H H^T on FP64 matrix
// cores, K + lambda I eliminated in place (L D L^T, no square roots), two
// triangular solves, a = H^T z.  All in FP64 (the BF16 path stages the
// BF16-rounded h), so no refinement step is needed.
// NT = 16 serves rows of at most 16 nonzeros, NT = 32 the rest; the LDS of a
// workgroup is sized for its NT, and the BF16 path keeps h as BF16 bits, so
// the common short rows take ~6 KB of LDS at R = 128 (several waves per SIMD)
// instead of the 25 KB a 32-row FP32 stage needs.
template <int R, int P, int NT>
struct ThinSmem {
  typedef typename std::conditional<P == kPrecBF16, uint16_t, double>::type HT;
  static constexpr int HS = R + 1;  // staged h stride (odd: spreads LDS banks)
  static constexpr int KS = NT + 1;
  HT H[NT * HS];
  double K[NT * KS];
  double x[NT];
  int bad;
  __device__ static double hval(HT v) {
    if constexpr (P == kPrecBF16)
      return (double)__uint_as_float((uint32_t)v << 16);
    else
      return v;
  }
};

template <int R, int N, int P, int NT>
__global__ void __launch_bounds__(64) k_rows_thin(MixedArgs a, uint64_t base) {
  typedef ThinSmem<R, P, NT> TS;
  __shared__ TS sm;
  const uint4 it = a.thin[base + blockIdx.x];
  const int l = threadIdx.x;
  const int n = (int)(it.z - it.y);
  constexpr int nt = NT;  // system size handled (zero-padded)
  // 1. stage h for the row's nonzeros (rows >= n zero) and x.  Lane p < n
  // loads nonzero p's ids once; the gathers of all nonzeros are then issued
  // back to back (independent loads, one memory latency for the row).
  constexpr int FPL = R / 64 > 0 ? R / 64 : 1;  // features per lane
  uint32_t myid[N - 1];
#pragma unroll
  for (int s = 0; s < N - 1; ++s) myid[s] = l < n ? a.oidx[s][it.y + l] : 0u;
  if (l < NT) sm.x[l] = l < n ? (double)a.vals[it.y + l] : 0.0;
#pragma unroll 8
  for (int p = 0; p < nt; ++p) {
    uint32_t id[N - 1];
#pragma unroll
    for (int s = 0; s < N - 1; ++s) id[s] = (uint32_t)__shfl((int)myid[s], p, 64);
#pragma unroll
    for (int j = 0; j < FPL; ++j) {
      const int f = l + 64 * j;
      if (f >= R) continue;
      typename TS::HT h = 0;
      if (p < n) {
        if constexpr (P == kPrecBF16) {
          float hf = 1.f;
#pragma unroll
          for (int s = 0; s < N - 1; ++s) hf *= __uint_as_float((uint32_t)a.fb[s][(size_t)id[s] * R + f] << 16);
          h = f2bf(hf);
        } else {
          double hd = 1.0;
#pragma unroll
          for (int s = 0; s < N - 1; ++s) hd *= (double)a.fr[s][(size_t)id[s] * R + f];
          h = hd;
        }
      }
      sm.H[p * TS::HS + f] = h;
    }
  }
  if (l == 0) sm.bad = 0;
  wave_sync();
  // 2. K = H H^T (lower 16x16 tiles) on v_mfma_f64_16x16x4f64
  f64x4 k00 = {0.0, 0.0, 0.0, 0.0}, k10 = k00, k11 = k00;
  for (int f0 = 0; f0 < R; f0 += 4) {
    const int f = f0 + (l >> 4);
    const double h0 = TS::hval(sm.H[(l & 15) * TS::HS + f]);
    k00 = __builtin_amdgcn_mfma_f64_16x16x4f64(h0, h0, k00, 0, 0, 0);
    if constexpr (NT == 32) {
      const double h1 = TS::hval(sm.H[(16 + (l & 15)) * TS::HS + f]);
      k10 = __builtin_amdgcn_mfma_f64_16x16x4f64(h1, h0, k10, 0, 0, 0);
      k11 = __builtin_amdgcn_mfma_f64_16x16x4f64(h1, h1, k11, 0, 0, 0);
    }
  }
  const double lam = a.lam;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int r = row16(v, l), c = l & 15;
    // padded rows/columns (>= n) become the identity
    auto put = [&](int rr, int cc, double val) {
      const bool pad = rr >= n || cc >= n;
      sm.K[rr * TS::KS + cc] = pad ? (rr == cc ? 1.0 : 0.0) : val + (rr == cc ? lam : 0.0);
    };
    put(r, c, k00[v]);
    if constexpr (NT == 32) {
      put(16 + r, c, k10[v]);
      put(16 + r, 16 + c, k11[v]);
    }
  }
  wave_sync();
  // 3. K = L D L^T in place: after step k, K[i][j] (i > j) = l_ij d_j and
  // K[j][j] = d_j.  Lane l updates column c = l % 32 for rows of its half.
  // The padding (rows and columns >= n) is the identity with no coupling, so
  // steps k >= n - 1 change nothing: the elimination and both solves stop at
  // the row's own size.
  {
    const int c = l & 31, half = l >> 5;
    bool ok = true;
    for (int k = 0; k < n - 1; ++k) {
      const double dk = sm.K[k * TS::KS + k];
      ok = ok && dk > 0.0;
      const double lck = (c < NT ? sm.K[c * TS::KS + k] : 0.0) * rcp_f64(dk > 0.0 ? dk : 1.0);
      double aik[16];
#pragma unroll
      for (int e = 0; e < 16; ++e) aik[e] = (16 * half + e < NT) ? sm.K[(16 * half + e) * TS::KS + k] : 0.0;
      wave_sync();
      if (c > k && c < nt) {
#pragma unroll
        for (int e = 0; e < 16; ++e) {
          const int i = 16 * half + e;
          if (i >= c && i < nt) sm.K[i * TS::KS + c] = __builtin_fma(-aik[e], lck, sm.K[i * TS::KS + c]);
        }
      }
      wave_sync();
    }
    ok = ok && sm.K[(n - 1) * TS::KS + n - 1] > 0.0;
    if (!ok && l == 0) sm.bad = 1;
  }
  wave_sync();
  // 4. z = (L D L^T)^-1 x: lane i < nt holds x_i
  double xi = l < nt ? sm.x[l] : 0.0;
  const double di = l < nt ? sm.K[l * TS::KS + l] : 1.0;
  const double rdi = rcp_f64(di > 0.0 ? di : 1.0);
  for (int j = 0; j < n - 1; ++j) {  // forward, unit lower L: l_ij = K[i][j] / d_j
    const double yj = __shfl(xi * rdi, j, 64);  // y_j / d_j
    if (l > j && l < nt) xi = __builtin_fma(-sm.K[l * TS::KS + j], yj, xi);
  }
  xi *= rdi;
  for (int j = n - 1; j > 0; --j) {  // backward, L^T: l_ji = K[j][i] / d_i
    const double zj = __shfl(xi, j, 64);
    if (l < j) xi = __builtin_fma(-sm.K[j * TS::KS + l] * rdi, zj, xi);
  }
  wave_sync();
  if (l < NT) sm.x[l] = xi;
  wave_sync();
  // 5. a = H^T z
  float *dst = a.out + (size_t)a.directory[it.x] * R;
  const bool ok = sm.bad == 0;
  for (int f = l; f < R; f += 64) {
    double acc = 0.0;
    for (int p = 0; p < n; ++p) acc = __builtin_fma(TS::hval(sm.H[p * TS::HS + f]), sm.x[p], acc);
    dst[f] = ok ? (float)acc : 0.f;
  }
  if (!ok && l == 0) atomicOr(a.flag, 1u);
}

// ------------------------------------------------------------- dual rows --
// Rows with kThinRow < n <= kDualRow nonzeros at R = 128, also in the dual
// form a = H^T (H H^T + lambda I)^-1 x, as an n x n system padded to 48 or 64:
// cheaper than the R x R primal solve (64^3 / 3 instead of R^3 / 3 FP64
// flops, one wave instead of a barrier-separated blocked Cholesky).  One wave
// per row: h staged in LDS (BF16 bits), K = H H^T tile by tile on
// v_mfma_f64_16x16x4f64, each 16 x 16 tile passed through a small LDS buffer
// into the registers of the lanes that own its columns (lane c holds column
// c of K), then the root-free right-looking L D L^T with one LDS publish per
// column (the forward solve folded in), the back solve from the owning lanes'
// registers, a = H^T z.  LDS is the h stage plus 3 KB; the 64 FP64 column
// entries per lane hold the kernel to one wave per SIMD.
template <int R, int P, int NT>
struct DualSmem {
  typedef typename std::conditional<P == kPrecBF16, uint16_t, double>::type HT;
  static constexpr int HS = R + 1;
  HT H[NT * HS];
  double T[16 * 17];     // one K tile on its way to the column owners
  double pub[NT + 2];    // column k and z_k / x_k
  double z[NT];
  __device__ static double hval(HT v) {
    if constexpr (P == kPrecBF16)
      return (double)__uint_as_float((uint32_t)v << 16);
    else
      return v;
  }
};

// NT = 48 serves rows of at most 48 nonzeros, NT = 64 the rest
template <int R, int N, int P, int NT>
__global__ void __launch_bounds__(64) k_rows_dual(MixedArgs a, uint64_t base) {
  // one lane owns one column: at most 64 columns per wave
  static_assert(R >= kDualRow && NT % 16 == 0 && NT <= 64, "dual systems of 48 or 64 rows");
  typedef DualSmem<R, P, NT> DS;
  __shared__ DS sm;
  const uint4 it = a.dual[base + blockIdx.x];
  const int l = threadIdx.x;
  const int n = (int)(it.z - it.y);
  // 1. stage h (rows >= n zero)
  constexpr int FPL = R / 64;
  uint32_t myid[N - 1];
#pragma unroll
  for (int s = 0; s < N - 1; ++s) myid[s] = l < n ? a.oidx[s][it.y + l] : 0u;
  const double xl = l < n ? (double)a.vals[it.y + l] : 0.0;
  const int lc = l < NT ? l : NT - 1;  // lanes past NT own no column
#pragma unroll 8
  for (int p = 0; p < NT; ++p) {
    uint32_t id[N - 1];
#pragma unroll
    for (int s = 0; s < N - 1; ++s) id[s] = (uint32_t)__shfl((int)myid[s], p, 64);
#pragma unroll
    for (int j = 0; j < FPL; ++j) {
      const int f = l + 64 * j;
      typename DS::HT h = 0;
      if (p < n) {
        if constexpr (P == kPrecBF16) {
          float hf = 1.f;
#pragma unroll
          for (int s = 0; s < N - 1; ++s) hf *= __uint_as_float((uint32_t)a.fb[s][(size_t)id[s] * R + f] << 16);
          h = f2bf(hf);
        } else {
          double hd = 1.0;
#pragma unroll
          for (int s = 0; s < N - 1; ++s) hd *= (double)a.fr[s][(size_t)id[s] * R + f];
          h = hd;
        }
      }
      sm.H[p * DS::HS + f] = h;
    }
  }
  wave_sync();
  // 2. K = H H^T + lambda I (lower 16x16 tiles; padded rows/columns ->
  // identity); lane c collects column c (rows >= c meaningful, the rest are
  // don't-care)
  const double lam = a.lam;
  double g[NT];
#pragma unroll
  for (int i = 0; i < NT; ++i) g[i] = 0.0;
#pragma unroll
  for (int bi = 0; bi < NT / 16; ++bi) {
#pragma unroll
    for (int bj = 0; bj <= bi; ++bj) {
      f64x4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 2
      for (int f0 = 0; f0 < R; f0 += 4) {
        const int f = f0 + (l >> 4);
        const double ha = DS::hval(sm.H[(16 * bi + (l & 15)) * DS::HS + f]);
        const double hb = DS::hval(sm.H[(16 * bj + (l & 15)) * DS::HS + f]);
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(ha, hb, acc, 0, 0, 0);
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = 16 * bi + row16(v, l), c = 16 * bj + (l & 15);
        const bool pad = r >= n || c >= n;
        sm.T[row16(v, l) * 17 + (l & 15)] = pad ? (r == c ? 1.0 : 0.0) : acc[v] + (r == c ? lam : 0.0);
      }
      wave_sync();
      if ((l >> 4) == bj) {  // (lanes >= NT never match: bj < NT / 16)
#pragma unroll
        for (int r = 0; r < 16; ++r) g[16 * bi + r] = sm.T[r * 17 + (l & 15)];
      }
      wave_sync();
    }
  }
  double b = xl;
  double *buf = sm.pub;  // column k (NT values), then z_k / x_k
  bool bad = false;
  double rd = 1.0;
#pragma unroll
  for (int k = 0; k < NT; ++k) {
    if (l == k) {
#pragma unroll
      for (int i = k; i < NT; ++i) buf[i] = g[i];
      buf[NT] = b;
    }
    wave_sync();
    double d = buf[k];
    if (!(d > 0.0)) {
      bad = true;
      d = 1.0;
    }
    const double rk = rcp_f64(d);
    const double zk = buf[NT];
    const double lck = (l > k && l < NT) ? buf[lc] * rk : 0.0;
#pragma unroll
    for (int i = k + 1; i < NT; ++i) g[i] = __builtin_fma(-buf[i], lck, g[i]);
    b = __builtin_fma(-lck, zk, b);
    if (l == k) rd = rk;
    wave_sync();
  }
  b *= rd;
#pragma unroll
  for (int k = NT - 1; k >= 0; --k) {
    if (l == k) buf[NT] = b;
    wave_sync();
    const double xk = buf[NT];
    const double lkc = l < k ? g[k] * rd : 0.0;
    b = __builtin_fma(-lkc, xk, b);
    wave_sync();
  }
  if (l < NT) sm.z[l] = b;
  wave_sync();
  // 4. a = H^T z
  float *dst = a.out + (size_t)a.directory[it.x] * R;
#pragma unroll
  for (int j = 0; j < FPL; ++j) {
    const int f = l + 64 * j;
    double acc = 0.0;
    for (int p = 0; p < n; ++p) acc = __builtin_fma(DS::hval(sm.H[p * DS::HS + f]), sm.z[p], acc);
    dst[f] = bad ? 0.f : (float)acc;
  }
  if (bad && l == 0) atomicOr(a.flag, 1u);
}

// ---------------------------------------------------------------- kernels --
// Gram + MTTKRP of one item into the accumulators and sm.vm (sm: MixedSmem,
// or GramSmem for the chunk kernel; .tiles stages the h rows)
template <int R, int N, class SM>
__device__ inline void gram_item(const MixedArgs &a, const uint4 &it, f32x16 *acc, SM &sm) {
  float mp[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  gram_chunk<R, N>(a, it.y, it.z, acc, mp, (uint32_t *)sm.tiles);
  reduce_m<R>(mp, sm);
}
template <int R, int N, class SM>
__device__ inline void gram_item(const MixedArgs &a, const uint4 &it, f64x4 *acc, SM &sm) {
  double mp[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  gram_f64<R, N>(a, it.y, it.z, acc, mp, sm.tiles);
  reduce_m<R>(mp, sm);
}

// the Gram phase's shared memory alone (chunks of long rows: no solve)
template <int R, int P>
struct GramSmem {
  static constexpr size_t STG = P == kPrecBF16 ? (size_t)R * kHTS * 4 : (size_t)32 * MixedSmem<R>::HS * 8;
  double tiles[STG / 8];
  double vm[R];
  double contrib[R / 32][R];
};

// accumulator slots of one wave and their place in a partial
template <int R, int P>
struct AccShape;
template <int R>
struct AccShape<R, kPrecBF16> {  // NW stripes of 32x32 blocks, 16 floats per lane each
  typedef f32x16 T;
  typedef float E;
  static constexpr int NSLOT = R / 32, NV = 16;
  static constexpr size_t PSZ = (size_t)R * R + R;  // elements per partial
  __device__ static bool used(int w, int j) { return j <= w; }
};
template <int R>
struct AccShape<R, kPrecF64> {  // round-robin 16x16 tiles, 4 doubles per lane each
  typedef f64x4 T;
  typedef double E;
  static constexpr int NSLOT = MixedSmem<R>::TPW, NV = 4;
  static constexpr size_t PSZ = (size_t)(R / 32) * MixedSmem<R>::TPW * 256 + R;
  __device__ static bool used(int w, int j) { return w + (R / 32) * j < MixedSmem<R>::NTILE; }
};

template <int R, int N, int P>
__global__ void __launch_bounds__(2 * R, 2) k_mixed_items(MixedArgs a) {
  typedef AccShape<R, P> S;
  __shared__ MixedSmem<R> sm;
  const uint4 it = a.items[blockIdx.x];
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
  typename S::T acc[S::NSLOT];
#pragma unroll
  for (int j = 0; j < S::NSLOT; ++j)
#pragma unroll
    for (int v = 0; v < S::NV; ++v) acc[j][v] = 0;
  gram_item<R, N>(a, it, acc, sm);
  if (it.w != kLightRow) {  // chunk of a heavy row: partial out
    typename S::E *Pp = (typename S::E *)a.part + (size_t)it.w * S::PSZ;
#pragma unroll
    for (int j = 0; j < S::NSLOT; ++j) {
      if (!S::used(w, j)) continue;
#pragma unroll
      for (int v = 0; v < S::NV; ++v) Pp[((w * S::NSLOT + j) * S::NV + v) * 64 + l] = acc[j][v];
    }
    if (tid < R) Pp[S::PSZ - R + tid] = (typename S::E)sm.vm[tid];
    return;
  }
  const bool ok = solve_row<R>(sm, acc, a.lam, a.refine);
  store_row<R>(a, sm, it.x, ok);
}

// chunks of long rows only: Gram + MTTKRP into the chunk's partial.  Without
// the solve's 79 KB of LDS tiles and its registers, several more workgroups
// fit on a CU, and the chunk Gram is bound by how many row gathers are in
// flight.
template <int R, int N, int P>
__global__ void __launch_bounds__(2 * R) k_mixed_chunks(MixedArgs a) {
  typedef AccShape<R, P> S;
  __shared__ GramSmem<R, P> sm;
  const uint4 it = a.items[blockIdx.x];
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
  typename S::T acc[S::NSLOT];
#pragma unroll
  for (int j = 0; j < S::NSLOT; ++j)
#pragma unroll
    for (int v = 0; v < S::NV; ++v) acc[j][v] = 0;
  gram_item<R, N>(a, it, acc, sm);
  typename S::E *Pp = (typename S::E *)a.part + (size_t)it.w * S::PSZ;
#pragma unroll
  for (int j = 0; j < S::NSLOT; ++j) {
    if (!S::used(w, j)) continue;
#pragma unroll
    for (int v = 0; v < S::NV; ++v) Pp[((w * S::NSLOT + j) * S::NV + v) * 64 + l] = acc[j][v];
  }
  if (tid < R) Pp[S::PSZ - R + tid] = (typename S::E)sm.vm[tid];
}

template <int R, int N, int P>
__global__ void __launch_bounds__(2 * R, 2) k_mixed_heavy(MixedArgs a) {
  typedef AccShape<R, P> S;
  __shared__ MixedSmem<R> sm;
  const uint32_t h = blockIdx.x;
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
  typename S::T acc[S::NSLOT];
#pragma unroll
  for (int j = 0; j < S::NSLOT; ++j)
#pragma unroll
    for (int v = 0; v < S::NV; ++v) acc[j][v] = 0;
  double mv = 0.0;
  uint32_t p0 = a.heavy_pb[h], p1 = a.heavy_pb[h + 1];
  const typename S::E *src = (const typename S::E *)a.part;
  if (a.heavy_gb && a.heavy_gb[h + 1] > a.heavy_gb[h]) {  // long row: its group sums
    p0 = a.heavy_gb[h];
    p1 = a.heavy_gb[h + 1];
    src = (const typename S::E *)a.gpart;
  }
  for (uint32_t pi = p0; pi < p1; ++pi) {
    const typename S::E *Pp = src + (size_t)pi * S::PSZ;
#pragma unroll
    for (int j = 0; j < S::NSLOT; ++j) {
      if (!S::used(w, j)) continue;
#pragma unroll
      for (int v = 0; v < S::NV; ++v) acc[j][v] += Pp[((w * S::NSLOT + j) * S::NV + v) * 64 + l];
    }
    if (tid < R) mv += (double)Pp[S::PSZ - R + tid];
  }
  if (tid < R) sm.vm[tid] = mv;
  __syncthreads();
  const bool ok = solve_row<R>(sm, acc, a.lam, a.refine);
  store_row<R>(a, sm, a.heavy_slot[h], ok);
}

// group sums of consecutive partials of one long row (fixed order, FP64
// accumulation, stored in the partials' own type), one workgroup per group
template <typename E>
__global__ void __launch_bounds__(256) k_mixed_group(const E *part, const uint2 *grp, uint64_t ngroups, size_t psz,
                                                     E *gpart) {
  const uint64_t g = blockIdx.x;
  if (g >= ngroups) return;
  const uint2 gr = grp[g];
  const E *P = part + (size_t)gr.x * psz;
  for (size_t e = threadIdx.x; e < psz; e += blockDim.x) {
    double acc = 0.0;
    for (uint32_t q = 0; q < gr.y; ++q) acc += (double)P[(size_t)q * psz + e];
    gpart[g * psz + e] = (E)acc;
  }
}

hipError_t launch_mixed_groups(hipStream_t st, const MixedArgs &a, int R) {
  if (a.ngroups == 0) return hipSuccess;
  const size_t bytes = mixed_partial_bytes(R, a.prec);
  if (a.prec == kPrecBF16)
    hipLaunchKernelGGL(k_mixed_group<float>, dim3((unsigned)a.ngroups), dim3(256), 0, st, (const float *)a.part,
                       a.grp, a.ngroups, bytes / sizeof(float), (float *)a.gpart);
  else
    hipLaunchKernelGGL(k_mixed_group<double>, dim3((unsigned)a.ngroups), dim3(256), 0, st, (const double *)a.part,
                       a.grp, a.ngroups, bytes / sizeof(double), (double *)a.gpart);
  return hipGetLastError();
}

__global__ void __launch_bounds__(256) k_to_bf16(const float *x, uint16_t *y, uint64_t n) {
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (uint64_t)gridDim.x * blockDim.x)
    y[p] = f2bf(x[p]);
}

#define CTC_ROW_LAUNCH(KERNEL, RR, NN, GRID)                                                             \
  do {                                                                                                  \
    if (a.prec == kPrecBF16)                                                                            \
      hipLaunchKernelGGL((KERNEL<RR, NN, kPrecBF16>), dim3((unsigned)(GRID)), dim3(2 * RR), 0, st, a); \
    else                                                                                                \
      hipLaunchKernelGGL((KERNEL<RR, NN, kPrecF64>), dim3((unsigned)(GRID)), dim3(2 * RR), 0, st, a);  \
  } while (0)

#define CTC_ROW_SWITCH(KERNEL, GRID)                   \
  do {                                                 \
    if (nmodes == 3) {                                 \
      if (R == 32) CTC_ROW_LAUNCH(KERNEL, 32, 3, GRID);   \
      else if (R == 64) CTC_ROW_LAUNCH(KERNEL, 64, 3, GRID); \
      else if (R == 128) CTC_ROW_LAUNCH(KERNEL, 128, 3, GRID); \
      else return hipErrorInvalidValue;                \
    } else if (nmodes == 4) {                          \
      if (R == 32) CTC_ROW_LAUNCH(KERNEL, 32, 4, GRID);   \
      else if (R == 64) CTC_ROW_LAUNCH(KERNEL, 64, 4, GRID); \
      else if (R == 128) CTC_ROW_LAUNCH(KERNEL, 128, 4, GRID); \
      else return hipErrorInvalidValue;                \
    } else {                                           \
      return hipErrorInvalidValue;                     \
    }                                                  \
  } while (0)

hipError_t launch_mixed_items(hipStream_t st, const MixedArgs &a, int nmodes, int R) {
  if (a.nitems == 0) return hipSuccess;
  CTC_ROW_SWITCH(k_mixed_items, a.nitems);
  return hipGetLastError();
}

hipError_t launch_mixed_chunks(hipStream_t st, const MixedArgs &a, int nmodes, int R) {
  if (a.nitems == 0) return hipSuccess;
  CTC_ROW_SWITCH(k_mixed_chunks, a.nitems);
  return hipGetLastError();
}

hipError_t launch_mixed_heavy(hipStream_t st, const MixedArgs &a, int nmodes, int R) {
  if (a.nheavy == 0) return hipSuccess;
  CTC_ROW_SWITCH(k_mixed_heavy, a.nheavy);
  return hipGetLastError();
}

// rows of at most 16 nonzeros first (NT = 16), then the rest (NT = 32)
#define CTC_THIN_LAUNCH(RR, NN)                                                                                    \
  do {                                                                                                            \
    const uint64_t n16 = a.nthin16, n32 = a.nthin - a.nthin16;                                                    \
    if (a.prec == kPrecBF16) {                                                                                    \
      if (n16) hipLaunchKernelGGL((k_rows_thin<RR, NN, kPrecBF16, 16>), dim3((unsigned)n16), dim3(64), 0, st, a, 0ull); \
      if (n32) hipLaunchKernelGGL((k_rows_thin<RR, NN, kPrecBF16, 32>), dim3((unsigned)n32), dim3(64), 0, st, a, n16);  \
    } else {                                                                                                      \
      if (n16) hipLaunchKernelGGL((k_rows_thin<RR, NN, kPrecF64, 16>), dim3((unsigned)n16), dim3(64), 0, st, a, 0ull);  \
      if (n32) hipLaunchKernelGGL((k_rows_thin<RR, NN, kPrecF64, 32>), dim3((unsigned)n32), dim3(64), 0, st, a, n16);   \
    }                                                                                                             \
  } while (0)

// rows of at most 48 nonzeros first (NT = 48), then the rest (NT = 64)
#define CTC_DUAL_LAUNCH(RR, NN)                                                                                   \
  do {                                                                                                           \
    const uint64_t n48 = a.ndual48, n64 = a.ndual - a.ndual48;                                                    \
    if (a.prec == kPrecBF16) {                                                                                   \
      if (n48) hipLaunchKernelGGL((k_rows_dual<RR, NN, kPrecBF16, 48>), dim3((unsigned)n48), dim3(64), 0, st, a, 0ull); \
      if (n64) hipLaunchKernelGGL((k_rows_dual<RR, NN, kPrecBF16, 64>), dim3((unsigned)n64), dim3(64), 0, st, a, n48);  \
    } else {                                                                                                     \
      if (n48) hipLaunchKernelGGL((k_rows_dual<RR, NN, kPrecF64, 48>), dim3((unsigned)n48), dim3(64), 0, st, a, 0ull);  \
      if (n64) hipLaunchKernelGGL((k_rows_dual<RR, NN, kPrecF64, 64>), dim3((unsigned)n64), dim3(64), 0, st, a, n48);   \
    }                                                                                                            \
  } while (0)

hipError_t launch_rows_dual(hipStream_t st, const MixedArgs &a, int nmodes, int R) {
  if (a.ndual == 0) return hipSuccess;
  if (R != 128) return hipErrorInvalidValue;  // dual rows are only formed at R = 128
  if (nmodes == 3) {
    CTC_DUAL_LAUNCH(128, 3);
  } else if (nmodes == 4) {
    CTC_DUAL_LAUNCH(128, 4);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
#undef CTC_DUAL_LAUNCH

hipError_t launch_rows_thin(hipStream_t st, const MixedArgs &a, int nmodes, int R) {
  if (a.nthin == 0) return hipSuccess;
  if (nmodes == 3) {
    if (R == 32) CTC_THIN_LAUNCH(32, 3);
    else if (R == 64) CTC_THIN_LAUNCH(64, 3);
    else if (R == 128) CTC_THIN_LAUNCH(128, 3);
    else return hipErrorInvalidValue;
  } else if (nmodes == 4) {
    if (R == 32) CTC_THIN_LAUNCH(32, 4);
    else if (R == 64) CTC_THIN_LAUNCH(64, 4);
    else if (R == 128) CTC_THIN_LAUNCH(128, 4);
    else return hipErrorInvalidValue;
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

size_t mixed_partial_bytes(int R, int prec) {
  const size_t NW = (size_t)R / 32, RT = (size_t)R / 16, NTILE = RT * (RT + 1) / 2;
  const size_t TPW = (NTILE + NW - 1) / NW;
  return prec == kPrecBF16 ? ((size_t)R * R + R) * sizeof(float) : (NW * TPW * 256 + R) * sizeof(double);
}

hipError_t launch_to_bf16(hipStream_t st, const float *x, uint16_t *y, uint64_t n) {
  if (n == 0) return hipSuccess;
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)blocks), dim3(256), 0, st, x, y, n);
  return hipGetLastError();
}

}  // namespace ctc
