/*
 * oracle.c — CPU restatement of abovekumo/cuTC (reference) for the tensor
 * completion hot path.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Reference files restated here (paths relative to the reference root,
 * SG = "software/single GPU/sgpu", NC = "software/multiple GPU_NCCL"):
 *   SG/io.c:58-106, 234-308     .tns text parsing and dims / index base
 *   SG/ciss.c:68-195            CISS (TB-COO) layout
 *   SG/als.cu:186-317, 467-491, 676-716   ALS per-mode update
 *   SG/sgd.cu:21-143, 264-324   SGD update rule
 *   SG/ccd.cu:532-554, NC/ccd.cu:28-180   canonical CCD++
 *   SG/completion.c:123-214, 296-362      loss / RMSE / Frobenius / convergence
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* .tns reader                                                               */
/* ------------------------------------------------------------------------ */

/* Lines are data lines when getline returned more than one byte and the first
 * byte is not '#' (SG/io.c:93, :251, :272). */
static int is_data_line(const char *line, ssize_t read) {
  return read > 1 && line[0] != '#';
}

int orc_tns_read(const char *path, orc_coo *out) {
  memset(out, 0, sizeof(*out));
  FILE *fin = fopen(path, "r");
  if (!fin) return 2; /* SG/io.c:199-202 returns NULL */
  char *line = NULL;
  size_t len = 0;
  ssize_t read;
  /* mode count = whitespace-separated tokens of the first data line minus one
   * (SG/io.c:250-261; strtok on " \t"). */
  uint32_t nmodes = 0;
  while ((read = getline(&line, &len, fin)) != -1) {
    if (is_data_line(line, read)) {
      char *save = NULL;
      for (char *p = strtok_r(line, " \t", &save); p; p = strtok_r(NULL, " \t", &save))
        ++nmodes;
      break;
    }
  }
  if (nmodes < 2) {
    free(line);
    fclose(fin);
    return 3;
  }
  --nmodes;
  if (nmodes > ORC_MAX_MODES) {
    free(line);
    fclose(fin);
    return 3;
  }
  uint64_t dims[ORC_MAX_MODES], offset[ORC_MAX_MODES];
  for (uint32_t m = 0; m < nmodes; ++m) {
    dims[m] = 0;
    offset[m] = 1; /* SG/io.c:263-266 */
  }
  uint64_t nnz = 0;
  rewind(fin);
  while ((read = getline(&line, &len, fin)) != -1) {
    if (!is_data_line(line, read)) continue;
    char *ptr = line;
    for (uint32_t m = 0; m < nmodes; ++m) {
      uint64_t ind = strtoull(ptr, &ptr, 10);
      if (ind > dims[m]) dims[m] = ind;     /* dims = max index, :278 */
      if (ind < offset[m]) offset[m] = ind; /* base = min index, :281 */
    }
    strtod(ptr, &ptr);
    ++nnz;
  }
  for (uint32_t m = 0; m < nmodes; ++m) {
    if (offset[m] != 0 && offset[m] != 1) { /* SG/io.c:291-296 */
      free(line);
      fclose(fin);
      return 4;
    }
    if (offset[m] == 0) ++dims[m]; /* SG/io.c:299-304 */
  }
  if (nnz == 0) {
    free(line);
    fclose(fin);
    return 3;
  }
  out->nmodes = nmodes;
  out->nnz = nnz;
  for (uint32_t m = 0; m < nmodes; ++m) {
    out->dims[m] = dims[m];
    out->offset[m] = offset[m];
    out->ind[m] = (uint32_t *)malloc(nnz * sizeof(uint32_t));
  }
  out->vals = (double *)malloc(nnz * sizeof(double));
  rewind(fin);
  uint64_t n = 0;
  while ((read = getline(&line, &len, fin)) != -1) {
    if (!is_data_line(line, read)) continue;
    char *ptr = line;
    for (uint32_t m = 0; m < nmodes; ++m) {
      uint64_t ind = strtoul(ptr, &ptr, 10); /* SG/io.c:96 */
      out->ind[m][n] = (uint32_t)(ind - offset[m]);
    }
    out->vals[n++] = strtod(ptr, &ptr); /* SG/io.c:98 */
  }
  free(line);
  fclose(fin);
  return 0;
}

void orc_coo_free(orc_coo *t) {
  for (uint32_t m = 0; m < ORC_MAX_MODES; ++m) {
    free(t->ind[m]);
    t->ind[m] = NULL;
  }
  free(t->vals);
  t->vals = NULL;
}

/* ------------------------------------------------------------------------ */
/* ordering                                                                  */
/* ------------------------------------------------------------------------ */

static void counting_pass(const uint32_t *key, uint64_t dim, uint64_t nnz,
                          const uint32_t *in, uint32_t *out, uint64_t *count) {
  memset(count, 0, (dim + 1) * sizeof(uint64_t));
  for (uint64_t p = 0; p < nnz; ++p) count[key[in[p]] + 1]++;
  for (uint64_t d = 0; d < dim; ++d) count[d + 1] += count[d];
  for (uint64_t p = 0; p < nnz; ++p) out[count[key[in[p]]]++] = in[p];
}

int orc_mode_order(const orc_coo *t, uint32_t mode, uint32_t *perm, uint64_t *rowptr) {
  if (mode >= t->nmodes) return 1;
  uint64_t nnz = t->nnz;
  uint64_t maxdim = 0;
  for (uint32_t m = 0; m < t->nmodes; ++m)
    if (t->dims[m] > maxdim) maxdim = t->dims[m];
  uint32_t *tmp = (uint32_t *)malloc(nnz * sizeof(uint32_t));
  uint64_t *count = (uint64_t *)malloc((maxdim + 1) * sizeof(uint64_t));
  for (uint64_t p = 0; p < nnz; ++p) perm[p] = (uint32_t)p;
  /* least significant key first: modes mode+N-1, ..., mode+1, mode */
  uint32_t *a = perm, *b = tmp;
  for (int s = (int)t->nmodes - 1; s >= 0; --s) {
    uint32_t m = (mode + (uint32_t)s) % t->nmodes;
    counting_pass(t->ind[m], t->dims[m], nnz, a, b, count);
    uint32_t *x = a;
    a = b;
    b = x;
  }
  if (a != perm) memcpy(perm, a, nnz * sizeof(uint32_t));
  if (rowptr) {
    uint64_t dim = t->dims[mode];
    memset(rowptr, 0, (dim + 1) * sizeof(uint64_t));
    for (uint64_t p = 0; p < nnz; ++p) rowptr[t->ind[mode][p] + 1]++;
    for (uint64_t d = 0; d < dim; ++d) rowptr[d + 1] += rowptr[d];
  }
  free(tmp);
  free(count);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* the reference's tile-reduction defects (SG/als.cu:226-312) as weights     */
/* ------------------------------------------------------------------------ */

int orc_quirk_weights(const orc_coo *t, uint32_t mode, const uint32_t *perm, uint32_t tile_len,
                      uint32_t flags, uint32_t *w) {
  if (mode >= t->nmodes || tile_len == 0 || tile_len > 64) return 1;
  const uint64_t n = t->nnz;
  const uint32_t *row = t->ind[mode];
  for (uint64_t b = 0; b < n; b += tile_len) {
    const uint64_t e = b + tile_len < n ? b + tile_len : n;
    const uint32_t L = (uint32_t)(e - b);
    int cont[64];
    uint32_t iters = 0; /* continuation lanes 1..L-1 (itercounter) */
    for (uint32_t l = 0; l < L; ++l) {
      const uint64_t p = b + l;
      cont[l] = p > 0 && row[perm[p]] == row[perm[p - 1]];
      if (l > 0) iters += (uint32_t)cont[l];
      w[perm[p]] = 1;
    }
    if (flags & 1) /* Q1: lane 0 through the lane before the next head */
      for (uint32_t l = 0; l < L && (l == 0 || cont[l]); ++l) w[perm[b + l]] = 0;
    if ((flags & 2) && L == tile_len && cont[L - 1]) { /* Q2 */
      uint32_t h = L - 1;
      while (h > 0 && cont[h]) --h;
      if (!((flags & 1) && h == 0)) w[perm[b + L - 1]] += iters - (L - 1 - h);
    }
  }
  if (flags & 16) { /* Q3: the CCD grid, tilenum / 256 + 1 blocks of 256 / tile_len warps (SG/ccd.cu:520) */
    const uint64_t tiles = n / tile_len + 1;
    const uint64_t covered = (tiles / 256 + 1) * (256 / tile_len);
    for (uint64_t p = covered * tile_len; p < n; ++p) w[perm[p]] = 0;
  }
  return 0;
}

/* The reference's CCD update of one mode (update_frac_gpu_as +
 * update_ccd_gpu, SG/ccd.cu:82-206,318-348), in FP64: for every row i with a
 * nonzero in t, num_f = sum w_p e_p h_f and den_f = sum w_p h_f^2 over its
 * nonzeros, e_p = x_p - <F_mode(i,:), h> from the row's values before the
 * update (x truncated toward zero when trunc_vals), then F_mode(i,f) = num_f /
 * (lambda + den_f) for every f; rows without nonzeros are kept. */
int orc_ccd_jacobi_mode(const orc_coo *t, uint32_t mode, uint32_t R, double **F, double lambda,
                        const uint32_t *w, int trunc_vals, const uint32_t *perm, const uint64_t *rowptr) {
  if (mode >= t->nmodes || R == 0 || R > 512) return 1;
  double h[512], num[512], den[512];
  for (uint64_t i = 0; i < t->dims[mode]; ++i) {
    if (rowptr[i] == rowptr[i + 1]) continue;
    double *a = F[mode] + (size_t)i * R;
    for (uint32_t f = 0; f < R; ++f) num[f] = den[f] = 0.0;
    for (uint64_t q = rowptr[i]; q < rowptr[i + 1]; ++q) {
      const uint64_t p = perm[q];
      const double wp = w ? (double)w[p] : 1.0;
      if (wp == 0.0) continue;
      for (uint32_t f = 0; f < R; ++f) h[f] = 1.0;
      for (uint32_t m = 0; m < t->nmodes; ++m) {
        if (m == mode) continue;
        const double *row = F[m] + (size_t)t->ind[m][p] * R;
        for (uint32_t f = 0; f < R; ++f) h[f] *= row[f];
      }
      double e = trunc_vals ? trunc(t->vals[p]) : t->vals[p];
      for (uint32_t f = 0; f < R; ++f) e -= a[f] * h[f];
      for (uint32_t f = 0; f < R; ++f) {
        num[f] += wp * e * h[f];
        den[f] += wp * h[f] * h[f];
      }
    }
    for (uint32_t f = 0; f < R; ++f) a[f] = num[f] / (lambda + den[f]);
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* reference CISS layout (SG/ciss.c:68-195)                                  */
/* ------------------------------------------------------------------------ */

int orc_ciss_reference_layout(const orc_coo *t, uint32_t mode, uint32_t tile_len,
                              int truncate_vals, uint64_t *directory,
                              uint64_t *dcounter, double *entries,
                              uint64_t *dlength, uint64_t *size) {
  if (mode >= t->nmodes || t->nmodes != 3 || tile_len == 0 || t->nnz == 0) return 1;
  const uint64_t nnz = t->nnz;
  const uint32_t m1 = (mode + 1) % 3, m2 = (mode + 2) % 3;
  /* raw ids as the reference holds them: 1-based (the reference's kernels
   * subtract one, SG/als.cu:256-258). */
#define RAW(mm, p) ((uint64_t)t->ind[(mm)][(p)] + 1u)

  /* directory + dcounter (SG/ciss.c:89-115).  The reference accumulates
   * dcounter in a scratch of length nnz indexed up to dlength; the values it
   * keeps are dcounter[0] = 0, dcounter[c] = index of the last nonzero of
   * slot c-1, dcounter[dlength] = nnz - 1. */
  uint64_t counter = 1;
  uint64_t last = RAW(mode, 0);
  directory[0] = last;
  dcounter[0] = 0;
  for (uint64_t p = 1; p < nnz; ++p) {
    if (RAW(mode, p) != last) {
      directory[counter] = RAW(mode, p);
      dcounter[counter] = p - 1;
      ++counter;
      last = RAW(mode, p);
    }
  }
  /* dcounter[1] is the running count after slot 0, which the reference's
   * accumulation leaves equal to the last index of slot 0 (SG/ciss.c:104-107). */
  dcounter[counter] = nnz - 1;
  *dlength = counter;

  /* tile records (SG/ciss.c:121-188) */
  const uint64_t L = tile_len;
  const uint64_t tiles = nnz / L + 1;
  *size = tiles + nnz + 1;
  uint64_t position = 0, slot_l = 0;
  uint64_t p = 0;
  while (p < nnz + 1) {
    uint64_t bitmap = 1;
    const uint64_t slot_s = slot_l;
    uint64_t running = directory[slot_s];
    double *tile = entries + position;
    tile[0] = -1.0 * (double)slot_s; /* header[0] = -(first slot); -0.0 for slot 0 */
    for (uint64_t lc = 1; lc <= L; ++lc) {
      if (p == nnz) { /* sentinel record, SG/ciss.c:144-151 */
        tile[lc * 3] = -1;
        tile[lc * 3 + 1] = -1;
        tile[lc * 3 + 2] = -1;
        ++p;
        break;
      }
      const uint64_t ti = RAW(mode, p);
      if (ti == running) {
        bitmap = (bitmap << 1) + 1;
      } else {
        bitmap = bitmap << 1;
        running = ti;
        ++slot_l;
        tile[1] = -1.0 * (double)slot_l; /* written only on a row change, :171 */
      }
      tile[lc * 3] = (double)RAW(m1, p);
      tile[lc * 3 + 1] = (double)RAW(m2, p);
      double v = t->vals[p];
      if (truncate_vals) v = (double)(uint64_t)v; /* idx_t tmpv, quirk Q5 */
      tile[lc * 3 + 2] = v;
      tile[2] = (double)bitmap;
      ++p;
    }
    position += (L + 1) * 3;
  }
#undef RAW
  return 0;
}

/* ------------------------------------------------------------------------ */
/* dense helpers                                                             */
/* ------------------------------------------------------------------------ */

int orc_cholesky(double *A, uint32_t n) {
  for (uint32_t i = 0; i < n; ++i) {
    for (uint32_t j = 0; j <= i; ++j) {
      double s = A[(size_t)i * n + j];
      for (uint32_t k = 0; k < j; ++k) s -= A[(size_t)i * n + k] * A[(size_t)j * n + k];
      if (i == j) {
        if (!(s > 0.0)) return 7;
        A[(size_t)i * n + i] = sqrt(s);
      } else {
        A[(size_t)i * n + j] = s / A[(size_t)j * n + j];
      }
    }
    for (uint32_t j = i + 1; j < n; ++j) A[(size_t)i * n + j] = 0.0;
  }
  return 0;
}

void orc_cholesky_solve(const double *L, uint32_t n, double *b) {
  for (uint32_t i = 0; i < n; ++i) { /* L y = b */
    double s = b[i];
    for (uint32_t k = 0; k < i; ++k) s -= L[(size_t)i * n + k] * b[k];
    b[i] = s / L[(size_t)i * n + i];
  }
  for (int i = (int)n - 1; i >= 0; --i) { /* L^T x = y */
    double s = b[i];
    for (uint32_t k = (uint32_t)i + 1; k < n; ++k) s -= L[(size_t)k * n + i] * b[k];
    b[i] = s / L[(size_t)i * n + i];
  }
}

/* float -> bfloat16 -> float, round to nearest even (finite inputs). */
float orc_bf16_round(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  u &= 0xFFFF0000u;
  float y;
  memcpy(&y, &u, 4);
  return y;
}

/* The BF16-operand ALS update (the mixed-precision path's arithmetic, no
 * counterpart in the reference): factor rows of the other modes are rounded
 * to BF16, h is their product (exact in FP32 for up to three BF16 factors),
 * G = sum bf16(h) bf16(h)^T and m = sum x bf16(h) are accumulated in FP64
 * here, and (G + lambda I) a = m is solved in FP64.  Rows with no nonzero
 * become 0.  Returns 0, or 7 if a system is not positive definite. */
int orc_als_mode_bf16(const orc_coo *t, uint32_t mode, uint32_t R, float **F, double lambda,
                      const uint32_t *perm, const uint64_t *rowptr, int nthreads) {
  if (mode >= t->nmodes || R == 0 || R > 512) return 1;
  const int64_t dim = (int64_t)t->dims[mode];
  int status = 0;
  (void)nthreads;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
#endif
  for (int64_t i = 0; i < dim; ++i) {
    float *dst = F[mode] + (size_t)i * R;
    if (rowptr[i] == rowptr[i + 1]) {
      for (uint32_t f = 0; f < R; ++f) dst[f] = 0.f;
      continue;
    }
    double *G = (double *)calloc((size_t)R * R, sizeof(double));
    double *rhs = (double *)calloc(R, sizeof(double));
    float h[512], hb[512];
    for (uint64_t q = rowptr[i]; q < rowptr[i + 1]; ++q) {
      const uint64_t p = perm[q];
      for (uint32_t f = 0; f < R; ++f) h[f] = 1.f;
      for (uint32_t m = 0; m < t->nmodes; ++m) {
        if (m == mode) continue;
        const float *row = F[m] + (size_t)t->ind[m][p] * R;
        for (uint32_t f = 0; f < R; ++f) h[f] *= orc_bf16_round(row[f]);
      }
      for (uint32_t f = 0; f < R; ++f) hb[f] = orc_bf16_round(h[f]);
      const double x = t->vals[p];
      for (uint32_t a = 0; a < R; ++a) {
        for (uint32_t b = 0; b <= a; ++b) G[(size_t)a * R + b] += (double)hb[a] * (double)hb[b];
        rhs[a] += x * (double)hb[a];
      }
    }
    for (uint32_t a = 0; a < R; ++a) {
      for (uint32_t b = 0; b < a; ++b) G[(size_t)b * R + a] = G[(size_t)a * R + b];
      G[(size_t)a * R + a] += lambda;
    }
    if (orc_cholesky(G, R) != 0) {
#ifdef _OPENMP
#pragma omp atomic write
#endif
      status = 7;
      for (uint32_t f = 0; f < R; ++f) dst[f] = 0.f;
    } else {
      orc_cholesky_solve(G, R, rhs);
      for (uint32_t f = 0; f < R; ++f) dst[f] = (float)rhs[f];
    }
    free(G);
    free(rhs);
  }
  return status;
}

/* The FP64 ALS update of selected rows only (same rule as orc_als_mode_f64,
 * SG/als.cu:691-716), from float factors: out[r * R + f] for rows[r].  For
 * checking a full-size update on a sample of rows.  Returns 0 or 7. */
/* Rows `rows` of mode `mode`'s loss gradient in FP64 from float factors, as
 * orc_grad_f64 computes it: g_i = -sum over the row's nonzeros p of e_p h_p
 * + lambda c_i a_i, e_p = x_p - <F_0(i0,:), F_1(i1,:), ...>, h_p the product
 * of the other modes' rows, c_i the row's nonzero count (reg_per_occurrence)
 * or 1.  perm / rowptr: orc_mode_order.  out: nrows x R.  Returns 0, or 1 on
 * bad arguments. */
int orc_grad_rows_f64(const orc_coo *t, uint32_t mode, uint32_t R, float *const *F, double lambda,
                      int reg_per_occurrence, const uint32_t *perm, const uint64_t *rowptr,
                      const uint32_t *rows, uint64_t nrows, double *out, int nthreads) {
  if (mode >= t->nmodes || R == 0 || R > 512) return 1;
  (void)nthreads;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads > 0 ? nthreads : 1)
#endif
  for (int64_t r = 0; r < (int64_t)nrows; ++r) {
    const uint32_t i = rows[r];
    double *g = out + (size_t)r * R;
    const float *a = F[mode] + (size_t)i * R;
    double h[512];
    for (uint32_t f = 0; f < R; ++f) g[f] = 0.0;
    for (uint64_t q = rowptr[i]; q < rowptr[i + 1]; ++q) {
      const uint64_t p = perm[q];
      for (uint32_t f = 0; f < R; ++f) h[f] = 1.0;
      for (uint32_t m = 0; m < t->nmodes; ++m) {
        if (m == mode) continue;
        const float *row = F[m] + (size_t)t->ind[m][p] * R;
        for (uint32_t f = 0; f < R; ++f) h[f] *= (double)row[f];
      }
      double e = t->vals[p];
      for (uint32_t f = 0; f < R; ++f) e -= (double)a[f] * h[f];
      for (uint32_t f = 0; f < R; ++f) g[f] -= e * h[f];
    }
    const double c = reg_per_occurrence ? (double)(rowptr[i + 1] - rowptr[i]) : 1.0;
    for (uint32_t f = 0; f < R; ++f) g[f] += lambda * c * (double)a[f];
  }
  return 0;
}

int orc_als_rows_f64(const orc_coo *t, uint32_t mode, uint32_t R, float *const *F, double lambda,
                     const uint32_t *perm, const uint64_t *rowptr, const uint32_t *rows, uint64_t nrows,
                     double *out, int nthreads) {
  if (mode >= t->nmodes || R == 0 || R > 512) return 1;
  int status = 0;
  (void)nthreads;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads > 0 ? nthreads : 1)
#endif
  for (int64_t r = 0; r < (int64_t)nrows; ++r) {
    const uint32_t i = rows[r];
    double *dst = out + (size_t)r * R;
    double *G = (double *)calloc((size_t)R * R, sizeof(double));
    double *rhs = (double *)calloc(R, sizeof(double));
    double h[512];
    for (uint64_t q = rowptr[i]; q < rowptr[i + 1]; ++q) {
      const uint64_t p = perm[q];
      for (uint32_t f = 0; f < R; ++f) h[f] = 1.0;
      for (uint32_t m = 0; m < t->nmodes; ++m) {
        if (m == mode) continue;
        const float *row = F[m] + (size_t)t->ind[m][p] * R;
        for (uint32_t f = 0; f < R; ++f) h[f] *= (double)row[f];
      }
      for (uint32_t a = 0; a < R; ++a) {
        for (uint32_t b = 0; b <= a; ++b) G[(size_t)a * R + b] += h[a] * h[b];
        rhs[a] += t->vals[p] * h[a];
      }
    }
    for (uint32_t a = 0; a < R; ++a) {
      for (uint32_t b = 0; b < a; ++b) G[(size_t)b * R + a] = G[(size_t)a * R + b];
      G[(size_t)a * R + a] += lambda;
    }
    if (rowptr[i] == rowptr[i + 1]) {
      for (uint32_t f = 0; f < R; ++f) dst[f] = 0.0;
    } else if (orc_cholesky(G, R) != 0) {
#ifdef _OPENMP
#pragma omp atomic write
#endif
      status = 7;
      for (uint32_t f = 0; f < R; ++f) dst[f] = 0.0;
    } else {
      orc_cholesky_solve(G, R, rhs);
      for (uint32_t f = 0; f < R; ++f) dst[f] = rhs[f];
    }
    free(G);
    free(rhs);
  }
  return status;
}

int orc_converge(orc_conv *c, double converge_rmse, uint64_t epoch, int *improved) {
  *improved = 0;
  if (converge_rmse - c->best_rmse < -c->tolerance) { /* SG/completion.c:332-341 */
    c->nbadepochs = 0;
    c->best_rmse = converge_rmse;
    c->bestepoch = epoch;
    *improved = 1;
    return 0;
  }
  c->nbadepochs += 1; /* :343-346 */
  return c->nbadepochs == c->max_badepochs;
}

/* ------------------------------------------------------------------------ */
/* CSF (SG/csf.c:11-331)                                                     */
/* ------------------------------------------------------------------------ */

/* stable counting sort of perm[0..n) by key[perm[.]] in [0, dim) */
static int csf_sort_by(const uint32_t *key, uint64_t dim, uint64_t n, uint32_t *perm, uint32_t *tmp) {
  uint64_t *cnt = (uint64_t *)calloc(dim + 1, sizeof(uint64_t));
  if (!cnt) return 1;
  for (uint64_t q = 0; q < n; ++q) cnt[key[perm[q]] + 1]++;
  for (uint64_t d = 0; d < dim; ++d) cnt[d + 1] += cnt[d];
  for (uint64_t q = 0; q < n; ++q) tmp[cnt[key[perm[q]]]++] = perm[q];
  memcpy(perm, tmp, n * sizeof(uint32_t));
  free(cnt);
  return 0;
}

int orc_csf_build(const orc_coo *t, uint32_t mode, orc_csf *out) {
  memset(out, 0, sizeof(*out));
  if (t->nmodes != 3 || mode > 2) return 1;
  /* level order: the reference's firstrank / secondrank / thirdrank
   * (SG/csf.c:30-34, SG/mttkrp.c:27-49) */
  static const uint32_t lv[3][3] = {{0, 1, 2}, {1, 0, 2}, {2, 0, 1}};
  const uint32_t l0 = lv[mode][0], l1 = lv[mode][1], l2 = lv[mode][2];
  const uint64_t n = t->nnz;
  out->mode = mode;
  out->level_mode[0] = l0;
  out->level_mode[1] = l1;
  out->level_mode[2] = l2;
  out->nnz = n;
  for (int m = 0; m < 3; ++m) out->dims[m] = t->dims[m];
  uint32_t *perm = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
  uint32_t *tmp = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
  for (uint64_t q = 0; q < n; ++q) perm[q] = (uint32_t)q;
  /* lexicographic (level 0, level 1), leaves in input order: LSD, least
   * significant key first (p_rearrange, SG/csf.c:148-186) */
  csf_sort_by(t->ind[l1], t->dims[l1], n, perm, tmp);
  csf_sort_by(t->ind[l0], t->dims[l0], n, perm, tmp);
  free(tmp);
  /* count nodes and fibers (p_modecounter, SG/csf.c:11-146) */
  uint64_t n0 = 0, n1 = 0;
  for (uint64_t q = 0; q < n; ++q) {
    const uint32_t p = perm[q];
    const int new0 = q == 0 || t->ind[l0][p] != t->ind[l0][perm[q - 1]];
    const int new1 = new0 || t->ind[l1][p] != t->ind[l1][perm[q - 1]];
    n0 += new0;
    n1 += new1;
  }
  out->nfibs[0] = n0;
  out->nfibs[1] = n1;
  out->nfibs[2] = n;
  out->fids[0] = (uint32_t *)malloc((n0 ? n0 : 1) * sizeof(uint32_t));
  out->fids[1] = (uint32_t *)malloc((n1 ? n1 : 1) * sizeof(uint32_t));
  out->fids[2] = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
  out->fptr[0] = (uint64_t *)malloc((n0 + 1) * sizeof(uint64_t));
  out->fptr[1] = (uint64_t *)malloc((n1 + 1) * sizeof(uint64_t));
  out->vals = (double *)malloc((n ? n : 1) * sizeof(double));
  uint64_t a = 0, b = 0;
  for (uint64_t q = 0; q < n; ++q) {
    const uint32_t p = perm[q];
    const int new0 = q == 0 || t->ind[l0][p] != t->ind[l0][perm[q - 1]];
    const int new1 = new0 || t->ind[l1][p] != t->ind[l1][perm[q - 1]];
    if (new0) {
      out->fids[0][a] = t->ind[l0][p];
      out->fptr[0][a++] = b;
    }
    if (new1) {
      out->fids[1][b] = t->ind[l1][p];
      out->fptr[1][b++] = q;
    }
    out->fids[2][q] = t->ind[l2][p];
    out->vals[q] = t->vals[p];
  }
  out->fptr[0][n0] = n1;
  out->fptr[1][n1] = n;
  free(perm);
  /* 64 groups of ceil(n0 / 64) level-0 nodes (SG/csf.c:75-92) */
  const uint64_t per = (n0 + 63) / 64;
  for (int g = 0; g < 65; ++g) out->blockref[g] = n0;
  if (n0 < 64) {
    for (uint64_t g = 0; g <= n0; ++g) out->blockref[g] = g;
  } else {
    out->blockref[0] = 0;
    for (int g = 1; g < 65; ++g) {
      if (per * (uint64_t)g < n0) {
        out->blockref[g] = per * (uint64_t)g;
      } else {
        out->blockref[g] = n0;
        break;
      }
    }
  }
  return 0;
}

void orc_csf_free(orc_csf *c) {
  for (int l = 0; l < 3; ++l) free(c->fids[l]);
  for (int l = 0; l < 2; ++l) free(c->fptr[l]);
  free(c->vals);
  memset(c, 0, sizeof(*c));
}

/* ------------------------------------------------------------------------ */
/* precision-templated kernels                                               */
/* ------------------------------------------------------------------------ */
#define REAL double
#define SFX(name) name##_f64
#include "oracle_impl.h"
#undef REAL
#undef SFX

#define REAL float
#define SFX(name) name##_f32
#include "oracle_impl.h"
#undef REAL
#undef SFX
