/*
 * oracle_impl.h — precision-templated body of the oracle (included twice by
 * oracle.c with REAL = double and REAL = float).  TEST INFRASTRUCTURE ONLY.
 */

/* prediction <F_0(i0,:), F_1(i1,:), ...> for nonzero p, accumulated in REAL
 * (SG/completion.c:29-51 with the 0-based row index the kernels use,
 * SG/als.cu:256-258; quirk Q6 is not reproduced). */
static inline REAL SFX(orc_predict)(const orc_coo *t, uint32_t R, REAL *const *F, uint64_t p) {
  REAL est = 0;
  for (uint32_t f = 0; f < R; ++f) {
    REAL v = F[0][(size_t)t->ind[0][p] * R + f];
    for (uint32_t m = 1; m < t->nmodes; ++m) v *= F[m][(size_t)t->ind[m][p] * R + f];
    est += v;
  }
  return est;
}

double SFX(orc_loss_sq)(const orc_coo *t, uint32_t R, REAL *const *F, int nthreads) {
  double loss = 0.0;
  (void)nthreads;
#ifdef _OPENMP
#pragma omp parallel for reduction(+ : loss) schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
#endif
  for (int64_t p = 0; p < (int64_t)t->nnz; ++p) {
    double err = t->vals[p] - (double)SFX(orc_predict)(t, R, F, (uint64_t)p);
    loss += err * err;
  }
  return loss;
}

double SFX(orc_frob_sq)(uint32_t nmodes, const uint64_t *dims, uint32_t R,
                        REAL *const *F, double lambda) {
  double reg = 0.0;
  for (uint32_t m = 0; m < nmodes; ++m) {
    double acc = 0.0;
    for (size_t x = 0; x < (size_t)dims[m] * R; ++x) acc += (double)F[m][x] * (double)F[m][x];
    reg += lambda * acc;
  }
  return reg;
}

/* h = hadamard product of the rows of every mode except `mode` for nonzero p */
static inline void SFX(orc_hrow)(const orc_coo *t, uint32_t mode, uint32_t R, REAL *const *F,
                                 uint64_t p, REAL *h) {
  int first = 1;
  for (uint32_t m = 0; m < t->nmodes; ++m) {
    if (m == mode) continue;
    const REAL *row = F[m] + (size_t)t->ind[m][p] * R;
    if (first) {
      for (uint32_t f = 0; f < R; ++f) h[f] = row[f];
      first = 0;
    } else {
      for (uint32_t f = 0; f < R; ++f) h[f] *= row[f];
    }
  }
}

int SFX(orc_mttkrp)(const orc_coo *t, uint32_t mode, uint32_t R, REAL *const *F,
                    const uint32_t *perm, const uint64_t *rowptr, REAL *out, int nthreads) {
  if (mode >= t->nmodes || R == 0) return 1;
  const int64_t dim = (int64_t)t->dims[mode];
  (void)nthreads;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads > 0 ? nthreads : 1)
#endif
  for (int64_t i = 0; i < dim; ++i) {
    REAL h[512];
    REAL *o = out + (size_t)i * R;
    for (uint32_t f = 0; f < R; ++f) o[f] = 0;
    for (uint64_t q = rowptr[i]; q < rowptr[i + 1]; ++q) {
      uint64_t p = perm[q];
      SFX(orc_hrow)(t, mode, R, F, p, h);
      REAL x = (REAL)t->vals[p];
      for (uint32_t f = 0; f < R; ++f) o[f] += x * h[f];
    }
  }
  return 0;
}

int SFX(orc_gram_rows)(const orc_coo *t, uint32_t mode, uint32_t R, REAL *const *F,
                       const uint32_t *perm, const uint64_t *rowptr, uint64_t row0,
                       uint64_t nrows, double *G) {
  if (mode >= t->nmodes || R == 0 || R > 512) return 1;
  REAL h[512];
  for (uint64_t r = 0; r < nrows; ++r) {
    uint64_t i = row0 + r;
    double *g = G + (size_t)r * R * R;
    for (size_t x = 0; x < (size_t)R * R; ++x) g[x] = 0.0;
    for (uint64_t q = rowptr[i]; q < rowptr[i + 1]; ++q) {
      SFX(orc_hrow)(t, mode, R, F, perm[q], h);
      for (uint32_t a = 0; a < R; ++a)
        for (uint32_t b = 0; b < R; ++b) g[(size_t)a * R + b] += (double)(h[a] * h[b]);
    }
  }
  return 0;
}

int SFX(orc_als_mode)(const orc_coo *t, uint32_t mode, uint32_t R, REAL **F, double lambda,
                      const uint32_t *perm, const uint64_t *rowptr, int nthreads) {
  if (mode >= t->nmodes || R == 0 || R > 512) return 1;
  const int64_t dim = (int64_t)t->dims[mode];
  int status = 0;
  /* the reference zeroes the factor (SG/als.cu:699) and then solves only the
   * rows present in the directory (SG/als.cu:481-490, 715-716), so a row
   * with no nonzero stays 0.  The other modes' factors are read while
   * F[mode] is rewritten, which is safe because they are different arrays. */
  (void)nthreads;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
#endif
  for (int64_t i = 0; i < dim; ++i) {
    REAL *dst = F[mode] + (size_t)i * R;
    if (rowptr[i] == rowptr[i + 1]) {
      for (uint32_t f = 0; f < R; ++f) dst[f] = 0;
      continue;
    }
    REAL h[512];
    REAL gacc[512 * 2];
    double *G = (double *)malloc((size_t)R * R * sizeof(double));
    double *rhs = (double *)malloc((size_t)R * sizeof(double));
    REAL *g = (REAL *)malloc((size_t)R * R * sizeof(REAL));
    for (size_t x = 0; x < (size_t)R * R; ++x) g[x] = 0;
    for (uint32_t f = 0; f < R; ++f) gacc[f] = 0;
    for (uint64_t q = rowptr[i]; q < rowptr[i + 1]; ++q) {
      uint64_t p = perm[q];
      SFX(orc_hrow)(t, mode, R, F, p, h);
      REAL x = (REAL)t->vals[p];
      /* lower triangle of h h^T (SG/als.cu:271-291) and x * h (:295-312) */
      for (uint32_t a = 0; a < R; ++a)
        for (uint32_t b = 0; b <= a; ++b) g[(size_t)a * R + b] += h[a] * h[b];
      for (uint32_t f = 0; f < R; ++f) gacc[f] += x * h[f];
    }
    for (uint32_t a = 0; a < R; ++a) {
      for (uint32_t b = 0; b <= a; ++b) {
        G[(size_t)a * R + b] = (double)g[(size_t)a * R + b];
        G[(size_t)b * R + a] = (double)g[(size_t)a * R + b];
      }
      G[(size_t)a * R + a] += lambda; /* p_update_matrice, SG/als.cu:484-487 */
      rhs[a] = (double)gacc[a];
    }
    if (orc_cholesky(G, R) != 0) {
#ifdef _OPENMP
#pragma omp atomic write
#endif
      status = 7;
      for (uint32_t f = 0; f < R; ++f) dst[f] = 0;
    } else {
      orc_cholesky_solve(G, R, rhs);
      for (uint32_t f = 0; f < R; ++f) dst[f] = (REAL)rhs[f];
    }
    free(G);
    free(rhs);
    free(g);
  }
  return status;
}

void SFX(orc_sgd_epoch)(const orc_coo *t, const uint32_t *order, uint32_t R, REAL **F,
                        double lr, double lambda, double scale) {
  const uint32_t N = t->nmodes;
  const REAL step = (REAL)(lr * scale);
  const REAL lam = (REAL)lambda;
  REAL buf[ORC_MAX_MODES][512];
  for (uint64_t q = 0; q < t->nnz; ++q) {
    uint64_t p = order ? order[q] : q;
    REAL *rows[ORC_MAX_MODES];
    for (uint32_t m = 0; m < N; ++m) {
      rows[m] = F[m] + (size_t)t->ind[m][p] * R;
      for (uint32_t f = 0; f < R; ++f) buf[m][f] = rows[m][f]; /* SG/sgd.cu:75-84 */
    }
    REAL pred = 0; /* :86-89 */
    for (uint32_t f = 0; f < R; ++f) {
      REAL v = buf[0][f];
      for (uint32_t m = 1; m < N; ++m) v *= buf[m][f];
      pred += v;
    }
    const REAL err = (REAL)t->vals[p] - pred; /* :90 */
    for (uint32_t f = 0; f < R; ++f) {          /* :92-99 */
      for (uint32_t m = 0; m < N; ++m) {
        REAL others = 1;
        for (uint32_t o = 0; o < N; ++o)
          if (o != m) others *= buf[o][f];
        REAL mod = err * others - lam * buf[m][f];
        rows[m][f] += step * mod;
      }
    }
  }
}

/* One epoch of the bulk-synchronous SGD schedule (libctc's default SGD).
 *
 * The reference's kernel gives every thread one tile of nonzeros in mode-0
 * order and lets all of them add their updates with atomicAdd while the
 * others read (SG/sgd.cu:21-143, :96-98); its multi-GPU version gives each
 * replica a tile range and averages the replicas after the sweep
 * (NC/sgd.cu:229-262, launched at :512, divides every replica by the GPU
 * count; :523-525 sum the replicas with ncclAllReduce).  This schedule fixes which values each update
 * reads, so that one restatement describes every run:
 *   - walker w owns positions [w*span, (w+1)*span) of the walk `order`;
 *   - sub-step s runs positions [w*span + s*sub, +sub) of every walker's span,
 *     all from the factors as they stood when the sub-step began;
 *   - a walker keeps its own copy of a row along a run of consecutive
 *     nonzeros with the same index in that mode (read from the sub-step's
 *     factors where the run starts; a later run on the same row in the same
 *     sub-step starts from them again), applies the reference step
 *     (SG/sgd.cu:86-99) to its copies, and hands in the run's delta when the
 *     run ends;
 *   - when the sub-step ends, a cold row adds the deltas of all its runs (the
 *     reference's atomicAdd) and a hot row adds their mean (replica
 *     averaging: it becomes the mean of the runs' copies).  A row of mode m is
 *     hot when it has more than hot_rows nonzeros in t and more than 4x
 *     t->nnz / dims[m].
 * With span >= nnz and sub = 1 this is the sequential rule of orc_sgd_epoch.
 * nthreads > 1 runs the walkers of a sub-step in parallel (the deltas are
 * then added in FP64 in thread order, a difference of FP64 rounding only). */
void SFX(orc_sgd_sync_epoch)(const orc_coo *t, const uint32_t *order, uint32_t R, REAL **F,
                             double lr, double lambda, double scale, uint64_t span, uint64_t sub,
                             uint64_t hot_rows, int bf16, int nthreads) {
  const uint32_t N = t->nmodes;
  const uint64_t n = t->nnz;
  if (n == 0 || span == 0 || sub == 0 || R > 512) return;
  const REAL step = (REAL)(lr * scale);
  const REAL lam = (REAL)lambda;
  const uint64_t W = (n + span - 1) / span;
  const uint64_t nsub = (span + sub - 1) / sub;
  int nth = nthreads > 0 ? nthreads : 1;
  double *D[ORC_MAX_MODES];
  uint32_t *nrun[ORC_MAX_MODES];
  unsigned char *hot[ORC_MAX_MODES];
  for (uint32_t m = 0; m < N; ++m) {
    D[m] = (double *)calloc((size_t)t->dims[m] * R, sizeof(double));
    nrun[m] = (uint32_t *)calloc((size_t)t->dims[m], sizeof(uint32_t));
    hot[m] = (unsigned char *)calloc((size_t)t->dims[m], 1);
    uint64_t *cnt = (uint64_t *)calloc((size_t)t->dims[m], sizeof(uint64_t));
    for (uint64_t p = 0; p < n; ++p) cnt[t->ind[m][p]]++;
    const double rel = 4.0 * (double)n / (double)t->dims[m];
    for (uint64_t i = 0; i < t->dims[m]; ++i) hot[m][i] = cnt[i] > hot_rows && (double)cnt[i] > rel;
    free(cnt);
  }
  for (uint64_t s = 0; s < nsub; ++s) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nth)
#endif
    for (int64_t w = 0; w < (int64_t)W; ++w) {
      const uint64_t s1 = ((uint64_t)w + 1) * span < n ? ((uint64_t)w + 1) * span : n;
      const uint64_t e0 = (uint64_t)w * span + s * sub;
      const uint64_t e1 = e0 + sub < s1 ? e0 + sub : s1;
      REAL v[ORC_MAX_MODES][512], del[ORC_MAX_MODES][512];
      uint32_t cur[ORC_MAX_MODES];
      int live[ORC_MAX_MODES] = {0, 0, 0, 0};
      for (uint64_t q = e0; q < e1; ++q) {
        const uint64_t p = order ? order[q] : q;
        for (uint32_t m = 0; m < N; ++m) {
          const uint32_t ix = t->ind[m][p];
          if (live[m] && ix == cur[m]) continue;
          if (live[m]) { /* the run on cur[m] ends: hand in its delta */
            double *d = D[m] + (size_t)cur[m] * R;
            for (uint32_t f = 0; f < R; ++f) {
              const double dv = bf16 && m > 0 ? (double)orc_bf16_round((float)del[m][f]) : (double)del[m][f];
#ifdef _OPENMP
#pragma omp atomic
#endif
              d[f] += dv;
            }
#ifdef _OPENMP
#pragma omp atomic
#endif
            nrun[m][cur[m]] += 1;
          }
          live[m] = 1;
          cur[m] = ix;
          for (uint32_t f = 0; f < R; ++f) {
            v[m][f] = F[m][(size_t)ix * R + f];
            del[m][f] = 0;
          }
        }
        REAL pred = 0; /* SG/sgd.cu:86-89 */
        for (uint32_t f = 0; f < R; ++f) {
          REAL x = v[0][f];
          for (uint32_t m = 1; m < N; ++m) x *= v[m][f];
          pred += x;
        }
        const REAL err = (REAL)t->vals[p] - pred; /* :90 */
        for (uint32_t f = 0; f < R; ++f) {          /* :92-99, pre-update values */
          REAL d[ORC_MAX_MODES];
          for (uint32_t m = 0; m < N; ++m) {
            REAL others = 1;
            for (uint32_t o = 0; o < N; ++o)
              if (o != m) others *= v[o][f];
            d[m] = step * (err * others - lam * v[m][f]);
          }
          for (uint32_t m = 0; m < N; ++m) {
            v[m][f] += d[m];
            del[m][f] += d[m];
          }
        }
      }
      for (uint32_t m = 0; m < N; ++m) {
        if (!live[m]) continue;
        double *d = D[m] + (size_t)cur[m] * R;
        for (uint32_t f = 0; f < R; ++f) {
          const double dv = bf16 && m > 0 ? (double)orc_bf16_round((float)del[m][f]) : (double)del[m][f];
#ifdef _OPENMP
#pragma omp atomic
#endif
          d[f] += dv;
        }
#ifdef _OPENMP
#pragma omp atomic
#endif
        nrun[m][cur[m]] += 1;
      }
    }
    /* merge: every row that took a run this sub-step */
    for (uint32_t m = 0; m < N; ++m) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nth)
#endif
      for (int64_t i = 0; i < (int64_t)t->dims[m]; ++i) {
        const uint32_t c = nrun[m][i];
        if (!c) continue;
        const double sc = hot[m][i] ? 1.0 / (double)c : 1.0;
        REAL *g = F[m] + (size_t)i * R;
        double *d = D[m] + (size_t)i * R;
        for (uint32_t f = 0; f < R; ++f) {
          g[f] = (REAL)((double)g[f] + d[f] * sc);
          d[f] = 0.0;
        }
        nrun[m][i] = 0;
      }
    }
  }
  for (uint32_t m = 0; m < N; ++m) {
    free(D[m]);
    free(nrun[m]);
    free(hot[m]);
  }
}

int SFX(orc_mttkrp_csf)(const orc_csf *c, uint32_t R, REAL *const *F, REAL *out, int nthreads) {
  if (R == 0 || R > 512) return 1;
  const REAL *B = F[c->level_mode[1]], *Cm = F[c->level_mode[2]];
  memset(out, 0, (size_t)c->dims[c->mode] * R * sizeof(REAL));
  (void)nthreads;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
#endif
  for (int g = 0; g < 64; ++g) {
    REAL sum[512];
    for (uint64_t a = c->blockref[g]; a < c->blockref[g + 1]; ++a) {
      REAL *o = out + (size_t)c->fids[0][a] * R;
      for (uint64_t f = c->fptr[0][a]; f < c->fptr[0][a + 1]; ++f) {
        for (uint32_t r = 0; r < R; ++r) sum[r] = 0; /* SG/mttkrp.c:67-78 */
        for (uint64_t q = c->fptr[1][f]; q < c->fptr[1][f + 1]; ++q) {
          const REAL x = (REAL)c->vals[q];
          const REAL *cr = Cm + (size_t)c->fids[2][q] * R;
          for (uint32_t r = 0; r < R; ++r) sum[r] += x * cr[r];
        }
        const REAL *br = B + (size_t)c->fids[1][f] * R; /* :79-82 */
        for (uint32_t r = 0; r < R; ++r) o[r] += sum[r] * br[r];
      }
    }
  }
  return 0;
}

void SFX(orc_residual)(const orc_coo *t, uint32_t R, REAL *const *F, REAL *res) {
  for (uint64_t p = 0; p < t->nnz; ++p)
    res[p] = (REAL)t->vals[p] - SFX(orc_predict)(t, R, F, p);
}

/* rank-one component f at nonzero p */
static inline REAL SFX(orc_comp)(const orc_coo *t, uint32_t R, REAL *const *F, uint32_t f,
                                 uint64_t p) {
  REAL v = F[0][(size_t)t->ind[0][p] * R + f];
  for (uint32_t m = 1; m < t->nmodes; ++m) v *= F[m][(size_t)t->ind[m][p] * R + f];
  return v;
}

int SFX(orc_ccd_epoch)(const orc_coo *t, uint32_t R, REAL **F, double lambda, uint32_t inner,
                       REAL *res, uint32_t *const *perms, uint64_t *const *rowptrs,
                       int nthreads) {
  const uint32_t N = t->nmodes;
  (void)nthreads;
  for (uint32_t f = 0; f < R; ++f) {
    /* add the current rank-one component back to the residual */
    for (uint64_t p = 0; p < t->nnz; ++p) res[p] += SFX(orc_comp)(t, R, F, f, p);
    for (uint32_t it = 0; it < inner; ++it) {
      for (uint32_t m = 0; m < N; ++m) {
        const int64_t dim = (int64_t)t->dims[m];
        const uint32_t *perm = perms[m];
        const uint64_t *rowptr = rowptrs[m];
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads > 0 ? nthreads : 1)
#endif
        for (int64_t i = 0; i < dim; ++i) {
          REAL num = 0, den = 0;
          for (uint64_t q = rowptr[i]; q < rowptr[i + 1]; ++q) {
            uint64_t p = perm[q];
            REAL h = 1;
            for (uint32_t o = 0; o < N; ++o)
              if (o != m) h *= F[o][(size_t)t->ind[o][p] * R + f];
            num += res[p] * h;
            den += h * h;
          }
          /* a_if <- num / (lambda + den) (SG/ccd.cu:345, p_compute_newcol) */
          F[m][(size_t)i * R + f] = (REAL)((double)num / (lambda + (double)den));
        }
      }
    }
    /* subtract the updated component */
    for (uint64_t p = 0; p < t->nnz; ++p) res[p] -= SFX(orc_comp)(t, R, F, f, p);
  }
  return 0;
}

double SFX(orc_grad)(const orc_coo *t, uint32_t R, REAL *const *F, double lambda,
                     int reg_per_occurrence, REAL **G) {
  const uint32_t N = t->nmodes;
  double loss = 0.0;
  for (uint32_t m = 0; m < N; ++m)
    for (size_t x = 0; x < (size_t)t->dims[m] * R; ++x) G[m][x] = 0;
  REAL h[512];
  for (uint64_t p = 0; p < t->nnz; ++p) {
    const REAL e = (REAL)t->vals[p] - SFX(orc_predict)(t, R, F, p);
    loss += 0.5 * (double)e * (double)e;
    for (uint32_t m = 0; m < N; ++m) {
      SFX(orc_hrow)(t, m, R, F, p, h);
      REAL *g = G[m] + (size_t)t->ind[m][p] * R;
      const REAL *a = F[m] + (size_t)t->ind[m][p] * R;
      for (uint32_t f = 0; f < R; ++f) {
        g[f] -= e * h[f];
        if (reg_per_occurrence) g[f] += (REAL)lambda * a[f];
      }
      if (reg_per_occurrence)
        for (uint32_t f = 0; f < R; ++f) loss += 0.5 * lambda * (double)a[f] * (double)a[f];
    }
  }
  if (!reg_per_occurrence) {
    for (uint32_t m = 0; m < N; ++m) {
      for (size_t x = 0; x < (size_t)t->dims[m] * R; ++x) {
        G[m][x] += (REAL)lambda * F[m][x];
        loss += 0.5 * lambda * (double)F[m][x] * (double)F[m][x];
      }
    }
  }
  return loss;
}
