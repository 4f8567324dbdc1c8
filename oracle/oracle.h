/*
 * oracle.h — CPU restatement of abovekumo/cuTC's tensor-completion rules.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it may be loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by
 * nothing else.  The product path (cutc_amd / libctc) never links or calls it.
 *
 * Every routine exists twice: *_f64 (the FP64 oracle, single-threaded unless
 * nthreads > 1 is passed where the result does not depend on thread count)
 * and *_f32 (the same source in FP32, used as the timed CPU baseline).
 *
 * Index conventions: indices are 0-based uint32, factors are row-major
 * dims[m] x R arrays.  "sorted by mode m" means lexicographic order over the
 * cyclic mode sequence (m, m+1, ..., m+N-1) mod N, the order the reference's
 * CISS builder expects its input in (SG/ciss.c:74, software/scripts/stensor.py:123-124).
 */
#ifndef CTC_ORACLE_H
#define CTC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_MODES 4

typedef struct orc_coo {
  uint32_t nmodes;
  uint64_t nnz;
  uint64_t dims[ORC_MAX_MODES];
  uint32_t *ind[ORC_MAX_MODES];
  double *vals;
  uint64_t offset[ORC_MAX_MODES]; /* index base found per mode by the text reader */
} orc_coo;

/* ---- .tns text reader (SG/io.c:58-106, 234-308) ---------------------------
 * Returns 0 on success, 2 if the file cannot be opened (reference returns
 * NULL, SG/io.c:199-202), 4 if a mode's smallest index is neither 0 nor 1
 * (reference exit(1), SG/io.c:291-296), 3 on an empty tensor. */
int orc_tns_read(const char *path, orc_coo *out);
void orc_coo_free(orc_coo *t);

/* ---- ordering ------------------------------------------------------------
 * perm[p] = position in t of the p-th nonzero in mode-m order (stable LSD
 * counting sort, least significant key first).  rowptr has dims[m]+1 entries
 * (CSR over all rows, empty rows included). */
int orc_mode_order(const orc_coo *t, uint32_t mode, uint32_t *perm, uint64_t *rowptr);

/* ---- the reference's tile-reduction defects as weights ---------------------
 * How many times each nonzero's term reaches its row's sums in the
 * reference's warp-per-tile reductions (tiles of tile_len = 32 consecutive
 * nonzeros of the mode-m order perm, one lane each; SG/als.cu:226-312):
 *   flags & 1 (Q1): lane 0's mybit is forced to 1 and only lanes with
 *     mybit = 0 write (SG/als.cu:245-249,285-288,307-310), so the tile's
 *     first row segment is never written: weight 0;
 *   flags & 2 (Q2): __shfl_down_sync from the last lane returns its own value
 *     (SG/als.cu:279-284), so in a full tile whose last lane continues a row
 *     segment that starts at lane h, the head receives the last lane's term
 *     itercounter - (tile_len - 1 - h) extra times, itercounter being the
 *     number of continuation lanes 1..tile_len-1 (SG/als.cu:233).
 *   flags & 16 (Q3, CCD only): the CCD launch has tilenum / 256 + 1 blocks of
 *     256 / tile_len warps, one tile each, tilenum = nnz / tile_len + 1
 *     (SG/ccd.cu:520): the tiles beyond have weight 0.
 * A partial last tile's last lane reads an inactive lane (undefined in the
 * reference); it gets no extra.  w[p] is indexed by the nonzero's position in
 * t.  Returns 0, or 1 on bad arguments. */
int orc_quirk_weights(const orc_coo *t, uint32_t mode, const uint32_t *perm, uint32_t tile_len,
                      uint32_t flags, uint32_t *w);
/* The reference's CCD update of one mode in FP64 (SG/ccd.cu:82-206,318-348;
 * quirk Q4: no add-back, every column at once): for each row i with a nonzero
 * in t, e_p = x_p - <F_mode(i,:), h_p> with the row's values before the update,
 * F_mode(i,f) = sum w_p e_p h_pf / (lambda + sum w_p h_pf^2) over the row's
 * nonzeros (w NULL: weights 1; x truncated toward zero when trunc_vals); rows
 * without nonzeros are kept.  perm / rowptr: orc_mode_order.  Returns 0, or 1
 * on bad arguments. */
int orc_ccd_jacobi_mode(const orc_coo *t, uint32_t mode, uint32_t R, double **F, double lambda,
                        const uint32_t *w, int trunc_vals, const uint32_t *perm, const uint64_t *rowptr);

/* ---- reference CISS (TB-COO) layout (SG/ciss.c:68-195) --------------------
 * Input: t, already sorted by `mode` (the reference requires it, SG/ciss.c:74).
 * Writes the reference's arrays exactly as ciss_alloc fills them:
 *   directory[dlength]  raw (1-based) ids of the distinct mode indices
 *   dcounter[dlength+1] cumulative last-nonzero index per directory slot
 *   entries[size*3]     tile records; fields ciss_alloc leaves unwritten are
 *                       left as the caller initialised them.
 * tile_len is DEFAULT_T_TILE_LENGTH (SG/base.h:62).  truncate_vals != 0
 * reproduces the reference's storage of values through an integer (quirk Q5,
 * SG/ciss.c:123,157,176).  Capacities: directory/dcounter >= nnz+1, entries
 * >= 3*(nnz/tile_len + 1 + nnz + 1).  Returns 0, or 1 on bad arguments. */
int orc_ciss_reference_layout(const orc_coo *t, uint32_t mode, uint32_t tile_len,
                              int truncate_vals, uint64_t *directory,
                              uint64_t *dcounter, double *entries,
                              uint64_t *dlength, uint64_t *size);

/* ---- CSF (compressed sparse fiber) layout, 3-way (SG/csf.c:11-331) --------
 * Level order of the CSF for `mode` (the reference's p_rearrange /
 * p_modecounter, SG/csf.c:60-195): mode 0 (i, j, k), mode 1 (j, i, k),
 * mode 2 (k, i, j).  fids[0] are the distinct level-0 indices, fptr[0][n] the
 * first level-1 fiber of level-0 node n; fids[1] the level-1 index of each
 * fiber, fptr[1][f] its first nonzero; fids[2] and vals the nonzeros.  Ids are
 * 0-based (the reference stores 1-based ids).  blockref[0..64] splits the
 * level-0 nodes into 64 groups of ceil(nfibs[0] / 64) nodes
 * (SG/csf.c:75-92): blockref[g] is the first node of group g, and the entries
 * after the last group hold nfibs[0] (the reference leaves them 0; either way
 * those groups are empty); nodes [blockref[g], blockref[g + 1]) of a group are
 * processed by one thread.  Leaves keep the input order within a
 * fiber (the reference's quicksort leaves it unspecified). */
typedef struct orc_csf {
  uint32_t mode;
  uint32_t level_mode[3];
  uint64_t nnz;
  uint64_t dims[3];
  uint64_t nfibs[3];
  uint64_t *fptr[2];
  uint32_t *fids[3];
  double *vals;
  uint64_t blockref[65];
} orc_csf;
/* Returns 0, or 1 if t is not 3-way. */
int orc_csf_build(const orc_coo *t, uint32_t mode, orc_csf *out);
void orc_csf_free(orc_csf *c);

/* ---- precision-templated kernels ------------------------------------------ */
#define ORC_DECLARE(REAL, SFX)                                                         \
  /* loss = sum (x - <a_i,b_j,c_k,...>)^2 over t (SG/completion.c:166-189) */          \
  double orc_loss_sq##SFX(const orc_coo *t, uint32_t R, REAL *const *F, int nthreads); \
  /* lambda * sum_m ||F_m||_F^2 (SG/completion.c:193-214) */                           \
  double orc_frob_sq##SFX(uint32_t nmodes, const uint64_t *dims, uint32_t R,          \
                          REAL *const *F, double lambda);                              \
  /* out[i,:] = sum_{nnz in row i of mode} x * (hadamard of other rows)              \
   * (SG/als.cu:295-312 intent) */                                                     \
  int orc_mttkrp##SFX(const orc_coo *t, uint32_t mode, uint32_t R, REAL *const *F,    \
                      const uint32_t *perm, const uint64_t *rowptr, REAL *out,         \
                      int nthreads);                                                   \
  /* per-row Gram G_i = sum h h^T (full R x R, SG/als.cu:271-291 intent) */            \
  int orc_gram_rows##SFX(const orc_coo *t, uint32_t mode, uint32_t R, REAL *const *F, \
                         const uint32_t *perm, const uint64_t *rowptr, uint64_t row0,  \
                         uint64_t nrows, double *G);                                   \
  /* one ALS mode update: F[mode] <- rows of (G_i + lambda I)^-1 m_i; rows with no   \
   * nonzero become 0.  Returns 0, or 7 if a system is not positive definite.        \
   * (SG/als.cu:691-716) */                                                            \
  int orc_als_mode##SFX(const orc_coo *t, uint32_t mode, uint32_t R, REAL **F,        \
                        double lambda, const uint32_t *perm, const uint64_t *rowptr,   \
                        int nthreads);                                                 \
  /* one SGD epoch, sequential over `order` (SG/sgd.cu:21-143 rule: all rows          \
   * updated from their pre-update values, step lr*scale) */                          \
  void orc_sgd_epoch##SFX(const orc_coo *t, const uint32_t *order, uint32_t R,         \
                          REAL **F, double lr, double lambda, double scale);          \
  /* one SGD epoch of the bulk-synchronous walker schedule (libctc's default):      \
   * walkers of `span` positions of `order`, sub-steps of `sub` nonzeros from the      \
   * sub-step's factors, per-run deltas summed (cold rows) or averaged (rows with     \
   * more than hot_rows and 4x the mean nonzeros); bf16 != 0: the mixed-precision   \
   * variant (run deltas of modes 1.. rounded through FP32 to BF16 before they are    \
   * summed); see oracle_impl.h */                                                    \
  void orc_sgd_sync_epoch##SFX(const orc_coo *t, const uint32_t *order, uint32_t R,    \
                               REAL **F, double lr, double lambda, double scale,       \
                               uint64_t span, uint64_t sub, uint64_t hot_rows,         \
                               int bf16, int nthreads);                                \
  /* MTTKRP of the CSF's mode into out (dims[mode] x R, overwritten), the         \
   * reference's serial mttkrp_csf (SG/mttkrp.c:16-93, with the free of the       \
   * undeclared `sumbuffer` at :91 read as the free of `sum`): per level-1 fiber   \
   * s = sum of x * C(k,:) over its nonzeros, then out(i,:) += s * B(j,:); the     \
   * 64 groups of level-0 nodes run on nthreads threads */                          \
  int orc_mttkrp_csf##SFX(const orc_csf *c, uint32_t R, REAL *const *F, REAL *out,    \
                          int nthreads);                                               \
  /* residual r = x - prediction (CCD++ set-up) */                                     \
  void orc_residual##SFX(const orc_coo *t, uint32_t R, REAL *const *F, REAL *res);     \
  /* one CCD++ epoch (canonical, SG/ccd.cu:532-554, NC/ccd.cu:28-180).               \
   * res holds x - prediction on entry and on exit.  Returns 0. */                     \
  int orc_ccd_epoch##SFX(const orc_coo *t, uint32_t R, REAL **F, double lambda,       \
                         uint32_t inner, REAL *res, uint32_t *const *perms,            \
                         uint64_t *const *rowptrs, int nthreads);                      \
  /* gradient of 0.5*sum e^2 + 0.5*lambda*sum_i c_i ||F_m(i,:)||^2;                   \
   * reg_per_occurrence: c_i = nnz in row i (the summed reference SGD step,          \
   * SG/sgd.cu:93-95), else c_i = 1.  Returns the loss value. */                       \
  double orc_grad##SFX(const orc_coo *t, uint32_t R, REAL *const *F, double lambda,    \
                       int reg_per_occurrence, REAL **G);

ORC_DECLARE(double, _f64)
ORC_DECLARE(float, _f32)

/* ---- dense helpers -------------------------------------------------------- */
/* In-place lower Cholesky of an n x n row-major SPD matrix (Cholesky-
 * Banachiewicz, Golub & Van Loan 4.2).  Returns 0, or 7 if not SPD. */
int orc_cholesky(double *A, uint32_t n);
/* Solve L L^T x = b in place (b -> x) given the lower factor from orc_cholesky. */
void orc_cholesky_solve(const double *L, uint32_t n, double *b);

/* ---- BF16-operand ALS (mixed-precision path) ------------------------------ */
/* float -> bfloat16 -> float, round to nearest even */
float orc_bf16_round(float x);
/* ALS mode update with BF16-rounded factor rows and BF16-rounded h in the Gram and m,
 * FP64 accumulation and solve; float factors.  Returns 0 or 7 (not SPD). */
int orc_als_mode_bf16(const orc_coo *t, uint32_t mode, uint32_t R, float **F, double lambda,
                      const uint32_t *perm, const uint64_t *rowptr, int nthreads);

/* Rows of one mode's loss gradient in FP64 from float factors (sampled check
 * of a full-size gradient), as orc_grad_f64: out[r * R + f] for rows[r].
 * Returns 0, or 1 on bad arguments. */
int orc_grad_rows_f64(const orc_coo *t, uint32_t mode, uint32_t R, float *const *F, double lambda,
                      int reg_per_occurrence, const uint32_t *perm, const uint64_t *rowptr,
                      const uint32_t *rows, uint64_t nrows, double *out, int nthreads);

/* FP64 ALS update of selected rows from float factors (sampled check of a
 * full-size update): out[r * R + f] for rows[r].  Returns 0 or 7. */
int orc_als_rows_f64(const orc_coo *t, uint32_t mode, uint32_t R, float *const *F, double lambda,
                     const uint32_t *perm, const uint64_t *rowptr, const uint32_t *rows, uint64_t nrows,
                     double *out, int nthreads);

/* ---- convergence (SG/completion.c:296-362) -------------------------------- */
typedef struct orc_conv {
  double best_rmse;
  double tolerance;
  uint64_t nbadepochs;
  uint64_t bestepoch;
  uint64_t max_badepochs;
} orc_conv;
/* Returns 1 when converged; *improved is set when the model is a new best. */
int orc_converge(orc_conv *c, double converge_rmse, uint64_t epoch, int *improved);

#ifdef __cplusplus
}
#endif
#endif
