/*
 * ctc.h — C ABI of libctc, the MI355X (gfx950) tensor-completion library.
 *
 * This is the drop-in boundary for abovekumo/cuTC's hot path.  Each entry
 * point names the reference interface it replaces (paths relative to the
 * reference root; SG = "software/single GPU/sgpu", GA =
 * "software/autotuning/performance_GA").  All functions return ctc_status and
 * never exit the process (the reference exit()s on every error, SG/base.h:79-90).
 *
 * Plain C types only: no HIP or torch types appear in any signature.  Device
 * pointers are passed as void* / float* and streams as void* (hipStream_t).
 *
 * Threading: a ctc_context owns one HIP stream; calls on one context (and the
 * models created from it) must not overlap.  Different contexts are
 * independent.
 */
#ifndef CTC_H
#define CTC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CTC_MAX_MODES 4
#define CTC_TILE_LEN 64 /* nonzeros per CISS tile: one wave64 load */

typedef enum ctc_status {
  CTC_OK = 0,
  CTC_ERR_INVALID_ARG = 1,
  CTC_ERR_IO = 2,         /* file cannot be opened (SG/io.c:199-202 returns NULL) */
  CTC_ERR_PARSE = 3,      /* malformed or empty tensor file */
  CTC_ERR_INDEX_BASE = 4, /* index base other than 0/1 */
  CTC_ERR_NOMEM = 5,
  CTC_ERR_HIP = 6,        /* HIP runtime error (SG/base.h:79-90 exit(-1)) */
  CTC_ERR_NOT_SPD = 7,    /* an ALS normal system is not positive definite
                             (potrf info, ignored by SG/als.cu:715) */
  CTC_ERR_UNSUPPORTED = 8,
  CTC_ERR_NO_DEVICE = 9
} ctc_status;

const char *ctc_status_string(ctc_status s);
/* Detail of the last error raised on the calling thread ("" if none). */
const char *ctc_last_error(void);
/* Library version string. */
const char *ctc_version(void);

/* ------------------------------------------------------------------------ */
/* COO tensors — replaces sptensor_t (SG/sptensor.h:14-26), tt_read         */
/* (SG/sptensor.c:64-70 -> SG/io.c:195-216), tt_free (SG/sptensor.c:145-159) */
/* ------------------------------------------------------------------------ */
typedef struct ctc_coo {
  uint32_t nmodes;
  uint64_t nnz;
  uint64_t dims[CTC_MAX_MODES];
  uint32_t *ind[CTC_MAX_MODES]; /* 0-based indices, nnz each */
  double *vals;                 /* nnz values */
  uint32_t base[CTC_MAX_MODES]; /* index base found in the file (0 or 1) */
} ctc_coo;

/* Allocates a zero-filled tensor with the given shape. */
ctc_status ctc_coo_alloc(uint32_t nmodes, uint64_t nnz, const uint64_t *dims, ctc_coo **out);
/* Reads a .tns text file: one "i j k ... value" per line, '#' comments and
 * blank lines skipped, per-mode index base 0 or 1, dims = largest index
 * (SG/io.c:58-106, 234-308).  Replaces tt_read. */
ctc_status ctc_coo_read(const char *path, ctc_coo **out);
/* Writes a .tns text file with the given index base; decimals < 0 writes
 * values with %.17g, otherwise with %.<decimals>f (software/scripts/tvtensor.py:76-82
 * uses %.1f). */
ctc_status ctc_coo_write(const ctc_coo *t, const char *path, uint32_t base, int decimals);
/* Writes a SPLATT binary coordinate file (tt_write_binary_file, SG/io.c:369-420;
 * header SG/io.c:116-152): 4-byte indices and values where they are exact,
 * 8-byte otherwise; indices stored with the given base.  ctc_coo_read reads
 * files ending in .bin in this format (SG/io.c:162-189, 430-513, with the
 * narrow-index path that the reference cannot take fixed). */
ctc_status ctc_coo_write_binary(const ctc_coo *t, const char *path, uint32_t base);
/* Data preparation, replacing software/scripts/{tvtensor,stensor}.py:
 * ctc_coo_sort — stable lexicographic sort by the given modes (first key most
 * significant); stensor.py:36-37 writes the training tensor sorted by (1, 2)
 * and (2, 0), tvtensor.py:64 sorts by (0, 1).
 * ctc_coo_split — seeded 80/10/10 split of the nonzeros taken in (0, 1)
 * order, one draw from 0..9 per nonzero: 0-7 train, 8 test, 9 validation
 * (tvtensor.py:73-81, which draws from an unseeded generator). */
ctc_status ctc_coo_sort(const ctc_coo *t, const uint32_t *keys, uint32_t nkeys, ctc_coo **out);
ctc_status ctc_coo_split(const ctc_coo *t, uint64_t seed, ctc_coo **train, ctc_coo **validation,
                         ctc_coo **test);
void ctc_coo_free(ctc_coo *t);

/* ------------------------------------------------------------------------ */
/* Synthetic tensors (the build's own seeded generator; replaces the         */
/* reference's unseeded data-prep scripts software/scripts/tvtensor.py:56-82)*/
/* ------------------------------------------------------------------------ */
typedef enum ctc_value_map {
  CTC_VALUES_RAW = 0,       /* CP model value + noise */
  CTC_VALUES_AFFINE_1_5 = 1, /* CP value mapped affinely onto [1,5], + noise */
  CTC_VALUES_INT_1_5 = 2     /* as AFFINE_1_5, then rounded and clipped to {1..5} */
} ctc_value_map;

typedef struct ctc_synth_spec {
  uint32_t nmodes;
  uint64_t dims[CTC_MAX_MODES];
  uint64_t nnz;                   /* distinct coordinates */
  double zipf[CTC_MAX_MODES];     /* Zipf exponent per mode; 0 = uniform */
  uint32_t true_rank;             /* rank of the CP ground truth (entries U(0,1)) */
  int32_t value_map;              /* ctc_value_map */
  double noise_std;
  double test_fraction;           /* exact count round(fraction * nnz) goes to test */
  uint64_t seed;
} ctc_synth_spec;

/* Generates train (sorted by mode 0) and test tensors.  test may be NULL when
 * test_fraction == 0. */
ctc_status ctc_synth_generate(const ctc_synth_spec *spec, ctc_coo **train, ctc_coo **test);

/* ------------------------------------------------------------------------ */
/* CISS (TB-COO) — replaces ciss_t / ciss_alloc (SG/ciss.h:19-29,            */
/* SG/ciss.c:68-195) and ciss_free (SG/ciss.c:254-260)                       */
/* ------------------------------------------------------------------------ */
typedef struct ctc_ciss ctc_ciss;

/* Host view of a built CISS.  Nonzeros are in lexicographic order of modes
 * (mode, mode+1, ..., mode+N-1) mod N.  Tile t covers nonzeros
 * [64 t, min(64 t + 64, nnz)); bit e of tile_mask[t] is set when nonzero
 * 64 t + e starts a new directory slot (bit 0 is also set for the first
 * nonzero of the tensor); tile_slot[t] is the slot of the tile's first
 * nonzero. */
typedef struct ctc_ciss_view {
  uint32_t nmodes, mode;
  uint64_t nnz, ntiles, dlength;
  uint64_t dims[CTC_MAX_MODES];
  const uint32_t *directory;                 /* dlength row ids (0-based) */
  const uint32_t *rowptr;                    /* dlength + 1 offsets */
  const uint32_t *oidx[CTC_MAX_MODES - 1];   /* indices of modes mode+1, mode+2, ... */
  const float *vals;
  const uint32_t *tile_slot;                 /* ntiles */
  const uint64_t *tile_mask;                 /* ntiles */
  const uint32_t *perm;                      /* source COO position of each nonzero */
} ctc_ciss_view;

/* Builds the CISS of `mode`; the input need not be sorted (the reference
 * requires pre-sorted input, SG/ciss.c:74, quirk Q12). */
ctc_status ctc_ciss_build(const ctc_coo *t, uint32_t mode, ctc_ciss **out);
ctc_status ctc_ciss_get_view(const ctc_ciss *c, ctc_ciss_view *view);
/* Emits the reference's own array-of-doubles layout (SG/ciss.c:89-188) for
 * a reference tile length L (1..52): directory (1-based ids), dcounter and
 * entries, with fields ciss_alloc leaves unwritten untouched.  src must be
 * the COO the CISS was built from.  Capacities: directory >= dlength,
 * dcounter >= dlength + 1, entries >= 3 * (nnz / L + 1 + nnz + 1). */
ctc_status ctc_ciss_reference_layout(const ctc_ciss *c, const ctc_coo *src, uint32_t tile_len,
                                     int truncate_vals, uint64_t *directory, uint64_t *dcounter,
                                     double *entries, uint64_t *dlength, uint64_t *size);
void ctc_ciss_free(ctc_ciss *c);
/* How many times each nonzero's term reaches its row's sums in the
 * reference's warp-per-tile reductions of `mode` (tiles of 32 nonzeros of the
 * mode's order, SG/als.cu:226-312, SG/ccd.cu:120-195) under the ctc_ref_quirk
 * bits CTC_QUIRK_TILE_HEAD / CTC_QUIRK_LAST_LANE / CTC_QUIRK_CCD_GRID of
 * `flags`: weights[p] for the nonzero at position p of t (1 everywhere with
 * none of them; other bits are ignored).  Replaces nothing in the reference
 * (its defects are implicit in p_mttkrp_gpu_as and update_frac_gpu_as); ALS
 * and CCD with ctc_options.ref_quirks run on each mode's nonzeros repeated by
 * these weights. */
ctc_status ctc_ref_quirk_weights(const ctc_coo *t, uint32_t mode, uint32_t flags, uint32_t *weights);

/* ------------------------------------------------------------------------ */
/* Device context                                                            */
/* ------------------------------------------------------------------------ */
typedef struct ctc_context ctc_context;

ctc_status ctc_context_create(int device, ctc_context **out);
/* Releases the caller's handle.  Models and batches created on the context
 * hold it too; it is freed when the last of them is destroyed, so the
 * handles may be destroyed in any order. */
void ctc_context_destroy(ctc_context *ctx);
/* The context's hipStream_t. */
ctc_status ctc_context_stream(ctc_context *ctx, void **stream);
/* Waits for all work queued on the context's stream. */
ctc_status ctc_context_synchronize(ctc_context *ctx);
/* Per-kernel timing with HIP events on the launch stream (off by default).
 * enabled = 1 times every launch; n > 1 times every n-th launch of each
 * kernel name only (counted from this call), so that the events, each a
 * packet between two launches, barely perturb a run of short kernels; the
 * statistics then hold the sampled launches.  Negative values are rejected. */
ctc_status ctc_context_set_profiling(ctc_context *ctx, int enabled);
ctc_status ctc_context_kernel_stats(ctc_context *ctx, const char *kernel, uint64_t *launches,
                                    double *total_ms);
ctc_status ctc_context_reset_stats(ctc_context *ctx);
/* Measured device-to-device copy bandwidth (GB/s, read + write) of the
 * library's streaming copy kernel (k_copy) over `bytes` bytes. */
ctc_status ctc_context_copy_bandwidth(ctc_context *ctx, uint64_t bytes, double *gbps);

/* ------------------------------------------------------------------------ */
/* Solver options — the reference's compile-time constants (SG/base.h:17-64, */
/* SG/tc.c:22) as run-time fields                                            */
/* ------------------------------------------------------------------------ */
typedef enum ctc_algorithm { CTC_ALS = 0, CTC_SGD = 1, CTC_CCD = 2 } ctc_algorithm; /* SG/tc.c:13-15 */
typedef enum ctc_lr_schedule {
  CTC_LR_BOLD_DRIVER = 0, /* x1.05 on improvement else x0.5 (SG/sgd.cu:314-323) */
  CTC_LR_EXP_DECAY = 1,   /* lr *= lr_decay after every epoch */
  CTC_LR_CONSTANT = 2
} ctc_lr_schedule;
typedef enum ctc_precision {
  CTC_PREC_FP32 = 0,  /* FP32 factors and arithmetic; FP64 solve and RMSE */
  CTC_PREC_MIXED = 1  /* ALS: BF16 MFMA Gram/MTTKRP, FP32 accumulate; FP64 Cholesky + 1 refinement.
                         SGD (CTC_SGD_SYNC only): the run deltas of modes 1.. are handed in
                         (staged) in BF16, halving the schedule's dominant traffic; factors,
                         the walkers' arithmetic and the sums of deltas stay FP32 / FP64 */
} ctc_precision;
typedef enum ctc_mode_order { CTC_ORDER_FIXED = 0, CTC_ORDER_ROTATE = 1 } ctc_mode_order;
/* Diagnostic reproduction of the reference's ALS and CCD kernels (bits of
 * ctc_options.ref_quirks; 0 = the intended rule).  The reference reduces
 * every tile of 32 nonzeros of a mode's order with one warp, in ALS
 * (SG/als.cu:226-312) and CCD (SG/ccd.cu:120-195) alike:
 *   CTC_QUIRK_TILE_HEAD   (Q1) the tile's first row segment is not written
 *                              (lane 0's mybit forced to 1, SG/als.cu:245-249);
 *   CTC_QUIRK_LAST_LANE   (Q2) lane 31's term is re-added by the shuffle that
 *                              reads lane 31 itself (SG/als.cu:279-284);
 *   CTC_QUIRK_TRUNC_VALUES (Q5) values stored through an integer
 *                              (SG/ciss.c:123,157,176);
 * and, for CCD only:
 *   CTC_QUIRK_NO_ADDBACK  (Q4) the reference's CCD update in place of CCD++:
 *                              per mode, every column of a row from one residual
 *                              e = x - <current factors>, a_ir = sum e h_r /
 *                              (lambda + sum h_r^2), rows of the mode's
 *                              directory only (SG/ccd.cu:157-195,318-348);
 *                              the modes in the ALS order (ctc_mode_order);
 *   CTC_QUIRK_CCD_GRID    (Q3) only the first 8 * (floor(T / 256) + 1) of the
 *                              T = floor(nnz / 32) + 1 tiles contribute, the
 *                              tiles the CCD launch covers (SG/ccd.cu:520).
 * ALS uses Q1, Q2, Q5 and ignores the CCD bits; CCD with any bit requires
 * Q4 (the reference's CCD has no other form); SGD accepts none. */
typedef enum ctc_ref_quirk {
  CTC_QUIRK_TILE_HEAD = 1,
  CTC_QUIRK_LAST_LANE = 2,
  CTC_QUIRK_TRUNC_VALUES = 4,
  CTC_QUIRK_NO_ADDBACK = 8,
  CTC_QUIRK_CCD_GRID = 16
} ctc_ref_quirk;
/* How concurrent SGD walkers combine updates of a factor row they share.  A
 * walker keeps the rows of its current nonzero in registers; "racy" rows are
 * stored after every nonzero (concurrent writers overwrite each other),
 * "atomic" rows are re-read every nonzero and every update is added with a
 * float atomic (the reference's atomicAdd, SG/sgd.cu:96-98). */
/* FP32 ALS kernels.  SPANS: waves walk fixed spans of the CISS and write
 * per-(span, row) FP64 Gram pieces, a second kernel sums a row's pieces and
 * solves (ranks 8-64).  ROWS: one workgroup per row forms its Gram on FP64
 * matrix cores and solves it in place with a blocked FP64 Cholesky, long
 * rows cut into chunks (ranks 32-128).  AUTO: ROWS for rank >= 64 (at rank 32
 * SPANS measured faster on the 100M-nonzero benchmark tensor). */
typedef enum ctc_als_kernel { CTC_ALS_AUTO = 0, CTC_ALS_SPANS = 1, CTC_ALS_ROWS = 2 } ctc_als_kernel;

typedef enum ctc_sgd_update {
  CTC_SGD_AUTO = 0,       /* walked-mode rows racy; a row of another mode is racy when
                             it has more than sgd_hot_rows nonzeros and more than 4x its
                             mode's mean (its atomics would serialise on one L2
                             channel); otherwise a walker applies its
                             updates to its register copy of the row and adds their sum
                             with atomics when its run of nonzeros on that row ends */
  CTC_SGD_ATOMIC = 1,     /* walked-mode rows racy, every other row atomic at every nonzero */
  CTC_SGD_RACY = 2,       /* every row racy (fastest; loses updates on shared rows) */
  CTC_SGD_NEXT_RACY = 3,  /* walked mode and the next mode racy, other modes atomic */
  CTC_SGD_HOT_ATOMIC = 4, /* as AUTO, but a row that is not hot is re-read and updated
                             with atomics at every nonzero */
  CTC_SGD_SYNC = 5        /* the default: bulk-synchronous and deterministic.  Walker w
                             owns nonzeros [w*tile_len, (w+1)*tile_len) of the mode-0
                             order; sub-step s runs nonzeros [w*tile_len + s*sgd_substep,
                             +sgd_substep) of every walker from the factors as they stood
                             when the sub-step began.  A walker keeps its copy of a row
                             along a run of equal indices and hands in the run's delta;
                             at the end of the sub-step a row adds the sum of its runs'
                             deltas (the reference's atomicAdd, SG/sgd.cu:96-98), or their
                             mean if it is hot (more than sgd_hot_rows nonzeros and more
                             than 4x its mode's mean; replica averaging,
                             NC/sgd.cu:229-262,523-525).  Every sum is taken in a fixed
                             order (a row's runs in FP64 in walk order; a mode-0 row that
                             several walkers share in 64-bit fixed point), so every run of
                             a sweep gives the same bits. */
} ctc_sgd_update;

typedef struct ctc_options {
  int32_t algorithm;       /* ctc_algorithm; SG/tc.c:22 */
  uint32_t rank;           /* DEFAULT_NFACTORS, SG/base.h:21 */
  double lambda;           /* *_REGULARIZATION, SG/base.h:39-41 */
  double learning_rate;    /* LEARN_RATE, SG/base.h:58 */
  int32_t lr_schedule;     /* ctc_lr_schedule */
  double lr_decay;         /* factor for CTC_LR_EXP_DECAY */
  double sgd_scale;        /* SGD_MODIFICATION*, SG/base.h:42-44 */
  uint32_t max_epochs;     /* DEFAULT_MAX_ITERATE, SG/base.h:22 */
  double tolerance;        /* TOLERANCE, SG/base.h:36 */
  uint32_t max_bad_epochs; /* MAX_BADEPOCHS, SG/base.h:32 */
  uint32_t tile_len;       /* nonzeros per wave work unit, multiple of 64
                              (the unit DEFAULT_T_TILE_LENGTH sized, SG/base.h:62);
                              0 = the kernel family's measured default (the ALS
                              default: spans and FP32 rows 1024, BF16 rows 4096,
                              8192 at rank 128) */
  uint32_t block;          /* threads per workgroup (DEFAULT_BLOCKSIZE, SG/base.h:61) */
  uint64_t seed;           /* replaces srand(time(NULL)), SG/completion.c:426 */
  int32_t mode_order;      /* ctc_mode_order; SG/als.cu:684-685 rotates at random */
  int32_t precision;       /* ctc_precision */
  uint32_t inner_iters;    /* CCD++ inner iterations, NUM_INNER SG/base.h:30 */
  int32_t refine;          /* iterative-refinement steps in the ALS solve (FP32 path: 0/1; mixed: at least 1) */
  int32_t eval_every_epoch;/* evaluate RMSE after every epoch (convergence needs it) */
  int32_t verbose;         /* print the reference's progress line per epoch */
  int32_t sgd_update;      /* ctc_sgd_update */
  int32_t spread;          /* 1: walkers visit spans in a strided order that
                              spreads a hot row's spans over the sweep */
  uint32_t sgd_hot_rows;   /* CTC_SGD_SYNC / CTC_SGD_AUTO hot-row threshold (nonzeros per
                              row; 0 = 65536); a row is hot only if it also has more than
                              4x its mode's mean */
  int32_t als_kernel;      /* ctc_als_kernel (FP32 ALS) */
  uint32_t sgd_substep;    /* CTC_SGD_SYNC: nonzeros per walker per sub-step (0 = 32);
                              at most tile_len */
  uint32_t ref_quirks;     /* ctc_ref_quirk bits (ALS and CCD; 0 = the intended rule) */
} ctc_options;

/* Reference defaults for an algorithm (SG/base.h, SG/completion.c:460-480). */
ctc_status ctc_options_default(int32_t algorithm, ctc_options *out);

/* ------------------------------------------------------------------------ */
/* Device-resident model: training CISS per mode, validation/test tensors    */
/* and factor matrices.  Replaces the per-call device set-up of tc_als /     */
/* tc_sgd / tc_ccd (SG/als.cu:529-652, SG/sgd.cu:175-254, SG/ccd.cu:388-502) */
/* ------------------------------------------------------------------------ */
typedef struct ctc_model ctc_model;

/* validation and test may be NULL.  The tensors are copied; the caller keeps
 * ownership of its ctc_coo objects. */
ctc_status ctc_model_create(ctc_context *ctx, const ctc_coo *train, const ctc_coo *validation,
                            const ctc_coo *test, uint32_t rank, ctc_model **out);
void ctc_model_destroy(ctc_model *m);
/* Factor entries U(-1,1) from a seeded generator (matrix_randomize,
 * SG/matrixprocess.c:156-165, 568-577). */
ctc_status ctc_model_init_random(ctc_model *m, uint64_t seed);
/* Host row-major dims[mode] x rank float arrays. */
ctc_status ctc_model_set_factor(ctc_model *m, uint32_t mode, const float *host);
/* best != 0 reads the best-validation copy (best_mats, SG/completion.c:337-341). */
ctc_status ctc_model_get_factor(ctc_model *m, uint32_t mode, int best, float *host);
/* Device pointer of a factor (row-major float, dims[mode] x rank). */
ctc_status ctc_model_factor_device(ctc_model *m, uint32_t mode, float **dev);
/* Builds (and uploads) the CISS copies an algorithm needs ahead of time, so
 * that solve timings exclude it. */
ctc_status ctc_model_prepare(ctc_model *m, const ctc_options *opt);

/* ------------------------------------------------------------------------ */
/* Solvers — replace tc_main_ciss (SG/completion.h:71-77) and tc_als /       */
/* tc_sgd / tc_ccd (SG/als.cuh:20-33, SG/sgd.cuh:18-32, SG/ccd.cuh:20-33)    */
/* ------------------------------------------------------------------------ */
typedef struct ctc_epoch_stats {
  double loss;      /* sum of squared training errors */
  double frob;      /* lambda * sum ||F_m||^2 */
  double rmse_tr, rmse_vl, rmse_te;
  double epoch_ms;  /* device time of the sweep (0 for the initial entry) */
  double lr;        /* learning rate used in the epoch (SGD) */
} ctc_epoch_stats;

typedef struct ctc_result {
  uint32_t epochs;        /* epochs run */
  uint32_t best_epoch;
  double best_rmse;
  double ms_per_epoch;    /* median epoch time over epochs 2.. (or all if fewer) */
  double nnz_per_s;       /* train nnz / (ms_per_epoch / 1000) */
  uint64_t train_nnz;
  int32_t converged;
  ctc_epoch_stats *history; /* caller-owned, history_cap entries; entry 0 = initial */
  uint32_t history_cap;
} ctc_result;

/* Runs the epoch loop of opt->algorithm with the reference's convergence
 * rule (SG/completion.c:296-362).  Replaces tc_main_ciss's dispatch. */
ctc_status ctc_solve(ctc_model *m, const ctc_options *opt, ctc_result *res);

/* Single steps (each one full sweep on the device, synchronised). */
ctc_status ctc_als_mode_update(ctc_model *m, uint32_t mode, const ctc_options *opt); /* SG/als.cu:691-716 */
ctc_status ctc_sgd_epoch(ctc_model *m, const ctc_options *opt, double lr);           /* SG/sgd.cu:286 */
ctc_status ctc_ccd_epoch(ctc_model *m, const ctc_options *opt);                      /* SG/ccd.cu:558-592 */
/* MTTKRP of the current factors for `mode` into a host dims[mode] x rank array. */
ctc_status ctc_mttkrp(ctc_model *m, uint32_t mode, float *host_out);
/* Loss and RMSEs of the current factors on the device (FP64 accumulation);
 * any output may be NULL; RMSE of an absent tensor is 0.
 * Replaces tc_loss_sq / tc_rmse / tc_frob_sq (SG/completion.c:123-214). */
ctc_status ctc_evaluate(ctc_model *m, double lambda, double *loss, double *frob, double *rmse_tr,
                        double *rmse_vl, double *rmse_te);

/* Gradient (adjoint) of 0.5*sum_train e^2 + 0.5*lambda*sum_m sum_i c_i ||F_m(i,:)||^2
 * with c_i = 1, or c_i = nonzeros in row i when reg_per_occurrence != 0 (the
 * summed SGD step of SG/sgd.cu:93-95).  factors/grads are device arrays
 * (row-major dims[m] x rank float), one per mode; stream is a hipStream_t
 * (NULL = the context stream).  *loss receives the loss value.
 * No reference counterpart: the gradient is the direction the reference's SGD
 * kernel (SG/sgd.cu:92-99) steps along, summed over nonzeros. */
ctc_status ctc_grad(ctc_model *m, const float *const *factors, float *const *grads,
                    double lambda, int reg_per_occurrence, double *loss, void *stream);
/* Allocation-free ctc_grad.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving or training loops).  Same gradients as ctc_grad; the loss is
 * written to the device double loss_device (may be NULL).  It calls no
 * hipMalloc/hipFree, issues no memset and does not synchronise; its partial
 * sums live in a fixed device array inside the library (1 MiB per GPU, held
 * until the process ends), so calls that overlap on one GPU race.  Every
 * mode's CISS must already be built (ctc_model_prepare with ALS or CCD++
 * options, or one ctc_grad call), else CTC_ERR_INVALID_ARG; a tensor whose
 * partial sums would not fit the array (more than 16384 blocks per walk even
 * with the longest walker spans) is rejected with CTC_ERR_INVALID_ARG. */
ctc_status ctc_grad_noalloc(ctc_model *m, const float *const *factors, float *const *grads,
                            double lambda, int reg_per_occurrence, double *loss_device, void *stream);

/* ------------------------------------------------------------------------ */
/* Reference-shaped entry points (drop-in for callers of the reference's C   */
/* API).  ctc_matrix mirrors ordi_matrix (SG/matrixprocess.h:11-18).         */
/* ------------------------------------------------------------------------ */
typedef struct ctc_matrix {
  uint64_t I;   /* rows */
  uint64_t J;   /* columns (rank) */
  double *values; /* row-major */
  int32_t row_col_flag;
} ctc_matrix;

/* tc_main_ciss (SG/completion.c:422-497): reads nothing, allocates random
 * factors, runs the algorithm with the reference's defaults (SG/base.h) and
 * the given overrides (opt may be NULL), prints the reference's progress line. */
ctc_status ctc_tc_main_ciss(const ctc_coo *traina, const ctc_coo *trainb, const ctc_coo *trainc,
                            const ctc_coo *validation, const ctc_coo *test, int algorithm_index,
                            const ctc_options *opt);
/* tc_als / tc_sgd / tc_ccd (SG/als.cuh:20-33, SG/sgd.cuh:18-32, SG/ccd.cuh:20-33):
 * same argument meaning; mats are updated in place and best_mats receives the
 * best-validation model.  trainb/trainc may be NULL (the CISS copies are built
 * internally from traina). */
ctc_status ctc_tc_als(const ctc_coo *traina, const ctc_coo *trainb, const ctc_coo *trainc,
                      const ctc_coo *validation, const ctc_coo *test, ctc_matrix **mats,
                      ctc_matrix **best_mats, int algorithm_index, double regularization_index,
                      double *best_rmse, double *tolerance, uint64_t *nbadepochs,
                      uint64_t *bestepochs, uint64_t *max_badepochs);
ctc_status ctc_tc_sgd(const ctc_coo *traina, const ctc_coo *trainb, const ctc_coo *trainc,
                      const ctc_coo *validation, const ctc_coo *test, ctc_matrix **mats,
                      ctc_matrix **best_mats, int algorithm_index, double regularization_index,
                      double learning_rate, double *best_rmse, double *tolerance,
                      uint64_t *nbadepochs, uint64_t *bestepochs, uint64_t *max_badepochs);
ctc_status ctc_tc_ccd(const ctc_coo *traina, const ctc_coo *trainb, const ctc_coo *trainc,
                      const ctc_coo *validation, const ctc_coo *test, ctc_matrix **mats,
                      ctc_matrix **best_mats, int algorithm_index, double regularization_index,
                      double *best_rmse, double *tolerance, uint64_t *nbadepochs,
                      uint64_t *bestepochs, uint64_t *max_badepochs);

/* ------------------------------------------------------------------------ */
/* Autotuning — replaces the genetic algorithm over (blocksize, tile length) */
/* of GA/fitness.c:31-100 with a measured search over (tile_len, block).     */
/* ------------------------------------------------------------------------ */
typedef struct ctc_tune_result {
  uint32_t tile_len;
  uint32_t block;
  double ms;       /* best measured sweep time */
  uint32_t trials; /* configurations measured */
  uint32_t sgd_substep; /* bulk-synchronous SGD: the tuned sub-step (else opt's) */
} ctc_tune_result;

/* Measured search over the work-unit shape of opt->algorithm's sweep, the
 * GA's fitness (mean epoch time over (blocksize, tile length),
 * GA/fitness.c:31-100): ALS / CCD++ / lock-free SGD over tile_len 64..4096 and
 * block 64..256; the default SGD schedule over walker spans 1024..8192,
 * sub-steps 16..64 and blocks 64 / 256 (span and sub-step define that
 * schedule, so the tuned shape is what later solves with it compute).  The
 * factors and the best-model copy are restored afterwards. */

ctc_status ctc_autotune(ctc_model *m, const ctc_options *opt, ctc_tune_result *out);

/* ------------------------------------------------------------------------ */
/* Batched ALS: many independent completion problems solved together.        */
/* No counterpart in the reference, which solves one tensor per process      */
/* (SG/tc.c:40-50); its per-problem rules are those of tc_als.               */
/* ------------------------------------------------------------------------ */
typedef struct ctc_batch ctc_batch;

/* Problems (all with the same mode count; rank 8, 16, 32 or 64 each) are
 * grouped by rank.  The problems of a rank class form one block-diagonal
 * tensor (problem p's rows follow those of the problems before it), so each
 * ALS phase is one launch over every problem of the class.  test may be NULL
 * or hold NULL entries.  The tensors are copied. */
ctc_status ctc_batch_create(ctc_context *ctx, uint32_t nproblems, const ctc_coo *const *train,
                            const ctc_coo *const *test, const uint32_t *ranks, ctc_batch **out);
void ctc_batch_destroy(ctc_batch *b);
ctc_status ctc_batch_info(const ctc_batch *b, uint32_t *nproblems, uint32_t *nclasses, uint64_t *train_nnz);
/* rank and problem count of rank class `cls` (classes in increasing rank) */
ctc_status ctc_batch_class_rank(const ctc_batch *b, uint32_t cls, uint32_t *rank, uint32_t *nproblems);
/* factor `mode` of problem `problem`, row-major FP32, rank columns and
 * max(train dims[mode], test dims[mode]) rows */
ctc_status ctc_batch_set_factor(ctc_batch *b, uint32_t problem, uint32_t mode, const float *host);
ctc_status ctc_batch_get_factor(ctc_batch *b, uint32_t problem, uint32_t mode, float *host);
/* Measured (tile_len, block) per rank class for an ALS sweep in the precision
 * each class runs under opt (ctc_autotune on each class); results has room
 * for 4 entries.  Later ctc_batch_als calls use
 * each class's tuned pair in place of opt's.  Factors are left unchanged. */
ctc_status ctc_batch_autotune(ctc_batch *b, const ctc_options *opt, ctc_tune_result *results,
                              uint32_t *nclasses);
/* `iters` ALS iterations (modes 0..N-1 in turn) on every problem.  With
 * opt->precision = CTC_PREC_MIXED the classes of rank 32 and 64 run the
 * BF16-operand row-fused path (FP64 solve, one refinement) and those of rank
 * 8 and 16 the FP32 kernels; CTC_PREC_FP32 runs every class in FP32.
 * use_graph != 0: the run is captured once into a hipGraph (one branch per
 * rank class, forked from and joined to the context's stream) and replayed
 * by later calls with the same iters and options; 0: launched directly.
 * ms receives the device time of the run.  Returns CTC_ERR_NOT_SPD if some
 * row's normal system was not positive definite (that row becomes 0). */
ctc_status ctc_batch_als(ctc_batch *b, const ctc_options *opt, uint32_t iters, int32_t use_graph, double *ms);
/* per-problem train and test RMSE (arrays of nproblems; either may be NULL) */
ctc_status ctc_batch_rmse(ctc_batch *b, double *rmse_tr, double *rmse_te);

#ifdef __cplusplus
}
#endif
#endif /* CTC_H */
