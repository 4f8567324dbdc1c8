// ctc_cli.cpp — command-line front end of libctc, a drop-in for the
// reference's `tc` program (SG/tc.c:17-60):
//
//   ctc <train_a.tns> <train_b.tns> <train_c.tns> <validate.tns> <test.tns> [flags]
//
// The three training files are the reference's mode-sorted copies of one
// tensor (software/scripts/stensor.py:123-124); libctc sorts internally, so
// only the first is read and the other two may be any path or "-".  The
// compile-time constants of SG/base.h become flags:
//   --alg als|sgd|ccd (SG/tc.c:22)       --rank R (:21)      --lambda L (:39-41)
//   --lr LR (:58)       --sgd-scale S (:42-44)               --max-epochs E (:22)
//   --tol T (:36)       --max-bad-epochs B (:32)             --tile-len N (:62)
//   --block B (:61)     --seed S (replaces srand(time))      --mode-order fixed|rotate
//   --lr-schedule bold|exp|const  --lr-decay G   --inner I (NUM_INNER, :30)
//   --autotune (replaces the GA of software/autotuning)      --json (summary line)
//   --dump-factors PREFIX (writes PREFIX<mode>.txt)
//   --precision fp32|mixed (BF16-operand ALS; SGD with BF16-staged run deltas)
//   --refine K   --sgd-update P  --sgd-hot-rows N
//   --sgd-substep N (nonzeros per walker per sub-step of the default SGD schedule)
//   --ref-quirks B (the reference's kernel defects: ALS and CCD bits 1 Q1, 2 Q2,
//                   4 Q5; CCD bits 8 Q4, its update form, and 16 Q3, its grid)
//
//   ctc prep <tensor> <prefix> [--seed S] [--decimals D] [--binary]  (data preparation)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ctc.h"

static void usage() {
  fprintf(stderr,
          "usage: ctc <train_a.tns> <train_b.tns> <train_c.tns> <validate.tns> <test.tns>\n"
          "           [--alg als|sgd|ccd] [--rank R] [--lambda L] [--lr LR] [--sgd-scale S]\n"
          "           [--max-epochs E] [--tol T] [--max-bad-epochs B] [--tile-len N] [--block B]\n"
          "           [--seed S] [--mode-order fixed|rotate] [--lr-schedule bold|exp|const]\n"
          "           [--lr-decay G] [--inner I] [--precision fp32|mixed] [--refine K]\n"
          "           [--sgd-update sync|auto|atomic|racy|next-racy|hot-atomic] [--sgd-hot-rows N]\n"
          "           [--sgd-substep N] [--ref-quirks B]\n"
          "           [--autotune] [--json] [--dump-factors PREFIX]\n");
}

static int die(ctc_status s, const char *what) {
  fprintf(stderr, "ctc: %s: %s (%s)\n", what, ctc_status_string(s), ctc_last_error());
  return (int)s;
}

// `ctc prep <tensor> <prefix> [--seed S] [--decimals D] [--binary]`: the
// reference's data preparation in one step.  Splits the tensor 80/10/10
// (software/scripts/tvtensor.py) into <prefix>_tr, _v and _te, and writes the
// training tensor also sorted by modes (1, 2) and (2, 0) as <prefix>_tr2 and
// _tr3 (software/scripts/stensor.py), as .tns text with %.<D>f values
// (default 1, as the scripts) or as .bin files.
static int prep(int argc, char **argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: ctc prep <tensor.tns|.bin> <out_prefix> [--seed S] [--decimals D] [--binary]\n");
    return 1;
  }
  uint64_t seed = 0;
  int decimals = 1;
  bool binary = false;
  for (int i = 4; i < argc; ++i) {
    if (!strcmp(argv[i], "--seed") && i + 1 < argc) seed = strtoull(argv[++i], nullptr, 0);
    else if (!strcmp(argv[i], "--decimals") && i + 1 < argc) decimals = atoi(argv[++i]);
    else if (!strcmp(argv[i], "--binary")) binary = true;
    else {
      fprintf(stderr, "ctc prep: unknown argument %s\n", argv[i]);
      return 1;
    }
  }
  ctc_coo *t = nullptr, *tr = nullptr, *va = nullptr, *te = nullptr, *tr2 = nullptr, *tr3 = nullptr;
  ctc_status s;
  if ((s = ctc_coo_read(argv[2], &t)) != CTC_OK) return die(s, argv[2]);
  if ((s = ctc_coo_split(t, seed, &tr, &va, &te)) != CTC_OK) return die(s, "split");
  const uint32_t k12[2] = {1, 2}, k20[2] = {2, 0};
  if ((s = ctc_coo_sort(tr, k12, 2, &tr2)) != CTC_OK || (s = ctc_coo_sort(tr, k20, 2, &tr3)) != CTC_OK)
    return die(s, "sort");
  const char *names[5] = {"_tr", "_v", "_te", "_tr2", "_tr3"};
  ctc_coo *parts[5] = {tr, va, te, tr2, tr3};
  for (int k = 0; k < 5; ++k) {
    std::string path = std::string(argv[3]) + names[k] + (binary ? ".bin" : ".tns");
    s = binary ? ctc_coo_write_binary(parts[k], path.c_str(), t->base[0])
               : ctc_coo_write(parts[k], path.c_str(), t->base[0], decimals);
    if (s != CTC_OK) return die(s, path.c_str());
    printf("%s: %llu nonzeros\n", path.c_str(), (unsigned long long)parts[k]->nnz);
  }
  for (ctc_coo *c : {t, tr, va, te, tr2, tr3}) ctc_coo_free(c);
  return 0;
}

int main(int argc, char **argv) {
  if (argc > 1 && !strcmp(argv[1], "prep")) return prep(argc, argv);
  std::vector<const char *> pos;
  int alg = CTC_SGD;  // SG/tc.c:22 default
  std::vector<std::pair<std::string, std::string>> kv;
  bool autotune = false, json = false;
  std::string dump;
  for (int i = 1; i < argc; ++i) {
    if (strncmp(argv[i], "--", 2) == 0) {
      std::string k = argv[i] + 2;
      if (k == "autotune") {
        autotune = true;
      } else if (k == "json") {
        json = true;
      } else if (k == "help") {
        usage();
        return 0;
      } else {
        if (i + 1 >= argc) {
          usage();
          return 1;
        }
        kv.emplace_back(k, argv[++i]);
      }
    } else {
      pos.push_back(argv[i]);
    }
  }
  if (pos.size() != 5) {
    usage();
    return 1;
  }
  for (auto &p : kv)
    if (p.first == "alg") alg = p.second == "als" ? CTC_ALS : p.second == "ccd" ? CTC_CCD : CTC_SGD;
  ctc_options o;
  ctc_options_default(alg, &o);
  for (auto &p : kv) {
    const std::string &k = p.first;
    const char *v = p.second.c_str();
    if (k == "alg") continue;
    else if (k == "rank") o.rank = (uint32_t)atoi(v);
    else if (k == "lambda") o.lambda = atof(v);
    else if (k == "lr") o.learning_rate = atof(v);
    else if (k == "sgd-scale") o.sgd_scale = atof(v);
    else if (k == "max-epochs") o.max_epochs = (uint32_t)atoi(v);
    else if (k == "tol") o.tolerance = atof(v);
    else if (k == "max-bad-epochs") o.max_bad_epochs = (uint32_t)atoi(v);
    else if (k == "tile-len") o.tile_len = (uint32_t)atoi(v);
    else if (k == "block") o.block = (uint32_t)atoi(v);
    else if (k == "seed") o.seed = strtoull(v, nullptr, 0);
    else if (k == "mode-order" || k == "lr-schedule" || k == "precision") {
      static const struct { const char *flag, *name; int32_t v; } names[] = {
          {"mode-order", "fixed", CTC_ORDER_FIXED}, {"mode-order", "rotate", CTC_ORDER_ROTATE},
          {"lr-schedule", "bold", CTC_LR_BOLD_DRIVER}, {"lr-schedule", "exp", CTC_LR_EXP_DECAY},
          {"lr-schedule", "const", CTC_LR_CONSTANT}, {"precision", "fp32", CTC_PREC_FP32},
          {"precision", "mixed", CTC_PREC_MIXED}};
      int found = -1;
      for (const auto &q : names)
        if (k == q.flag && !strcmp(v, q.name)) found = q.v;
      if (found < 0) {
        fprintf(stderr, "ctc: unknown --%s %s\n", k.c_str(), v);
        return 1;
      }
      (k == "mode-order" ? o.mode_order : k == "lr-schedule" ? o.lr_schedule : o.precision) = found;
    } else if (k == "lr-decay") o.lr_decay = atof(v);
    else if (k == "inner") o.inner_iters = (uint32_t)atoi(v);
    else if (k == "refine") o.refine = atoi(v);
    else if (k == "sgd-update") {
      static const struct { const char *name; int32_t v; } pol[] = {
          {"sync", CTC_SGD_SYNC},   {"auto", CTC_SGD_AUTO},           {"atomic", CTC_SGD_ATOMIC},
          {"racy", CTC_SGD_RACY},   {"next-racy", CTC_SGD_NEXT_RACY}, {"hot-atomic", CTC_SGD_HOT_ATOMIC}};
      int found = -1;
      for (const auto &q : pol)
        if (!strcmp(v, q.name)) found = q.v;
      if (found < 0) {
        fprintf(stderr, "ctc: unknown --sgd-update %s\n", v);
        return 1;
      }
      o.sgd_update = found;
    } else if (k == "sgd-substep")
      o.sgd_substep = (uint32_t)atoi(v);
    else if (k == "sgd-hot-rows") o.sgd_hot_rows = (uint32_t)strtoul(v, nullptr, 0);
    else if (k == "ref-quirks") o.ref_quirks = (uint32_t)strtoul(v, nullptr, 0);
    else if (k == "dump-factors") dump = v;
    else {
      fprintf(stderr, "ctc: unknown flag --%s\n", k.c_str());
      return 1;
    }
  }
  o.verbose = 1;

  ctc_coo *train = nullptr, *valid = nullptr, *test = nullptr;
  ctc_status s;
  if ((s = ctc_coo_read(pos[0], &train)) != CTC_OK) return die(s, pos[0]);
  if (strcmp(pos[3], "-") != 0 && (s = ctc_coo_read(pos[3], &valid)) != CTC_OK) return die(s, pos[3]);
  if (strcmp(pos[4], "-") != 0 && (s = ctc_coo_read(pos[4], &test)) != CTC_OK) return die(s, pos[4]);
  // validation/test dims may be smaller (SG/io.c: dims = largest index); widen to train's
  for (ctc_coo *t : {valid, test}) {
    if (!t) continue;
    for (uint32_t m = 0; m < t->nmodes && m < train->nmodes; ++m)
      if (t->dims[m] < train->dims[m]) t->dims[m] = train->dims[m];
  }
  ctc_context *ctx = nullptr;
  if ((s = ctc_context_create(0, &ctx)) != CTC_OK) return die(s, "context");
  ctc_model *m = nullptr;
  if ((s = ctc_model_create(ctx, train, valid, test, o.rank, &m)) != CTC_OK) return die(s, "model");
  if ((s = ctc_model_init_random(m, o.seed)) != CTC_OK) return die(s, "init");
  if (autotune) {
    ctc_tune_result tr;
    if ((s = ctc_model_prepare(m, &o)) != CTC_OK) return die(s, "prepare");
    if ((s = ctc_autotune(m, &o, &tr)) != CTC_OK) return die(s, "autotune");
    printf("autotune: tile_len %u block %u (%.3f ms per sweep, %u trials)\n", tr.tile_len, tr.block, tr.ms,
           tr.trials);
    o.tile_len = tr.tile_len;
    o.block = tr.block;
  }
  std::vector<ctc_epoch_stats> hist(o.max_epochs + 1);
  ctc_result r{};
  r.history = hist.data();
  r.history_cap = (uint32_t)hist.size();
  if ((s = ctc_solve(m, &o, &r)) != CTC_OK) return die(s, "solve");
  printf("best epoch %u  best RMSE %0.5e  %.3f ms/epoch  %.3e nnz/s\n", r.best_epoch, r.best_rmse,
         r.ms_per_epoch, r.nnz_per_s);
  if (json)
    printf("{\"algorithm\": %d, \"rank\": %u, \"epochs\": %u, \"best_epoch\": %u, \"best_rmse\": %.9g, "
           "\"ms_per_epoch\": %.6g, \"nnz_per_s\": %.6g, \"rmse_te\": %.9g}\n",
           alg, o.rank, r.epochs, r.best_epoch, r.best_rmse, r.ms_per_epoch, r.nnz_per_s,
           hist[r.epochs].rmse_te);
  if (!dump.empty()) {
    for (uint32_t md = 0; md < train->nmodes; ++md) {
      std::vector<float> f(train->dims[md] * o.rank);
      if ((s = ctc_model_get_factor(m, md, 1, f.data())) != CTC_OK) return die(s, "factor");
      std::string path = dump + std::to_string(md) + ".txt";
      FILE *fp = fopen(path.c_str(), "w");
      if (!fp) return die(CTC_ERR_IO, path.c_str());
      for (uint64_t i = 0; i < train->dims[md]; ++i) {
        for (uint32_t c = 0; c < o.rank; ++c) fprintf(fp, c ? " %.9g" : "%.9g", f[i * o.rank + c]);
        fputc('\n', fp);
      }
      fclose(fp);
    }
  }
  ctc_model_destroy(m);
  ctc_context_destroy(ctx);
  ctc_coo_free(train);
  ctc_coo_free(valid);
  ctc_coo_free(test);
  return 0;
}
