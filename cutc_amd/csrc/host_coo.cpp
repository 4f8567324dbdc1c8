// host_coo.cpp — errors, host threading helpers, COO tensors, .tns I/O and
// the seeded synthetic-tensor generator of libctc.
#include <algorithm>
#include <atomic>
#include <charconv>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <numeric>
#include <thread>

#include "ctc_internal.h"
#include "host_parallel.h"

namespace ctc {

static thread_local std::string g_last_error;

void set_error(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

ctc_status fail(ctc_status s, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return s;
}

unsigned host_threads() {
  const char *env = getenv("CTC_HOST_THREADS");
  if (env && atoi(env) > 0) return (unsigned)atoi(env);
  env = getenv("OMP_NUM_THREADS");
  unsigned hw = std::thread::hardware_concurrency();
  if (hw == 0) hw = 1;
  if (env && atoi(env) > 0) return std::min<unsigned>((unsigned)atoi(env), hw);
  return std::min<unsigned>(hw, 32);
}

void counting_sort(const uint32_t *key, uint64_t dim, std::vector<uint32_t> &order,
                   std::vector<uint32_t> &tmp) {
  const uint64_t n = order.size();
  tmp.resize(n);
  unsigned T = host_threads();
  if (n < (1u << 20) || (uint64_t)T * dim > (1ull << 28)) T = 1;
  std::vector<std::vector<uint64_t>> hist(T, std::vector<uint64_t>(dim + 1, 0));
  auto chunk = [&](unsigned t, uint64_t &b, uint64_t &e) {
    b = n * t / T;
    e = n * (t + 1) / T;
  };
  run_threads(T, [&](unsigned t) {
    uint64_t b, e;
    chunk(t, b, e);
    auto &h = hist[t];
    for (uint64_t p = b; p < e; ++p) h[key[order[p]]]++;
  });
  // exclusive offsets: bucket-major, thread-minor keeps the sort stable
  uint64_t run = 0;
  for (uint64_t d = 0; d < dim; ++d) {
    for (unsigned t = 0; t < T; ++t) {
      uint64_t c = hist[t][d];
      hist[t][d] = run;
      run += c;
    }
  }
  run_threads(T, [&](unsigned t) {
    uint64_t b, e;
    chunk(t, b, e);
    auto &h = hist[t];
    for (uint64_t p = b; p < e; ++p) tmp[h[key[order[p]]]++] = order[p];
  });
  order.swap(tmp);
}

}  // namespace ctc

using namespace ctc;

// -------------------------------------------------------------------------
// status strings
// -------------------------------------------------------------------------
extern "C" const char *ctc_status_string(ctc_status s) {
  switch (s) {
    case CTC_OK: return "ok";
    case CTC_ERR_INVALID_ARG: return "invalid argument";
    case CTC_ERR_IO: return "cannot open file";
    case CTC_ERR_PARSE: return "malformed or empty tensor";
    case CTC_ERR_INDEX_BASE: return "index base must be 0 or 1";
    case CTC_ERR_NOMEM: return "out of memory";
    case CTC_ERR_HIP: return "HIP runtime error";
    case CTC_ERR_NOT_SPD: return "normal-equation system not positive definite";
    case CTC_ERR_UNSUPPORTED: return "unsupported configuration";
    case CTC_ERR_NO_DEVICE: return "no GPU device";
  }
  return "unknown status";
}

extern "C" const char *ctc_last_error(void) { return g_last_error.c_str(); }
extern "C" const char *ctc_version(void) { return "ctc 0.1 (gfx950)"; }

// -------------------------------------------------------------------------
// COO alloc / free
// -------------------------------------------------------------------------
extern "C" ctc_status ctc_coo_alloc(uint32_t nmodes, uint64_t nnz, const uint64_t *dims,
                                    ctc_coo **out) {
  if (!out || nmodes < 2 || nmodes > CTC_MAX_MODES)
    return fail(CTC_ERR_INVALID_ARG, "ctc_coo_alloc: nmodes must be 2..%d", CTC_MAX_MODES);
  if (nnz > 0xFFFFFFFFull)
    return fail(CTC_ERR_UNSUPPORTED, "ctc_coo_alloc: nnz above 2^32-1 is not supported");
  ctc_coo *t = (ctc_coo *)calloc(1, sizeof(ctc_coo));
  if (!t) return fail(CTC_ERR_NOMEM, "ctc_coo_alloc");
  t->nmodes = nmodes;
  t->nnz = nnz;
  size_t n = nnz ? nnz : 1;
  for (uint32_t m = 0; m < nmodes; ++m) {
    t->dims[m] = dims ? dims[m] : 0;
    t->ind[m] = (uint32_t *)calloc(n, sizeof(uint32_t));
    t->base[m] = 0;
  }
  t->vals = (double *)calloc(n, sizeof(double));
  bool ok = t->vals != nullptr;
  for (uint32_t m = 0; m < nmodes; ++m) ok = ok && t->ind[m];
  if (!ok) {
    ctc_coo_free(t);
    return fail(CTC_ERR_NOMEM, "ctc_coo_alloc: %llu nonzeros", (unsigned long long)nnz);
  }
  *out = t;
  return CTC_OK;
}

extern "C" void ctc_coo_free(ctc_coo *t) {
  if (!t) return;
  for (uint32_t m = 0; m < CTC_MAX_MODES; ++m) free(t->ind[m]);
  free(t->vals);
  free(t);
}

// -------------------------------------------------------------------------
// .tns reader (semantics of SG/io.c:58-106, 234-308)
// -------------------------------------------------------------------------
namespace {

inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\f' || c == '\v'; }

// A data line is a line whose first byte is not '#' and that holds a
// non-blank character (SG/io.c:93 skips empty and '#' lines).
inline bool data_line(const char *b, const char *e) {
  if (b >= e || *b == '#') return false;
  for (const char *p = b; p < e; ++p)
    if (!is_space(*p)) return true;
  return false;
}

struct LineChunk {
  const char *b, *e;
  uint64_t lines = 0, first = 0;
};

}  // namespace

// -------------------------------------------------------------------------
// SPLATT binary coordinate files (SG/io.h:79-84, SG/io.c:116-189,369-513):
//   int32 magic (0 = coordinate), uint64 index width (4 or 8), uint64 value
//   width (4 or 8), then in the index width: nmodes, dims[nmodes], nnz, the
//   nmodes index arrays of nnz entries each; then nnz values.
// Indices are stored as the writer held them (1-based for the reference's
// text-read tensors); the reader applies the text reader's base rule.
// -------------------------------------------------------------------------
namespace {

bool has_suffix(const char *path, const char *suf) {
  const size_t n = strlen(path), k = strlen(suf);
  return n >= k && strcmp(path + n - k, suf) == 0;
}

ctc_status read_words(FILE *f, uint64_t width, uint64_t count, uint64_t *dst, const char *path) {
  if (width == 8) {
    if (count && fread(dst, 8, count, f) != count) return fail(CTC_ERR_PARSE, "'%s': truncated file", path);
    return CTC_OK;
  }
  std::vector<uint32_t> buf(std::min<uint64_t>(count, 1u << 20));
  for (uint64_t n = 0; n < count;) {
    const uint64_t k = std::min<uint64_t>(buf.size(), count - n);
    if (fread(buf.data(), 4, k, f) != k) return fail(CTC_ERR_PARSE, "'%s': truncated file", path);
    for (uint64_t i = 0; i < k; ++i) dst[n + i] = buf[i];
    n += k;
  }
  return CTC_OK;
}

ctc_status read_binary(const char *path, ctc_coo **out) {
  FILE *f = fopen(path, "rb");
  if (!f) return fail(CTC_ERR_IO, "failed to open '%s'", path);
  std::unique_ptr<FILE, int (*)(FILE *)> guard(f, fclose);
  int32_t magic = -1;
  uint64_t iw = 0, vw = 0;
  if (fread(&magic, 4, 1, f) != 1 || fread(&iw, 8, 1, f) != 1 || fread(&vw, 8, 1, f) != 1)
    return fail(CTC_ERR_PARSE, "'%s': truncated binary header", path);
  if (magic != 0) return fail(CTC_ERR_PARSE, "'%s': not a binary coordinate tensor (magic %d)", path, magic);
  if ((iw != 4 && iw != 8) || (vw != 4 && vw != 8))
    return fail(CTC_ERR_PARSE, "'%s': index width %llu / value width %llu", path, (unsigned long long)iw,
                (unsigned long long)vw);
  uint64_t hdr[1 + CTC_MAX_MODES + 1];
  CTC_TRY(read_words(f, iw, 1, hdr, path));
  const uint64_t nmodes = hdr[0];
  if (nmodes < 2 || nmodes > CTC_MAX_MODES) return fail(CTC_ERR_UNSUPPORTED, "'%s': %llu modes", path, (unsigned long long)nmodes);
  CTC_TRY(read_words(f, iw, nmodes + 1, hdr + 1, path));
  const uint64_t nnz = hdr[1 + nmodes];
  uint64_t dims[CTC_MAX_MODES] = {0, 0, 0, 0};
  for (uint64_t m = 0; m < nmodes; ++m) dims[m] = hdr[1 + m];
  ctc_coo *t = nullptr;
  CTC_TRY(ctc_coo_alloc((uint32_t)nmodes, nnz, dims, &t));
  std::unique_ptr<ctc_coo, void (*)(ctc_coo *)> tg(t, ctc_coo_free);
  std::vector<uint64_t> idx(std::min<uint64_t>(nnz, 1u << 22));
  for (uint64_t m = 0; m < nmodes; ++m) {
    uint64_t lo = ~0ull, hi = 0;
    for (uint64_t n = 0; n < nnz;) {
      const uint64_t k = std::min<uint64_t>(idx.size(), nnz - n);
      CTC_TRY(read_words(f, iw, k, idx.data(), path));
      for (uint64_t i = 0; i < k; ++i) {
        lo = std::min(lo, idx[i]);
        hi = std::max(hi, idx[i]);
        t->ind[m][n + i] = (uint32_t)idx[i];
      }
      n += k;
    }
    // the text reader's rule (SG/io.c:263-300): base 0 if some index is 0, else 1
    const uint32_t base = nnz && lo == 0 ? 0u : 1u;
    if (hi - base >= 0xFFFFFFFFull) return fail(CTC_ERR_UNSUPPORTED, "'%s': index beyond 32 bits", path);
    if (base)
      for (uint64_t n = 0; n < nnz; ++n) t->ind[m][n] -= 1;
    t->base[m] = base;
    if (nnz && hi + 1 - base > t->dims[m]) t->dims[m] = hi + 1 - base;
  }
  if (vw == 8) {
    if (nnz && fread(t->vals, 8, nnz, f) != nnz) return fail(CTC_ERR_PARSE, "'%s': truncated values", path);
  } else {
    std::vector<float> fb(std::min<uint64_t>(nnz, 1u << 20));
    for (uint64_t n = 0; n < nnz;) {
      const uint64_t k = std::min<uint64_t>(fb.size(), nnz - n);
      if (fread(fb.data(), 4, k, f) != k) return fail(CTC_ERR_PARSE, "'%s': truncated values", path);
      for (uint64_t i = 0; i < k; ++i) t->vals[n + i] = fb[i];
      n += k;
    }
  }
  *out = tg.release();
  return CTC_OK;
}

}  // namespace

extern "C" ctc_status ctc_coo_write_binary(const ctc_coo *t, const char *path, uint32_t base) {
  if (!t || !path || base > 1) return fail(CTC_ERR_INVALID_ARG, "ctc_coo_write_binary");
  FILE *f = fopen(path, "wb");
  if (!f) return fail(CTC_ERR_IO, "failed to open '%s' for writing", path);
  std::unique_ptr<FILE, int (*)(FILE *)> guard(f, fclose);
  // narrowest exact widths (SG/io.c:116-152 picks them the same way)
  uint64_t iw = 4, vw = 4;
  if (t->nnz >= 0xFFFFFFFFull) iw = 8;
  for (uint32_t m = 0; m < t->nmodes; ++m)
    if (t->dims[m] + base >= 0xFFFFFFFFull) iw = 8;
  for (uint64_t p = 0; p < t->nnz && vw == 4; ++p)
    if ((double)(float)t->vals[p] != t->vals[p]) vw = 8;
  const int32_t magic = 0;
  bool ok = fwrite(&magic, 4, 1, f) == 1 && fwrite(&iw, 8, 1, f) == 1 && fwrite(&vw, 8, 1, f) == 1;
  auto put = [&](uint64_t v) {
    if (iw == 8) {
      ok = ok && fwrite(&v, 8, 1, f) == 1;
    } else {
      const uint32_t w = (uint32_t)v;
      ok = ok && fwrite(&w, 4, 1, f) == 1;
    }
  };
  put(t->nmodes);
  for (uint32_t m = 0; m < t->nmodes; ++m) put(t->dims[m]);
  put(t->nnz);
  std::vector<uint32_t> b32;
  std::vector<uint64_t> b64;
  for (uint32_t m = 0; m < t->nmodes && ok; ++m) {
    if (iw == 8) {
      b64.resize(t->nnz);
      for (uint64_t p = 0; p < t->nnz; ++p) b64[p] = (uint64_t)t->ind[m][p] + base;
      ok = ok && fwrite(b64.data(), 8, t->nnz, f) == t->nnz;
    } else {
      b32.resize(t->nnz);
      for (uint64_t p = 0; p < t->nnz; ++p) b32[p] = t->ind[m][p] + base;
      ok = ok && fwrite(b32.data(), 4, t->nnz, f) == t->nnz;
    }
  }
  if (vw == 8) {
    ok = ok && fwrite(t->vals, 8, t->nnz, f) == t->nnz;
  } else {
    std::vector<float> fv(t->nnz);
    for (uint64_t p = 0; p < t->nnz; ++p) fv[p] = (float)t->vals[p];
    ok = ok && fwrite(fv.data(), 4, t->nnz, f) == t->nnz;
  }
  if (!ok) return fail(CTC_ERR_IO, "write error on '%s'", path);
  return CTC_OK;
}

extern "C" ctc_status ctc_coo_read(const char *path, ctc_coo **out) {
  if (!path || !out) return fail(CTC_ERR_INVALID_ARG, "ctc_coo_read: null argument");
  // file type by extension, as SG/io.c:21-52 decides (.bin: binary coordinates)
  if (has_suffix(path, ".bin")) return read_binary(path, out);
  FILE *f = fopen(path, "rb");
  if (!f) return fail(CTC_ERR_IO, "failed to open '%s'", path);
  fseek(f, 0, SEEK_END);
  long sz = ftell(f);
  fseek(f, 0, SEEK_SET);
  std::vector<char> buf((size_t)sz + 1);
  size_t got = sz > 0 ? fread(buf.data(), 1, (size_t)sz, f) : 0;
  fclose(f);
  buf[got] = '\n';
  const char *begin = buf.data(), *end = buf.data() + got;

  // mode count from the first data line: tokens - 1 (SG/io.c:250-261)
  uint32_t nmodes = 0;
  for (const char *p = begin; p < end;) {
    const char *le = (const char *)memchr(p, '\n', (size_t)(end - p));
    if (!le) le = end;
    if (data_line(p, le)) {
      uint32_t tok = 0;
      const char *q = p;
      while (q < le) {
        while (q < le && is_space(*q)) ++q;
        if (q >= le) break;
        ++tok;
        while (q < le && !is_space(*q)) ++q;
      }
      nmodes = tok ? tok - 1 : 0;
      break;
    }
    p = le + 1;
  }
  if (nmodes < 2 || nmodes > CTC_MAX_MODES)
    return fail(CTC_ERR_PARSE, "'%s': cannot determine 2..%d modes from the first line", path,
                CTC_MAX_MODES);

  // split into chunks at line boundaries, count data lines per chunk
  unsigned T = host_threads();
  if (got < (1u << 22)) T = 1;
  std::vector<LineChunk> ch(T);
  for (unsigned t = 0; t < T; ++t) {
    const char *b = begin + got * t / T;
    const char *e = begin + got * (t + 1) / T;
    if (t > 0) {
      const char *nl = (const char *)memchr(b - 1, '\n', (size_t)(end - (b - 1)));
      b = nl ? nl + 1 : end;
    }
    if (t + 1 < T) {
      const char *nl = (const char *)memchr(e - 1, '\n', (size_t)(end - (e - 1)));
      e = nl ? nl + 1 : end;
    } else {
      e = end;
    }
    if (b > e) b = e;
    ch[t].b = b;
    ch[t].e = e;
  }
  run_threads(T, [&](unsigned t) {
    uint64_t n = 0;
    for (const char *p = ch[t].b; p < ch[t].e;) {
      const char *le = (const char *)memchr(p, '\n', (size_t)(ch[t].e - p));
      if (!le) le = ch[t].e;
      if (data_line(p, le)) ++n;
      p = le + 1;
    }
    ch[t].lines = n;
  });
  uint64_t nnz = 0;
  for (auto &c : ch) {
    c.first = nnz;
    nnz += c.lines;
  }
  if (nnz == 0) return fail(CTC_ERR_PARSE, "'%s' holds no nonzeros", path);
  ctc_coo *t = nullptr;
  CTC_TRY(ctc_coo_alloc(nmodes, nnz, nullptr, &t));
  std::vector<std::vector<uint64_t>> mx(T, std::vector<uint64_t>(nmodes, 0)),
      mn(T, std::vector<uint64_t>(nmodes, 1));
  std::atomic<int> bad{0};
  run_threads(T, [&](unsigned th) {
    uint64_t n = ch[th].first;
    for (const char *p = ch[th].b; p < ch[th].e;) {
      const char *le = (const char *)memchr(p, '\n', (size_t)(ch[th].e - p));
      if (!le) le = ch[th].e;
      if (data_line(p, le)) {
        const char *q = p;
        for (uint32_t m = 0; m < nmodes; ++m) {
          while (q < le && is_space(*q)) ++q;
          uint64_t v = 0;
          auto r = std::from_chars(q, le, v);
          if (r.ec != std::errc()) bad = 1;
          q = r.ptr;
          if (v > 0xFFFFFFFFull) bad = 2;
          t->ind[m][n] = (uint32_t)v;
          mx[th][m] = std::max(mx[th][m], v);  // dims = max index (SG/io.c:278)
          mn[th][m] = std::min(mn[th][m], v);  // base = min(1, min index) (:281)
        }
        while (q < le && is_space(*q)) ++q;
        double v = 0;
        auto r = std::from_chars(q, le, v);
        if (r.ec != std::errc()) bad = 1;
        t->vals[n] = v;
        ++n;
      }
      p = le + 1;
    }
  });
  if (bad) {
    ctc_coo_free(t);
    return fail(CTC_ERR_PARSE, "'%s': malformed line (or index above 2^32-1)", path);
  }
  for (uint32_t m = 0; m < nmodes; ++m) {
    uint64_t dmax = 0, base = 1;
    for (unsigned th = 0; th < T; ++th) {
      dmax = std::max(dmax, mx[th][m]);
      base = std::min(base, mn[th][m]);
    }
    // base is 0 or 1 by construction: the reference's check at
    // SG/io.c:291-296 can never fire because the minimum starts at 1.
    t->base[m] = (uint32_t)base;
    t->dims[m] = base == 0 ? dmax + 1 : dmax;  // SG/io.c:299-304
  }
  if (t->nnz > 0) {
    uint32_t b[CTC_MAX_MODES];
    for (uint32_t m = 0; m < nmodes; ++m) b[m] = t->base[m];
    parallel_for(nnz, host_threads(), [&](uint64_t lo, uint64_t hi, unsigned) {
      for (uint32_t m = 0; m < nmodes; ++m)
        if (b[m])
          for (uint64_t p = lo; p < hi; ++p) t->ind[m][p] -= b[m];
    });
  }
  *out = t;
  return CTC_OK;
}

extern "C" ctc_status ctc_coo_write(const ctc_coo *t, const char *path, uint32_t base,
                                    int decimals) {
  if (!t || !path || base > 1) return fail(CTC_ERR_INVALID_ARG, "ctc_coo_write");
  FILE *f = fopen(path, "w");
  if (!f) return fail(CTC_ERR_IO, "failed to open '%s' for writing", path);
  char fmt[16];
  if (decimals < 0)
    snprintf(fmt, sizeof fmt, "%%.17g\n");
  else
    snprintf(fmt, sizeof fmt, "%%.%df\n", decimals);
  for (uint64_t p = 0; p < t->nnz; ++p) {
    for (uint32_t m = 0; m < t->nmodes; ++m) fprintf(f, "%u ", t->ind[m][p] + base);
    fprintf(f, fmt, t->vals[p]);
  }
  if (fclose(f) != 0) return fail(CTC_ERR_IO, "write error on '%s'", path);
  return CTC_OK;
}

// -------------------------------------------------------------------------
// synthetic generator
// -------------------------------------------------------------------------
namespace {

inline uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

inline double u01(uint64_t x) { return (double)(x >> 11) * 0x1.0p-53; }

// xoshiro256** (Blackman & Vigna), seeded by splitmix64
struct Xoshiro {
  uint64_t s[4];
  explicit Xoshiro(uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) {
      x += 0x9e3779b97f4a7c15ull;
      s[i] = splitmix64(x);
    }
  }
  static inline uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
  inline uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  inline double uniform() { return u01(next()); }
};

// inverse-CDF sampler of Zipf(s) over ranks 1..n (rank 1 -> index 0)
struct IndexSampler {
  uint64_t n = 0;
  double s = 0;
  std::vector<double> cdf;
  void init(uint64_t dim, double expo) {
    n = dim;
    s = expo;
    cdf.clear();
    if (expo > 0) {
      cdf.resize(dim);
      double acc = 0;
      for (uint64_t r = 0; r < dim; ++r) {
        acc += std::pow((double)(r + 1), -expo);
        cdf[r] = acc;
      }
      for (auto &c : cdf) c /= acc;
      cdf.back() = 1.0;
    }
  }
  inline uint64_t draw(Xoshiro &g) const {
    if (s <= 0) return (uint64_t)(g.uniform() * (double)n) % n;
    double u = g.uniform();
    return (uint64_t)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin()) % n;
  }
};

void radix_sort_u64(std::vector<uint64_t> &a, uint64_t maxkey) {
  const uint64_t n = a.size();
  if (n < 2) return;
  std::vector<uint64_t> tmp(n);
  int bits = 1;
  while (bits < 64 && (maxkey >> bits)) ++bits;
  unsigned T = host_threads();
  if (n < (1u << 20)) T = 1;
  const int D = 11;
  const uint64_t B = 1ull << D;
  std::vector<uint64_t> hist((size_t)T * B);
  for (int shift = 0; shift < bits; shift += D) {
    std::fill(hist.begin(), hist.end(), 0);
    run_threads(T, [&](unsigned t) {
      uint64_t lo = n * t / T, hi = n * (t + 1) / T;
      uint64_t *h = &hist[(size_t)t * B];
      for (uint64_t p = lo; p < hi; ++p) h[(a[p] >> shift) & (B - 1)]++;
    });
    uint64_t run = 0;
    for (uint64_t d = 0; d < B; ++d)
      for (unsigned t = 0; t < T; ++t) {
        uint64_t c = hist[(size_t)t * B + d];
        hist[(size_t)t * B + d] = run;
        run += c;
      }
    run_threads(T, [&](unsigned t) {
      uint64_t lo = n * t / T, hi = n * (t + 1) / T;
      uint64_t *h = &hist[(size_t)t * B];
      for (uint64_t p = lo; p < hi; ++p) tmp[h[(a[p] >> shift) & (B - 1)]++] = a[p];
    });
    a.swap(tmp);
  }
}

}  // namespace

extern "C" ctc_status ctc_synth_generate(const ctc_synth_spec *sp, ctc_coo **train,
                                         ctc_coo **test) {
  if (!sp || !train) return fail(CTC_ERR_INVALID_ARG, "ctc_synth_generate: null argument");
  const uint32_t N = sp->nmodes;
  if (N < 2 || N > CTC_MAX_MODES) return fail(CTC_ERR_INVALID_ARG, "synth: nmodes 2..4");
  long double cells = 1;
  for (uint32_t m = 0; m < N; ++m) {
    if (sp->dims[m] == 0 || sp->dims[m] > 0xFFFFFFFFull)
      return fail(CTC_ERR_INVALID_ARG, "synth: dims must be 1..2^32-1");
    cells *= (long double)sp->dims[m];
  }
  if (cells >= 1.8e19L) return fail(CTC_ERR_UNSUPPORTED, "synth: coordinate space above 2^64");
  if (sp->nnz == 0 || (long double)sp->nnz > cells * 0.5L)
    return fail(CTC_ERR_INVALID_ARG, "synth: nnz must be in 1..cells/2");
  if (sp->test_fraction < 0 || sp->test_fraction >= 1)
    return fail(CTC_ERR_INVALID_ARG, "synth: test_fraction must be in [0,1)");
  const uint64_t nnz = sp->nnz;
  uint64_t maxkey = 1;
  {
    uint64_t k = 1;
    for (uint32_t m = 0; m < N; ++m) k *= sp->dims[m];
    maxkey = k - 1;
  }
  IndexSampler smp[CTC_MAX_MODES];
  for (uint32_t m = 0; m < N; ++m) smp[m].init(sp->dims[m], sp->zipf[m]);
  const uint64_t seed_coord = splitmix64(sp->seed ^ 0xC00Dull);
  const uint64_t seed_truth = splitmix64(sp->seed ^ 0x7F07ull);
  const uint64_t seed_noise = splitmix64(sp->seed ^ 0x9015Eull);
  const uint64_t seed_split = splitmix64(sp->seed ^ 0x5917ull);

  // 1) distinct coordinates: sample, sort, de-duplicate, top up until exact
  std::vector<uint64_t> keys;
  const uint64_t CH = 1ull << 20;
  for (uint64_t round = 0; keys.size() < nnz; ++round) {
    if (round > 200) return fail(CTC_ERR_INVALID_ARG, "synth: cannot place %llu distinct nonzeros",
                                 (unsigned long long)nnz);
    uint64_t need = nnz - keys.size();
    std::vector<uint64_t> fresh(need);
    uint64_t nch = (need + CH - 1) / CH;
    parallel_for(nch, host_threads(), [&](uint64_t lo, uint64_t hi, unsigned) {
      for (uint64_t c = lo; c < hi; ++c) {
        Xoshiro g(splitmix64(seed_coord ^ (round << 40) ^ c));
        uint64_t b = c * CH, e = std::min(need, b + CH);
        for (uint64_t p = b; p < e; ++p) {
          uint64_t key = 0;
          for (uint32_t m = 0; m < N; ++m) key = key * sp->dims[m] + smp[m].draw(g);
          fresh[p] = key;
        }
      }
    });
    radix_sort_u64(fresh, maxkey);
    fresh.erase(std::unique(fresh.begin(), fresh.end()), fresh.end());
    std::vector<uint64_t> merged;
    merged.reserve(keys.size() + fresh.size());
    std::merge(keys.begin(), keys.end(), fresh.begin(), fresh.end(), std::back_inserter(merged));
    merged.erase(std::unique(merged.begin(), merged.end()), merged.end());
    keys.swap(merged);
  }

  // 2) decode coordinates and evaluate the CP ground truth
  const uint32_t TR = std::max<uint32_t>(1, sp->true_rank);
  std::vector<std::vector<float>> truth(N);
  for (uint32_t m = 0; m < N; ++m) {
    truth[m].resize(sp->dims[m] * TR);
    parallel_for(sp->dims[m] * TR, host_threads(), [&](uint64_t lo, uint64_t hi, unsigned) {
      for (uint64_t x = lo; x < hi; ++x)
        truth[m][x] = (float)u01(splitmix64(seed_truth ^ ((uint64_t)m << 58) ^ x));
    });
  }
  std::vector<uint32_t> coord[CTC_MAX_MODES];
  for (uint32_t m = 0; m < N; ++m) coord[m].resize(nnz);
  std::vector<double> val(nnz);
  unsigned T = host_threads();
  std::vector<double> tmin(T, 1e300), tmax(T, -1e300);
  parallel_for(nnz, T, [&](uint64_t lo, uint64_t hi, unsigned th) {
    for (uint64_t p = lo; p < hi; ++p) {
      uint64_t k = keys[p];
      for (int m = (int)N - 1; m >= 0; --m) {
        coord[m][p] = (uint32_t)(k % sp->dims[m]);
        k /= sp->dims[m];
      }
      double v = 0;
      for (uint32_t r = 0; r < TR; ++r) {
        double prod = 1;
        for (uint32_t m = 0; m < N; ++m) prod *= truth[m][(uint64_t)coord[m][p] * TR + r];
        v += prod;
      }
      val[p] = v;
      tmin[th] = std::min(tmin[th], v);
      tmax[th] = std::max(tmax[th], v);
    }
  });
  double vmin = *std::min_element(tmin.begin(), tmin.end());
  double vmax = *std::max_element(tmax.begin(), tmax.end());
  const double span = vmax > vmin ? vmax - vmin : 1.0;
  parallel_for(nnz, T, [&](uint64_t lo, uint64_t hi, unsigned) {
    for (uint64_t p = lo; p < hi; ++p) {
      double v = val[p];
      if (sp->value_map != CTC_VALUES_RAW) v = 1.0 + 4.0 * (v - vmin) / span;
      if (sp->noise_std > 0) {
        // Box-Muller on two counter-based uniforms of this coordinate
        uint64_t h = splitmix64(seed_noise ^ keys[p]);
        double u1 = u01(h), u2 = u01(splitmix64(h));
        if (u1 < 1e-300) u1 = 1e-300;
        v += sp->noise_std * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
      }
      if (sp->value_map == CTC_VALUES_INT_1_5) v = std::min(5.0, std::max(1.0, std::nearbyint(v)));
      val[p] = v;
    }
  });

  // 3) exact-size test split: the ntest coordinates with the smallest hash
  const uint64_t ntest = (uint64_t)std::llround(sp->test_fraction * (double)nnz);
  std::vector<uint8_t> is_test(nnz, 0);
  if (ntest > 0) {
    std::vector<uint64_t> score(nnz);
    parallel_for(nnz, T, [&](uint64_t lo, uint64_t hi, unsigned) {
      for (uint64_t p = lo; p < hi; ++p) score[p] = splitmix64(seed_split ^ keys[p]);
    });
    std::vector<uint64_t> s2 = score;
    std::nth_element(s2.begin(), s2.begin() + (ntest - 1), s2.end());
    const uint64_t thr = s2[ntest - 1];
    uint64_t taken = 0;
    for (uint64_t p = 0; p < nnz; ++p)
      if (score[p] < thr) {
        is_test[p] = 1;
        ++taken;
      }
    for (uint64_t p = 0; p < nnz && taken < ntest; ++p)
      if (score[p] == thr) {
        is_test[p] = 1;
        ++taken;
      }
  }
  ctc_coo *tr = nullptr, *te = nullptr;
  CTC_TRY(ctc_coo_alloc(N, nnz - ntest, sp->dims, &tr));
  if (ntest > 0 && test) {
    ctc_status s = ctc_coo_alloc(N, ntest, sp->dims, &te);
    if (s != CTC_OK) {
      ctc_coo_free(tr);
      return s;
    }
  }
  uint64_t a = 0, b = 0;
  for (uint64_t p = 0; p < nnz; ++p) {
    ctc_coo *dst = is_test[p] ? te : tr;
    uint64_t &q = is_test[p] ? b : a;
    if (!dst) {
      ++q;
      continue;
    }
    for (uint32_t m = 0; m < N; ++m) dst->ind[m][q] = coord[m][p];
    dst->vals[q] = val[p];
    ++q;
  }
  *train = tr;
  if (test) *test = te;
  return CTC_OK;
}

// -------------------------------------------------------------------------
// data preparation (software/scripts/tvtensor.py, stensor.py)
// -------------------------------------------------------------------------
namespace {

// gathers the nonzeros of t in the given order into a new tensor
ctc_status gather(const ctc_coo *t, const uint32_t *order, uint64_t n, ctc_coo **out) {
  CTC_TRY(ctc_coo_alloc(t->nmodes, n, t->dims, out));
  for (uint32_t m = 0; m < t->nmodes; ++m) {
    (*out)->base[m] = t->base[m];
    for (uint64_t p = 0; p < n; ++p) (*out)->ind[m][p] = t->ind[m][order[p]];
  }
  for (uint64_t p = 0; p < n; ++p) (*out)->vals[p] = t->vals[order[p]];
  return CTC_OK;
}

// stable lexicographic order of the nonzeros by keys[0], keys[1], ...
ctc_status sorted_order(const ctc_coo *t, const uint32_t *keys, uint32_t nkeys, std::vector<uint32_t> &order) {
  order.resize(t->nnz);
  std::iota(order.begin(), order.end(), 0u);
  std::vector<uint32_t> tmp;
  for (uint32_t k = nkeys; k-- > 0;) {  // least significant key first
    if (keys[k] >= t->nmodes) return fail(CTC_ERR_INVALID_ARG, "sort key %u of a %u-mode tensor", keys[k], t->nmodes);
    ctc::counting_sort(t->ind[keys[k]], t->dims[keys[k]], order, tmp);
  }
  return CTC_OK;
}

}  // namespace

extern "C" ctc_status ctc_coo_sort(const ctc_coo *t, const uint32_t *keys, uint32_t nkeys, ctc_coo **out) {
  if (!t || !out || (nkeys && !keys) || nkeys > CTC_MAX_MODES) return fail(CTC_ERR_INVALID_ARG, "ctc_coo_sort");
  std::vector<uint32_t> order;
  CTC_TRY(sorted_order(t, keys, nkeys, order));
  return gather(t, order.data(), t->nnz, out);
}

extern "C" ctc_status ctc_coo_split(const ctc_coo *t, uint64_t seed, ctc_coo **train, ctc_coo **validation,
                                    ctc_coo **test) {
  if (!t || !train || !validation || !test || t->nmodes < 2) return fail(CTC_ERR_INVALID_ARG, "ctc_coo_split");
  *train = *validation = *test = nullptr;
  // nonzeros in (mode 0, mode 1) order (tvtensor.py:64), then one draw from
  // 0..9 each: 0-7 train, 8 test, 9 validation (tvtensor.py:73-81)
  const uint32_t keys[2] = {0, 1};
  std::vector<uint32_t> order;
  CTC_TRY(sorted_order(t, keys, 2, order));
  std::vector<uint32_t> parts[3];
  Xoshiro rng(seed ^ 0x5EED5917ull);
  for (uint32_t p : order) {
    const uint64_t d = ((rng.next() >> 32) * 10) >> 32;
    parts[d < 8 ? 0 : (d == 8 ? 2 : 1)].push_back(p);
  }
  ctc_status s = gather(t, parts[0].data(), parts[0].size(), train);
  if (s == CTC_OK) s = gather(t, parts[1].data(), parts[1].size(), validation);
  if (s == CTC_OK) s = gather(t, parts[2].data(), parts[2].size(), test);
  if (s != CTC_OK) {
    ctc_coo_free(*train);
    ctc_coo_free(*validation);
    ctc_coo_free(*test);
    *train = *validation = *test = nullptr;
  }
  return s;
}
