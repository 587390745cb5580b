/*
 * csaw_oracle.c -- CPU oracle for the C-SAW sampling / random-walk path.
 *
 * TEST INFRASTRUCTURE ONLY (see csaw_oracle.h).  Plain C99 + GCC __int128,
 * built with -ffp-contract=off and without fast-math so every float
 * operation is a single IEEE binary32 operation in program order.
 */
#include "csaw_oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Counter domain tags (bits 28..31 of counter word 3). */
enum { DOM_SEEDS = 1, DOM_RMAT = 2, DOM_WEIGHTS = 3, DOM_SELECT = 4, DOM_WALK = 5, DOM_MDRW = 6 };

/* ------------------------------------------------------------------------ */
/* Philox4x32-10, Salmon, Moraes, Dror, Shaw, "Parallel Random Numbers: As  */
/* Easy as 1, 2, 3", SC'11.  Multipliers and Weyl increments as published. */
/* ------------------------------------------------------------------------ */
void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline void draw(uint64_t key, uint32_t a, uint32_t b, uint32_t c, uint32_t d, uint32_t o[4]) {
  uint32_t ctr[4] = {a, b, c, d};
  uint32_t k[2] = {(uint32_t)key, (uint32_t)(key >> 32)};
  orc_philox(ctr, k, o);
}

static inline uint64_t t64(const uint32_t o[4]) { return ((uint64_t)o[1] << 32) | o[0]; }

static inline uint64_t mulhi64(uint64_t a, uint64_t b) {
  return (uint64_t)(((unsigned __int128)a * b) >> 64);
}

void orc_free(void *p) { free(p); }

/* ------------------------------------------------------------------------ */
/* RMAT generator (synthetic input; the reference reads beg.bin/csr.bin,    */
/* graph.hpp:18-74, and ships no graph).                                     */
/* ------------------------------------------------------------------------ */
static uint32_t thr(double x) {
  if (x <= 0.0) return 0;
  if (x >= 1.0) return 0xFFFFFFFFu;
  return (uint32_t)(x * 4294967296.0);
}

static void radix_sort_u64(uint64_t *a, uint64_t *tmp, int64_t n, int key_bits) {
  int64_t *cnt = (int64_t *)malloc(sizeof(int64_t) * 65536);
  for (int shift = 0; shift < key_bits; shift += 16) {
    memset(cnt, 0, sizeof(int64_t) * 65536);
    for (int64_t i = 0; i < n; ++i) cnt[(a[i] >> shift) & 0xFFFF]++;
    int64_t s = 0;
    for (int i = 0; i < 65536; ++i) { int64_t c = cnt[i]; cnt[i] = s; s += c; }
    for (int64_t i = 0; i < n; ++i) tmp[cnt[(a[i] >> shift) & 0xFFFF]++] = a[i];
    memcpy(a, tmp, sizeof(uint64_t) * n);
  }
  free(cnt);
}

int orc_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t key,
             int64_t *V_out, int64_t *E_out, int64_t **off_out, int32_t **adj_out) {
  if (scale < 1 || scale > 30 || edge_factor < 1) return -1;
  const int64_t V = (int64_t)1 << scale;
  const int64_t M = (int64_t)edge_factor << scale;
  const uint32_t ta = thr(a), tab = thr(a + b), tabc = thr(a + b + c);
  uint64_t *keys = (uint64_t *)malloc(sizeof(uint64_t) * 2 * M);
  uint64_t *tmp = (uint64_t *)malloc(sizeof(uint64_t) * 2 * M);
  if (!keys || !tmp) { free(keys); free(tmp); return -2; }
  int64_t n = 0;
  for (int64_t e = 0; e < M; ++e) {
    uint32_t src = 0, dst = 0, o[4];
    for (int l = 0; l < scale; ++l) {
      if ((l & 3) == 0) draw(key, (uint32_t)e, (uint32_t)(e >> 32), (uint32_t)(l >> 2), (uint32_t)DOM_RMAT << 28, o);
      uint32_t r = o[l & 3];
      uint32_t bit = 1u << (scale - 1 - l);
      if (r < ta) {
      } else if (r < tab) {
        dst |= bit;
      } else if (r < tabc) {
        src |= bit;
      } else {
        src |= bit;
        dst |= bit;
      }
    }
    if (src == dst) continue;
    keys[n++] = ((uint64_t)src << 32) | dst;
    keys[n++] = ((uint64_t)dst << 32) | src;
  }
  radix_sort_u64(keys, tmp, n, 64);
  free(tmp);
  int64_t *off = (int64_t *)calloc((size_t)V + 1, sizeof(int64_t));
  int64_t E = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (i > 0 && keys[i] == keys[i - 1]) continue;
    keys[E++] = keys[i];
  }
  int32_t *adj = (int32_t *)malloc(sizeof(int32_t) * (E ? E : 1));
  for (int64_t i = 0; i < E; ++i) {
    off[(keys[i] >> 32) + 1]++;
    adj[i] = (int32_t)(uint32_t)keys[i];
  }
  for (int64_t v = 0; v < V; ++v) off[v + 1] += off[v];
  free(keys);
  *V_out = V;
  *E_out = E;
  *off_out = off;
  *adj_out = adj;
  return 0;
}

/* Seeds, A.8: seed_i = mulhi64(philox(i), V). */
void orc_seeds(uint64_t key, uint32_t request, int64_t V, int64_t n, int32_t *out) {
  for (int64_t i = 0; i < n; ++i) {
    uint32_t o[4];
    draw(key, (uint32_t)i, (uint32_t)(i >> 32), request, (uint32_t)DOM_SEEDS << 28, o);
    out[i] = (int32_t)mulhi64(t64(o), (uint64_t)V);
  }
}

/* Edge weights uniform in (0, 1]: ((o0 >> 8) + 1) * 2^-24, exact in binary32. */
void orc_weights(uint64_t key, int64_t E, float *w) {
  for (int64_t e = 0; e < E; ++e) {
    uint32_t o[4];
    draw(key, (uint32_t)e, (uint32_t)(e >> 32), 0, (uint32_t)DOM_WEIGHTS << 28, o);
    w[e] = (float)((o[0] >> 8) + 1u) * 5.9604644775390625e-08f;
  }
}

/* Sequential FP32 CDF in CSR order: c_i = fl32(c_{i-1} + w_i), c_{-1} = 0.
 * This is C-SAW's probability cache idea (functions.cuh:190-208, 329-341)
 * with a fixed accumulation order instead of the truncating parallel scan. */
void orc_fp32_cdf(int64_t V, const int64_t *off, const float *w, float *cdf) {
  for (int64_t v = 0; v < V; ++v) {
    float acc = 0.0f;
    for (int64_t i = off[v]; i < off[v + 1]; ++i) {
      acc = acc + w[i];
      cdf[i] = acc;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Inverse transform sampling.                                              */
/* ------------------------------------------------------------------------ */
static inline int64_t deg(const orc_graph *g, int64_t v) { return g->off[v + 1] - g->off[v]; }

/* EdgeBias hook defaults (api.cuh:13-20): degree of the neighbour; the
 * uniform policy is the api.cuh "return 1" alternative (api.cuh:18-19). */
static inline uint32_t edge_bias(const orc_graph *g, int32_t bias, int32_t u) {
  if (bias == ORC_BIAS_UNIFORM) return 1u;
  return (uint32_t)deg(g, u);
}

/* Update hook default (api.cuh:23-27): the selected vertex itself. */
static inline int32_t update_hook(int32_t selected, int32_t source) {
  (void)source;
  return selected;
}

/* first i in [0,n) with C[i] > x (binary_search, functions.cuh:11-43, made
 * an exact upper bound over an inclusive CDF). */
static inline int64_t upper_bound_u64(const uint64_t *C, int64_t n, uint64_t x) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (C[mid] > x) hi = mid; else lo = mid + 1;
  }
  return lo;
}

static inline int64_t upper_bound_f32(const float *C, int64_t n, float x) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (C[mid] > x) hi = mid; else lo = mid + 1;
  }
  return lo;
}

typedef struct {
  int mode; /* 0 uniform over d, 1 integer CDF, 2 fp32 CDF */
  int64_t d;
  uint64_t T;
  float Tf;
  const uint64_t *C;
  const float *Cf;
} its_t;

/* Builds the ITS state for vertex v (ITS, functions.cuh:160-176): integer
 * biases get an exact inclusive uint64 CDF; all-zero weights fall back to
 * uniform over the d neighbours. */
static void its_build(const orc_graph *g, int32_t bias, int64_t v, uint64_t *scratch, its_t *s) {
  const int64_t o = g->off[v];
  s->d = deg(g, v);
  s->mode = 0;
  if (bias == ORC_BIAS_UNIFORM) return;
  if (bias == ORC_BIAS_WEIGHT) {
    s->Cf = g->cdf + o;
    s->Tf = s->Cf[s->d - 1];
    if (s->Tf > 0.0f) s->mode = 2;
    return;
  }
  uint64_t acc = 0;
  for (int64_t i = 0; i < s->d; ++i) {
    acc += edge_bias(g, bias, g->adj[o + i]);
    scratch[i] = acc;
  }
  s->C = scratch;
  s->T = acc;
  if (acc > 0) s->mode = 1;
}

static inline int64_t its_pick(const its_t *s, const uint32_t o[4]) {
  const uint64_t t = t64(o);
  if (s->mode == 1) return upper_bound_u64(s->C, s->d, mulhi64(t, s->T));
  if (s->mode == 2) {
    float u = (float)(o[0] >> 8) * 5.9604644775390625e-08f;
    float x = u * s->Tf;
    return upper_bound_f32(s->Cf, s->d, x);
  }
  return (int64_t)mulhi64(t, (uint64_t)s->d);
}

/* Region-remap mode (heur_normalize, functions.cuh:743-868; the reference's
 * default collision policy, NORMALIZE at main.cu:26, 80-82).  Its draws live
 * in the normalised CDF p_i = fl32(fl64(C_i) / fl64(T)) (ITS's division by
 * the total, functions.cuh:171-174, over the inclusive CDF; uniform:
 * p_i = fl32((i + 1) / d)), and a draw r in [0, 1) picks the first i with
 * p_i > r, clamped to d - 1 (the reference's search could return d). */
static inline float remap_p(const its_t *s, int64_t i) {
  if (s->mode == 1) return (float)((double)s->C[i] / (double)s->T);
  if (s->mode == 2) return (float)((double)s->Cf[i] / (double)s->Tf);
  return (float)((double)(i + 1) / (double)s->d);
}

static inline int64_t remap_pick(const its_t *s, float r) {
  int64_t lo = 0, hi = s->d;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (remap_p(s, mid) > r) hi = mid; else lo = mid + 1;
  }
  return lo < s->d ? lo : s->d - 1;
}

typedef struct {
  uint64_t *cdf;   /* maxdeg */
  uint32_t *stamp; /* maxdeg */
  uint32_t cur;
} work_t;

static int64_t max_degree(const orc_graph *g) {
  int64_t m = 1;
  for (int64_t v = 0; v < g->V; ++v) {
    int64_t d = deg(g, v);
    if (d > m) m = d;
  }
  return m;
}

static void work_init(work_t *w, int64_t maxdeg) {
  w->cdf = (uint64_t *)malloc(sizeof(uint64_t) * maxdeg);
  w->stamp = (uint32_t *)calloc((size_t)maxdeg, sizeof(uint32_t));
  w->cur = 0;
}
static void work_free(work_t *w) { free(w->cdf); free(w->stamp); }

/* Per-source selection (canonical form of heur_normalize/select,
 * functions.cuh:306-410 and 743-868):
 *  - d == 0: nothing (main.cu:99-104);
 *  - without replacement and d <= F: every neighbour in CSR order
 *    (functions.cuh:859-867, which copied F entries even when d < F);
 *  - otherwise slots j = 0..F-1 in order; attempt a draws
 *    philox(key; (j<<16)|a, pos, sample, DOM_SELECT<<28 | request<<8 | hop)
 *    and is accepted unless already taken (bitmap_search, functions.cuh:45-70);
 *    after ORC_RESELECT attempts (the 400-try bound, functions.cuh:844) the
 *    first untaken index scanning upward cyclically from the last draw wins;
 *  - dedup == 2: the same slots with heur_normalize's region remap on
 *    collisions (above). */
static int32_t select_core(const orc_graph *g, int32_t bias, int32_t dedup, int32_t F, int64_t v,
                           uint64_t key, uint32_t request, uint32_t hop, uint32_t sample, uint32_t pos,
                           int32_t *idx, work_t *w) {
  const int64_t d = deg(g, v);
  if (d == 0 || F <= 0) return 0;
  if (dedup && d <= F) {
    for (int64_t i = 0; i < d; ++i) idx[i] = (int32_t)i;
    return (int32_t)d;
  }
  its_t s;
  its_build(g, bias, v, w->cdf, &s);
  const uint32_t c3 = ((uint32_t)DOM_SELECT << 28) | (request << 8) | (hop & 0xFF);
  uint32_t o[4];
  if (!dedup) {
    for (int32_t j = 0; j < F; ++j) {
      draw(key, (uint32_t)j << 16, pos, sample, c3, o);
      idx[j] = (int32_t)its_pick(&s, o);
    }
    return F;
  }
  if (++w->cur == 0) { memset(w->stamp, 0, sizeof(uint32_t) * (size_t)d); w->cur = 1; }
  if (dedup == 2) {
    /* region remap: per slot, bounds lb = 0, hb = 1; each attempt draws r,
     * picks; on a collision with the picked interval [A, B) it remaps r into
     * the rest of [lb, hb) with heur_normalize's float sequence
     * (functions.cuh:813-843, the reads of degree_l at selected - 1 and
     * selected + 1 taken as A = p_{i-1} (0 for i = 0) and B = p_i) and
     * picks again; after ORC_RESELECT attempts the cyclic fallback. */
    for (int32_t j = 0; j < F; ++j) {
      float lb = 0.0f, hb = 1.0f;
      int64_t i = 0;
      int found = 0;
      for (uint32_t a = 0; a < ORC_RESELECT && !found; ++a) {
        draw(key, ((uint32_t)j << 16) | a, pos, sample, c3, o);
        float r = (float)(o[0] >> 8) * 5.9604644775390625e-08f;
        i = remap_pick(&s, r);
        if (w->stamp[i] != w->cur) { found = 1; break; }
        const float A = i > 0 ? remap_p(&s, i - 1) : 0.0f;
        const float B = remap_p(&s, i);
        const float lambda = (A - lb) / (A - lb + hb - B);
        const float delta = (B - A) / (hb - lb);
        r = r / lambda;
        if (r < A) {
          hb = A;
        } else {
          r = r + delta;
          lb = B;
        }
        i = remap_pick(&s, r);
        if (w->stamp[i] != w->cur) found = 1;
      }
      if (!found) {
        do { i = (i + 1 == d) ? 0 : i + 1; } while (w->stamp[i] == w->cur);
      }
      w->stamp[i] = w->cur;
      idx[j] = (int32_t)i;
    }
    return F;
  }
  for (int32_t j = 0; j < F; ++j) {
    int64_t i = 0;
    int found = 0;
    for (uint32_t a = 0; a < ORC_RESELECT; ++a) {
      draw(key, ((uint32_t)j << 16) | a, pos, sample, c3, o);
      i = its_pick(&s, o);
      if (w->stamp[i] != w->cur) { found = 1; break; }
    }
    if (!found) {
      do { i = (i + 1 == d) ? 0 : i + 1; } while (w->stamp[i] == w->cur);
    }
    w->stamp[i] = w->cur;
    idx[j] = (int32_t)i;
  }
  return F;
}

int32_t orc_select_source(const orc_graph *g, int32_t bias, int32_t dedup, int32_t fanout,
                          int32_t v, uint64_t key, uint32_t request, uint32_t hop,
                          uint32_t sample, uint32_t pos, int32_t *idx) {
  work_t w;
  int64_t d = deg(g, v);
  work_init(&w, d > 0 ? d : 1);
  int32_t k = select_core(g, bias, dedup, fanout, v, key, request, hop, sample, pos, idx, &w);
  work_free(&w);
  return k;
}

int64_t orc_sample_capacity(const orc_request *r) {
  int64_t total = 0, n = r->samples;
  for (int h = 0; h < r->depth; ++h) {
    n *= r->fanout[h];
    total += n;
  }
  return total;
}

static int valid_request(const orc_graph *g, const orc_request *r, const int32_t *seeds) {
  if (r->samples < 0 || r->depth < 0 || r->depth > ORC_MAX_HOPS) return 0;
  if (r->bias == ORC_BIAS_WEIGHT && !g->cdf) return 0;
  for (int h = 0; h < r->depth; ++h)
    if (r->fanout[h] < 0) return 0;
  for (int32_t s = 0; s < r->samples; ++s)
    if (seeds[s] < 0 || seeds[s] >= g->V) return 0;
  return 1;
}

/* Sequential, hop-synchronous restatement of the persistent queue loop
 * (main.cu:21-93 + frontier(), functions.cuh:971-1011): hop h expands the
 * frontier in position order; every emitted vertex joins frontier h+1. */
static int64_t sample_sequential(const orc_graph *g, const orc_request *r, const int32_t *seeds,
                                 int32_t *src, int32_t *dst, int32_t *smp, int64_t *hop_off) {
  const int32_t S = r->samples;
  int64_t n = S;
  int64_t *fo = (int64_t *)malloc(sizeof(int64_t) * ((size_t)S + 1));
  int32_t *fv0 = (int32_t *)malloc(sizeof(int32_t) * (S ? S : 1));
  int32_t *fs0 = (int32_t *)malloc(sizeof(int32_t) * (S ? S : 1));
  for (int32_t s = 0; s < S; ++s) { fv0[s] = seeds[s]; fs0[s] = s; fo[s] = s; }
  fo[S] = S;
  const int32_t *fv = fv0, *fs = fs0;
  work_t w;
  work_init(&w, max_degree(g));
  int32_t maxF = 1;
  for (int h = 0; h < r->depth; ++h) if (r->fanout[h] > maxF) maxF = r->fanout[h];
  int32_t *idx = (int32_t *)malloc(sizeof(int32_t) * maxF);
  int64_t E = 0;
  for (int h = 0; h < r->depth; ++h) {
    const int64_t hop_start = E;
    int64_t *ho = hop_off + (int64_t)h * (S + 1);
    int32_t next_s = 0;
    for (int64_t e = 0; e < n; ++e) {
      const int32_t s = fs[e];
      while (next_s <= s) ho[next_s++] = E;
      const int32_t v = fv[e];
      const int32_t k = select_core(g, r->bias, r->dedup, r->fanout[h], v, r->key, r->request, (uint32_t)h,
                                    (uint32_t)s, (uint32_t)(e - fo[s]), idx, &w);
      const int64_t o = g->off[v];
      for (int32_t j = 0; j < k; ++j) {
        src[E] = v;
        dst[E] = update_hook(g->adj[o + idx[j]], v);
        smp[E] = s;
        ++E;
      }
    }
    while (next_s <= S) ho[next_s++] = E;
    for (int32_t s = 0; s <= S; ++s) fo[s] = ho[s] - hop_start;
    fv = dst + hop_start;
    fs = smp + hop_start;
    n = E - hop_start;
  }
  free(idx);
  work_free(&w);
  free(fo);
  free(fv0);
  free(fs0);
  return E;
}

/* Per-sample expansion for the OpenMP baseline: the same select_core calls
 * with the same (hop, sample, position) coordinates, so the assembled
 * result is byte-identical to sample_sequential. */
typedef struct {
  int32_t *v;  /* expanded vertex list of every hop, concatenated */
  int32_t *p;  /* parent vertex per entry */
  int64_t cap, len;
  int64_t cnt[ORC_MAX_HOPS];
} sample_buf;

static void sbuf_push(sample_buf *b, int32_t parent, int32_t v) {
  if (b->len == b->cap) {
    b->cap = b->cap ? b->cap * 2 : 64;
    b->v = (int32_t *)realloc(b->v, sizeof(int32_t) * b->cap);
    b->p = (int32_t *)realloc(b->p, sizeof(int32_t) * b->cap);
  }
  b->p[b->len] = parent;
  b->v[b->len] = v;
  b->len++;
}

static int64_t sample_parallel(const orc_graph *g, const orc_request *r, const int32_t *seeds,
                               int32_t *src, int32_t *dst, int32_t *smp, int64_t *hop_off, int nthreads) {
  const int32_t S = r->samples;
  sample_buf *bufs = (sample_buf *)calloc((size_t)(S ? S : 1), sizeof(sample_buf));
  const int64_t maxdeg = max_degree(g);
  int32_t maxF = 1;
  for (int h = 0; h < r->depth; ++h) if (r->fanout[h] > maxF) maxF = r->fanout[h];
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads)
#endif
  {
    work_t w;
    work_init(&w, maxdeg);
    int32_t *idx = (int32_t *)malloc(sizeof(int32_t) * maxF);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
    for (int32_t s = 0; s < S; ++s) {
      sample_buf *b = &bufs[s];
      int64_t fbeg = 0, fend = 0; /* frontier = b->v[fbeg..fend) (hop 0: the seed) */
      int32_t seed = seeds[s];
      for (int h = 0; h < r->depth; ++h) {
        int64_t before = b->len;
        int64_t n = (h == 0) ? 1 : fend - fbeg;
        for (int64_t e = 0; e < n; ++e) {
          const int32_t v = (h == 0) ? seed : b->v[fbeg + e];
          const int32_t k = select_core(g, r->bias, r->dedup, r->fanout[h], v, r->key, r->request,
                                        (uint32_t)h, (uint32_t)s, (uint32_t)e, idx, &w);
          const int64_t o = g->off[v];
          for (int32_t j = 0; j < k; ++j) sbuf_push(b, v, update_hook(g->adj[o + idx[j]], v));
        }
        b->cnt[h] = b->len - before;
        fbeg = before;
        fend = b->len;
      }
    }
    free(idx);
    work_free(&w);
  }
  /* hop-major offsets */
  int64_t E = 0;
  for (int h = 0; h < r->depth; ++h) {
    int64_t *ho = hop_off + (int64_t)h * (S + 1);
    for (int32_t s = 0; s < S; ++s) { ho[s] = E; E += bufs[s].cnt[h]; }
    ho[S] = E;
  }
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 64)
#endif
  for (int32_t s = 0; s < S; ++s) {
    int64_t pos = 0;
    for (int h = 0; h < r->depth; ++h) {
      int64_t at = hop_off[(int64_t)h * (S + 1) + s];
      for (int64_t i = 0; i < bufs[s].cnt[h]; ++i, ++pos) {
        src[at + i] = bufs[s].p[pos];
        dst[at + i] = bufs[s].v[pos];
        smp[at + i] = s;
      }
    }
    free(bufs[s].v);
    free(bufs[s].p);
  }
  free(bufs);
  return E;
}

int64_t orc_sample(const orc_graph *g, const orc_request *r, const int32_t *seeds,
                   int32_t *src, int32_t *dst, int32_t *smp, int64_t *hop_off, int nthreads) {
  if (!valid_request(g, r, seeds)) return -1;
  if (nthreads <= 1) return sample_sequential(g, r, seeds, src, dst, smp, hop_off);
  return sample_parallel(g, r, seeds, src, dst, smp, hop_off, nthreads);
}

/* Random walk (FrontierSize 1, NeighborSize 1, Depth = length; the
 * reference's README.md:34 "Depth/Length"): one ITS draw per step with
 * counter (step, walker lo, walker hi, DOM_WALK<<28 | request). */
int64_t orc_walk(const orc_graph *g, int32_t bias, uint64_t key, uint32_t request,
                 const int32_t *seeds, int64_t nwalkers, int32_t len, int32_t *W, int nthreads) {
  if (bias == ORC_BIAS_WEIGHT && !g->cdf) return -1;
  for (int64_t w = 0; w < nwalkers; ++w)
    if (seeds[w] < 0 || seeds[w] >= g->V) return -1;
  const int64_t maxdeg = max_degree(g);
  int64_t steps = 0;
  if (nthreads < 1) nthreads = 1;
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads) reduction(+ : steps)
#endif
  {
    uint64_t *scratch = (uint64_t *)malloc(sizeof(uint64_t) * maxdeg);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 256)
#endif
    for (int64_t w = 0; w < nwalkers; ++w) {
      int32_t *row = W + w * ((int64_t)len + 1);
      int32_t v = seeds[w];
      row[0] = v;
      for (int32_t t = 0; t < len; ++t) {
        if (v >= 0 && deg(g, v) == 0) v = -1;
        if (v < 0) { row[t + 1] = -1; continue; }
        its_t s;
        its_build(g, bias, v, scratch, &s);
        uint32_t o[4];
        draw(key, (uint32_t)t, (uint32_t)w, (uint32_t)(w >> 32), ((uint32_t)DOM_WALK << 28) | request, o);
        v = g->adj[g->off[v] + its_pick(&s, o)];
        row[t + 1] = v;
        ++steps;
      }
    }
    free(scratch);
  }
  return steps;
}

/* VertexBias hook default (api.cuh:4-11): the vertex degree. */
static inline uint32_t vertex_bias(const orc_graph *g, int32_t v) { return (uint32_t)deg(g, v); }

/* Multi-dimensional random walk (check_layer, main.cu:95-194, and ITS_MDRW,
 * functions.cuh:1014-1060): per step pick a frontier slot proportionally to
 * VertexBias, then a uniform neighbour of it (rand_integer, functions.cuh:210);
 * a degree-0 pick consumes the step (main.cu:162-163). */
int64_t orc_mdrw(const orc_graph *g, uint64_t key, uint32_t request, int32_t samples, int32_t fs,
                 int32_t depth, const int32_t *seeds, int32_t *src, int32_t *dst, int64_t *off) {
  if (fs < 1 || depth < 0 || samples < 0) return -1;
  for (int64_t i = 0; i < (int64_t)samples * fs; ++i)
    if (seeds[i] < 0 || seeds[i] >= g->V) return -1;
  int32_t *front = (int32_t *)malloc(sizeof(int32_t) * fs);
  uint64_t *C = (uint64_t *)malloc(sizeof(uint64_t) * fs);
  const uint32_t c3 = ((uint32_t)DOM_MDRW << 28) | request;
  int64_t E = 0;
  for (int32_t s = 0; s < samples; ++s) {
    off[s] = E;
    memcpy(front, seeds + (int64_t)s * fs, sizeof(int32_t) * fs);
    for (int32_t t = 0; t < depth; ++t) {
      uint64_t acc = 0;
      for (int32_t q = 0; q < fs; ++q) { acc += vertex_bias(g, front[q]); C[q] = acc; }
      uint32_t o[4];
      draw(key, (uint32_t)t, 0u, (uint32_t)s, c3, o);
      const int64_t q = acc ? upper_bound_u64(C, fs, mulhi64(t64(o), acc))
                            : (int64_t)mulhi64(t64(o), (uint64_t)fs);
      const int32_t u = front[q];
      const int64_t d = deg(g, u);
      if (d == 0) continue;
      draw(key, (uint32_t)t, 1u, (uint32_t)s, c3, o);
      const int32_t nb = g->adj[g->off[u] + (int64_t)mulhi64(t64(o), (uint64_t)d)];
      src[E] = u;
      dst[E] = nb;
      ++E;
      front[q] = nb;
    }
  }
  off[samples] = E;
  free(front);
  free(C);
  return E;
}
