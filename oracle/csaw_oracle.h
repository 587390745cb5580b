/*
 * csaw_oracle.h -- CPU oracle for the C-SAW sampling / random-walk path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a single-threaded C restatement of
 * C-SAW's selection logic (inverse-transform sampling over per-neighbour
 * biases, binary search, taken-set dedup with bounded reselect, frontier
 * expansion, the VertexBias/EdgeBias/Update hooks) under the canonical,
 * deterministic semantics documented in DESIGN.md ("Canonical semantics").
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker or as the timed CPU baseline.  The
 * product library never links or calls anything here.
 *
 * Reference functions restated (paths relative to the reference root):
 *   non-stream/functions.cuh:160-176  ITS (CDF build)      -> its_int / its_f32
 *   non-stream/functions.cuh:112-158  gpu_prefix (scan)     -> build_int_cdf
 *   non-stream/functions.cuh:11-43    binary_search         -> upper_bound_*
 *   non-stream/functions.cuh:45-70    bitmap_search         -> taken set
 *   non-stream/functions.cuh:743-868  heur_normalize/select -> select_source
 *   non-stream/functions.cuh:872-892  get_neighbors         -> select_source
 *   non-stream/functions.cuh:971-1011 frontier              -> orc_sample
 *   non-stream/functions.cuh:1014-1060 ITS_MDRW, main.cu:95-194 -> orc_mdrw
 *   non-stream/api.cuh:4-27           VertexBias/EdgeBias/Update defaults
 *   streaming/streaming_sampling.cu:251-253 degree = off[v+1]-off[v]
 *
 * The reference draws with cuRAND XORWOW seeded by thread id; that stream is
 * replaced by counter-based Philox4x32-10 keyed by logical coordinates, so the
 * exact edges the reference would emit are not reproducible (parity against
 * the reference's own stream is unpinned by design).
 */
#ifndef CSAW_ORACLE_H
#define CSAW_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_BIAS_UNIFORM = 0, ORC_BIAS_DEGREE = 1, ORC_BIAS_WEIGHT = 2, ORC_BIAS_USER = 3 };
enum { ORC_MAX_HOPS = 8, ORC_RESELECT = 400 };

typedef struct {
  int64_t V, E;
  const int64_t *off; /* V+1 */
  const int32_t *adj; /* E   */
  const float *cdf;   /* E, sequential FP32 CDF per vertex (ORC_BIAS_WEIGHT), may be NULL otherwise */
} orc_graph;

typedef struct {
  int32_t samples;
  int32_t depth;
  int32_t fanout[ORC_MAX_HOPS];
  int32_t bias;
  int32_t dedup; /* 0: with replacement, 1: without replacement per source */
  uint32_t request;
  uint64_t key;
} orc_request;

/* Philox4x32-10 (Salmon et al., SC'11). */
void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* RMAT: ef*2^scale generated pairs, symmetrised, self loops and duplicates
 * removed, sorted CSR.  Returns 0 on success; *off / *adj are malloc'd. */
int orc_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t key,
             int64_t *V, int64_t *E, int64_t **off, int32_t **adj);
void orc_free(void *p);

void orc_seeds(uint64_t key, uint32_t request, int64_t V, int64_t n, int32_t *out);
void orc_weights(uint64_t key, int64_t E, float *w);
void orc_fp32_cdf(int64_t V, const int64_t *off, const float *w, float *cdf);

/* Upper bound on the number of edges orc_sample can emit. */
int64_t orc_sample_capacity(const orc_request *r);

/* Multi-hop neighbour sampling.  Output is hop-major, then sample, then
 * parent position, then slot.  hop_off has depth*(samples+1) entries:
 * hop_off[h*(S+1)+s] is the absolute index of sample s's first hop-h edge,
 * hop_off[h*(S+1)+S] the end of hop h.  Returns the number of edges.
 * nthreads == 1 runs the sequential hop-synchronous restatement; nthreads > 1
 * runs the same selection per sample under OpenMP and assembles the same
 * layout (the CPU baseline). */
int64_t orc_sample(const orc_graph *g, const orc_request *r, const int32_t *seeds,
                   int32_t *src, int32_t *dst, int32_t *smp, int64_t *hop_off, int nthreads);

/* Random walk: W[w*(len+1)+t]; -1 once the walk hits a vertex of degree 0.
 * Returns the number of steps that moved (non -1 entries after column 0). */
int64_t orc_walk(const orc_graph *g, int32_t bias, uint64_t key, uint32_t request,
                 const int32_t *seeds, int64_t nwalkers, int32_t len, int32_t *W, int nthreads);

/* Multi-dimensional random walk (FrontierSize > 1).  seeds has samples*fs
 * entries.  Output per sample, compacted: off has samples+1 entries.
 * Returns the number of edges. */
int64_t orc_mdrw(const orc_graph *g, uint64_t key, uint32_t request, int32_t samples, int32_t fs,
                 int32_t depth, const int32_t *seeds, int32_t *src, int32_t *dst, int64_t *off);

/* Expose the single-source selection for unit tests: returns k and writes
 * the chosen neighbour indices (positions within v's list) to idx. */
int32_t orc_select_source(const orc_graph *g, int32_t bias, int32_t dedup, int32_t fanout,
                          int32_t v, uint64_t key, uint32_t request, uint32_t hop,
                          uint32_t sample, uint32_t pos, int32_t *idx);

#ifdef __cplusplus
}
#endif
#endif
