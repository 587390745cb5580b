/*
 * csaw.h -- C ABI of the MI355X-native C-SAW sampling / random-walk engine.
 *
 * Plain C: pointers, sizes, status codes; no HIP or torch types.  Every
 * function returns a csaw_status (0 = OK) unless stated otherwise and never
 * calls exit().  Host pointers are read synchronously.  Device work runs on
 * the library's own stream unless a stream is passed (void* = hipStream_t).
 *
 * Reference surfaces replaced (paths relative to the C-SAW reference root):
 *  - struct arguments Sampler(beg, csr, n_blocks, n_threads, n_subgraph,
 *    FrontierSize, NeighborSize, Depth, args, rank)   non-stream/sampler.cuh:1-8,
 *    non-stream/main.cu:196-308, streaming/streaming_sampling.cu:427-728
 *  - class Sampling / Si (samples, FrontierSize, NeighborSize, Depth, output
 *    edge list Si{vertex,edge,start})                 non-stream/sample_class.cuh:76-90,146-185
 *  - graph<...>(beg_file, csr_file, weight_file)      non-stream/graph.hpp:4-111
 *  - gpu_graph(ginst)                                 non-stream/gpu_graph.cuh:10-65
 *  - VertexBias / EdgeBias / Update hooks             non-stream/api.cuh:4-27
 *    (device side: c-saw_amd/csrc/csaw_api.h; selected here by csaw_bias)
 */
#ifndef CSAW_H
#define CSAW_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CSAW_MAX_HOPS 8
#define CSAW_MAX_FANOUT 64

/* Status codes (replace H_ERR/HRR print-and-exit, util.h:5-16, herror.h:2-10). */
typedef enum csaw_status {
  CSAW_OK = 0,
  CSAW_ERR_INVALID_ARG = 1,
  CSAW_ERR_INVALID_GRAPH = 2,
  CSAW_ERR_OOM = 3,
  CSAW_ERR_BUDGET_TOO_SMALL = 4,
  CSAW_ERR_HIP = 5,
  CSAW_ERR_CAPACITY = 6,
  CSAW_ERR_NO_DEVICE = 7,
  CSAW_ERR_UNSUPPORTED = 8,
  CSAW_ERR_IO = 9
} csaw_status;

/* Bias policy: which EdgeBias hook the ITS select uses (api.cuh:13-20). */
typedef enum csaw_bias {
  CSAW_BIAS_UNIFORM = 0, /* EdgeBias = 1                                   */
  CSAW_BIAS_DEGREE = 1,  /* EdgeBias = degree of the neighbour (integer)   */
  CSAW_BIAS_WEIGHT = 2,  /* FP32 edge weight, sequential CDF in CSR order  */
  CSAW_BIAS_USER = 3     /* the compiled-in user hooks of csaw_api.h       */
} csaw_bias;

/* Kernel variants (granularity x taken-set representation).  Every variant
 * returns bit-identical edge lists; they differ only in speed. */
typedef enum csaw_variant {
  CSAW_VAR_THREAD = 0,      /* one lane per frontier entry                  */
  CSAW_VAR_WAVE_BITMAP = 1, /* one wave64 per entry, LDS bitmap taken-set    */
  CSAW_VAR_WAVE_HASH = 2,   /* one wave64 per entry, register/LDS list set   */
  CSAW_VAR_BLOCK_BITMAP = 3,/* one 256-lane workgroup per entry, bitmap      */
  CSAW_VAR_BLOCK_HASH = 4,  /* one 256-lane workgroup per entry, list set    */
  CSAW_NUM_VARIANTS = 5
} csaw_variant;

typedef enum csaw_sched {
  CSAW_SCHED_HEURISTIC = 0, /* C-SAW degree thresholds (util.h:19-20)         */
  CSAW_SCHED_MLP = 1,       /* learned cost model, heuristic as fallback      */
  CSAW_SCHED_FORCED = 2,    /* one variant for every chunk it is valid for    */
  CSAW_SCHED_RANDOM = 3     /* random valid variant per chunk (training data) */
} csaw_sched;

typedef enum csaw_mode {
  CSAW_MODE_PLAIN = 0,  /* one request at a time, host-synchronised hops     */
  CSAW_MODE_BATCHED = 1 /* all requests per launch, device-side sizes        */
} csaw_mode;

/* Human-readable text of the last error on this thread. */
const char *csaw_last_error(void);

/* ---------------------------------------------------------------- graph */
typedef struct csaw_graph csaw_graph;

/* Uploads a host CSR.  offsets: V+1 entries, int32 or int64 per
 * offsets_are_i64; indices: E int32 neighbour ids; weights: E floats or NULL.
 * Validates monotone offsets, offsets[0]==0, offsets[V]==E, ids in [0,V). */
int csaw_graph_create(const void *offsets, int offsets_are_i64, const int32_t *indices,
                      const float *weights_or_null, int64_t V, int64_t E, csaw_graph **out);
/* Generates an RMAT graph on the GPU (ef*2^scale Philox-drawn pairs,
 * symmetrised, self loops and duplicates removed, sorted lists) and, when
 * with_weights, FP32 weights uniform in (0,1] per CSR entry. */
int csaw_graph_create_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t key,
                           int with_weights, csaw_graph **out);
/* Reads the reference's on-disk CSR (beg.bin: V+1 int64 offsets, csr.bin:
 * E int64 ids; graph.hpp:18-74) and uploads it; weight_file may be NULL. */
int csaw_graph_load_bin(const char *beg_file, const char *csr_file, const char *weight_file,
                        csaw_graph **out);
/* Writes a host CSR in that format (int64 offsets and ids). */
int csaw_csr_write_bin(const char *beg_file, const char *csr_file, const int64_t *offsets,
                       const int32_t *indices, int64_t V, int64_t E);
/* Converts a SNAP text edge list ('#' / '%' comment lines, "src dst" per
 * line, extra columns ignored) to beg.bin / csr.bin: ids kept, V = largest
 * id + 1, optional symmetrisation, self loops and repeated edges dropped,
 * neighbour lists sorted (the reference relied on an external converter,
 * README.md:14-22). */
int csaw_snap_to_bin(const char *snap_file, int symmetrize, const char *beg_file, const char *csr_file,
                     int64_t *V, int64_t *E);
int csaw_graph_info(const csaw_graph *g, int64_t *V, int64_t *E, int64_t *max_degree);
/* Copies the CSR back to host buffers (any may be NULL). */
int csaw_graph_copy_csr(const csaw_graph *g, int64_t *offsets, int32_t *indices, float *weights);
int csaw_graph_destroy(csaw_graph *g);

/* Seeds: seed_i = mulhi64(philox(key; i, 0, request, DOM_SEEDS), V), written
 * to host (the reference drew mt19937 seeds in [1,10000], main.cu:257-262). */
int csaw_seeds(const csaw_graph *g, uint64_t key, uint32_t request, int64_t n, int32_t *out);
/* The same seeds computed on the host for a vertex count V. */
int csaw_seeds_v(int64_t V, uint64_t key, uint32_t request, int64_t n, int32_t *out);

/* ------------------------------------------------------------- sampling */
/* The Sampling-class parameter surface (sample_class.cuh:146-185 and the
 * Sampler() argument list, sampler.cuh:3-8). */
typedef struct csaw_params {
  int32_t samples;       /* n_subgraph                                     */
  int32_t frontier_size; /* FrontierSize (1: neighbour sampling / walk;    */
                         /*  >1: multi-dimensional random walk)            */
  int32_t neighbor_size; /* NeighborSize (fanout when fanouts == NULL)     */
  int32_t depth;         /* Depth                                          */
  const int32_t *fanouts;/* optional per-hop fanouts, `depth` entries      */
  int32_t bias;          /* csaw_bias                                      */
  int32_t dedup;         /* 1: no repeated neighbour per source; 0: with   */
                         /*  replacement; 2: no repeats, collisions        */
                         /*  resolved by heur_normalize's region remap     */
                         /*  (the reference's default, functions.cuh:743-  */
                         /*  868) instead of redraws                       */
  uint64_t seed;         /* Philox key                                     */
  uint32_t request_id;   /* counter word; keeps batched == plain           */
  int32_t sched;         /* csaw_sched                                     */
  int32_t variant;       /* csaw_variant for CSAW_SCHED_FORCED             */
  int32_t n_blocks;      /* launch overrides (0 = automatic)               */
  int32_t n_threads;
} csaw_params;

typedef struct csaw_stats {
  int64_t edges;          /* sampled edges (or moved walk steps)           */
  double kernel_ms;       /* hipEvent time of the device work              */
  double algo_bytes;      /* algorithmic bytes of the expansion trace      */
  int64_t expansions;     /* frontier entries expanded                     */
  int64_t variant_chunks[CSAW_NUM_VARIANTS]; /* chunks run per variant     */
} csaw_stats;

/* Output edge list Si, hop-major: hop h, then sample, then parent position,
 * then slot.  hop_offsets has depth*(samples+1) entries (absolute indices). */
typedef struct csaw_edges csaw_edges;

/* Sampler(): seeds are host int32 of samples*frontier_size entries. */
int csaw_sample(csaw_graph *g, const csaw_params *p, const int32_t *seeds, csaw_edges **out,
                csaw_stats *stats);
int csaw_edges_info(const csaw_edges *e, int64_t *n_edges, int32_t *depth, int32_t *samples);
int csaw_edges_copy(const csaw_edges *e, int32_t *src, int32_t *dst, int32_t *sample,
                    int64_t *hop_offsets);
void csaw_edges_free(csaw_edges *e);

/* Random walk: walks[w*(len+1)+t], -1 after a degree-0 vertex. */
int csaw_random_walk(csaw_graph *g, int32_t bias, uint64_t seed, uint32_t request_id,
                     const int32_t *seeds, int64_t n_walkers, int32_t len, int32_t *walks,
                     csaw_stats *stats);

/* Launch shapes (workgroups per CU per kernel family), chosen by
 * csaw_autotune.  Replaces the fixed <<<n_blocks, n_threads>>> of the
 * reference's run.sh:2 (400 x 128). */
typedef struct csaw_tuning {
  int32_t count_per_cu;
  int32_t expand_per_cu;
  int32_t thread_per_cu;
  int32_t block_per_cu;
  int32_t grid_override; /* fixed grid for the plain expand kernel (0: auto) */
  /* Tile shapes (0: default).  chunk_min: smallest chunk of frontier
   * entries a queue hands out (a power of two, 8 .. 1024; chunks grow to
   * whole 1024-entry buckets when the hop is large).  group_min: smallest
   * sub-wave group of lanes that expands one entry (4 .. 32; a group is at
   * least the entry's fanout).  The edges do not depend on either. */
  int32_t chunk_min;
  int32_t group_min;
} csaw_tuning;

/* Learned cost model (16 -> 64 -> 64 -> CSAW_NUM_VARIANTS, BF16 weights,
 * FP32 accumulate); trained and evaluated by the library's HIP kernels. */
typedef struct csaw_mlp csaw_mlp;
int csaw_mlp_num_params(void);
int csaw_mlp_create(uint64_t init_key, csaw_mlp **out);
int csaw_mlp_destroy(csaw_mlp *m);
/* FP32 master parameters (csaw_mlp_num_params floats). */
int csaw_mlp_get_params(const csaw_mlp *m, float *params);
int csaw_mlp_set_params(csaw_mlp *m, const float *params);
/* Forward on the GPU: X[n][16] -> Y[n][CSAW_NUM_VARIANTS]. */
int csaw_mlp_forward(csaw_mlp *m, const float *X, int64_t n, float *Y);
/* Masked mean-squared-error loss and its gradient with respect to the
 * (BF16-rounded) parameters, on the GPU.  T, M: [n][CSAW_NUM_VARIANTS]. */
int csaw_mlp_grad(csaw_mlp *m, const float *X, const float *T, const float *M, int64_t n, float *loss,
                  float *grad);
/* Minibatch SGD on the GPU (forward, backward and update kernels). */
int csaw_mlp_train(csaw_mlp *m, const float *X, const float *T, const float *M, int64_t n, int32_t epochs,
                   int32_t batch, float lr, uint64_t shuffle_key, float *final_loss);

/* Multi-dimensional random walk (FrontierSize > 1; check_layer, main.cu:95-194):
 * seeds has samples*frontier_size entries; src/dst hold samples*depth
 * edges; offsets[samples+1] delimit each sample's edges. */
int csaw_mdrw(csaw_graph *g, uint64_t seed, uint32_t request_id, int32_t samples, int32_t frontier_size,
              int32_t depth, const int32_t *seeds, int32_t *src, int32_t *dst, int64_t *offsets,
              csaw_stats *stats);

/* csaw_mdrw over caller-owned DEVICE buffers that allocates no memory: no
 * counterpart in the reference.  For callers that cannot allocate (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * d_seeds: samples*frontier_size ids; d_src, d_dst: samples*depth entries;
 * d_offsets: samples+1 (the edges are compacted as csaw_mdrw returns them).
 * Work is enqueued on `stream` (NULL: the graph's stream).  Scratch is a
 * static 32 MiB device array per GPU, which limits samples*(8*depth+4) to
 * about 32 MiB (CSAW_ERR_CAPACITY beyond).  Calls that overlap on one GPU
 * race on that scratch.  sync != 0 waits and reports a seed outside [0, V)
 * (CSAW_ERR_INVALID_ARG); such a seed's sample emits nothing in any case. */
int csaw_mdrw_noalloc(csaw_graph *g, uint64_t seed, uint32_t request_id, int32_t samples,
                      int32_t frontier_size, int32_t depth, const int32_t *d_seeds, int32_t *d_src,
                      int32_t *d_dst, int64_t *d_offsets, void *stream, int sync);

/* Reusable walk plan: buffers allocated once, seeds resident on the device. */
typedef struct csaw_walk_plan csaw_walk_plan;
int csaw_walk_plan_create(csaw_graph *g, int32_t bias, uint64_t seed, uint32_t request_id,
                          const int32_t *seeds, int64_t n_walkers, int32_t len, csaw_walk_plan **out);
int csaw_walk_plan_run(csaw_walk_plan *p, void *stream, int sync);
int csaw_walk_plan_result(const csaw_walk_plan *p, int32_t *walks, int64_t *steps, double *kernel_ms,
                          double *algo_bytes);
int csaw_walk_plan_destroy(csaw_walk_plan *p);

/* ------------------------------------------------ plans (batched, reusable) */
typedef struct csaw_request {
  int32_t samples;
  int32_t depth;
  int32_t fanouts[CSAW_MAX_HOPS];
  int32_t bias;
  int32_t dedup;         /* as csaw_params.dedup (0, 1 or 2)            */
  uint64_t seed;
  uint32_t request_id;
  const int32_t *seeds; /* host, `samples` entries */
} csaw_request;

typedef struct csaw_plan_opts {
  int32_t mode;      /* csaw_mode                                          */
  int32_t sched;     /* csaw_sched                                         */
  int32_t variant;   /* forced variant                                     */
  int32_t use_graph; /* capture the hop sequence in a hipGraph (batched)   */
  int32_t timing;    /* record per-kernel hipEvent times                   */
  int32_t log_chunks;/* record per-chunk timings (cost-model data)         */
  uint64_t explore_key; /* key of CSAW_SCHED_RANDOM                        */
  const struct csaw_mlp *mlp; /* cost model for CSAW_SCHED_MLP (or NULL)   */
  const struct csaw_tuning *tuning; /* launch shapes (or NULL: defaults)   */
  int32_t cdf_cache; /* integer-bias draws use the per-graph CDF cache (the */
                     /* reference's Cp, functions.cuh:329-341; off there)   */
  int32_t serial_expand; /* batched: run a hop's variant kernels one after */
                         /* another on one stream (default: side by side  */
                         /* on four streams)                               */
} csaw_plan_opts;

typedef struct csaw_plan csaw_plan;
int csaw_plan_create(csaw_graph *g, const csaw_request *reqs, int32_t n_req,
                     const csaw_plan_opts *opts, csaw_plan **out);
/* Runs every request.  stream: hipStream_t or NULL for the plan's stream.
 * Blocks until the device work is done when sync != 0. */
int csaw_plan_run(csaw_plan *p, void *stream, int sync);
int csaw_plan_stats(const csaw_plan *p, csaw_stats *stats);
int csaw_plan_request_info(const csaw_plan *p, int32_t req, int64_t *n_edges);
int csaw_plan_copy(const csaw_plan *p, int32_t req, int32_t *src, int32_t *dst, int32_t *sample,
                   int64_t *hop_offsets);
/* csaw_plan_copy into device buffers (src/dst/sample: csaw_plan_request_info
 * entries, hop_offsets: depth*(samples+1)), enqueued on stream (hipStream_t,
 * NULL: the plan's stream); blocks until done when sync != 0.  For callers
 * that keep the edge lists on the GPU (framework tensors); no counterpart in
 * the reference, whose Si stayed in device memory (README.md:36). */
int csaw_plan_copy_device(const csaw_plan *p, int32_t req, int32_t *src, int32_t *dst, int32_t *sample,
                          int64_t *hop_offsets, void *stream, int sync);
/* Order-sensitive 64-bit digest of request req's edge list (src, dst,
 * sample, position), computed on the device: equal edge lists give equal
 * digests whichever plan mode or variant produced them. */
int csaw_plan_digest(const csaw_plan *p, int32_t req, uint64_t *digest);
/* Average device time per launch of each kernel kind over the runs since
 * the last reset (timing != 0), and the algorithmic bytes per launch of
 * that kind (exact per-expansion byte model).  names: NUL-separated list. */
int csaw_plan_kernel_times(const csaw_plan *p, char *names, size_t names_len, double *avg_ms,
                           int64_t *launches, double *avg_bytes, int32_t max_kinds, int32_t *n_kinds);
int csaw_plan_reset_times(csaw_plan *p);
int csaw_plan_destroy(csaw_plan *p);
/* Per-chunk cost-model data of the last run (log_chunks != 0): features
 * [n][16], the variant that ran, log2 of its device ticks, a chunk key that
 * is stable across runs of the same plan (hop << 32 | block*128 + bucket). */
int csaw_plan_chunk_log(const csaw_plan *p, float *features, int32_t *variant, float *log2_ticks,
                        uint64_t *keys, int64_t max_n, int64_t *n);
/* Tries launch and tile shapes (coordinate descent over workgroups per CU
 * of each kernel family, chunk_min and group_min) on the given requests and
 * returns the fastest; best_ms is its mean device time per run. */
int csaw_autotune(csaw_graph *g, const csaw_request *reqs, int32_t n_req, const csaw_plan_opts *base,
                  csaw_tuning *best, double *best_ms);

/* --------------------------------------------- streaming (out-of-core graphs) */
/* A CSR kept in pinned host memory (the reference's streaming sampler,
 * streaming/streaming_sampling.cu:427-728, which copied the whole graph to
 * the device first; here only partition slices travel). */
typedef struct csaw_host_graph csaw_host_graph;
/* RMAT generated on the GPU, then moved to pinned host memory. */
int csaw_host_graph_create_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t key,
                                csaw_host_graph **out);
int csaw_host_graph_from_csr(const int64_t *offsets, const int32_t *indices, int64_t V, int64_t E,
                             csaw_host_graph **out);
/* offsets / indices: the pinned host arrays (valid until destroy). */
int csaw_host_graph_info(const csaw_host_graph *g, int64_t *V, int64_t *E, int64_t *max_degree,
                         const int64_t **offsets, const int32_t **indices);
int csaw_host_graph_destroy(csaw_host_graph *g);
/* Builds a packed copy of the neighbour ids in pinned memory (blocks of 64
 * entries: first id + zigzag differences in 1-4 bytes each), encoded on
 * `threads` host threads (0: 16); streamed runs then copy packed slices and
 * unpack them on the device.  packed_bytes: bytes of the packed copy. */
int csaw_host_graph_pack(csaw_host_graph *g, int32_t threads, int64_t *packed_bytes);

typedef struct csaw_stream_stats {
  int64_t edges;
  double device_ms;          /* device time of the whole streamed run       */
  double h2d_bytes;          /* adjacency bytes copied host -> device       */
  int32_t partitions;
  int64_t peak_device_bytes; /* budget charged at the peak                  */
  int64_t slice_bytes;       /* largest partition slice                     */
  double expand_ms;          /* device time of the per-partition work:      */
                             /*  binning, variant choice, hub CDFs and      */
                             /*  expansion (copy waits excluded)            */
  int64_t chunks;            /* chunks scheduled (batched mode)             */
  int64_t variant_chunks[CSAW_NUM_VARIANTS]; /* chunks run per variant      */
  int32_t direct_partitions; /* (hop, partition) pairs read in place        */
  int32_t reused_slices;     /* slices a hop found left on the device       */
  double copy_ms;            /* time the copy stream spent in slice copies  */
  double h2d_raw_bytes;      /* bytes the copied slices hold unpacked       */
  double hop_ms[CSAW_MAX_HOPS]; /* device time of each hop                  */
} csaw_stream_stats;

/* Where a partition's adjacency comes from in a hop. */
typedef enum csaw_slice_policy {
  CSAW_SLICE_AUTO = 0,   /* copy the slice, or read the pinned host lists in */
                         /*  place when the partition's frontier needs less */
                         /*  than half the slice's bytes                    */
  CSAW_SLICE_COPY = 1,   /* always copy slices (double-buffered)            */
  CSAW_SLICE_DIRECT = 2  /* always read in place                            */
} csaw_slice_policy;

/* How each partition's frontier entries are expanded. */
typedef struct csaw_stream_opts {
  int32_t mode;         /* CSAW_MODE_BATCHED: entries binned into chunks by  */
                        /*  degree class and each chunk run by the variant  */
                        /*  `sched` picks; CSAW_MODE_PLAIN: one wave per     */
                        /*  listed entry                                     */
  int32_t sched;        /* csaw_sched                                        */
  int32_t variant;      /* forced variant                                    */
  uint64_t explore_key; /* key of CSAW_SCHED_RANDOM                          */
  const struct csaw_mlp *mlp;       /* cost model for CSAW_SCHED_MLP or NULL */
  const struct csaw_tuning *tuning; /* launch shapes or NULL (defaults)      */
  int32_t slice_policy; /* csaw_slice_policy                                 */
  int32_t raw_slices;   /* copy unpacked slices even if the graph is packed  */
  float direct_factor;  /* CSAW_SLICE_AUTO reads a partition in place when   */
                        /*  factor x its estimated bus bytes < its slice's  */
                        /*  bytes (0: 2)                                     */
  int32_t bin_min_entries; /* batched: partitions with fewer frontier       */
                        /*  entries run one wave per entry, unbinned (0:     */
                        /*  16384; 1: bin every partition)                   */
  /* Per-chunk cost-model data (batched mode; each partition's chunks are
   * read back after it expands, so logging runs slower).  Rows as in
   * csaw_plan_chunk_log; keys are hop << 40 | partition << 32 | chunk key. */
  int32_t log_chunks;
  int64_t log_cap;            /* rows the buffers below hold                */
  float *log_features;        /* [log_cap][16]                              */
  int32_t *log_variant;
  float *log_log2_ticks;
  uint64_t *log_keys;
  int64_t *log_n;             /* rows produced (may exceed log_cap)         */
} csaw_stream_opts;

/* Samples one request (uniform, degree or user bias) streaming vertex-range
 * partitions through two device buffers under budget_bytes of device
 * memory; CSAW_ERR_BUDGET_TOO_SMALL when the resident state does not fit.
 * The edge list equals the in-memory csaw_sample result whatever the mode,
 * schedule or variant.  opts NULL: batched, heuristic schedule. */
int csaw_stream_sample_opts(csaw_host_graph *g, const csaw_request *req, int64_t budget_bytes,
                            int32_t n_partitions, const csaw_stream_opts *opts, csaw_edges **out,
                            csaw_stream_stats *stats);
/* csaw_stream_sample_opts with opts NULL. */
int csaw_stream_sample(csaw_host_graph *g, const csaw_request *req, int64_t budget_bytes,
                       int32_t n_partitions, csaw_edges **out, csaw_stream_stats *stats);
/* Wraps host arrays as a csaw_edges result (copies them). */
int csaw_edges_from_host(int32_t depth, int32_t samples, int64_t n, const int32_t *src, const int32_t *dst,
                         const int32_t *sample, const int64_t *hop_offsets, csaw_edges **out);

/* C4 request mix: n requests, seed counts log-uniform in [lo, hi], fanouts
 * {25,10} or {15,10,5}, bias uniform / degree / weight with p = 1/3 each,
 * all drawn from Philox(mix_key).  Writes the request headers (seeds are
 * generated with csaw_seeds by the caller); returns total seeds. */
int64_t csaw_request_mix(uint64_t mix_key, int32_t n, int32_t lo, int32_t hi, csaw_request *out);

#ifdef __cplusplus
}
#endif
#endif
