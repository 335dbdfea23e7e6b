/*
 * gbx.h -- C ABI of libgbx_hip.so, the MI355X (gfx950) backend for the
 * GraphBlox vertex-centric pull PageRank / SpMV path.
 *
 * Plain C: plain pointers and sizes, no HIP or torch types.  Every entry
 * point returns an int status (GBX_OK = 0, negative on error) and leaves a
 * message retrievable with gbx_last_error().  Host pointers are borrowed for
 * the duration of the call.  Device memory is owned by the objects below.
 * One context per host thread; objects are not thread safe.
 *
 * Reference interfaces each group replaces (paths relative to the GraphBlox
 * repository root):
 *   gbx_edges_*                       readEdgeListsbin (01_host/src/structures/edgeList.c:450-706)
 *   gbx_csr_build, gbx_graph_create*  sortRunAlgorithms (01_host/src/preprocess/sortRun.c:26-69),
 *                                     countSortEdgesBySource/-AndDestination (countsort.c:28-270),
 *                                     mapVerticesWithInOutDegree_base (01_host/src/structures/vertex.c:204-328),
 *                                     removeDulpicatesSelfLoopEdges (edgeList.c:255-321),
 *                                     graphCSRPreProcessingStep's forward+transpose pipeline
 *                                     (01_host/src/structures/graphCSR.c:243-327)
 *   gbx_pr_*                          pageRankPullGraphCSR (01_host/src/algorithms/openmp/pageRank.c:245-368)
 *                                     and its stats (pageRank.h:19-30)
 *   gbx_spmv_pull, gbx_spmv_fixed     SPMVPullGraphCSR (01_host/src/algorithms/openmp/SPMV.c:117-215)
 *                                     and its variants (SPMV.c:216-590)
 *   gbx_gather_accumulate_u32         the FPGA overlay's Single.PR integer pull
 *                                     (02_device/ol/GraphBlox/Single/topology.json:162-201,
 *                                     01_host/src/utils_fpga_cpp/buffer_mapping.PR.cpp:2-68)
 *   gbx_ppr_batched, gbx_graph_batch_*, gbx_pr_adjoint, gbx_rmat_*, gbx_spmv_pull_noalloc
 *                                     no reference counterpart (batched, differentiable,
 *                                     generator and allocation-free forms of the same path)
 */
#ifndef GBX_H
#define GBX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define GBX_OK 0
#define GBX_ERR_ARG (-1)     /* invalid argument (NULL, size, mode) */
#define GBX_ERR_NOMEM (-2)   /* device or host allocation failed */
#define GBX_ERR_HIP (-3)     /* HIP runtime error */
#define GBX_ERR_RANGE (-4)   /* vertex id >= 2^30 or beyond num_vertices */
#define GBX_ERR_STATE (-5)   /* object used in a state it does not support */
#define GBX_ERR_IO (-6)      /* a file could not be opened or read */

typedef struct gbx_ctx gbx_ctx;
typedef struct gbx_graph gbx_graph;
typedef struct gbx_pr gbx_pr;

/* Create a context on HIP device `device` (its own non-blocking stream). */
int gbx_init(int device, gbx_ctx **out);
int gbx_free(gbx_ctx *ctx);
/* Last error message for this context (never NULL). */
const char *gbx_last_error(const gbx_ctx *ctx);
const char *gbx_status_string(int status);
/* Wait for all work on the context's stream. */
int gbx_sync(gbx_ctx *ctx);
/* The context's hipStream_t, as an opaque pointer. */
void *gbx_stream(gbx_ctx *ctx);
/* Library build id string (architecture and version). */
const char *gbx_version(void);

/* ---------------------------------------------------------------- CSR build */
/* Sort modes of sortRunAlgorithms (sortRun.c:26-69): 0 radix by source,
 * 1 radix by source+destination, 2 count by source, 3 count by
 * source+destination.  0==2 and 1==3 produce identical arrays. */
#define GBX_SORT_RADIX_SRC 0
#define GBX_SORT_RADIX_SRC_DST 1
#define GBX_SORT_COUNT_SRC 2
#define GBX_SORT_COUNT_SRC_DST 3

/* One CSR from a host edge list, on the GPU, bit-exact with the reference:
 * stable sort (mode as above), then out_degree[v] / edges_idx[v]
 * (edges_idx[v] = 0 when out_degree[v] = 0, vertex.c:39-45).
 * w may be NULL; with move_weights the weights follow the sort
 * (`#if WEIGHTED`, countsort.c:118-120), else sorted_w receives w unchanged.
 * Outputs: sorted_src/sorted_dst/sorted_w [num_edges] (nullable),
 * out_degree/edges_idx [num_vertices] (nullable).
 * Ids must be < num_vertices <= 2^30. */
int gbx_csr_build(gbx_ctx *ctx, const uint32_t *src, const uint32_t *dst,
                  const float *w, uint32_t num_edges, uint32_t num_vertices,
                  int sort_mode, int move_weights, uint32_t *sorted_src,
                  uint32_t *sorted_dst, float *sorted_w, uint32_t *out_degree,
                  uint32_t *edges_idx);

/* radixSortEdgesByPageRank (reorder.c:165-221): labels[0..n) reordered by a
 * stable ascending sort on the fixed-point key FloatToFixed32SORT(rank of the
 * label's position) = (u32)(ranks[i] * 2^20 + 0.5f) (fixedPoint.h:38), on the
 * GPU.  ranks[i] is the key of the label found at labels[i] on entry (the
 * reference passes the identity realRanks).  Host arrays, in place. */
int gbx_rank_sort(gbx_ctx *ctx, const float *ranks, uint32_t num_vertices, uint32_t *labels);

/* ---------------------------------------------------------------- graphs */
#define GBX_GRAPH_DEDUPE 1u     /* -k: (src,dst) sort + removeDulpicatesSelfLoopEdges */
#define GBX_GRAPH_WEIGHTED 2u   /* WEIGHTED=1: weights move through the sorts */
#define GBX_GRAPH_SYMMETRIC 4u  /* -s: append the reversed edges (edgeList.c:560-578) */

/* Device-resident graph: forward CSR and transpose (in-)CSR built on the GPU
 * from a host edge list, in graphCSRPreProcessingStep's order
 * (graphCSR.c:243-327, DIRECTED=1).  num_vertices = 0 selects the
 * reference's rule N = max id + 1 (edgeList.c:671-679). */
int gbx_graph_create(gbx_ctx *ctx, const uint32_t *src, const uint32_t *dst,
                     const float *w, uint32_t num_edges, uint32_t num_vertices,
                     int sort_mode, unsigned flags, gbx_graph **out);

/* The relabel levels of graphCSRPreProcessingStep (arguments lmode, lmode_l2,
 * lmode_l3; graphCSR.c:268-290): after the sort (and -k dedupe) each nonzero
 * level renames every vertex as reorderGraphProcess (reorder.c:1556-1616)
 * orders them, on the GPU, and sorts the edges again: 1 out-degree, 2
 * in-degree, 3 in+out degree, 10 random degree (descending degree); 4/5 DBG
 * and 12/13 Corder (out/in degree buckets); 6/7 HUBSort; 8/9 HUBCluster; 11
 * the labels in file_labels (new id of each vertex, num_file_labels >= N,
 * each below N); other codes order by out-degree.  labels_out [N]
 * (nullable) receives the composed new id of every original vertex
 * (EdgeList.label_array).  With all levels 0 this is gbx_graph_create. */
typedef struct gbx_relabel {
    uint32_t lmode[3];
    const uint32_t *file_labels;
    uint32_t num_file_labels;
} gbx_relabel;

int gbx_graph_create_relabel(gbx_ctx *ctx, const uint32_t *src, const uint32_t *dst,
                             const float *w, uint32_t num_edges, uint32_t num_vertices,
                             int sort_mode, unsigned flags, const gbx_relabel *relabel,
                             uint32_t *labels_out, gbx_graph **out);

/* Seeded R-MAT graph generated on the GPU: (edge_factor << scale) edges,
 * quadrant probabilities a, b, c, d = 1-a-b-c, ids relabelled by a seeded
 * bijection of [0, 2^scale), self loops and duplicates removed (first
 * occurrence kept, generation order preserved), N = max id + 1.
 * weight_kind: 0 none (1.0f), 1 uniform(0,1), 2 lognormal(0,1); weights are
 * keyed by (weight_seed, kept-edge index). */
int gbx_graph_create_rmat(gbx_ctx *ctx, uint32_t scale, uint32_t edge_factor,
                          double a, double b, double c, uint64_t seed,
                          int weight_kind, uint64_t weight_seed, int sort_mode,
                          unsigned flags, gbx_graph **out);

/* The generated edge list itself (host copies), for checking the generator:
 * src/dst/w sized for edge_factor << scale entries; *num_edges receives the
 * kept count (after self-loop/duplicate removal when dedupe != 0). */
int gbx_rmat_edges(gbx_ctx *ctx, uint32_t scale, uint32_t edge_factor, double a,
                   double b, double c, uint64_t seed, int dedupe, int weight_kind,
                   uint64_t weight_seed, uint32_t *src, uint32_t *dst, float *w,
                   uint64_t *num_edges);

/* Device-resident graph from CSR arrays already built on the host in the
 * reference's layout (e.g. by graphCSRPreProcessingStep): forward and
 * transpose out_degree/edges_idx [num_vertices], src/dst [num_edges],
 * w [num_edges] (nullable).  Arrays are copied; nothing is re-sorted.  All
 * five transpose arguments NULL (the DIRECTED=0 GraphCSR, graphCSR.h:26-30):
 * the device builds the transpose from the forward arrays (swapped edges,
 * stable sort by the new source).  Checked before anything else uses them,
 * on every side given (GBX_ERR_ARG / GBX_ERR_RANGE otherwise): src
 * non-decreasing, every id below num_vertices, every edge e of source v
 * inside [edges_idx[v], edges_idx[v] + out_degree[v]) and the out-degrees
 * summing to num_edges. */
int gbx_graph_from_csr(gbx_ctx *ctx, uint32_t num_vertices, uint32_t num_edges,
                       const uint32_t *fwd_out_degree, const uint32_t *fwd_edges_idx,
                       const uint32_t *fwd_src, const uint32_t *fwd_dst, const float *fwd_w,
                       const uint32_t *inv_out_degree, const uint32_t *inv_edges_idx,
                       const uint32_t *inv_src, const uint32_t *inv_dst, const float *inv_w,
                       gbx_graph **out);

int gbx_graph_info(const gbx_graph *g, uint32_t *num_vertices, uint32_t *num_edges);

/* ---------------------------------------------------------------- edge-list ingest */
/* A device-resident edge list (struct EdgeList, edgeList.h:12-25): what
 * readEdgeListsbin (edgeList.c:450-706) makes of a .bin / .wbin file, built
 * on the GPU from the file's raw records.  Records are 2 u32 (src, dst), or
 * with weighted_records 2 u32 + f32 (a .wbin; the weight's bits as stored);
 * E = bytes / record size (a partial last record is ignored, as the
 * reference's division does).  inverse swaps src and dst of every record;
 * symmetric appends every (swapped) edge reversed at [E, 2E) (edgeList.c:
 * 551-578).  Weights are 1.0f without weighted_records.  N = max id + 1
 * (edgeList.c:671-679; ids must stay below 2^30, the column mask); the
 * largest weight (max_weight, starting at 0.0f, NaN never larger) and the
 * count come back from gbx_edges_info. */
typedef struct gbx_edges gbx_edges;

/* From records in host memory (for example a mapped file): copied to the
 * device through pinned staging buffers, 64 MiB at a time. */
int gbx_edges_from_records(gbx_ctx *ctx, const void *records, uint64_t num_bytes, int weighted_records,
                           int inverse, int symmetric, gbx_edges **out);
/* From a file: read in 64 MiB pieces into pinned buffers, each copied to
 * the device while the next is read. */
int gbx_edges_read_file(gbx_ctx *ctx, const char *path, int weighted_records, int inverse, int symmetric,
                        gbx_edges **out);
int gbx_edges_info(const gbx_edges *e, uint32_t *num_vertices, uint32_t *num_edges, float *max_weight);
/* Host copies of the edge arrays (any may be NULL), in file order. */
int gbx_edges_export(const gbx_edges *e, uint32_t *src, uint32_t *dst, float *w);
/* gbx_graph_create_relabel on the device edge list (flags: GBX_GRAPH_DEDUPE,
 * GBX_GRAPH_WEIGHTED; the symmetric copy is made at ingest).  The edge list
 * is consumed: its arrays become the graph's, and it holds no edges after
 * the call, whether or not the call succeeds (free it with gbx_edges_free).
 * labels_out [N from gbx_edges_info], nullable. */
int gbx_graph_create_from_edges(gbx_ctx *ctx, gbx_edges *e, int sort_mode, unsigned flags,
                                const gbx_relabel *relabel, uint32_t *labels_out, gbx_graph **out);
int gbx_edges_free(gbx_edges *e);

/* Copy the forward (inverse = 0) or transpose (inverse = 1) CSR arrays to
 * host, in the reference's layout (struct Vertex / struct EdgeList of
 * vertex.h:17-23, edgeList.h:12-25).  Any output may be NULL. */
int gbx_graph_export(const gbx_graph *g, int inverse, uint32_t *out_degree,
                     uint32_t *in_degree, uint32_t *edges_idx, uint32_t *src,
                     uint32_t *dst, float *w);

int gbx_graph_free(gbx_graph *g);

/* ---------------------------------------------------------------- PageRank */
#define GBX_PR_REF 0         /* pageRank.c semantics: r0 = 1-d, r' = (1-d) + d*sum, no dangling
                                redistribution, output r/N */
#define GBX_PR_STOCHASTIC 1  /* r0 = t, r' = (1-d) t + d (sum + D/N), D = dangling mass,
                                normalised by the previous sum in FP64 */
/* The reference's fixed-point pull PageRank, bit-exact (REF update, contributions
 * r/outdeg converted to fixed point and summed as integers; binned method only):
 * pageRankPullFixedPoint{64,32,16,8}BitGraphCSR (pageRank.c:496-1030), pushpull
 * codes 2 (also 3, the push form), 9, 10, 11. */
#define GBX_PR_FIXED64 2
#define GBX_PR_FIXED32 3
#define GBX_PR_FIXED16 4
#define GBX_PR_FIXED8 5
/* The reference's quantized pull PageRank, bit-exact (REF update; contributions
 * r/outdeg quantized each iteration with the scale (max - min) / (2^k - 1) of
 * their exact min and max, summed as integers; binned method only):
 * pageRankPullQuant{32,16,8}BitGraphCSR (pageRank.c:1159-1577), pushpull
 * codes 4, 12, 13. */
#define GBX_PR_QUANT32 6
#define GBX_PR_QUANT16 7
#define GBX_PR_QUANT8 8
/* The reference's data-driven pull PageRank (pageRankDataDrivenPullGraphCSR,
 * pageRank.c:1711-1870; pushpull code 6): worklist rows only (all in the
 * first iteration, then those with an in-neighbour updated in the previous
 * one); a row stores its new rank only when it changed by epsilon or more.
 * Unweighted FP32, pull method. */
#define GBX_PR_DATA_DRIVEN 9
/* The reference's data-driven push (pageRankDataDrivenPushGraphCSR,
 * pageRank.c:1872-2027; pushpull code 7) and pull-push (pageRankDataDriven
 * PullPushGraphCSR, pageRank.c:2029-2202; code 8) PageRank: residuals pushed
 * along out-edges over a frontier of vertices whose residual crossed epsilon.
 * The reference updates in place in vertex order (its result depends on the
 * thread schedule); here each iteration processes its frontier concurrently
 * with float atomics, so results are reproducible to the fixed point the
 * method converges to, not bit for bit.  Unweighted FP32. */
#define GBX_PR_DD_PUSH 10
#define GBX_PR_DD_PULL_PUSH 11

#define GBX_STOP_ACTIVE 0    /* stop when no vertex has |dr| >= epsilon (pageRank.c:327-339) */
#define GBX_STOP_L1 1        /* stop when sum |dr| < epsilon (normalised scale) */
#define GBX_STOP_NONE 2      /* run exactly max_iters */

#define GBX_F32 0            /* FP32 contributions and row sums */
#define GBX_BF16 1           /* BF16 contribution gathers, FP32 Kahan row sums */

#define GBX_METHOD_AUTO 0    /* binned for large graphs (tuning.binned_min_vertices), else pull */
#define GBX_METHOD_PULL 1    /* degree-binned pull: one gather of the source contribution per edge */
#define GBX_METHOD_BINNED 2  /* propagation blocking: contributions scattered into destination-range
                                bins with sequential writes, then summed per bin in LDS with exact
                                fixed-point integer adds (deterministic); REF and unweighted
                                STOCHASTIC only */

#define GBX_PULL_IN 0        /* rows = transpose CSR (DIRECTED=1, pageRank.c:266-268) */
#define GBX_PULL_FWD 1       /* rows = forward CSR (DIRECTED=0, pageRank.c:269-272) */

typedef struct gbx_pr_params {
    float damp;          /* damping factor, 0.85f in the reference (pageRank.h:13) */
    uint32_t max_iters;  /* arguments->iterations */
    double epsilon;      /* threshold for the stop rule */
    int mode;            /* GBX_PR_REF / GBX_PR_STOCHASTIC / GBX_PR_FIXED* / GBX_PR_QUANT* /
                            GBX_PR_DATA_DRIVEN / GBX_PR_DD_PUSH / GBX_PR_DD_PULL_PUSH */
    int stop;            /* GBX_STOP_* */
    int precision;       /* GBX_F32 / GBX_BF16 */
    int direction;       /* GBX_PULL_IN / GBX_PULL_FWD */
    int weighted;        /* 1: transition = edge weight / sum of the source's out-weights */
    int method;          /* GBX_METHOD_* */
} gbx_pr_params;

typedef struct gbx_pr_report {
    uint32_t iterations;   /* reference convention: index of the stopping iteration, or max_iters */
    uint32_t executed;     /* iterations actually executed */
    uint32_t active;       /* active vertices in the last executed iteration */
    uint32_t converged;    /* 1 if the stop rule fired */
    double error_total;    /* REF: sum|dr|/N; STOCHASTIC: sum|dr| of the last iteration */
    double rank_sum;       /* FP64 sum of the returned ranks */
    double time_total_s;   /* wall time of the iteration loop (device, HIP events) */
} gbx_pr_report;

/* Fill params with the reference defaults (damp 0.85, 20 iterations,
 * epsilon 1e-4, REF, ACTIVE stop, F32, PULL_IN, unweighted). */
void gbx_pr_params_default(gbx_pr_params *p);

/* One call: upload-free run on a device-resident graph to the stop rule,
 * ranks (normalised, original vertex order) copied to ranks_out[N].
 * teleport: NULL (uniform) or host array [N] summing to 1 (STOCHASTIC). */
int gbx_pr_pull(gbx_ctx *ctx, gbx_graph *g, const gbx_pr_params *p,
                const float *teleport, float *ranks_out, gbx_pr_report *rep);

/* Step-wise interface: create (builds or reuses the degree-binned plan),
 * reset (initial ranks), launch n iterations asynchronously (iterations after
 * the stop rule fired are no-ops on the device), read back. */
int gbx_pr_create(gbx_ctx *ctx, gbx_graph *g, const gbx_pr_params *p, gbx_pr **out);
int gbx_pr_reset(gbx_pr *pr, const float *teleport);
int gbx_pr_launch(gbx_pr *pr, uint32_t n_iters);
/* As gbx_pr_launch, with a HIP event pair around every iteration kernel on
 * the context stream; *kernel_ms_avg receives the mean kernel duration. */
int gbx_pr_launch_timed(gbx_pr *pr, uint32_t n_iters, double *kernel_ms_avg);
int gbx_pr_read(gbx_pr *pr, float *ranks_out, gbx_pr_report *rep);
int gbx_pr_free(gbx_pr *pr);

/* ---------------------------------------------------------------- batched */
/* Personalised PageRank for num_vectors (1..64) teleport vectors at once
 * (no reference counterpart).  Vector q is uniform over its seed set
 * seeds[seed_ptr[q] .. seed_ptr[q+1]) (original vertex ids); seed_ptr NULL
 * selects the uniform teleport for every vector.  Semantics per vector are
 * GBX_PR_STOCHASTIC; p must ask for STOCHASTIC, F32, unweighted, and stop
 * GBX_STOP_L1 (all vectors below epsilon) or GBX_STOP_NONE.  The returned
 * object is driven with gbx_pr_reset (teleport NULL), gbx_pr_launch,
 * gbx_pr_launch_timed and gbx_pr_read, where ranks_out is [N][num_vectors]
 * (vertex-major, original order) and rep points to num_vectors reports. */
int gbx_ppr_create(gbx_ctx *ctx, gbx_graph *g, const gbx_pr_params *p, uint32_t num_vectors,
                   const uint32_t *seed_ptr, const uint32_t *seeds, gbx_pr **out);
/* Shape of a batched object: vectors, lanes per row (the power of two >= k
 * divided by the vectors held per lane), hub rows and hub chunks. */
int gbx_ppr_info(const gbx_pr *pr, uint32_t *num_vectors, uint32_t *vector_width,
                 uint32_t *hub_rows, uint32_t *hub_chunks);

/* Replay n_iters iterations as one HIP graph (captured on first use for this
 * (n_iters, iteration parity, teleport given to the last gbx_pr_reset or
 * not) and cached in the object): same results as gbx_pr_launch, one launch
 * instead of one or two per iteration. */
int gbx_pr_launch_graph(gbx_pr *pr, uint32_t n_iters);

/* ---------------------------------------------------------------- adjoint */
/* Gradients of L = sum_v loss_weights[v] * p[v] through the weighted
 * stochastic PageRank fixed point p = (1-d) t + d (P^T p + (delta.p)/N),
 * P[u,v] = W[u,v] / sum_x W[u,x] (no reference counterpart).  The graph must
 * carry weights (GBX_GRAPH_WEIGHTED).  Implicit differentiation: the forward
 * pull runs to p's stop rule, then the adjoint fixed point
 *   lambda = g + d (P lambda + delta (1.lambda)/N)
 * is iterated as a pull over the forward CSR (rows = sources, row-normalised
 * weights) to the same rule; only the converged p is kept.  Outputs
 * (nullable): ranks_out = p [N]; lambda_out [N]; grad_w [E] = dL/dW for the
 * forward-CSR edges in gbx_graph_export(g, 0, ...) order,
 *   d p[u] (lambda[v] - (P lambda)[u]) / sum_x W[u,x];
 * grad_teleport [N] = (1-d) lambda (dL/dt).  teleport: NULL (uniform) or
 * host [N] summing to 1.  p: F32, stop GBX_STOP_L1 or NONE, max_iters per
 * fixed point; with GBX_STOP_L1 the forward stops at sum|dp| < epsilon
 * (sum p = 1) and the adjoint at sum|d lambda| < epsilon * sum|lambda|. */
int gbx_pr_adjoint(gbx_ctx *ctx, gbx_graph *g, const gbx_pr_params *p, const float *teleport,
                   const float *loss_weights, float *ranks_out, float *lambda_out, float *grad_w,
                   float *grad_teleport, gbx_pr_report *fwd_rep, gbx_pr_report *adj_rep);

/* ---------------------------------------------------------------- graph batches */
/* Many independent graphs as one device object, stored block-diagonally
 * (graph k's vertex v becomes vertex offsets[k] + v; no reference
 * counterpart).  The member graphs are copied; they may be freed afterwards.
 * gbx_pr_create on the batch runs every graph's own GBX_PR_STOCHASTIC
 * PageRank (teleport 1/N_k, its own dangling mass, normalisation and L1 stop
 * rule) in one launch per iteration; gbx_pr_read returns the ranks in the
 * batch's vertex order and one report per graph. */
int gbx_graph_batch_create(gbx_ctx *ctx, uint32_t num_graphs, gbx_graph *const *graphs,
                           gbx_graph **out);
/* Number of member graphs and their vertex offsets [num_graphs + 1]
 * (a single graph reports 1 and {0, N}). */
int gbx_graph_batch_info(const gbx_graph *g, uint32_t *num_graphs, uint32_t *vertex_offsets);

/* ---------------------------------------------------------------- tuning */
/* Launch shapes of the pull kernels.  Defaults come from measurements on
 * MI355X and can be overridden per context (and at gbx_init from the
 * environment: GBX_BIN_THRESHOLDS, GBX_HUB_CHUNK, GBX_LDS_BYTES,
 * GBX_BLOCKS_PER_CU, GBX_PPR_HUB, GBX_PPR_CHUNK, GBX_PPR_BLOCKS_PER_CU,
 * GBX_PPR_VPL, GBX_BINNED_MIN_VERTICES, GBX_BINNED_CHUNK_EDGES,
 * GBX_BINNED_BIN_EDGES, GBX_BINNED_BIN_ROWS, GBX_BINNED_CHUNK_ROWS,
 * GBX_BINNED_HOT_SOURCES, GBX_BINNED_EDGE_VALUES).  A change applies to plans and objects created afterwards;
 * results stay within the same tolerances for every setting. */
typedef struct gbx_tuning {
    uint32_t bin_thresholds[7];    /* degree thresholds: hub, then 64, 32, 16, 8, 4, 2 lanes per row;
                                      hub 0 = by the graph's (largest member's) edges E: 2048 below
                                      2^25, 4096 below 2^27, else 8192 */
    uint32_t hub_chunk;            /* edges per hub chunk (one block each); 0 = 4 x the hub threshold */
    uint32_t lds_bytes;            /* LDS bytes per block for the hottest contributions (0 = off) */
    uint32_t blocks_per_cu;        /* persistent blocks per CU, single-vector kernels */
    uint32_t ppr_hub;              /* batched form: rows of at least this degree are chunked */
    uint32_t ppr_chunk;            /* batched form: edges per chunk */
    uint32_t ppr_blocks_per_cu;    /* batched form and graph batches: blocks per CU */
    uint32_t ppr_vectors_per_lane; /* batched form: contributions per lane (1, 2 or 4) */
    uint32_t binned_min_vertices;  /* GBX_METHOD_AUTO picks the binned method from this many vertices */
    uint32_t binned_chunk_edges;   /* binned: out-edges per source chunk at most; 0 = E / CUs
                                      clamped to [65536, 524288] (one scatter block per chunk);
                                      with the run layout no limit (chunks cut by sources; the
                                      scatter's blocks take equal shares of all chunks' edges) */
    uint32_t binned_bin_edges;     /* binned: in-edges per destination bin (one accumulate unit each);
                                      0 = E / CUs clamped to [65536, 524288]; with the run
                                      layout 7/8 of E / CUs clamped to [65536, 1048576] */
    uint32_t binned_bin_rows;      /* binned: destination rows per bin (<= 20000: LDS int64 sums) */
    uint32_t binned_chunk_rows;    /* binned: sources per chunk (<= 32768: LDS contributions; 32736 with the run layout) */
    uint32_t binned_hot_sources;   /* binned: the first this many sources (plan order, the most
                                      gathered) are summed from LDS in the accumulate pass, their
                                      edges bypassing the scatter (0 = off; <= 65528) */
    uint32_t binned_edge_values;   /* binned, FP32/BF16 kinds: 0 = the scatter sums the edges of one
                                      source chunk into one row and stores one value per such run
                                      (default); 1 = one value per edge */
    uint32_t reserved[1];
} gbx_tuning;

int gbx_tuning_get(const gbx_ctx *ctx, gbx_tuning *t);
int gbx_tuning_set(gbx_ctx *ctx, const gbx_tuning *t);

/* Autotuner: times short runs of the iteration p selects on graph g (the
 * pull kernel or the binned iteration, num_vectors = 0) or of the batched
 * form (num_vectors > 0) over a small search of the tuning fields of that
 * iteration (coordinate descent from the current setting: bin thresholds,
 * hub chunk, LDS prefix and blocks per CU for the pull kernel; edges per
 * chunk and per bin, rows per bin and hot sources for the binned one; bin
 * thresholds, hub chunk and ppr_blocks_per_cu for a batch of graphs; lanes,
 * hub cut and blocks per CU for the batched form), confirms the pick against
 * the starting setting on longer runs (the start stays unless the pick is
 * faster), installs it in ctx and returns it in *best with its iteration
 * time. */
int gbx_autotune(gbx_ctx *ctx, gbx_graph *g, const gbx_pr_params *p, uint32_t num_vectors,
                 uint32_t iters_per_trial, gbx_tuning *best, double *best_ms);

/* Plan statistics: counts of rows per degree class and edges per class
 * (for reporting and the autotuner); arrays of length *n_classes (<= 16). */
int gbx_pr_plan_info(const gbx_pr *pr, uint32_t *n_classes, uint32_t *class_lanes,
                     uint32_t *class_rows, uint64_t *class_edges);

/* ---------------------------------------------------------------- SpMV */
/* SPMVPullGraphCSR semantics: x[v] = 1/outdeg[v] (0.001f when 0) unless
 * x_in is given; y starts at 0 and accumulates y += w[j] x[col[j]] over
 * `iterations` passes without reset; w = 0.0001f unless weighted.
 * Outputs in original vertex order; x_out/y_out [N] nullable;
 * *seconds (nullable) receives the device time of the iterations. */
int gbx_spmv_pull(gbx_ctx *ctx, gbx_graph *g, uint32_t iterations, int weighted,
                  int direction, const float *x_in, float *x_out, float *y_out,
                  double *seconds);

/* gbx_spmv_pull on caller-owned device vectors, allocating nothing (no
 * counterpart in the reference: for callers that cannot allocate device
 * memory, such as code under HIP graph capture or a serving loop).  x, y:
 * device arrays [N] in the graph's vertex order; y = the sum of
 * `iterations` passes of y += w[j] x[col[j]] from y = 0 (w = 0.0001f unless
 * weighted), the same values gbx_spmv_pull returns for that x.  Launched
 * on the context's stream and not waited for.  Needs the degree-binned
 * plan of `direction` to exist already (built by gbx_spmv_pull or
 * gbx_pr_create on this graph; GBX_ERR_STATE otherwise).  Hub chunks are
 * combined in a device array of the library (512 KiB per device, held
 * until the process ends): plans of more than 65536 hub chunks or hub rows
 * are refused (GBX_ERR_RANGE).  Calls that overlap on one GPU race on that
 * array. */
int gbx_spmv_pull_noalloc(gbx_ctx *ctx, gbx_graph *g, uint32_t iterations, int weighted,
                          int direction, const float *x, float *y);

/* The reference's fixed-point SpMV variants, bit-exact: bits 32 =
 * SPMVPullFixedPointGraphCSR (SPMV.c:313-449, rows as `direction`), bits 64
 * = SPMVPushFixedPointGraphCSR (SPMV.c:451-590; it sums the same integer
 * terms as a pull over the transpose rows).  Same input vector and
 * accumulation over `iterations` passes as gbx_spmv_pull; y_out receives
 * the reference's float conversion of the fixed-point sums. */
int gbx_spmv_fixed(gbx_ctx *ctx, gbx_graph *g, uint32_t iterations, int weighted,
                   int direction, int bits, float *x_out, float *y_out, double *seconds);

/* ---------------------------------------------------------------- integer gather */
/* The integer pull of the FPGA overlay's Single.PR program (topology.json:
 * 162-201, PR.ol; host mapping buffer_mapping.PR.cpp:2-68): for every row v
 * (the transpose CSR for GBX_PULL_IN, as the DIRECTED host maps it,
 * graphbloxenv.cpp:346-352; the forward CSR for GBX_PULL_FWD),
 *   aux2[N + v] = sum over v's edges (v, u) of aux1[N + u]   (mod 2^32),
 * ALU_ACC's unsigned 32-bit add (engine_alu_ops_kernel.sv:104-112).  aux1,
 * aux2: host u32[2N], the overlay's buffers 7 and 8
 * (test_GraphBlox_USER_MANAGED.cpp:775-804); aux2[0, N) is left unchanged.
 * With aux1[N + u] = 1 the result is every row's degree, the check the
 * reference's host runs (test_GraphBlox_USER_MANAGED.cpp:518-528). */
int gbx_gather_accumulate_u32(gbx_ctx *ctx, gbx_graph *g, int direction, const uint32_t *aux1,
                              uint32_t *aux2);

#ifdef __cplusplus
}
#endif
#endif /* GBX_H */
