/*
 * gbx_oracle.h -- CPU oracle for the pull PageRank / SpMV path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load this library, and only as the checker
 * (or the timed CPU baseline); the product library never links or calls it.
 *
 * Contents (each function cites the reference code it restates; all paths
 * are relative to the GraphBlox repository root):
 *   - stable counting sort of an edge list by source (and by destination
 *     first for the "src-dst" modes)          01_host/src/preprocess/countsort.c:28-270
 *   - vertex mapping (out_degree, edges_idx)  01_host/src/structures/vertex.c:204-328
 *   - adjacent duplicate / self-loop removal  01_host/src/structures/edgeList.c:255-321
 *   - FP32 OpenMP pull PageRank, reference semantics
 *                                             01_host/src/algorithms/openmp/pageRank.c:245-368
 *   - FP32 OpenMP pull SpMV                   01_host/src/algorithms/openmp/SPMV.c:117-215
 *   - FP64 PageRank oracles (reference-semantics and stochastic/personalised)
 *   - FP64 adjoint of the stochastic weighted PageRank fixed point
 *   - the reference's PageRank variants: fixed-point pull 64/32/16/8-bit,
 *     quantized pull 32/16/8-bit, data-driven pull, push and pull-push
 *                                             01_host/src/algorithms/openmp/pageRank.c:496-2202
 *   - the seeded counter-based R-MAT generator (no reference counterpart)
 */
#ifndef GBX_ORACLE_H
#define GBX_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* In-place stable counting sort.  mode: 0/2 = by source, 1/3 = by
 * destination then source (sortRun.c:26-69; radix modes give identical
 * output).  move_weights mirrors `#if WEIGHTED` at countsort.c:118-120. */
int orc_count_sort(uint32_t *src, uint32_t *dst, float *w, uint32_t num_edges,
                   uint32_t num_vertices, int mode, int move_weights);

/* out_degree[v] and edges_idx[v] from a source-sorted edge list; edges_idx
 * stays 0 for vertices without edges (vertex.c:39-45,276-285). */
int orc_map_vertices(const uint32_t *sorted_src, uint32_t num_edges,
                     uint32_t num_vertices, uint32_t *out_degree,
                     uint32_t *edges_idx);

/* Adjacent duplicate and self-loop removal on a (src,dst)-sorted list, in
 * place; returns the new edge count.  Weights pass through a uint32_t as in
 * edgeList.c:261,270. */
uint32_t orc_remove_duplicates(uint32_t *src, uint32_t *dst, float *w,
                               uint32_t num_edges);

/* FP32 pull PageRank with the reference's semantics (pageRank.c:245-368).
 * fwd_out_degree: forward out-degree (divisor); rows come from
 * (row_degree,row_idx,row_col).  ranks_out receives the normalised ranks.
 * iter_seconds (nullable, length max_iters) receives per-iteration wall time.
 * nthreads < 1 runs on every OpenMP thread (omp_get_max_threads()).
 */
int orc_pr_pull_f32(uint32_t N, const uint32_t *fwd_out_degree,
                    const uint32_t *row_degree, const uint32_t *row_idx,
                    const uint32_t *row_col, uint32_t max_iters, double epsilon,
                    int nthreads, float *ranks_out, uint32_t *iters_out,
                    double *err_out, double *iter_seconds);

/* FP32 pull SpMV (SPMV.c:117-215).  w may be NULL (constant 0.0001f).
 * x_out/y_out receive the input and accumulated output vectors. */
int orc_spmv_pull_f32(uint32_t N, const uint32_t *fwd_out_degree,
                      const uint32_t *row_degree, const uint32_t *row_idx,
                      const uint32_t *row_col, const float *w,
                      uint32_t iterations, int nthreads, float *x_out,
                      float *y_out, double *iter_seconds);

/* FP64 PageRank power iteration on an in-CSR (row_ptr has N+1 entries).
 *  mode 0 (REF): r0 = 1-d; r' = (1-d) + d*sum r[u]/outdeg[u]; stop when no
 *     |dr| >= epsilon (or max_iters); output r/N.
 *  mode 1 (STOCHASTIC): r0 = t; r' = (1-d)*t + d*(sum P[u,v] r[u] + D/N),
 *     D = dangling mass; stop when L1(dr) < epsilon (or max_iters).
 *  w (nullable) holds per-edge transition weights in row order (already
 *  row-normalised by the caller); without w, P[u,v] = 1/outdeg[u].
 *  teleport (nullable) = t (sums to 1); default uniform.
 */
int orc_pr_f64(uint32_t N, const uint32_t *outdeg, const uint64_t *row_ptr,
               const uint32_t *col, const double *w, const double *teleport,
               double damp, int mode, uint32_t max_iters, double epsilon,
               double *ranks_out, uint32_t *iters_out, double *err_out);

/* FP64 adjoint of the stochastic weighted fixed point
 *   p = (1-d) t + d (P^T p + (delta.p)/N 1),  P[u,v] = W[u,v]/sum_v W[u,v]
 * for loss L = sum_v g[v] p[v]:  lambda = g + d (P lambda + delta (1.lambda)/N)
 * solved by iteration to tol; then
 *   dL/dt[v]   = (1-d) lambda[v]
 *   dL/dW[e=(u,v)] = d p[u] (lambda[v] - sum_x P[u,x] lambda[x]) / S[u]
 * with S[u] = sum of u's raw out-weights.  Edges are given as an in-CSR
 * (row v lists sources u) with raw weights wraw in the same order.  The
 * forward stops at sum|dp| < tol, the adjoint at sum|d lambda| < tol *
 * sum|lambda| (or max_iters each). */
int orc_pr_adjoint_f64(uint32_t N, const uint64_t *row_ptr, const uint32_t *col,
                       const double *wraw, const double *teleport,
                       const double *g, double damp, uint32_t max_iters,
                       double tol, double *p_out, double *lambda_out,
                       double *grad_w_out, double *grad_t_out);

/* L = sum_v g[v] p[v] at the forward fixed point for raw weights wraw
 * (same model as orc_pr_adjoint_f64), iterating from p0 (NULL: teleport)
 * until sum|dp| < tol or max_iters; p_out [N] receives p.  Central finite
 * differences of this are the independent check of the adjoint. */
double orc_pr_loss_f64(uint32_t N, const uint64_t *row_ptr, const uint32_t *col,
                       const double *wraw, const double *teleport, const double *g,
                       double damp, const double *p0, uint32_t max_iters, double tol,
                       double *p_out);

/* L(w_plus) - L(w_minus) for the same model, both fixed points iterated
 * `iters` times from p0 (NULL: teleport) in x87 extended precision, so that
 * a central difference is not swamped by FP64 rounding at hub rows (sums of
 * 10^5 terms whose rounding changes with any perturbation). */
double orc_pr_loss_diff(uint32_t N, const uint64_t *row_ptr, const uint32_t *col,
                        const double *w_plus, const double *w_minus, const double *teleport,
                        const double *g, double damp, const double *p0, uint32_t iters);

/* The reference's fixed-point pull PageRank (pageRankPullFixedPoint{64,32,
 * 16,8}BitGraphCSR, pageRank.c:496-1030; pushpull codes 2, 9, 10, 11, and
 * 3 = the push form of 64): in-CSR rows, REF update, stop when no vertex
 * changes by epsilon or more; ranks divided by N at the end.  *iters_out
 * follows PageRankStats.iterations (the stopping iteration's index). */
int orc_pr_pull_fixed(uint32_t N, const uint32_t *outdeg, const uint64_t *row_ptr,
                      const uint32_t *col, int bits, uint32_t max_iters, double epsilon,
                      float *ranks_out, uint32_t *iters_out, double *err_out);

/* The reference's quantized pull PageRank (pageRankPullQuant{32,16,8}Bit
 * GraphCSR, pageRank.c:1159-1577; pushpull codes 4, 12, 13): per-iteration
 * min/max scale, zero point 0, exact uint64 sums of the quantized
 * contributions; same update, stop rule and outputs as orc_pr_pull_fixed. */
int orc_pr_pull_quant(uint32_t N, const uint32_t *outdeg, const uint64_t *row_ptr,
                      const uint32_t *col, int bits, uint32_t max_iters, double epsilon,
                      float *ranks_out, uint32_t *iters_out, double *err_out);

/* The reference's data-driven pull PageRank (pageRankDataDrivenPullGraphCSR,
 * pageRank.c:1711-1870; pushpull code 6): worklist rows only, conditional
 * update, the out-neighbours of updated rows form the next worklist.  row_ptr
 * and col: in-CSR; fwd_row_ptr and fwd_col: forward CSR (out-neighbours). */
int orc_pr_pull_dd(uint32_t N, const uint32_t *outdeg, const uint64_t *row_ptr,
                   const uint32_t *col, const uint64_t *fwd_row_ptr, const uint32_t *fwd_col,
                   uint32_t max_iters, double epsilon, float *ranks_out, uint32_t *iters_out,
                   double *err_out);

/* The reference's data-driven push (pull_push 0, pageRankDataDrivenPush
 * GraphCSR, pageRank.c:1872-2027, pushpull code 7) and pull-push (pull_push
 * 1, pageRankDataDrivenPullPushGraphCSR, pageRank.c:2029-2202, code 8)
 * PageRank, sequential and in place as one reference thread runs them:
 * residual pushes along out-edges drive the worklist. */
int orc_pr_dd_push(uint32_t N, const uint32_t *outdeg, const uint64_t *row_ptr,
                   const uint32_t *col, const uint64_t *fwd_row_ptr, const uint32_t *fwd_col,
                   int pull_push, uint32_t max_iters, double epsilon, float *ranks_out,
                   uint32_t *iters_out, double *err_out);

/* Seeded counter-based R-MAT generator (same integer arithmetic as the GPU
 * generator).  Writes num_edges = edge_factor << scale edges, ids permuted
 * by the seeded bijection on [0, 2^scale). */
int orc_rmat_generate(uint32_t scale, uint32_t edge_factor, double a, double b,
                      double c, uint64_t seed, uint32_t *src, uint32_t *dst);

/* The seeded bijection on [0, 2^scale) used for vertex relabelling. */
uint32_t orc_rmat_permute(uint32_t x, uint32_t scale, uint64_t seed);

/* Remove self loops and duplicate (src,dst) pairs, keeping the first
 * occurrence and the original order; returns the kept count. */
uint64_t orc_dedupe_keep_first(uint32_t *src, uint32_t *dst, float *w,
                               uint64_t num_edges);

#ifdef __cplusplus
}
#endif
#endif
