/*
 * ref_driver.c -- test-infrastructure driver around the GraphBlox reference's
 * own OpenMP host code, compiled from the reference sources by oracle/Makefile
 * into oracle/_ref/.  It is a checker only: nothing in the product links it.
 *
 * It exposes a flat C interface (plain pointers and sizes) so the Python
 * tests and the golden-vector generator can drive:
 *   - graphCSRPreProcessingStep  (01_host/src/structures/graphCSR.c:243-327),
 *     also with the relabel levels  (01_host/src/preprocess/reorder.c:1556-1716)
 *   - pageRankPullGraphCSR       (01_host/src/algorithms/openmp/pageRank.c:245-368)
 *   - SPMVPullGraphCSR           (01_host/src/algorithms/openmp/SPMV.c:117-215)
 * and to run pageRankPullGraphCSR on CSR arrays supplied by the caller
 * (used as the timed CPU baseline).
 *
 * The reference prints progress tables to stdout; every call here sends
 * stdout to /dev/null for its duration.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#include <unistd.h>
#include <fcntl.h>
#include <omp.h>

#include "graphConfig.h"
#include "edgeList.h"
#include "vertex.h"
#include "graphCSR.h"
#include "pageRank.h"
#include "SPMV.h"
#include "timer.h"

static int quiet_begin(void)
{
    fflush(stdout);
    int saved = dup(1);
    int devnull = open("/dev/null", O_WRONLY);
    if (devnull >= 0) {
        dup2(devnull, 1);
        close(devnull);
    }
    return saved;
}

static void quiet_end(int saved)
{
    fflush(stdout);
    if (saved >= 0) {
        dup2(saved, 1);
        close(saved);
    }
}

int refd_directed(void) { return DIRECTED; }
int refd_weighted(void) { return WEIGHTED; }

/* graphCSRPreProcessingStep on a .bin/.wbin file. */
void *refd_build(const char *fname, uint32_t sort, int dflag, int symmetric,
                 int weighted, int nthreads)
{
    struct Arguments *a = argumentsNew();
    a->fnameb = (char *)fname;
    a->sort = sort;
    a->dflag = dflag;
    if (dflag)
        a->sort = 1; /* GraphBlox.c:358-359: dedupe forces sort=1 */
    a->symmetric = (uint32_t)symmetric;
    a->weighted = (uint32_t)weighted;
    omp_set_num_threads(nthreads > 0 ? nthreads : 1);
    int s = quiet_begin();
    struct GraphCSR *g = graphCSRPreProcessingStep(a);
    quiet_end(s);
    free(a);
    return g;
}

/* graphCSRPreProcessingStep with the relabel levels (lmode, lmode_l2,
 * lmode_l3; graphCSR.c:268-290, reorderGraphProcess reorder.c:1556-1616);
 * fnamel: the label file of lmode 11 (may be NULL). */
void *refd_build_relabel(const char *fname, uint32_t sort, int dflag, int symmetric, int weighted,
                         int nthreads, uint32_t lmode, uint32_t lmode_l2, uint32_t lmode_l3,
                         const char *fnamel)
{
    struct Arguments *a = argumentsNew();
    a->fnameb = (char *)fname;
    a->fnamel = (char *)fnamel;
    a->sort = sort;
    a->dflag = dflag;
    if (dflag)
        a->sort = 1; /* GraphBlox.c:358-359 */
    a->symmetric = (uint32_t)symmetric;
    a->weighted = (uint32_t)weighted;
    a->lmode = lmode;
    a->lmode_l2 = lmode_l2;
    a->lmode_l3 = lmode_l3;
    omp_set_num_threads(nthreads > 0 ? nthreads : 1);
    int s = quiet_begin();
    struct GraphCSR *g = graphCSRPreProcessingStep(a);
    quiet_end(s);
    free(a);
    return g;
}

/* The forward edge list's label arrays (old id -> new id and back). */
int refd_labels(void *gp, uint32_t *label, uint32_t *inverse_label)
{
    struct GraphCSR *g = (struct GraphCSR *)gp;
    struct EdgeList *e = g->sorted_edges_array;
    uint32_t N = g->num_vertices;
    if (label) memcpy(label, e->label_array, (size_t)N * 4);
    if (inverse_label) memcpy(inverse_label, e->inverse_label_array, (size_t)N * 4);
    return 0;
}

uint32_t refd_num_vertices(void *gp) { return ((struct GraphCSR *)gp)->num_vertices; }
uint32_t refd_num_edges(void *gp) { return ((struct GraphCSR *)gp)->num_edges; }

/* Copy out the forward (inverse=0) or transpose (inverse=1) CSR arrays. */
int refd_export(void *gp, int inverse, uint32_t *out_degree, uint32_t *in_degree,
                uint32_t *edges_idx, uint32_t *src, uint32_t *dst, float *w)
{
    struct GraphCSR *g = (struct GraphCSR *)gp;
    struct Vertex *v = g->vertices;
    struct EdgeList *e = g->sorted_edges_array;
#if DIRECTED
    if (inverse) {
        v = g->inverse_vertices;
        e = g->inverse_sorted_edges_array;
    }
#else
    if (inverse)
        return -1;
#endif
    uint32_t N = g->num_vertices, E = g->num_edges;
    if (out_degree) memcpy(out_degree, v->out_degree, (size_t)N * 4);
    if (in_degree) memcpy(in_degree, v->in_degree, (size_t)N * 4);
    if (edges_idx) memcpy(edges_idx, v->edges_idx, (size_t)N * 4);
    if (src) memcpy(src, e->edges_array_src, (size_t)E * 4);
    if (dst) memcpy(dst, e->edges_array_dest, (size_t)E * 4);
    if (w) memcpy(w, e->edges_array_weight, (size_t)E * 4);
    return 0;
}

void refd_free(void *gp)
{
    int s = quiet_begin();
    graphCSRFree((struct GraphCSR *)gp);
    quiet_end(s);
}

int refd_pagerank(void *gp, uint32_t iterations, double epsilon, int nthreads,
                  float *ranks_out, uint32_t *iters_out, double *err_out,
                  double *time_out)
{
    struct Arguments *a = argumentsNew();
    a->iterations = iterations;
    a->epsilon = epsilon;
    a->ker_numThreads = nthreads > 0 ? nthreads : 1;
    omp_set_num_threads(a->ker_numThreads);
    int s = quiet_begin();
    struct PageRankStats *st = pageRankPullGraphCSR(a, (struct GraphCSR *)gp);
    quiet_end(s);
    if (ranks_out) memcpy(ranks_out, st->pageRanks, (size_t)st->num_vertices * 4);
    if (iters_out) *iters_out = st->iterations;
    if (err_out) *err_out = st->error_total;
    if (time_out) *time_out = st->time_total;
    freePageRankStats(st);
    free(a);
    return 0;
}

/* pageRankGraphCSR with a variant code (arguments->pushpull, pageRank.c:177-242). */
int refd_pagerank_variant(void *gp, int pushpull, uint32_t iterations, double epsilon, int nthreads,
                          float *ranks_out, uint32_t *iters_out, double *err_out)
{
    struct Arguments *a = argumentsNew();
    a->iterations = iterations;
    a->epsilon = epsilon;
    a->pushpull = pushpull;
    a->ker_numThreads = nthreads > 0 ? nthreads : 1;
    omp_set_num_threads(a->ker_numThreads);
    int s = quiet_begin();
    struct PageRankStats *st = pageRankGraphCSR(a, (struct GraphCSR *)gp);
    quiet_end(s);
    if (ranks_out) memcpy(ranks_out, st->pageRanks, (size_t)st->num_vertices * 4);
    if (iters_out) *iters_out = st->iterations;
    if (err_out) *err_out = st->error_total;
    freePageRankStats(st);
    free(a);
    return 0;
}

int refd_spmv(void *gp, uint32_t iterations, int nthreads, float *y_out,
              float *x_out, double *time_out)
{
    struct Arguments *a = argumentsNew();
    a->iterations = iterations;
    a->ker_numThreads = nthreads > 0 ? nthreads : 1;
    omp_set_num_threads(a->ker_numThreads);
    int s = quiet_begin();
    struct SPMVStats *st = SPMVPullGraphCSR(a, (struct GraphCSR *)gp);
    quiet_end(s);
    if (y_out) memcpy(y_out, st->vector_output, (size_t)st->num_vertices * 4);
    if (x_out) memcpy(x_out, st->vector_input, (size_t)st->num_vertices * 4);
    if (time_out) *time_out = st->time_total;
    freeSPMVStats(st);
    free(a);
    return 0;
}

/* SPMVGraphCSR (SPMV.c:89-116) by variant code: 0 pull, 1 push, 2 pull fixed
 * point, 3 push fixed point. */
int refd_spmv_variant(void *gp, int pushpull, uint32_t iterations, int nthreads, float *y_out,
                      float *x_out)
{
    struct Arguments *a = argumentsNew();
    a->iterations = iterations;
    a->pushpull = (uint32_t)pushpull;
    a->ker_numThreads = nthreads > 0 ? nthreads : 1;
    omp_set_num_threads(a->ker_numThreads);
    int s = quiet_begin();
    struct SPMVStats *st = SPMVGraphCSR(a, (struct GraphCSR *)gp);
    quiet_end(s);
    if (y_out) memcpy(y_out, st->vector_output, (size_t)st->num_vertices * 4);
    if (x_out) memcpy(x_out, st->vector_input, (size_t)st->num_vertices * 4);
    freeSPMVStats(st);
    free(a);
    return 0;
}

#if DIRECTED
/*
 * Run the reference pageRankPullGraphCSR on caller-owned CSR arrays (no file,
 * no CPU sort).  fwd_out_degree is the forward out-degree (read at
 * pageRank.c:297-299); in_degree/in_idx/in_src are the transpose CSR rows
 * (pageRank.c:266-268).  Arrays are borrowed; nothing is freed here.
 */
int refd_pagerank_on_csr(uint32_t N, uint32_t E, uint32_t *fwd_out_degree,
                         uint32_t *in_degree, uint32_t *in_idx, uint32_t *in_src,
                         uint32_t iterations, double epsilon, int nthreads,
                         float *ranks_out, uint32_t *iters_out, double *err_out,
                         double *time_out)
{
    struct Vertex fwd = {N, fwd_out_degree, in_degree, in_idx};
    struct Vertex inv = {N, in_degree, fwd_out_degree, in_idx};
    struct EdgeList inv_edges;
    memset(&inv_edges, 0, sizeof(inv_edges));
    inv_edges.num_edges = E;
    inv_edges.num_vertices = N;
    inv_edges.edges_array_dest = in_src;
    struct GraphCSR g;
    memset(&g, 0, sizeof(g));
    g.num_vertices = N;
    g.num_edges = E;
    g.vertices = &fwd;
    g.inverse_vertices = &inv;
    g.inverse_sorted_edges_array = &inv_edges;
    g.sorted_edges_array = &inv_edges; /* unused by the pull kernel */
    return refd_pagerank(&g, iterations, epsilon, nthreads, ranks_out,
                         iters_out, err_out, time_out);
}
#endif
