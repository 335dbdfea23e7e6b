/*
 * graphblox_hip.h -- the GraphBlox host API for the pull PageRank / SpMV
 * path, served by the MI355X backend (libgraphblox_hip.so over libgbx_hip.so).
 *
 * The structs below have the same field order and types as the reference's
 * own headers so that reference host code (the L2 dispatch in
 * 01_host/src/utils_graph/graphRun.c) can call these functions unchanged:
 *   struct Arguments      01_host/include/config/graphConfig.h:18-63
 *   struct EdgeList       01_host/include/structures/edgeList.h:12-25
 *   struct Vertex         01_host/include/structures/vertex.h:17-23
 *   struct GraphCSR       01_host/include/structures/graphCSR.h:14-31
 *   struct PageRankStats  01_host/include/algorithms/openmp/pageRank.h:19-30
 *   struct SPMVStats      01_host/include/algorithms/openmp/SPMV.h:17-26
 * GBX_WEIGHTED selects the WEIGHTED=1 layout (GraphCSR.max_weight) and the
 * weighted SpMV, as the reference's compile-time WEIGHTED switch does
 * (graphConfig.h:13).  GBX_DIRECTED (default 1) selects the layout and the
 * pull direction of the reference's DIRECTED switch (graphConfig.h:14):
 * with 1 the GraphCSR carries the transpose (inverse_*) and the pull walks
 * it (pageRank.c:266-268); with 0 the GraphCSR has the forward CSR only and
 * the pull walks the forward rows (pageRank.c:269-272).
 *
 * Ownership follows the reference: the callee allocates stats and arrays
 * (4096-byte aligned, myMalloc.c:24-39); the caller frees them with
 * freePageRankStats / freeSPMVStats / graphCSRFree.  I/O failures end the
 * process as err(1, ...) does in the reference (edgeList.c:466-486); a GPU
 * failure inside an algorithm prints the backend's message and returns NULL,
 * which the reference's dispatch treats as "not available".
 */
#ifndef GRAPHBLOX_HIP_H
#define GRAPHBLOX_HIP_H

#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef GBX_WEIGHTED
#define GBX_WEIGHTED 0
#endif
#ifndef GBX_DIRECTED
#define GBX_DIRECTED 1
#endif

/* layout of the reference's mt19937 state (mt19937.h): 624 words + index */
typedef struct gbx_mt19937_state {
    unsigned long mt[624];
    int mti;
} gbx_mt19937_state;

struct Arguments {
    int wflag;
    int xflag;
    int sflag;
    int Sflag;
    int dflag;
    uint32_t binSize;
    uint32_t verbosity;
    uint32_t iterations;
    uint32_t trials;
    double epsilon;
    uint32_t source;
    uint32_t algorithm;
    uint32_t datastructure;
    uint32_t pushpull;
    uint32_t sort;
    uint32_t lmode;
    uint32_t lmode_l2;
    uint32_t lmode_l3;
    uint32_t mmode;
    uint32_t symmetric;
    uint32_t weighted;
    uint32_t delta;
    int pre_numThreads;
    int algo_numThreads;
    int ker_numThreads;
    char *fnameb;
    char *fnamel;
    uint32_t fnameb_format;
    uint32_t convert_format;
    gbx_mt19937_state mt19937var;
    uint32_t cache_size;
    char *kernel_name;
    int cu_id;
    int device_index;
    char *xclbin_path;
    char *overlay_path;
    int ctrl_mode;
    int endian_read;
    int endian_write;
    int flush_cache;
    int bankGroupIndex;
};

struct EdgeList {
    uint32_t num_edges;
    uint32_t num_vertices;
    uint32_t avg_degree;
    float max_weight;
    float *edges_array_weight;
    uint32_t *edges_array_src;
    uint32_t *edges_array_dest;
    uint32_t *mask_array;
    uint32_t *label_array;
    uint32_t *inverse_label_array;
};

struct Vertex {
    uint32_t num_vertices;
    uint32_t *out_degree;
    uint32_t *in_degree;
    uint32_t *edges_idx;
};

struct GraphCSR {
    uint32_t num_edges;
    uint32_t num_vertices;
    uint32_t avg_degree;
#if GBX_WEIGHTED
    float max_weight;
#endif
    struct Vertex *vertices;
    struct EdgeList *sorted_edges_array;
#if GBX_DIRECTED
    struct Vertex *inverse_vertices;
    struct EdgeList *inverse_sorted_edges_array;
#endif
};

struct PageRankStats {
    float damp;
    float base_pr;
    uint32_t iterations;
    uint32_t num_vertices;
    uint32_t activeVertices;
    uint32_t *realRanks;
    float *pageRanks;
    double time_total;
    double error_total;
};

struct SPMVStats {
    uint32_t iterations;
    struct Arguments *arguments;
    uint32_t num_vertices;
    float *vector_output;
    float *vector_input;
    double time_total;
};

/* graphConfig.c: defaults of argumentsNew (iterations 20, epsilon 1e-4, ...) */
struct Arguments *argumentsNew(void);
void argumentsFree(struct Arguments *arguments);

/* edgeList.c:450-706: mmap'd .bin/.wbin reader, N = max id + 1 */
struct EdgeList *readEdgeListsbin(const char *fname, uint8_t inverse, uint32_t symmetric,
                                  uint32_t weighted);
void freeEdgeList(struct EdgeList *edgeList);

/* graphCSR.c:243-327: read, sort (arguments->sort; -k dedupe forces 1),
 * forward CSR, then the transpose -- the sorts run on the GPU. */
struct GraphCSR *graphCSRPreProcessingStep(struct Arguments *arguments);
void graphCSRFree(struct GraphCSR *graphCSR);

/* pageRank.c:47-88, 177-368 */
struct PageRankStats *newPageRankStatsGraphCSR(struct GraphCSR *graph);
void freePageRankStats(struct PageRankStats *stats);
struct PageRankStats *pageRankGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
struct PageRankStats *pageRankPullGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
/* pageRank.c:370-492 (push; same sums as the pull iteration) */
struct PageRankStats *pageRankPushGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
/* pageRank.c:496-1030 and 1032-1157: the fixed-point variants, bit-exact */
struct PageRankStats *pageRankPullFixedPoint64BitGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
struct PageRankStats *pageRankPushFixedPointGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
struct PageRankStats *pageRankPullFixedPoint32BitGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
struct PageRankStats *pageRankPullFixedPoint16BitGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
struct PageRankStats *pageRankPullFixedPoint8BitGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
/* pageRank.c:1159-1577: the quantized pull variants, bit-exact */
struct PageRankStats *pageRankPullQuant32BitGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
struct PageRankStats *pageRankPullQuant16BitGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
struct PageRankStats *pageRankPullQuant8BitGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
/* pageRank.c:1579-1709: push quantised (the 32-bit quantized terms, exact sums) */
struct PageRankStats *pageRankPushQuantGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
/* pageRank.c:1711-2202: the data-driven pull, push and pull-push variants
 * (GBX_DIRECTED=1 builds; the GBX_DIRECTED=0 build returns NULL for them) */
struct PageRankStats *pageRankDataDrivenPullGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
struct PageRankStats *pageRankDataDrivenPushGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
struct PageRankStats *pageRankDataDrivenPullPushGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);

/* graphCSR.c:457-744: the .csr binary file (writer and reader, in this
 * build's DIRECTED layout) and the text-fixture writer */
void writeToBinFileGraphCSR(const char *fname, struct GraphCSR *graph);
struct GraphCSR *readFromBinFileGraphCSR(const char *fname);
void writetoTextFilesGraphCSR(const char *fname, struct GraphCSR *graph);

/* reorder.c:165-221: labels sorted by rank (GPU), and graphRun.c:933-975 */
uint32_t *radixSortEdgesByPageRank(float *pageRanks, uint32_t *labels, uint32_t num_vertices);
struct PageRankStats *runPageRankAlgorithm(struct Arguments *arguments, void *graph);

/* graphStats.h:32-41, graphStats.c:589-1010: top-k similarity of two
 * PageRank runs' rankings */
struct PageRankCorrelationStats {
    uint32_t levenshtein_distance;
    float float_Kendall;
    float real_Kendall;
    uint32_t intersection;
    uint32_t mismatch;
    double avg_error_float;
    double avg_error_relative;
};
struct PageRankCorrelationStats collectStatsPageRank_topK(struct PageRankStats *ref_stats, struct PageRankStats *stats,
                                                          uint32_t *ref_rankedVertices_total,
                                                          uint32_t *ref_rankedVertices_inverse,
                                                          uint32_t *rankedVertices_inverse, uint32_t topk,
                                                          uint32_t num_vertices, FILE *fptr, uint32_t verbose);
void collectStatsPageRank(struct Arguments *arguments, struct PageRankStats *ref_stats, struct PageRankStats *stats,
                          uint32_t trial);

/* SPMV.c:27-215 */
struct SPMVStats *newSPMVStatsGraphCSR(struct GraphCSR *graph);
void freeSPMVStats(struct SPMVStats *stats);
struct SPMVStats *SPMVGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
struct SPMVStats *SPMVPullGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
/* SPMV.c:216-590: push (the pull's terms) and the fixed-point variants (bit-exact) */
struct SPMVStats *SPMVPushGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
struct SPMVStats *SPMVPullFixedPointGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
struct SPMVStats *SPMVPushFixedPointGraphCSR(struct Arguments *arguments, struct GraphCSR *graph);
/* graphRun.c:859-893: SPMVGraphCSR for the CSR data structures (0, 4 and
 * unknown codes); segments (1, 5) and 6 are not implemented (NULL) */
struct SPMVStats *runSPMVAlgorithm(struct Arguments *arguments, void *graph);

/* Backend housekeeping (no reference counterpart).
 * The device copy of a GraphCSR that this library did not build (one made
 * by the reference's own graphCSRPreProcessingStep, or read from a file) is
 * made on first use and cached; every call checks the cached copy against a
 * fingerprint of the host graph (sizes, the struct and array addresses and
 * 64 sampled entries of every array) and uploads it again on a mismatch, so
 * a graph freed by other code and a new one at the same address are told
 * apart.  After editing a graph's arrays in place, call
 * graphbloxHipForget(graph): an edit that misses every sampled entry keeps
 * the fingerprint.  graphbloxHipShutdown releases the process-wide HIP
 * context and every cached device graph. */
void graphbloxHipForget(const struct GraphCSR *graph);
void graphbloxHipShutdown(void);

#ifdef __cplusplus
}
#endif
#endif /* GRAPHBLOX_HIP_H */
