/*
 * graphblox_hip.c -- the GraphBlox host API for pull PageRank / SpMV on the
 * MI355X backend.  Each function keeps the name, arguments, ownership and
 * error behaviour of the reference function cited above it and runs the
 * compute through the C ABI of libgbx_hip.so (include/gbx.h).
 */
#define _GNU_SOURCE
#include "graphblox_hip.h"

#include <fcntl.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include "gbx.h"

#define GBX_HOST_ALIGN 4096 /* myMalloc.h:9-11 (ALIGNED, CACHELINE_BYTES) */

/* 4096-byte aligned allocation that ends the process on failure
 * (my_malloc, myMalloc.c:24-67). */
static void *host_malloc(size_t size)
{
    void *p = NULL;
    if (size == 0) size = 4;
    if (posix_memalign(&p, GBX_HOST_ALIGN, size) != 0 || !p) {
        fprintf(stderr, "graphblox_hip: out of host memory (%zu bytes)\n", size);
        exit(EXIT_FAILURE);
    }
    return p;
}

static void die_io(const char *what, const char *fname)
{
    fprintf(stderr, "graphblox_hip: %s: %s\n", what, fname);
    exit(1);
}

/* ------------------------------------------------------------ backend state */

static gbx_ctx *g_ctx = NULL;

/* The pull walks the transpose rows with DIRECTED=1 (pageRank.c:266-268),
 * the forward rows with DIRECTED=0 (pageRank.c:269-272).  The push forms
 * push along the forward out-edges under either setting (pageRank.c:424-437,
 * SPMV.c:232-281), so each vertex receives the sum over its in-edges: the
 * transpose rows' pull. */
#define PULL_DIRECTION (GBX_DIRECTED ? GBX_PULL_IN : GBX_PULL_FWD)
#define PUSH_DIRECTION GBX_PULL_IN

/* What a cached device graph was made from: sizes, the addresses of the
 * host structs and arrays, and a hash of 64 sampled entries of every array.
 * A graph freed by other code and a new one built at the same address
 * differ in at least one of them (unless every sampled entry agrees). */
enum { PRINT_PTRS = 14, PRINT_SAMPLES = 64 };
struct graph_print {
    uint32_t N, E;
    const void *ptr[PRINT_PTRS];
    uint64_t hash;
};

struct graph_entry {
    const struct GraphCSR *host;
    struct graph_print print;
    gbx_graph *dev;
    struct graph_entry *next;
};
static struct graph_entry *g_graphs = NULL;

static uint64_t hash_samples(uint64_t h, const void *arr, uint32_t n)
{
    const uint32_t *a = (const uint32_t *)arr;
    if (!a || n == 0) return h ^ 0x9e3779b97f4a7c15ull;
    const uint32_t k = n < PRINT_SAMPLES ? n : PRINT_SAMPLES;
    for (uint32_t i = 0; i < k; i++) {
        const uint32_t pos = k == 1 ? 0 : (uint32_t)(((uint64_t)i * (n - 1)) / (k - 1));
        h = (h ^ a[pos]) * 0x100000001b3ull; /* FNV-1a, one word at a time */
    }
    return h;
}

static void fingerprint(const struct GraphCSR *g, struct graph_print *fp)
{
    memset(fp, 0, sizeof(*fp));
    fp->N = g->num_vertices;
    fp->E = g->num_edges;
    const struct Vertex *vs[2] = {g->vertices, NULL};
    const struct EdgeList *es[2] = {g->sorted_edges_array, NULL};
#if GBX_DIRECTED
    vs[1] = g->inverse_vertices;
    es[1] = g->inverse_sorted_edges_array;
#endif
    uint64_t h = 0xcbf29ce484222325ull;
    int k = 0;
    for (int side = 0; side < 2; side++) {
        const struct Vertex *v = vs[side];
        const struct EdgeList *e = es[side];
        fp->ptr[k++] = v;
        fp->ptr[k++] = e;
        if (v) {
            fp->ptr[k++] = v->out_degree;
            fp->ptr[k++] = v->edges_idx;
            h = hash_samples(h, v->out_degree, g->num_vertices);
            h = hash_samples(h, v->edges_idx, g->num_vertices);
        } else {
            k += 2;
        }
        if (e) {
            fp->ptr[k++] = e->edges_array_src;
            fp->ptr[k++] = e->edges_array_dest;
            fp->ptr[k++] = GBX_WEIGHTED ? (const void *)e->edges_array_weight : NULL;
            h = hash_samples(h, e->edges_array_src, g->num_edges);
            h = hash_samples(h, e->edges_array_dest, g->num_edges);
            if (GBX_WEIGHTED) h = hash_samples(h, e->edges_array_weight, g->num_edges);
        } else {
            k += 3;
        }
    }
    fp->hash = h;
}

static int same_print(const struct graph_print *a, const struct graph_print *b)
{
    if (a->N != b->N || a->E != b->E || a->hash != b->hash) return 0;
    for (int i = 0; i < PRINT_PTRS; i++)
        if (a->ptr[i] != b->ptr[i]) return 0;
    return 1;
}

static gbx_ctx *backend(void)
{
    if (!g_ctx) {
        int dev = 0;
        const char *s = getenv("GBX_DEVICE");
        if (s) dev = atoi(s);
        int rc = gbx_init(dev, &g_ctx);
        if (rc != GBX_OK) {
            fprintf(stderr, "graphblox_hip: HIP backend unavailable: %s\n", gbx_last_error(NULL));
            g_ctx = NULL;
        }
    }
    return g_ctx;
}

static void unregister_graph(const struct GraphCSR *host)
{
    struct graph_entry **pp = &g_graphs;
    while (*pp) {
        if ((*pp)->host == host) {
            struct graph_entry *e = *pp;
            *pp = e->next;
            gbx_graph_free(e->dev);
            free(e);
            return;
        }
        pp = &(*pp)->next;
    }
}

static void register_graph(const struct GraphCSR *host, gbx_graph *dev)
{
    unregister_graph(host);
    struct graph_entry *e = host_malloc(sizeof(*e));
    e->host = host;
    fingerprint(host, &e->print);
    e->dev = dev;
    e->next = g_graphs;
    g_graphs = e;
}

/* The cached device copy of host, or NULL; a copy whose fingerprint no longer
 * matches the host graph is released. */
static gbx_graph *lookup_graph(const struct GraphCSR *host)
{
    for (struct graph_entry *e = g_graphs; e; e = e->next)
        if (e->host == host) {
            struct graph_print now;
            fingerprint(host, &now);
            if (same_print(&now, &e->print)) return e->dev;
            unregister_graph(host);
            return NULL;
        }
    return NULL;
}

void graphbloxHipForget(const struct GraphCSR *graph)
{
    if (graph) unregister_graph(graph);
}

void graphbloxHipShutdown(void)
{
    while (g_graphs) {
        struct graph_entry *e = g_graphs;
        g_graphs = e->next;
        gbx_graph_free(e->dev);
        free(e);
    }
    if (g_ctx) gbx_free(g_ctx);
    g_ctx = NULL;
}

/* Device graph for a host GraphCSR: the one built by graphCSRPreProcessingStep,
 * or an upload of arrays built elsewhere (gbx_graph_from_csr; with DIRECTED=0
 * the forward CSR only, whose transpose the device builds itself). */
static gbx_graph *device_graph(struct GraphCSR *graph)
{
    gbx_graph *g = lookup_graph(graph);
    if (g) return g;
    gbx_ctx *ctx = backend();
    if (!ctx) return NULL;
    struct Vertex *fv = graph->vertices;
    struct EdgeList *fe = graph->sorted_edges_array;
    if (!fv || !fe) {
        fprintf(stderr, "graphblox_hip: graph has no forward CSR\n");
        return NULL;
    }
    const float *fw = GBX_WEIGHTED ? fe->edges_array_weight : NULL;
#if GBX_DIRECTED
    struct Vertex *iv = graph->inverse_vertices;
    struct EdgeList *ie = graph->inverse_sorted_edges_array;
    if (!iv || !ie) {
        fprintf(stderr, "graphblox_hip: graph has no transpose CSR (DIRECTED=1 layout)\n");
        return NULL;
    }
    const float *iw = GBX_WEIGHTED ? ie->edges_array_weight : NULL;
    int rc = gbx_graph_from_csr(ctx, graph->num_vertices, graph->num_edges, fv->out_degree, fv->edges_idx,
                                fe->edges_array_src, fe->edges_array_dest, fw, iv->out_degree, iv->edges_idx,
                                ie->edges_array_src, ie->edges_array_dest, iw, &g);
#else
    int rc = gbx_graph_from_csr(ctx, graph->num_vertices, graph->num_edges, fv->out_degree, fv->edges_idx,
                                fe->edges_array_src, fe->edges_array_dest, fw, NULL, NULL, NULL, NULL, NULL, &g);
#endif
    if (rc != GBX_OK) {
        fprintf(stderr, "graphblox_hip: upload failed: %s\n", gbx_last_error(ctx));
        return NULL;
    }
    register_graph(graph, g);
    return g;
}

/* ------------------------------------------------------------ arguments */

/* argumentsNew, graphConfig.c:47-101 (defaults) */
struct Arguments *argumentsNew(void)
{
    struct Arguments *a = host_malloc(sizeof(*a));
    memset(a, 0, sizeof(*a));
    a->binSize = 512;
    a->iterations = 20;
    a->trials = 1;
    a->epsilon = 0.0001;
    a->cache_size = 32768;
    a->delta = 1;
    a->pre_numThreads = 1;
    a->algo_numThreads = 1;
    a->ker_numThreads = 1;
    a->fnameb_format = 1;
    a->convert_format = 1;
    a->flush_cache = 1;
    return a;
}

void argumentsFree(struct Arguments *arguments) { free(arguments); }

/* ------------------------------------------------------------ edge lists */

static struct EdgeList *new_edge_list(uint32_t num_edges)
{
    struct EdgeList *e = host_malloc(sizeof(*e));
    memset(e, 0, sizeof(*e));
    e->edges_array_src = host_malloc((size_t)num_edges * 4);
    e->edges_array_dest = host_malloc((size_t)num_edges * 4);
    e->edges_array_weight = host_malloc((size_t)num_edges * 4);
    for (uint32_t i = 0; i < num_edges; i++) e->edges_array_weight[i] = 1.0f; /* edgeList.c:108-114 */
    e->num_edges = num_edges;
    e->max_weight = 1.0f;
    return e;
}

static void label_arrays(struct EdgeList *e, uint32_t num_vertices)
{
    e->mask_array = host_malloc(((size_t)num_vertices + 1) * 4);
    e->label_array = host_malloc(((size_t)num_vertices + 1) * 4);
    e->inverse_label_array = host_malloc(((size_t)num_vertices + 1) * 4);
    for (uint32_t i = 0; i < num_vertices; i++) {
        e->mask_array[i] = 0;
        e->label_array[i] = i;
        e->inverse_label_array[i] = i;
    }
}

/* readEdgeListsbin, edgeList.c:450-706 */
struct EdgeList *readEdgeListsbin(const char *fname, uint8_t inverse, uint32_t symmetric, uint32_t weighted)
{
    int fd = open(fname, O_RDONLY);
    if (fd == -1) die_io("open", fname);
    struct stat fs;
    if (fstat(fd, &fs) == -1) die_io("stat", fname);
    const size_t rec = weighted ? 12 : 8;
    const uint64_t n64 = (uint64_t)fs.st_size / rec;
    if (n64 > 0xFFFFFFFFull || (symmetric && n64 > 0x7FFFFFFFull)) die_io("too many edges", fname);
    const uint32_t num_edges = (uint32_t)n64;
    const uint32_t *buf = NULL;
    void *map = NULL;
    if (fs.st_size > 0) {
        map = mmap(NULL, (size_t)fs.st_size, PROT_READ, MAP_PRIVATE, fd, 0);
        if (map == MAP_FAILED) die_io("mmap", fname);
        buf = (const uint32_t *)map;
    }
    const uint32_t total = symmetric ? 2 * num_edges : num_edges;
    struct EdgeList *e = new_edge_list(total);
    const size_t stride = rec / 4;
    uint32_t maxid = 0;
    float max_weight = 0.0f;
    for (uint32_t i = 0; i < num_edges; i++) {
        uint32_t s = buf[stride * i], d = buf[stride * i + 1];
        if (inverse) {
            uint32_t t = s;
            s = d;
            d = t;
        }
        e->edges_array_src[i] = s;
        e->edges_array_dest[i] = d;
        if (symmetric) {
            e->edges_array_src[i + num_edges] = d;
            e->edges_array_dest[i + num_edges] = s;
        }
#if GBX_WEIGHTED
        float w = 1.0f;
        if (weighted) memcpy(&w, &buf[stride * i + 2], 4);
        e->edges_array_weight[i] = w;
        if (symmetric) e->edges_array_weight[i + num_edges] = w;
        if (w > max_weight) max_weight = w;
#endif
        if (s > maxid) maxid = s;
        if (d > maxid) maxid = d;
    }
    e->num_vertices = maxid + 1; /* edgeList.c:679 */
    e->avg_degree = e->num_edges / e->num_vertices;
    e->max_weight = GBX_WEIGHTED ? max_weight : 1.0f;
    label_arrays(e, e->num_vertices);
    if (map) munmap(map, (size_t)fs.st_size);
    close(fd);
    return e;
}

void freeEdgeList(struct EdgeList *e)
{
    if (!e) return;
    free(e->edges_array_src);
    free(e->edges_array_dest);
    free(e->edges_array_weight);
    free(e->mask_array);
    free(e->label_array);
    free(e->inverse_label_array);
    free(e);
}

/* ------------------------------------------------------------ CSR */

static struct Vertex *new_vertex_array(uint32_t num_vertices)
{
    struct Vertex *v = host_malloc(sizeof(*v));
    v->num_vertices = num_vertices;
    v->out_degree = host_malloc(((size_t)num_vertices + 1) * 4);
    v->in_degree = host_malloc(((size_t)num_vertices + 1) * 4);
    v->edges_idx = host_malloc(((size_t)num_vertices + 1) * 4);
    v->out_degree[num_vertices] = v->in_degree[num_vertices] = v->edges_idx[num_vertices] = 0;
    return v;
}

static void free_vertex_array(struct Vertex *v)
{
    if (!v) return;
    free(v->out_degree);
    free(v->in_degree);
    free(v->edges_idx);
    free(v);
}

/* relabelEdgeListFromFile, reorder.c:1662-1716: the new id of each vertex,
 * one "%u" per line, num_vertices of them (a shorter file is an error here;
 * the reference leaves the rest of its array unset). */
static uint32_t *read_label_file(const char *fname, uint32_t num_vertices)
{
    if (!fname) die_io("open", "(no label file)");
    FILE *f = fopen(fname, "r");
    if (!f) die_io("open", fname);
    uint32_t *labels = host_malloc(((size_t)num_vertices + 1) * 4);
    for (uint32_t x = 0; x < num_vertices; x++)
        if (fscanf(f, "%u", &labels[x]) != 1) die_io("too few labels in", fname);
    fclose(f);
    return labels;
}

/* graphCSRPreProcessingStep, graphCSR.c:243-327: the reader
 * (readEdgeListsbin's records, inverse swap and symmetric append, max id),
 * both count sorts, the -k dedupe, the relabel levels (lmode, lmode_l2,
 * lmode_l3) and the vertex mapping run on the GPU.  With
 * DIRECTED=0 the GraphCSR holds the forward CSR only, whose in_degree the
 * reference leaves zero (vertex.c:39-45, written only under DIRECTED at
 * vertex.c:291-311); the device keeps its transpose as well. */
struct GraphCSR *graphCSRPreProcessingStep(struct Arguments *arguments)
{
    gbx_ctx *ctx = backend();
    if (!ctx) exit(EXIT_FAILURE);
    if (arguments->mmode)
        fprintf(stderr, "graphblox_hip: mask modes (mmode) are not provided by this backend; ignored\n");
    gbx_edges *in = NULL;
    if (gbx_edges_read_file(ctx, arguments->fnameb, arguments->weighted ? 1 : 0, 0, arguments->symmetric ? 1 : 0,
                            &in) != GBX_OK) {
        fprintf(stderr, "graphblox_hip: %s\n", gbx_last_error(ctx));
        exit(EXIT_FAILURE);
    }
    uint32_t in_vertices = 0;
    float in_max_weight = 1.0f;
    gbx_edges_info(in, &in_vertices, NULL, &in_max_weight);
    /* the relabel levels (graphCSR.c:268-290) run on the GPU inside the build */
    gbx_relabel rl;
    memset(&rl, 0, sizeof(rl));
    rl.lmode[0] = arguments->lmode;
    rl.lmode[1] = arguments->lmode_l2;
    rl.lmode[2] = arguments->lmode_l3;
    const int relabel = rl.lmode[0] || rl.lmode[1] || rl.lmode[2];
    uint32_t *file_labels = NULL;
    if (rl.lmode[0] == 11 || rl.lmode[1] == 11 || rl.lmode[2] == 11)
        file_labels = read_label_file(arguments->fnamel, in_vertices);
    rl.file_labels = file_labels;
    rl.num_file_labels = file_labels ? in_vertices : 0;
    uint32_t *labels = relabel ? host_malloc(((size_t)in_vertices + 1) * 4) : NULL;
    unsigned flags = 0;
    uint32_t sort = arguments->sort;
    if (arguments->dflag) {
        flags |= GBX_GRAPH_DEDUPE;
        sort = 1; /* GraphBlox.c:358-359 */
    }
    if (GBX_WEIGHTED) flags |= GBX_GRAPH_WEIGHTED;
    if (sort > 3) sort = 0; /* sortRun.c:59-64: unknown modes fall back to radix by source */
    gbx_graph *g = NULL;
    int rc = gbx_graph_create_from_edges(ctx, in, (int)sort, flags, relabel ? &rl : NULL, labels, &g);
    gbx_edges_free(in);
    free(file_labels);
    if (rc != GBX_OK) {
        fprintf(stderr, "graphblox_hip: CSR build failed: %s\n", gbx_last_error(ctx));
        exit(EXIT_FAILURE);
    }
    uint32_t N = 0, E = 0;
    gbx_graph_info(g, &N, &E);
    struct GraphCSR *csr = host_malloc(sizeof(*csr));
    memset(csr, 0, sizeof(*csr));
    csr->num_vertices = N;
    csr->num_edges = E;
    csr->avg_degree = E / N;
#if GBX_WEIGHTED
    csr->max_weight = in_max_weight;
#endif
    for (int inv = 0; inv < 1 + GBX_DIRECTED; inv++) {
        struct Vertex *v = new_vertex_array(N);
        struct EdgeList *el = new_edge_list(E);
        el->num_vertices = N;
        el->avg_degree = E / N;
        el->max_weight = GBX_WEIGHTED ? in_max_weight : 1.0f;
        label_arrays(el, N);
        if (labels) /* reorder.c:376-380: old id -> new id and back */
            for (uint32_t x = 0; x < N; x++) {
                el->label_array[x] = labels[x];
                el->inverse_label_array[labels[x]] = x;
            }
        rc = gbx_graph_export(g, inv, v->out_degree, v->in_degree, v->edges_idx, el->edges_array_src,
                              el->edges_array_dest, el->edges_array_weight);
        if (rc != GBX_OK) {
            fprintf(stderr, "graphblox_hip: CSR export failed: %s\n", gbx_last_error(ctx));
            exit(EXIT_FAILURE);
        }
        if (!GBX_DIRECTED) memset(v->in_degree, 0, (size_t)N * 4);
#if GBX_DIRECTED
        if (inv) {
            csr->inverse_vertices = v;
            csr->inverse_sorted_edges_array = el;
            continue;
        }
#endif
        csr->vertices = v;
        csr->sorted_edges_array = el;
    }
    free(labels);
    register_graph(csr, g);
    printf(" -----------------------------------------------------\n");
    printf("| %-51s | \n", "GraphCSR (HIP backend)");
    printf(" -----------------------------------------------------\n");
    printf("| %-30s | %-18u | \n", "Number of Vertices (V)", N);
    printf("| %-30s | %-18u | \n", "Number of Edges (E)", E);
    printf(" -----------------------------------------------------\n");
    return csr;
}

void graphCSRFree(struct GraphCSR *graph)
{
    if (!graph) return;
    unregister_graph(graph);
    free_vertex_array(graph->vertices);
    freeEdgeList(graph->sorted_edges_array);
#if GBX_DIRECTED
    free_vertex_array(graph->inverse_vertices);
    freeEdgeList(graph->inverse_sorted_edges_array);
#endif
    free(graph);
}

/* ------------------------------------------------------------ CSR files */

static char *path_with(const char *fname, const char *ext)
{
    const size_t n = strlen(fname), m = strlen(ext);
    char *p = host_malloc(n + m + 1);
    memcpy(p, fname, n);
    memcpy(p + n, ext, m + 1);
    return p;
}

static void write_all(FILE *f, const void *buf, size_t words, const char *path)
{
    if (words && fwrite(buf, 4, words, f) != words) die_io("write", path);
}

static void read_all(FILE *f, void *buf, size_t words, const char *path)
{
    if (words && fread(buf, 4, words, f) != words) die_io("read", path);
}

enum { CSR_IO_CHUNK = 1 << 16 }; /* records staged per fwrite / fread */

/* writeToBinFileGraphCSR, graphCSR.c:457-527: "<fname>.csr" holds num_edges
 * and num_vertices (u32) [and max_weight (f32) with WEIGHTED]; then for every
 * vertex the forward (out_degree, in_degree, edges_idx) followed by the
 * inverse triple when the graph has an inverse; then for every edge the
 * forward (src, dest [, weight]) followed by the inverse record. */
void writeToBinFileGraphCSR(const char *fname, struct GraphCSR *graph)
{
    char *path = path_with(fname, ".csr");
    FILE *f = fopen(path, "wb");
    if (!f) die_io("open", path);
    const uint32_t N = graph->num_vertices, E = graph->num_edges;
#if GBX_DIRECTED
    const int inv = graph->inverse_vertices != NULL;
#else
    const int inv = 0;
#endif
    write_all(f, &graph->num_edges, 1, path);
    write_all(f, &graph->num_vertices, 1, path);
#if GBX_WEIGHTED
    write_all(f, &graph->max_weight, 1, path);
#endif
    uint32_t *rec = host_malloc((size_t)CSR_IO_CHUNK * 6 * 4);
    for (uint32_t b = 0; b < N; b += CSR_IO_CHUNK) {
        const uint32_t n = N - b < CSR_IO_CHUNK ? N - b : CSR_IO_CHUNK;
        size_t k = 0;
        for (uint32_t i = b; i < b + n; i++) {
            const struct Vertex *v = graph->vertices;
            rec[k++] = v->out_degree[i];
            rec[k++] = v->in_degree[i];
            rec[k++] = v->edges_idx[i];
#if GBX_DIRECTED
            if (inv) {
                const struct Vertex *w = graph->inverse_vertices;
                rec[k++] = w->out_degree[i];
                rec[k++] = w->in_degree[i];
                rec[k++] = w->edges_idx[i];
            }
#endif
        }
        write_all(f, rec, k, path);
    }
    for (uint32_t b = 0; b < E; b += CSR_IO_CHUNK) {
        const uint32_t n = E - b < CSR_IO_CHUNK ? E - b : CSR_IO_CHUNK;
        size_t k = 0;
        for (uint32_t i = b; i < b + n; i++) {
            for (int side = 0; side < 1 + inv; side++) {
#if GBX_DIRECTED
                const struct EdgeList *e = side ? graph->inverse_sorted_edges_array : graph->sorted_edges_array;
#else
                const struct EdgeList *e = graph->sorted_edges_array;
#endif
                rec[k++] = e->edges_array_src[i];
                rec[k++] = e->edges_array_dest[i];
#if GBX_WEIGHTED
                memcpy(&rec[k++], &e->edges_array_weight[i], 4);
#endif
            }
        }
        write_all(f, rec, k, path);
    }
    free(rec);
    if (fclose(f) != 0) die_io("close", path);
    free(path);
}

/* readFromBinFileGraphCSR, graphCSR.c:638-744: the file written above (with
 * the inverse records in the DIRECTED=1 layout, without them in the
 * DIRECTED=0 layout), as a host GraphCSR; its device copy is made when an
 * algorithm first runs on it. */
struct GraphCSR *readFromBinFileGraphCSR(const char *fname)
{
    FILE *f = fopen(fname, "rb");
    if (!f) die_io("open", fname);
    uint32_t hdr[2];
    read_all(f, hdr, 2, fname);
    const uint32_t E = hdr[0], N = hdr[1];
    struct GraphCSR *csr = host_malloc(sizeof(*csr));
    memset(csr, 0, sizeof(*csr));
    csr->num_edges = E;
    csr->num_vertices = N;
    csr->avg_degree = N ? E / N : 0;
    float max_weight = 1.0f;
#if GBX_WEIGHTED
    read_all(f, &max_weight, 1, fname);
    csr->max_weight = max_weight;
#endif
    const int sides = 1 + GBX_DIRECTED;
    struct Vertex *vs[2] = {NULL, NULL};
    struct EdgeList *es[2] = {NULL, NULL};
    for (int side = 0; side < sides; side++) {
        vs[side] = new_vertex_array(N);
        struct EdgeList *el = new_edge_list(E);
        el->num_vertices = N;
        el->avg_degree = csr->avg_degree;
        el->max_weight = max_weight;
        label_arrays(el, N);
        es[side] = el;
    }
    csr->vertices = vs[0];
    csr->sorted_edges_array = es[0];
#if GBX_DIRECTED
    csr->inverse_vertices = vs[1];
    csr->inverse_sorted_edges_array = es[1];
#endif
    const size_t per_vertex = 3 * (size_t)sides, per_side = GBX_WEIGHTED ? 3 : 2;
    const size_t per_edge = per_side * (size_t)sides;
    uint32_t *rec = host_malloc((size_t)CSR_IO_CHUNK * 6 * 4);
    for (uint32_t b = 0; b < N; b += CSR_IO_CHUNK) {
        const uint32_t n = N - b < CSR_IO_CHUNK ? N - b : CSR_IO_CHUNK;
        read_all(f, rec, (size_t)n * per_vertex, fname);
        for (uint32_t i = 0; i < n; i++)
            for (int side = 0; side < sides; side++) {
                const uint32_t *r = rec + (size_t)i * per_vertex + 3 * side;
                vs[side]->out_degree[b + i] = r[0];
                vs[side]->in_degree[b + i] = r[1];
                vs[side]->edges_idx[b + i] = r[2];
            }
    }
    for (uint32_t b = 0; b < E; b += CSR_IO_CHUNK) {
        const uint32_t n = E - b < CSR_IO_CHUNK ? E - b : CSR_IO_CHUNK;
        read_all(f, rec, (size_t)n * per_edge, fname);
        for (uint32_t i = 0; i < n; i++) {
            const uint32_t *r = rec + (size_t)i * per_edge;
            for (int side = 0; side < sides; side++) {
                struct EdgeList *e = es[side];
                const uint32_t *q = r + side * per_side;
                e->edges_array_src[b + i] = q[0];
                e->edges_array_dest[b + i] = q[1];
#if GBX_WEIGHTED
                memcpy(&e->edges_array_weight[b + i], &q[2], 4);
#endif
            }
        }
    }
    free(rec);
    fclose(f);
    printf(" -----------------------------------------------------\n");
    printf("| %-51s | \n", "GraphCSR (read from .csr)");
    printf(" -----------------------------------------------------\n");
    printf("| %-30s | %-18u | \n", "Number of Vertices (V)", N);
    printf("| %-30s | %-18u | \n", "Number of Edges (E)", E);
    printf(" -----------------------------------------------------\n");
    return csr;
}

static FILE *open_text(const char *fname, const char *ext, char **path)
{
    *path = path_with(fname, ext);
    FILE *f = fopen(*path, "w");
    if (!f) die_io("open", *path);
    return f;
}

/* writetoTextFilesGraphCSR, graphCSR.c:529-635: the forward CSR as text, one
 * value per line: <fname>.out_degree (preceded by num_vertices), .in_degree,
 * .edges_idx, .edges_array_src (preceded by num_edges), .edges_array_dest
 * [and .edges_array_weight, "%f", with WEIGHTED]. */
void writetoTextFilesGraphCSR(const char *fname, struct GraphCSR *graph)
{
    static const char *ext[] = {".out_degree", ".in_degree", ".edges_idx", ".edges_array_src",
                                ".edges_array_dest", ".edges_array_weight"};
    const int nfiles = GBX_WEIGHTED ? 6 : 5;
    FILE *f[6] = {NULL};
    char *path[6] = {NULL};
    for (int k = 0; k < nfiles; k++) f[k] = open_text(fname, ext[k], &path[k]);
    const struct Vertex *v = graph->vertices;
    const struct EdgeList *e = graph->sorted_edges_array;
    fprintf(f[0], "%u\n", graph->num_vertices);
    for (uint32_t i = 0; i < graph->num_vertices; i++) {
        fprintf(f[0], "%u\n", v->out_degree[i]);
        fprintf(f[1], "%u\n", v->in_degree[i]);
        fprintf(f[2], "%u\n", v->edges_idx[i]);
    }
    fprintf(f[3], "%u\n", graph->num_edges);
    for (uint32_t i = 0; i < graph->num_edges; i++) {
        fprintf(f[3], "%u\n", e->edges_array_src[i]);
        fprintf(f[4], "%u\n", e->edges_array_dest[i]);
        if (GBX_WEIGHTED) fprintf(f[5], "%f\n", e->edges_array_weight[i]);
    }
    for (int k = 0; k < nfiles; k++) {
        if (fclose(f[k]) != 0) die_io("close", path[k]);
        free(path[k]);
    }
}

/* ------------------------------------------------------------ PageRank */

/* newPageRankStatsGraphCSR / freePageRankStats, pageRank.c:47-88 */
struct PageRankStats *newPageRankStatsGraphCSR(struct GraphCSR *graph)
{
    struct PageRankStats *s = host_malloc(sizeof(*s));
    s->damp = 0.85f; /* Damp, pageRank.h:13 */
    s->base_pr = 1.0f - s->damp;
    s->iterations = 0;
    s->num_vertices = graph->num_vertices;
    s->activeVertices = 0;
    s->time_total = 0.0;
    s->error_total = 0.0;
    s->realRanks = host_malloc((size_t)graph->num_vertices * 4);
    s->pageRanks = host_malloc((size_t)graph->num_vertices * 4);
    for (uint32_t v = 0; v < graph->num_vertices; v++) {
        s->pageRanks[v] = s->base_pr;
        s->realRanks[v] = v;
    }
    return s;
}

void freePageRankStats(struct PageRankStats *stats)
{
    if (!stats) return;
    free(stats->realRanks);
    free(stats->pageRanks);
    free(stats);
}

/* One PageRank run on the device in a given mode; stats as the reference
 * fills them (pageRank.c:339-365). */
static struct PageRankStats *run_pagerank_dir(struct Arguments *arguments, struct GraphCSR *graph, int mode,
                                              int direction, const char *name)
{
    gbx_graph *g = device_graph(graph);
    if (!g) return NULL;
    struct PageRankStats *stats = newPageRankStatsGraphCSR(graph);
    gbx_pr_params p;
    gbx_pr_params_default(&p);
    p.damp = stats->damp;
    p.max_iters = arguments->iterations;
    p.epsilon = arguments->epsilon;
    p.mode = mode;
    p.stop = GBX_STOP_ACTIVE;
    p.precision = GBX_F32;
    p.direction = direction;
    gbx_pr_report rep;
    int rc = gbx_pr_pull(g_ctx, g, &p, NULL, stats->pageRanks, &rep);
    if (rc != GBX_OK) {
        fprintf(stderr, "graphblox_hip: %s failed: %s\n", name, gbx_last_error(g_ctx));
        freePageRankStats(stats);
        return NULL;
    }
    stats->iterations = rep.iterations;
    stats->activeVertices = rep.active;
    stats->error_total = rep.error_total;
    stats->time_total = rep.time_total_s;
    printf(" -----------------------------------------------------\n");
    printf("| %-51s | \n", name);
    printf(" -----------------------------------------------------\n");
    printf("| %-10s | %-8s | %-15s | %-9s | \n", "Iterations", "PR Sum", "Error", "Time (S)");
    printf(" -----------------------------------------------------\n");
    printf("| %-10u | %-8lf | %-15.13lf | %-9f | \n", stats->iterations, rep.rank_sum, stats->error_total,
           stats->time_total);
    printf(" -----------------------------------------------------\n");
    return stats;
}

static struct PageRankStats *run_pagerank(struct Arguments *arguments, struct GraphCSR *graph, int mode,
                                          const char *name)
{
    return run_pagerank_dir(arguments, graph, mode, PULL_DIRECTION, name);
}

static struct PageRankStats *run_pagerank_push(struct Arguments *arguments, struct GraphCSR *graph, int mode,
                                               const char *name)
{
    return run_pagerank_dir(arguments, graph, mode, PUSH_DIRECTION, name);
}

/* pageRankPullGraphCSR, pageRank.c:245-368 (FP32; transpose rows with
 * DIRECTED=1, forward rows with DIRECTED=0) */
struct PageRankStats *pageRankPullGraphCSR(struct Arguments *arguments, struct GraphCSR *graph)
{
    return run_pagerank(arguments, graph, GBX_PR_REF, "Page Rank Pull (HIP)");
}

/* pageRankPushGraphCSR, pageRank.c:370-492: each vertex pushes r/outdeg along
 * its out-edges, which gives every vertex the sum over its in-edges that the
 * pull iteration gathers; the device runs that iteration. */
struct PageRankStats *pageRankPushGraphCSR(struct Arguments *arguments, struct GraphCSR *graph)
{
    return run_pagerank_push(arguments, graph, GBX_PR_REF, "Page Rank Push (HIP, pull sums)");
}

/* pageRankPullFixedPoint64BitGraphCSR, pageRank.c:496-630 (bit-exact) */
struct PageRankStats *pageRankPullFixedPoint64BitGraphCSR(struct Arguments *arguments, struct GraphCSR *graph)
{
    return run_pagerank(arguments, graph, GBX_PR_FIXED64, "Page Rank Pull Fixed Point 64-bit (HIP)");
}

/* pageRankPushFixedPointGraphCSR, pageRank.c:1032-1157: the push form sums the
 * same 64-bit integers as pageRankPullFixedPoint64BitGraphCSR (bit-exact). */
struct PageRankStats *pageRankPushFixedPointGraphCSR(struct Arguments *arguments, struct GraphCSR *graph)
{
    return run_pagerank_push(arguments, graph, GBX_PR_FIXED64, "Page Rank Push Fixed Point (HIP, pull sums)");
}

/* pageRankPullFixedPoint32BitGraphCSR, pageRank.c:632-765 (bit-exact) */
struct PageRankStats *pageRankPullFixedPoint32BitGraphCSR(struct Arguments *arguments, struct GraphCSR *graph)
{
    return run_pagerank(arguments, graph, GBX_PR_FIXED32, "Page Rank Pull Fixed Point 32-bit (HIP)");
}

/* pageRankPullFixedPoint16BitGraphCSR, pageRank.c:767-897 (bit-exact) */
struct PageRankStats *pageRankPullFixedPoint16BitGraphCSR(struct Arguments *arguments, struct GraphCSR *graph)
{
    return run_pagerank(arguments, graph, GBX_PR_FIXED16, "Page Rank Pull Fixed Point 16-bit (HIP)");
}

/* pageRankPullFixedPoint8BitGraphCSR, pageRank.c:899-1030 (bit-exact) */
struct PageRankStats *pageRankPullFixedPoint8BitGraphCSR(struct Arguments *arguments, struct GraphCSR *graph)
{
    return run_pagerank(arguments, graph, GBX_PR_FIXED8, "Page Rank Pull Fixed Point 8-bit (HIP)");
}

/* pageRankPullQuant32BitGraphCSR, pageRank.c:1159-1297 (bit-exact) */
struct PageRankStats *pageRankPullQuant32BitGraphCSR(struct Arguments *arguments, struct GraphCSR *graph)
{
    return run_pagerank(arguments, graph, GBX_PR_QUANT32, "Page Rank Pull Quant_32 (HIP)");
}

/* pageRankPullQuant16BitGraphCSR, pageRank.c:1299-1436 (bit-exact) */
struct PageRankStats *pageRankPullQuant16BitGraphCSR(struct Arguments *arguments, struct GraphCSR *graph)
{
    return run_pagerank(arguments, graph, GBX_PR_QUANT16, "Page Rank Pull Quant_16 (HIP)");
}

/* pageRankPullQuant8BitGraphCSR, pageRank.c:1438-1577 (bit-exact) */
struct PageRankStats *pageRankPullQuant8BitGraphCSR(struct Arguments *arguments, struct GraphCSR *graph)
{
    return run_pagerank(arguments, graph, GBX_PR_QUANT8, "Page Rank Pull Quant_8 (HIP)");
}

/* pageRankPushQuantGraphCSR, pageRank.c:1579-1709: its float atomics add
 * scale * (q - zero) per edge, and (q - zero) mod 2^32 is the 32-bit pull
 * quantization of the same contribution, so it sums the terms of
 * pageRankPullQuant32BitGraphCSR in float; this runs the exact-sum form */
struct PageRankStats *pageRankPushQuantGraphCSR(struct Arguments *arguments, struct GraphCSR *graph)
{
    return run_pagerank_push(arguments, graph, GBX_PR_QUANT32, "Page Rank Push Quant_32 (HIP, pull sums)");
}

/* The data-driven variants take their worklist from the out-neighbours of
 * the transpose rows; with DIRECTED=0 the reference runs them over the
 * forward arrays for both (pageRank.c:1742-1748), which this backend does
 * not provide: NULL, the reference's "not available" (graphRun.c:948-957). */
static struct PageRankStats *run_data_driven(struct Arguments *arguments, struct GraphCSR *graph, int mode,
                                             const char *name)
{
    if (!GBX_DIRECTED) {
        printf(" -----------------------------------------------------\n");
        printf("| %-51s | \n", "data-driven PageRank: DIRECTED=1 builds only");
        printf(" -----------------------------------------------------\n");
        return NULL;
    }
    return run_pagerank(arguments, graph, mode, name);
}

/* pageRankDataDrivenPullGraphCSR, pageRank.c:1711-1870 */
struct PageRankStats *pageRankDataDrivenPullGraphCSR(struct Arguments *arguments, struct GraphCSR *graph)
{
    return run_data_driven(arguments, graph, GBX_PR_DATA_DRIVEN, "Page Rank Pull DD (HIP)");
}

/* pageRankDataDrivenPushGraphCSR, pageRank.c:1872-2027 (frontier iteration:
 * the same fixed point, not the same order of in-place updates) */
struct PageRankStats *pageRankDataDrivenPushGraphCSR(struct Arguments *arguments, struct GraphCSR *graph)
{
    return run_data_driven(arguments, graph, GBX_PR_DD_PUSH, "Page Rank Push DD (HIP)");
}

/* pageRankDataDrivenPullPushGraphCSR, pageRank.c:2029-2202 (frontier iteration) */
struct PageRankStats *pageRankDataDrivenPullPushGraphCSR(struct Arguments *arguments, struct GraphCSR *graph)
{
    return run_data_driven(arguments, graph, GBX_PR_DD_PULL_PUSH, "Page Rank Pull-Push DD (HIP)");
}

/* pageRankGraphCSR, pageRank.c:177-242: the variant selected by
 * arguments->pushpull; push quantised (5) runs the 32-bit quantized pull
 * (pageRankPushQuantGraphCSR adds the same scale * q terms with float
 * atomics). */
struct PageRankStats *pageRankGraphCSR(struct Arguments *arguments, struct GraphCSR *graph)
{
    switch (arguments->pushpull) {
    case 1: return pageRankPushGraphCSR(arguments, graph);
    case 2: return pageRankPullFixedPoint64BitGraphCSR(arguments, graph);
    case 3: return pageRankPushFixedPointGraphCSR(arguments, graph);
    case 9: return pageRankPullFixedPoint32BitGraphCSR(arguments, graph);
    case 10: return pageRankPullFixedPoint16BitGraphCSR(arguments, graph);
    case 11: return pageRankPullFixedPoint8BitGraphCSR(arguments, graph);
    case 4: return pageRankPullQuant32BitGraphCSR(arguments, graph);
    case 12: return pageRankPullQuant16BitGraphCSR(arguments, graph);
    case 13: return pageRankPullQuant8BitGraphCSR(arguments, graph);
    case 5: return pageRankPushQuantGraphCSR(arguments, graph);
    case 6: return pageRankDataDrivenPullGraphCSR(arguments, graph);
    case 7: return pageRankDataDrivenPushGraphCSR(arguments, graph);
    case 8: return pageRankDataDrivenPullPushGraphCSR(arguments, graph);
    default: return pageRankPullGraphCSR(arguments, graph);
    }
}

/* radixSortEdgesByPageRank, reorder.c:165-221: labels stably sorted by the
 * ascending fixed-point key of pageRanks, on the GPU; in place, returned. */
uint32_t *radixSortEdgesByPageRank(float *pageRanks, uint32_t *labels, uint32_t num_vertices)
{
    gbx_ctx *ctx = backend();
    if (!ctx) exit(EXIT_FAILURE);
    if (gbx_rank_sort(ctx, pageRanks, num_vertices, labels) != GBX_OK) {
        fprintf(stderr, "graphblox_hip: rank sort failed: %s\n", gbx_last_error(ctx));
        exit(EXIT_FAILURE);
    }
    return labels;
}

/* runPageRankAlgorithm, graphRun.c:933-975: the CSR data structures (0, 4 and
 * unknown codes) run pageRankGraphCSR; segments (1, 5) and 6 are not
 * implemented, as in the reference.  The ranking realRanks is then sorted by
 * rank. */
struct PageRankStats *runPageRankAlgorithm(struct Arguments *arguments, void *graph)
{
    struct PageRankStats *stats = NULL;
    switch (arguments->datastructure) {
    case 1: case 5: case 6:
        printf(" -----------------------------------------------------\n");
        printf("| %-51s | \n", "PageRank NOT YET IMPLEMENTED");
        printf(" -----------------------------------------------------\n");
        break;
    default: stats = pageRankGraphCSR(arguments, (struct GraphCSR *)graph); break;
    }
    if (stats) stats->realRanks = radixSortEdgesByPageRank(stats->pageRanks, stats->realRanks, stats->num_vertices);
    return stats;
}

/* ------------------------------------------------------------ SpMV */

/* newSPMVStatsGraphCSR / freeSPMVStats, SPMV.c:27-68 */
struct SPMVStats *newSPMVStatsGraphCSR(struct GraphCSR *graph)
{
    struct SPMVStats *s = host_malloc(sizeof(*s));
    s->iterations = 0;
    s->arguments = NULL;
    s->num_vertices = graph->num_vertices;
    s->time_total = 0.0;
    s->vector_output = host_malloc((size_t)graph->num_vertices * 4);
    s->vector_input = host_malloc((size_t)graph->num_vertices * 4);
    for (uint32_t v = 0; v < graph->num_vertices; v++) s->vector_output[v] = s->vector_input[v] = 0.0f;
    return s;
}

void freeSPMVStats(struct SPMVStats *stats)
{
    if (!stats) return;
    free(stats->vector_output);
    free(stats->vector_input);
    free(stats);
}

/* One SpMV variant on the device copy of graph: bits 0 = the float sums,
 * 32 / 64 = the fixed-point forms; rows in `direction`. */
static struct SPMVStats *run_spmv(struct Arguments *arguments, struct GraphCSR *graph, int bits, int direction,
                                  const char *name)
{
    gbx_graph *g = device_graph(graph);
    if (!g) return NULL;
    struct SPMVStats *stats = newSPMVStatsGraphCSR(graph);
    stats->arguments = arguments;
    double secs = 0.0;
    int rc = bits ? gbx_spmv_fixed(g_ctx, g, arguments->iterations, GBX_WEIGHTED, direction, bits,
                                   stats->vector_input, stats->vector_output, &secs)
                  : gbx_spmv_pull(g_ctx, g, arguments->iterations, GBX_WEIGHTED, direction, NULL,
                                  stats->vector_input, stats->vector_output, &secs);
    if (rc != GBX_OK) {
        fprintf(stderr, "graphblox_hip: %s failed: %s\n", name, gbx_last_error(g_ctx));
        freeSPMVStats(stats);
        return NULL;
    }
    stats->iterations = arguments->iterations;
    stats->time_total = secs;
    double sum = 0.0;
    for (uint32_t v = 0; v < graph->num_vertices; v++)
        sum += ((int)(stats->vector_output[v] * 100 + .5) / 100.0); /* SPMV.c:196-200 */
    printf(" -----------------------------------------------------\n");
    printf("| %-15s | %-15s | %-15s | \n", "Iterations", "Sum", "Time (S)");
    printf(" -----------------------------------------------------\n");
    printf("| %-15u | %-15lf | %-15f | \n", stats->iterations, sum, stats->time_total);
    printf(" -----------------------------------------------------\n");
    return stats;
}

/* SPMVPullGraphCSR, SPMV.c:117-215 (transpose rows and weights with
 * DIRECTED=1, forward ones with DIRECTED=0; weighted when GBX_WEIGHTED) */
struct SPMVStats *SPMVPullGraphCSR(struct Arguments *arguments, struct GraphCSR *graph)
{
    return run_spmv(arguments, graph, 0, PULL_DIRECTION, "SPMVPullGraphCSR");
}

/* SPMVPushGraphCSR, SPMV.c:216-311: the terms of the transpose rows' pull
 * (it pushes along the forward out-edges under either DIRECTED setting) */
struct SPMVStats *SPMVPushGraphCSR(struct Arguments *arguments, struct GraphCSR *graph)
{
    return run_spmv(arguments, graph, 0, PUSH_DIRECTION, "SPMVPushGraphCSR");
}

/* SPMVPullFixedPointGraphCSR, SPMV.c:313-449 (bit-exact) */
struct SPMVStats *SPMVPullFixedPointGraphCSR(struct Arguments *arguments, struct GraphCSR *graph)
{
    return run_spmv(arguments, graph, 32, PULL_DIRECTION, "SPMVPullFixedPointGraphCSR");
}

/* SPMVPushFixedPointGraphCSR, SPMV.c:451-590 (bit-exact) */
struct SPMVStats *SPMVPushFixedPointGraphCSR(struct Arguments *arguments, struct GraphCSR *graph)
{
    return run_spmv(arguments, graph, 64, PUSH_DIRECTION, "SPMVPushFixedPointGraphCSR");
}

/* SPMVGraphCSR, SPMV.c:89-116: the variant selected by arguments->pushpull */
struct SPMVStats *SPMVGraphCSR(struct Arguments *arguments, struct GraphCSR *graph)
{
    switch (arguments->pushpull) {
    case 1: return SPMVPushGraphCSR(arguments, graph);
    case 2: return SPMVPullFixedPointGraphCSR(arguments, graph);
    case 3: return SPMVPushFixedPointGraphCSR(arguments, graph);
    default: return SPMVPullGraphCSR(arguments, graph);
    }
}

/* runSPMVAlgorithm, graphRun.c:859-893: the CSR data structures (0, 4 and
 * unknown codes) run SPMVGraphCSR; segments (1, 5) and 6 are not
 * implemented, as in the reference. */
struct SPMVStats *runSPMVAlgorithm(struct Arguments *arguments, void *graph)
{
    switch (arguments->datastructure) {
    case 1: case 5: case 6:
        printf(" -----------------------------------------------------\n");
        printf("| %-51s | \n", "SPMV NOT YET IMPLEMENTED");
        printf(" -----------------------------------------------------\n");
        return NULL;
    default: return SPMVGraphCSR(arguments, (struct GraphCSR *)graph);
    }
}
