/* graphstats_hip.c -- the reference's PageRank top-k similarity statistics
 * (01_host/src/utils_graph/graphStats.c:36-1010): how far the ranking of
 * one PageRank run (stats) is from a reference run (ref_stats), over the
 * top-k vertices of the rank-sorted realRanks arrays (the last k entries:
 * runPageRankAlgorithm sorts them ascending by rank).
 *
 * Host code: k is at most 10000 (O(k^2) pair and edit-distance loops of a
 * reporting tool).  The definitions below follow the reference's, including
 * its rules where they differ from the textbook ones (each is noted). */
#define _GNU_SOURCE
#include "graphblox_hip.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static void *stats_malloc(size_t n)
{
    void *p = malloc(n ? n : 1);
    if (!p) {
        fprintf(stderr, "graphblox_hip: stats: out of memory\n");
        exit(EXIT_FAILURE);
    }
    return p;
}

/* The minimum of three edit costs as graphStats.c:36-47 picks it: a if it is
 * strictly below both others, else b if strictly below both, else c -- so a
 * tie between a and b below c yields c. */
static int pick3(int a, int b, int c)
{
    if (a < b && a < c) return a;
    if (b < a && b < c) return b;
    return c;
}

/* levenshtein_distance (graphStats.c:49-190): edit distance between two id
 * sequences, unit costs, built column by column with pick3 as the minimum;
 * two columns of the table suffice for the distance. */
static uint32_t edit_distance(const uint32_t *a, uint32_t la, const uint32_t *b, uint32_t lb)
{
    if (la == 0) return lb;
    if (lb == 0) return la;
    uint32_t *prev = stats_malloc(((size_t)la + 1) * 4), *cur = stats_malloc(((size_t)la + 1) * 4);
    for (uint32_t i = 0; i <= la; i++) prev[i] = i;
    for (uint32_t j = 1; j <= lb; j++) {
        cur[0] = j;
        for (uint32_t i = 1; i <= la; i++) {
            const uint32_t del = cur[i - 1] + 1, ins = prev[i] + 1;
            const uint32_t sub = prev[i - 1] + (a[i - 1] == b[j - 1] ? 0u : 1u);
            cur[i] = (uint32_t)pick3((int)del, (int)ins, (int)sub);
        }
        uint32_t *t = prev;
        prev = cur;
        cur = t;
    }
    const uint32_t d = prev[la];
    free(prev);
    free(cur);
    return d;
}

/* kendallSmallN (graphStats.c:419-480): Kendall's tau-b over all pairs,
 * concordant minus discordant over sqrt(pairs not tied in x) * sqrt(pairs
 * not tied in y), in float; ties counted as the reference counts them. */
static float kendall_pairs(const float *x, const float *y, int n)
{
    int s = 0, tx = 0, ty = 0;
    for (int i = 0; i < n; i++) {
        for (int j = i + 1; j < n; j++) {
            const int dy = (y[i] > y[j]) - (y[i] < y[j]);
            const int dx = (x[i] > x[j]) - (x[i] < x[j]);
            if (dy != 0) {
                if (dx != 0) s += dx * dy;
                else tx++;
            } else {
                ty++;
                if (x[i] == x[j]) tx++;
            }
        }
    }
    const int pairs = n * (n - 1) / 2;
    if (tx < pairs && ty < pairs) return s / (sqrtf((float)(pairs - tx)) * sqrtf((float)(pairs - ty)));
    return 0.0f;
}

/* intersection_topK (graphStats.c:571-587): how many of the reference's top-k
 * vertices (the last k of ref_total) are also in this run's top k (their
 * position in this run, run_pos, is at least n - k). */
static uint32_t topk_intersection(const uint32_t *ref_total, const uint32_t *run_pos, uint32_t n, uint32_t k)
{
    if (k > n) k = n;
    uint32_t hits = 0;
    for (uint32_t v = n - k; v < n; v++) hits += run_pos[ref_total[v]] >= n - k;
    return hits;
}

/* avg_mismatch_ranks_real_topK (graphStats.c:510-527): vertices of the
 * reference's top k whose rank position differs between the runs. */
static uint32_t topk_mismatch(const uint32_t *ref_pos, const uint32_t *run_pos, const uint32_t *ref_order, uint32_t n,
                              uint32_t k)
{
    if (k > n) k = n;
    uint32_t m = 0;
    for (uint32_t v = n - k; v < n; v++) m += run_pos[ref_order[v]] != ref_pos[ref_order[v]];
    return m;
}

/* avg_error_ranks_real_topK (graphStats.c:530-546): mean over the reference's
 * top k of |position difference| / (reference position + 1).  The reference
 * takes the absolute value with the integer abs(); positions are integers
 * below 2^31, so this is the same value. */
static double topk_position_error(const uint32_t *ref_pos, const uint32_t *run_pos, const uint32_t *ref_order,
                                  uint32_t n, uint32_t k)
{
    if (k > n) k = n;
    double e = 0.0;
    for (uint32_t v = n - k; v < n; v++) {
        const uint32_t u = ref_order[v];
        e += fabs((double)run_pos[u] - (double)ref_pos[u]) / ((double)ref_pos[u] + 1);
    }
    return e / k;
}

/* avg_error_ranks_float_topK (graphStats.c:548-568): mean over the
 * reference's top k of the relative rank-value error, vertices with a
 * non-positive reference rank adding nothing (but counting in k). */
static double topk_value_error(const float *ref_rank, const float *run_rank, const uint32_t *ref_order, uint32_t n,
                               uint32_t k)
{
    if (k > n) k = n;
    double e = 0.0;
    for (uint32_t v = n - k; v < n; v++) {
        const uint32_t u = ref_order[v];
        if ((double)ref_rank[u] > 0.0) e += fabs((double)run_rank[u] - (double)ref_rank[u]) / (double)ref_rank[u];
    }
    return e / k;
}

static void print_topk_block(FILE *f, uint32_t k, const struct PageRankCorrelationStats *c, int precise)
{
    fprintf(f, "\n-----------------------------------------------------\n");
    fprintf(f, "topk:         %u \n", k);
    fprintf(f, "-----------------------------------------------------\n");
    fprintf(f, "levenshtein_distance: %u \n", c->levenshtein_distance);
    fprintf(f, "Rank float Kendall:   %lf\n", c->float_Kendall);
    fprintf(f, "Rank real  Kendall:   %lf\n", c->real_Kendall);
    fprintf(f, "intersection:         %u \n", c->intersection);
    fprintf(f, "mismatch:             %u \n", c->mismatch);
    if (precise) {
        fprintf(f, "avg_error_float:      %.22lf\n", c->avg_error_float);
        fprintf(f, "avg_error_relative:   %.22lf\n", c->avg_error_relative);
    } else {
        fprintf(f, "avg_error_float:      %lf\n", c->avg_error_float);
        fprintf(f, "avg_error_relative:   %lf\n", c->avg_error_relative);
    }
    fprintf(f, "-----------------------------------------------------\n");
}

/* collectStatsPageRank_topK, graphStats.c:589-691.  The top-k sequences are
 * the last topk entries of each run's realRanks.  As in the reference, the
 * reference run's values in the float Kendall tau are read at this run's
 * top-k vertices (ref_stats->pageRanks[stats->realRanks[v]]). */
struct PageRankCorrelationStats collectStatsPageRank_topK(struct PageRankStats *ref_stats, struct PageRankStats *stats,
                                                          uint32_t *ref_rankedVertices_total,
                                                          uint32_t *ref_rankedVertices_inverse,
                                                          uint32_t *rankedVertices_inverse, uint32_t topk,
                                                          uint32_t num_vertices, FILE *fptr, uint32_t verbose)
{
    if (topk > num_vertices) topk = num_vertices;
    const uint32_t first = num_vertices - topk;
    uint32_t *run_ids = stats_malloc((size_t)topk * 4), *ref_ids = stats_malloc((size_t)topk * 4);
    float *run_val = stats_malloc((size_t)topk * 4), *ref_val = stats_malloc((size_t)topk * 4);
    float *run_id_f = stats_malloc((size_t)topk * 4), *ref_id_f = stats_malloc((size_t)topk * 4);
    for (uint32_t u = 0; u < topk; u++) {
        const uint32_t v = first + u;
        run_ids[u] = stats->realRanks[v];
        ref_ids[u] = ref_stats->realRanks[v];
        run_val[u] = stats->pageRanks[stats->realRanks[v]];
        ref_val[u] = ref_stats->pageRanks[stats->realRanks[v]];
        run_id_f[u] = (float)stats->realRanks[v];
        ref_id_f[u] = (float)ref_stats->realRanks[v];
    }
    struct PageRankCorrelationStats c;
    memset(&c, 0, sizeof(c));
    c.levenshtein_distance = edit_distance(ref_ids, topk, run_ids, topk);
    c.float_Kendall = kendall_pairs(ref_val, run_val, (int)topk);
    c.real_Kendall = kendall_pairs(ref_id_f, run_id_f, (int)topk);
    c.intersection = topk_intersection(ref_rankedVertices_total, rankedVertices_inverse, num_vertices, topk);
    c.avg_error_float = topk_value_error(ref_stats->pageRanks, stats->pageRanks, ref_stats->realRanks, num_vertices,
                                         topk);
    c.avg_error_relative = topk_position_error(ref_rankedVertices_inverse, rankedVertices_inverse,
                                               ref_stats->realRanks, num_vertices, topk);
    c.mismatch = topk_mismatch(ref_rankedVertices_inverse, rankedVertices_inverse, ref_stats->realRanks, num_vertices,
                               topk);
    if (verbose > 0) {
        print_topk_block(stdout, topk, &c, 1);
        if (fptr) print_topk_block(fptr, topk, &c, 0);
    }
    free(run_ids);
    free(ref_ids);
    free(run_val);
    free(ref_val);
    free(run_id_f);
    free(ref_id_f);
    return c;
}

static void add_stats(struct PageRankCorrelationStats *sum, const struct PageRankCorrelationStats *c)
{
    sum->levenshtein_distance += c->levenshtein_distance;
    sum->float_Kendall += c->float_Kendall;
    sum->real_Kendall += c->real_Kendall;
    sum->intersection += c->intersection;
    sum->mismatch += c->mismatch;
    sum->avg_error_float += c->avg_error_float;
    sum->avg_error_relative += c->avg_error_relative;
}

static void print_table(FILE *f, const uint32_t *ks, const struct PageRankCorrelationStats *c, uint32_t n)
{
    static const char *rule = "----------------------------------------------------------------------------------"
                              "------------------------\n";
    fputs(rule, f);
    fprintf(f, "Top K                 ");
    for (uint32_t x = 0; x < n; x++) fprintf(f, "%-14u ", ks[x]);
    fprintf(f, "\n");
    fputs(rule, f);
    fprintf(f, "levenshtein_distance  ");
    for (uint32_t x = 0; x < n; x++) fprintf(f, "%-14u ", c[x].levenshtein_distance);
    fprintf(f, "\nRank float Kendall    ");
    for (uint32_t x = 0; x < n; x++) fprintf(f, "%-14lf ", c[x].float_Kendall);
    fprintf(f, "\nRank real  Kendall    ");
    for (uint32_t x = 0; x < n; x++) fprintf(f, "%-14lf ", c[x].real_Kendall);
    fprintf(f, "\nintersection          ");
    for (uint32_t x = 0; x < n; x++) fprintf(f, "%-14u ", c[x].intersection);
    fprintf(f, "\nmismatch              ");
    for (uint32_t x = 0; x < n; x++) fprintf(f, "%-14u ", c[x].mismatch);
    fprintf(f, "\navg_error_float       ");
    for (uint32_t x = 0; x < n; x++) fprintf(f, "%-14lf ", c[x].avg_error_float);
    fprintf(f, "\navg_error_relative    ");
    for (uint32_t x = 0; x < n; x++) fprintf(f, "%-14lf ", c[x].avg_error_relative);
    fprintf(f, "\n");
    fputs(rule, f);
}

static void print_avg(FILE *f, const char *head, const struct PageRankCorrelationStats *a,
                      const struct Arguments *arguments, const struct PageRankStats *stats)
{
    fprintf(f, "-----------------------------------------------------\n");
    fputs(head, f);
    fprintf(f, "-----------------------------------------------------\n");
    fprintf(f, "levenshtein_distance: %u \n", a->levenshtein_distance);
    fprintf(f, "Rank float Kendall:   %lf\n", a->float_Kendall);
    fprintf(f, "Rank real  Kendall:   %lf\n", a->real_Kendall);
    fprintf(f, "intersection:         %u \n", a->intersection);
    fprintf(f, "mismatch:             %u \n", a->mismatch);
    fprintf(f, "avg_error_float:      %lf\n", a->avg_error_float);
    fprintf(f, "avg_error_relative:   %lf\n", a->avg_error_relative);
    fprintf(f, "-----------------------------------------------------\n");
    fprintf(f, "-----------------------------------------------------\n");
    fprintf(f, "numThreads:           %u \n", arguments->pre_numThreads);
    fprintf(f, "Time (S):             %lf\n", stats->time_total);
    fprintf(f, "Iterations:           %u \n", stats->iterations);
    fprintf(f, "-----------------------------------------------------\n");
}

/* collectStatsPageRank, graphStats.c:691-1010: the top-k statistics for
 * k = 30, 100, 300, 1000, 5000, 10000 (those not above N), their averages
 * over bins of 1000 positions (verbosity > 1), and k = binSize, appended to
 * "<fnameb>_<algorithm>_<datastructure>_<trial>_<pushpull>.stats".  The
 * per-vertex rank listing the reference adds at verbosity > 2 is not
 * written. */
void collectStatsPageRank(struct Arguments *arguments, struct PageRankStats *ref_stats, struct PageRankStats *stats,
                          uint32_t trial)
{
    static const uint32_t ks[] = {30, 100, 300, 1000, 5000, 10000};
    const uint32_t N = ref_stats->num_vertices;
    uint32_t *run_pos = stats_malloc((size_t)N * 4 + 4), *ref_pos = stats_malloc((size_t)N * 4 + 4);
    uint32_t *ref_total = stats_malloc((size_t)N * 4 + 4);
    for (uint32_t v = 0; v < N; v++) {
        run_pos[stats->realRanks[v]] = v;
        ref_pos[ref_stats->realRanks[v]] = v;
        ref_total[v] = ref_stats->realRanks[v];
    }
    char *fname = stats_malloc(strlen(arguments->fnameb) + 64);
    sprintf(fname, "%s_%u_%u_%u_%u.stats", arguments->fnameb, arguments->algorithm, arguments->datastructure, trial,
            arguments->pushpull);
    FILE *fptr = fopen(fname, "a+");
    if (!fptr) {
        fprintf(stderr, "graphblox_hip: open: %s\n", fname);
        exit(1);
    }
    uint32_t topk = arguments->binSize < N ? arguments->binSize : N;
    struct PageRankCorrelationStats table[6];
    uint32_t nk = 0;
    while (nk < 6 && ks[nk] <= N) {
        table[nk] = collectStatsPageRank_topK(ref_stats, stats, ref_total, ref_pos, run_pos, ks[nk], N, fptr, 0);
        nk++;
    }
    if (arguments->verbosity > 0) {
        print_table(stdout, ks, table, nk);
        print_table(fptr, ks, table, nk);
    }
    uint32_t chunk_x = 1000, chunk_num = (N + chunk_x - 1) / chunk_x;
    struct PageRankCorrelationStats sum, avg;
    memset(&sum, 0, sizeof(sum));
    if (arguments->verbosity > 1) {
        if (chunk_num == 1) {
            chunk_x = N;
            const struct PageRankCorrelationStats c =
                collectStatsPageRank_topK(ref_stats, stats, ref_total, ref_pos, run_pos, chunk_x, N, fptr, 1);
            add_stats(&sum, &c);
        } else {
            for (uint32_t x = 0; x < chunk_num; x++) {
                const struct PageRankCorrelationStats c = collectStatsPageRank_topK(
                    ref_stats, stats, ref_total, ref_pos, run_pos, chunk_x, N - chunk_x * x, fptr, 0);
                add_stats(&sum, &c);
            }
        }
    }
    avg.levenshtein_distance = chunk_num ? sum.levenshtein_distance / chunk_num : 0;
    avg.float_Kendall = chunk_num ? sum.float_Kendall / chunk_num : 0.0f;
    avg.real_Kendall = chunk_num ? sum.real_Kendall / chunk_num : 0.0f;
    avg.intersection = chunk_num ? sum.intersection / chunk_num : 0;
    avg.mismatch = chunk_num ? sum.mismatch / chunk_num : 0;
    avg.avg_error_float = chunk_num ? sum.avg_error_float / chunk_num : 0.0;
    avg.avg_error_relative = chunk_num ? sum.avg_error_relative / chunk_num : 0.0;
    (void)collectStatsPageRank_topK(ref_stats, stats, ref_total, ref_pos, run_pos, topk, N, fptr, 1);
    char head[128];
    snprintf(head, sizeof(head), "Avg (Sum(bin)*n)/n:    (Sum(%u)*%u)/%u \n", chunk_x, chunk_num, chunk_num);
    print_avg(stdout, head, &avg, arguments, stats);
    if (arguments->verbosity > 0) {
        snprintf(head, sizeof(head), "Avg (Sum_n(bins))/n:    (Sum_%u(%u)/%u \n", chunk_num, chunk_x, chunk_num);
        print_avg(fptr, head, &avg, arguments, stats);
    }
    fclose(fptr);
    free(fname);
    free(run_pos);
    free(ref_pos);
    free(ref_total);
}
