/*
 * gbx_oracle.c -- CPU oracle for the pull PageRank / SpMV path.
 *
 * TEST INFRASTRUCTURE ONLY (see gbx_oracle.h): a clean-room restatement of
 * the GraphBlox reference algorithms used to check the HIP path and to time
 * the CPU baseline.  Never linked into the product library.
 */
#include "gbx_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define VERTEX_ID_MASK 0x3FFFFFFFu /* reorder.h:23,30 (EXTRACT_VALUE) */

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* One stable counting-sort pass keyed on key[] (countsort.c:28-140 with P=1;
 * the reference's P-thread version produces the same order because its scan
 * is vertex-major, thread-minor over contiguous chunks, countsort.c:93-104). */
static int stable_pass(uint32_t *key, uint32_t *other, float *w,
                       uint32_t num_edges, uint32_t num_vertices,
                       int move_weights, int key_is_src)
{
    uint32_t *count = calloc((size_t)num_vertices + 1, sizeof(uint32_t));
    uint32_t *k2 = malloc((size_t)num_edges * sizeof(uint32_t) + 4);
    uint32_t *o2 = malloc((size_t)num_edges * sizeof(uint32_t) + 4);
    float *w2 = move_weights ? malloc((size_t)num_edges * sizeof(float) + 4) : NULL;
    if (!count || !k2 || !o2 || (move_weights && !w2)) {
        free(count); free(k2); free(o2); free(w2);
        return -1;
    }
    (void)key_is_src;
    for (uint32_t i = 0; i < num_edges; i++)
        count[key[i]]++;
    uint32_t base = 0;
    for (uint32_t v = 0; v < num_vertices; v++) {
        uint32_t c = count[v];
        count[v] = base;
        base += c;
    }
    for (uint32_t i = 0; i < num_edges; i++) {
        uint32_t pos = count[key[i]]++;
        k2[pos] = key[i];
        o2[pos] = other[i];
        if (move_weights) w2[pos] = w[i];
    }
    memcpy(key, k2, (size_t)num_edges * 4);
    memcpy(other, o2, (size_t)num_edges * 4);
    if (move_weights) memcpy(w, w2, (size_t)num_edges * 4);
    free(count); free(k2); free(o2); free(w2);
    return 0;
}

int orc_count_sort(uint32_t *src, uint32_t *dst, float *w, uint32_t num_edges,
                   uint32_t num_vertices, int mode, int move_weights)
{
    if (num_edges == 0) return 0;
    if (move_weights && !w) return -1;
    /* countSortEdgesBySourceAndDestination = by destination, then by source
     * (countsort.c:262-270). */
    if (mode == 1 || mode == 3) {
        if (stable_pass(dst, src, w, num_edges, num_vertices, move_weights, 0))
            return -1;
    }
    return stable_pass(src, dst, w, num_edges, num_vertices, move_weights, 1);
}

int orc_map_vertices(const uint32_t *sorted_src, uint32_t num_edges,
                     uint32_t num_vertices, uint32_t *out_degree,
                     uint32_t *edges_idx)
{
    memset(out_degree, 0, (size_t)num_vertices * 4);
    memset(edges_idx, 0, (size_t)num_vertices * 4);
    for (uint32_t i = 0; i < num_edges; i++) {
        uint32_t v = sorted_src[i] & VERTEX_ID_MASK;
        out_degree[v]++;
        if (i == 0 || sorted_src[i] != sorted_src[i - 1])
            edges_idx[v] = i;
    }
    return 0;
}

uint32_t orc_remove_duplicates(uint32_t *src, uint32_t *dst, float *w,
                               uint32_t num_edges)
{
    uint32_t i = 0, j = 0;
    /* edgeList.c:266-274 skips a leading run of self loops. */
    while (i < num_edges && src[i] == dst[i]) i++;
    if (i == num_edges) return 0;
    for (; i < num_edges; i++) {
        uint32_t s = src[i], d = dst[i];
        uint32_t wi = (uint32_t)w[i]; /* edgeList.c:261,270: uint32_t temp */
        if (s == d) continue;
        if (j > 0 && src[j - 1] == s && dst[j - 1] == d) continue;
        src[j] = s;
        dst[j] = d;
        w[j] = (float)wi;
        j++;
    }
    return j;
}

int orc_pr_pull_f32(uint32_t N, const uint32_t *fwd_out_degree,
                    const uint32_t *row_degree, const uint32_t *row_idx,
                    const uint32_t *row_col, uint32_t max_iters, double epsilon,
                    int nthreads, float *ranks_out, uint32_t *iters_out,
                    double *err_out, double *iter_seconds)
{
    /* pageRank.c:54-55,68: damp = 0.85f, base_pr = 1.0f - damp, r0 = base_pr */
    const float damp = 0.85f;
    const float base_pr = 1.0f - damp;
    float *r = ranks_out;
    float *next = malloc((size_t)N * sizeof(float) + 4);
    float *contrib = malloc((size_t)N * sizeof(float) + 4);
    if (!next || !contrib) { free(next); free(contrib); return -1; }
    if (nthreads < 1) nthreads = omp_get_max_threads(); /* 0: every OpenMP thread */
    double error_total = 0.0;
    uint32_t it;

#pragma omp parallel for num_threads(nthreads)
    for (uint32_t v = 0; v < N; v++) {
        r[v] = base_pr;
        next[v] = 0.0f;
    }

    for (it = 0; it < max_iters; it++) {
        double t0 = now_s();
        uint32_t active = 0;
        error_total = 0.0;
        /* phase 1: contribution r/outdeg (pageRank.c:295-302) */
#pragma omp parallel for num_threads(nthreads)
        for (uint32_t v = 0; v < N; v++)
            contrib[v] = fwd_out_degree[v] ? r[v] / (float)fwd_out_degree[v] : 0.0f;
        /* phase 2: sequential float gather per row (pageRank.c:304-317) */
#pragma omp parallel for schedule(dynamic, 1024) num_threads(nthreads)
        for (uint32_t v = 0; v < N; v++) {
            float acc = 0.0f;
            uint32_t beg = row_idx[v], end = beg + row_degree[v];
            for (uint32_t j = beg; j < end; j++)
                acc += contrib[row_col[j] & VERTEX_ID_MASK];
            next[v] = acc;
        }
        /* phase 3: update, error and active count (pageRank.c:320-334) */
#pragma omp parallel for num_threads(nthreads) reduction(+ : error_total, active)
        for (uint32_t v = 0; v < N; v++) {
            float prev = r[v];
            float nr = base_pr + damp * next[v];
            r[v] = nr;
            next[v] = 0.0f;
            double e = fabs(nr - prev);
            error_total += e / N;
            if (e >= epsilon) active++;
        }
        if (iter_seconds) iter_seconds[it] = now_s() - t0;
        if (active == 0) break; /* pageRank.c:338-339 */
    }
    /* pageRank.c:344-350 */
#pragma omp parallel for num_threads(nthreads)
    for (uint32_t v = 0; v < N; v++)
        r[v] = r[v] / N;
    if (iters_out) *iters_out = it;
    if (err_out) *err_out = error_total;
    free(next);
    free(contrib);
    return 0;
}

int orc_spmv_pull_f32(uint32_t N, const uint32_t *fwd_out_degree,
                      const uint32_t *row_degree, const uint32_t *row_idx,
                      const uint32_t *row_col, const float *w,
                      uint32_t iterations, int nthreads, float *x_out,
                      float *y_out, double *iter_seconds)
{
    if (nthreads < 1) nthreads = omp_get_max_threads(); /* 0: every OpenMP thread */
    /* SPMV.c:156-163 */
#pragma omp parallel for num_threads(nthreads)
    for (uint32_t v = 0; v < N; v++) {
        x_out[v] = fwd_out_degree[v] ? 1.0f / fwd_out_degree[v] : 0.001f;
        y_out[v] = 0.0f;
    }
    for (uint32_t it = 0; it < iterations; it++) {
        double t0 = now_s();
        /* SPMV.c:167-194: y accumulates across iterations, never reset */
#pragma omp parallel for schedule(dynamic, 1024) num_threads(nthreads)
        for (uint32_t v = 0; v < N; v++) {
            uint32_t beg = row_idx[v], end = beg + row_degree[v];
            for (uint32_t j = beg; j < end; j++) {
                float weight = w ? w[j] : 0.0001f;
                y_out[v] += weight * x_out[row_col[j] & VERTEX_ID_MASK];
            }
        }
        if (iter_seconds) iter_seconds[it] = now_s() - t0;
    }
    return 0;
}

int orc_pr_f64(uint32_t N, const uint32_t *outdeg, const uint64_t *row_ptr,
               const uint32_t *col, const double *w, const double *teleport,
               double damp, int mode, uint32_t max_iters, double epsilon,
               double *ranks_out, uint32_t *iters_out, double *err_out)
{
    /* OpenMP over vertices; every row sum is sequential in edge order, so a
     * vertex's value does not depend on the thread count (only the FP64
     * dangling-mass and error reductions are reassociated) */
    double *r = ranks_out;
    double *c = malloc((size_t)N * sizeof(double) + 8);
    double *nx = malloc((size_t)N * sizeof(double) + 8);
    if (!c || !nx) { free(c); free(nx); return -1; }
    const double base = 1.0 - damp;
    uint32_t it = 0;
    double err = 0.0;
#pragma omp parallel for schedule(static)
    for (uint32_t v = 0; v < N; v++)
        r[v] = mode == 0 ? base : (teleport ? teleport[v] : 1.0 / N);
    for (it = 0; it < max_iters; it++) {
        double dang = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : dang)
        for (uint32_t v = 0; v < N; v++) {
            if (outdeg[v]) {
                c[v] = w ? r[v] : r[v] / outdeg[v];
            } else {
                c[v] = 0.0;
                dang += r[v];
            }
        }
        uint32_t active = 0;
        err = 0.0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : err, active)
        for (uint32_t v = 0; v < N; v++) {
            double acc = 0.0;
            for (uint64_t j = row_ptr[v]; j < row_ptr[v + 1]; j++)
                acc += w ? w[j] * c[col[j]] : c[col[j]];
            double nr;
            if (mode == 0)
                nr = base + damp * acc;
            else
                nr = base * (teleport ? teleport[v] : 1.0 / N) + damp * (acc + dang / N);
            double e = fabs(nr - r[v]);
            nx[v] = nr;
            if (mode == 0) {
                err += e / N;
                if (e >= epsilon) active++;
            } else {
                err += e;
            }
        }
        memcpy(r, nx, (size_t)N * sizeof(double));
        if (mode == 0 ? active == 0 : err < epsilon) {
            it++;
            break;
        }
    }
    if (mode == 0)
#pragma omp parallel for schedule(static)
        for (uint32_t v = 0; v < N; v++) r[v] /= N;
    if (iters_out) *iters_out = it;
    if (err_out) *err_out = err;
    free(c);
    free(nx);
    return 0;
}

/* Row sums S[u] of the raw out-weights and P[j] = wraw[j] / S[col[j]] for an
 * in-CSR (row v lists its sources u = col[j]). */
static int transition_f64(uint32_t N, const uint64_t *row_ptr, const uint32_t *col,
                          const double *wraw, double **S_out, double **P_out)
{
    const uint64_t E = row_ptr[N];
    double *S = calloc((size_t)N + 1, sizeof(double));
    double *P = malloc((size_t)E * sizeof(double) + 8);
    if (!S || !P) { free(S); free(P); return -1; }
    for (uint64_t j = 0; j < E; j++) S[col[j]] += wraw[j];
#pragma omp parallel for schedule(dynamic, 1024)
    for (uint32_t v = 0; v < N; v++)
        for (uint64_t j = row_ptr[v]; j < row_ptr[v + 1]; j++) P[j] = wraw[j] / S[col[j]];
    *S_out = S;
    *P_out = P;
    return 0;
}

/* Forward fixed point p = (1-d) t + d (P^T p + (delta.p)/N) by pull
 * iterations from p (in/out) until sum|dp| < tol or max_iters. */
static uint32_t forward_f64(uint32_t N, const uint64_t *row_ptr, const uint32_t *col,
                            const double *S, const double *P, const double *teleport,
                            double damp, uint32_t max_iters, double tol, double *p, double *tmp)
{
    uint32_t it = 0;
    while (it < max_iters) {
        double dang = 0.0, err = 0.0;
#pragma omp parallel for reduction(+ : dang)
        for (uint32_t u = 0; u < N; u++)
            if (S[u] == 0.0) dang += p[u];
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : err)
        for (uint32_t v = 0; v < N; v++) {
            double acc = 0.0;
            for (uint64_t j = row_ptr[v]; j < row_ptr[v + 1]; j++) acc += P[j] * p[col[j]];
            tmp[v] = (1.0 - damp) * (teleport ? teleport[v] : 1.0 / N) + damp * (acc + dang / N);
            err += fabs(tmp[v] - p[v]);
        }
        memcpy(p, tmp, (size_t)N * sizeof(double));
        it++;
        if (err < tol) break;
    }
    return it;
}

int orc_pr_adjoint_f64(uint32_t N, const uint64_t *row_ptr, const uint32_t *col,
                       const double *wraw, const double *teleport,
                       const double *g, double damp, uint32_t max_iters,
                       double tol, double *p_out, double *lambda_out,
                       double *grad_w_out, double *grad_t_out)
{
    const uint64_t E = row_ptr[N];
    double *S = NULL, *P = NULL;
    if (transition_f64(N, row_ptr, col, wraw, &S, &P)) return -1;
    double *Plam = malloc((size_t)N * sizeof(double) + 8);
    double *tmp = malloc((size_t)N * sizeof(double) + 8);
    uint64_t *fptr = calloc((size_t)N + 2, sizeof(uint64_t));
    uint32_t *fcol = malloc((size_t)E * sizeof(uint32_t) + 4);
    double *fval = malloc((size_t)E * sizeof(double) + 8);
    if (!Plam || !tmp || !fptr || !fcol || !fval) {
        free(S); free(P); free(Plam); free(tmp); free(fptr); free(fcol); free(fval);
        return -1;
    }
    /* forward fixed point */
    for (uint32_t v = 0; v < N; v++) p_out[v] = teleport ? teleport[v] : 1.0 / N;
    forward_f64(N, row_ptr, col, S, P, teleport, damp, max_iters, tol, p_out, tmp);
    /* the forward CSR (row u lists P[u, v]) for a pull form of P lambda */
    for (uint64_t j = 0; j < E; j++) fptr[col[j] + 2]++;
    for (uint32_t u = 0; u < N; u++) fptr[u + 2] += fptr[u + 1];
    for (uint32_t v = 0; v < N; v++)
        for (uint64_t j = row_ptr[v]; j < row_ptr[v + 1]; j++) {
            const uint64_t k = fptr[col[j] + 1]++;
            fcol[k] = v;
            fval[k] = P[j];
        }
    /* adjoint fixed point lambda = g + d (P lambda + delta (1.lambda)/N),
     * stopped at sum|d lambda| < tol * sum|lambda| */
    for (uint32_t v = 0; v < N; v++) lambda_out[v] = g[v];
    for (uint32_t it = 0; it < max_iters; it++) {
        double sum = 0.0, err = 0.0, mag = 0.0;
#pragma omp parallel for reduction(+ : sum)
        for (uint32_t v = 0; v < N; v++) sum += lambda_out[v];
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : err, mag)
        for (uint32_t u = 0; u < N; u++) {
            double acc = 0.0;
            for (uint64_t k = fptr[u]; k < fptr[u + 1]; k++) acc += fval[k] * lambda_out[fcol[k]];
            const double nl = g[u] + damp * (acc + (S[u] == 0.0 ? sum / N : 0.0));
            err += fabs(nl - lambda_out[u]);
            mag += fabs(nl);
            tmp[u] = nl;
        }
        memcpy(lambda_out, tmp, (size_t)N * sizeof(double));
        if (err < tol * mag) break;
    }
    /* (P lambda)[u] for the weight gradient */
#pragma omp parallel for schedule(dynamic, 1024)
    for (uint32_t u = 0; u < N; u++) {
        double acc = 0.0;
        for (uint64_t k = fptr[u]; k < fptr[u + 1]; k++) acc += fval[k] * lambda_out[fcol[k]];
        Plam[u] = acc;
    }
    if (grad_w_out)
#pragma omp parallel for schedule(dynamic, 1024)
        for (uint32_t v = 0; v < N; v++)
            for (uint64_t j = row_ptr[v]; j < row_ptr[v + 1]; j++) {
                const uint32_t u = col[j];
                grad_w_out[j] = damp * p_out[u] * (lambda_out[v] - Plam[u]) / S[u];
            }
    if (grad_t_out)
        for (uint32_t v = 0; v < N; v++) grad_t_out[v] = (1.0 - damp) * lambda_out[v];
    free(S); free(P); free(Plam); free(tmp); free(fptr); free(fcol); free(fval);
    return 0;
}

/* The forward fixed point in x87 extended precision (64-bit significand)
 * for raw weights w, from p0, for exactly `iters` pull iterations. */
static int forward_ld(uint32_t N, const uint64_t *row_ptr, const uint32_t *col, const double *w,
                      const double *teleport, double damp, const double *p0, uint32_t iters,
                      long double *p)
{
    const uint64_t E = row_ptr[N];
    long double *S = calloc((size_t)N + 1, sizeof(long double));
    double *P = malloc((size_t)E * sizeof(double) + 8);
    long double *tmp = malloc((size_t)N * sizeof(long double) + 16);
    if (!S || !P || !tmp) { free(S); free(P); free(tmp); return -1; }
    for (uint64_t j = 0; j < E; j++) S[col[j]] += w[j];
#pragma omp parallel for schedule(dynamic, 1024)
    for (uint32_t v = 0; v < N; v++)
        for (uint64_t j = row_ptr[v]; j < row_ptr[v + 1]; j++) P[j] = (double)((long double)w[j] / S[col[j]]);
    const long double d = damp, base = 1.0L - (long double)damp;
    for (uint32_t v = 0; v < N; v++) p[v] = p0 ? p0[v] : (teleport ? teleport[v] : 1.0L / N);
    for (uint32_t it = 0; it < iters; it++) {
        long double dang = 0.0L;
        for (uint32_t u = 0; u < N; u++)
            if (S[u] == 0.0L) dang += p[u];
#pragma omp parallel for schedule(dynamic, 1024)
        for (uint32_t v = 0; v < N; v++) {
            long double acc = 0.0L;
            for (uint64_t j = row_ptr[v]; j < row_ptr[v + 1]; j++) acc += (long double)P[j] * p[col[j]];
            tmp[v] = base * (teleport ? (long double)teleport[v] : 1.0L / N) + d * (acc + dang / N);
        }
        memcpy(p, tmp, (size_t)N * sizeof(long double));
    }
    free(S); free(P); free(tmp);
    return 0;
}

double orc_pr_loss_diff(uint32_t N, const uint64_t *row_ptr, const uint32_t *col,
                        const double *w_plus, const double *w_minus, const double *teleport,
                        const double *g, double damp, const double *p0, uint32_t iters)
{
    long double *pp = malloc((size_t)N * sizeof(long double) + 16);
    long double *pm = malloc((size_t)N * sizeof(long double) + 16);
    if (!pp || !pm || forward_ld(N, row_ptr, col, w_plus, teleport, damp, p0, iters, pp) ||
        forward_ld(N, row_ptr, col, w_minus, teleport, damp, p0, iters, pm)) {
        free(pp); free(pm);
        return NAN;
    }
    long double dl = 0.0L;
    for (uint32_t v = 0; v < N; v++) dl += (long double)g[v] * (pp[v] - pm[v]);
    free(pp); free(pm);
    return (double)dl;
}

double orc_pr_loss_f64(uint32_t N, const uint64_t *row_ptr, const uint32_t *col,
                       const double *wraw, const double *teleport, const double *g,
                       double damp, const double *p0, uint32_t max_iters, double tol,
                       double *p_out)
{
    double *S = NULL, *P = NULL;
    if (transition_f64(N, row_ptr, col, wraw, &S, &P)) return NAN;
    double *tmp = malloc((size_t)N * sizeof(double) + 8);
    if (!tmp) { free(S); free(P); return NAN; }
    for (uint32_t v = 0; v < N; v++) p_out[v] = p0 ? p0[v] : (teleport ? teleport[v] : 1.0 / N);
    forward_f64(N, row_ptr, col, S, P, teleport, damp, max_iters, tol, p_out, tmp);
    double L = 0.0;
    for (uint32_t v = 0; v < N; v++) L += g[v] * p_out[v];
    free(S); free(P); free(tmp);
    return L;
}

/* ---- fixed-point pull PageRank (pageRank.c:496-1030, fixedPoint.h)
 *
 * Per iteration: c[u] = conv_bits(r[u] / outdeg[u]) (0 for outdeg 0), an
 * exact uint64 sum of c over the in-edges of each row, then
 * r' = base + d * back_bits(sum).  The conversions follow fixedPoint.h:
 *   64: DoubleToFixed64 = (u64)((double)x * 2^32 + 0.5), back = (double)s / 2^32,
 *       update in double then rounded to float (pageRank.c:577);
 *   32: FloatToFixed32 = (u32)(x * 65536.0f + 0.5f), back = (float)s / 65536.0f;
 *   16: FloatToFixed16 = (u16)(x * 16384.0f + 0.5f), back = (float)s / 16384.0f;
 *    8: FloatToFixed8  = (u8)(x * 64.0f + 0.5f),     back = (float)s / 64.0f;
 * float to u16/u8 goes through a 32-bit signed truncation and keeps the low
 * bits, as x86-64 code compiled from the reference does; the push form of
 * the 64-bit variant (pageRank.c:1032-1157) sums the same integers. */
static uint64_t fixed_conv(float x, int bits)
{
    switch (bits) {
    case 64: return (uint64_t)((double)x * 4294967296.0 + 0.5);
    case 32: return (uint64_t)(uint32_t)(int64_t)(x * 65536.0f + 0.5f);
    case 16: return (uint64_t)(uint16_t)(int32_t)(x * 16384.0f + 0.5f);
    default: return (uint64_t)(uint8_t)(int32_t)(x * 64.0f + 0.5f);
    }
}

static float fixed_update(float base, float damp, uint64_t s, int bits)
{
    switch (bits) {
    case 64: return (float)((double)base + (double)damp * ((double)s / 4294967296.0));
    case 32: return base + damp * ((float)s / 65536.0f);
    case 16: return base + damp * ((float)s / 16384.0f);
    default: return base + damp * ((float)s / 64.0f);
    }
}

int orc_pr_pull_fixed(uint32_t N, const uint32_t *outdeg, const uint64_t *row_ptr,
                      const uint32_t *col, int bits, uint32_t max_iters, double epsilon,
                      float *ranks_out, uint32_t *iters_out, double *err_out)
{
    if (bits != 64 && bits != 32 && bits != 16 && bits != 8) return -1;
    uint64_t *c = malloc((size_t)N * sizeof(uint64_t) + 8);
    if (!c) return -1;
    const float damp = 0.85f, base = 1.0f - damp; /* pageRank.c:55 */
    float *r = ranks_out;
    for (uint32_t v = 0; v < N; v++) r[v] = base;
    uint32_t it = 0;
    double err = 0.0;
    for (it = 0; it < max_iters; it++) {
#pragma omp parallel for
        for (uint32_t v = 0; v < N; v++)
            c[v] = outdeg[v] ? fixed_conv(r[v] / (float)outdeg[v], bits) : 0;
        uint32_t active = 0;
        err = 0.0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : err, active)
        for (uint32_t v = 0; v < N; v++) {
            uint64_t s = 0;
            for (uint64_t j = row_ptr[v]; j < row_ptr[v + 1]; j++) s += c[col[j] & VERTEX_ID_MASK];
            const float prev = r[v];
            const float next = fixed_update(base, damp, s, bits);
            r[v] = next;
            const double e = fabs((double)next - (double)prev);
            err += e / N;
            if (e >= epsilon) active++;
        }
        if (active == 0) {
            it++;
            break;
        }
    }
    for (uint32_t v = 0; v < N; v++) r[v] /= N;
    if (iters_out) *iters_out = it < max_iters ? it - 1 : it;
    if (err_out) *err_out = err;
    free(c);
    return 0;
}

/* ---- quantized pull PageRank (pageRank.c:1159-1577, quantization.h)
 *
 * Per iteration: c[u] = r[u] / outdeg[u] in float (0 for outdeg 0); the
 * exact min and max of c over all vertices (getMinMax_k); the scale
 * GetScale_k = (max - min in double) / RANGE_MAX_k (1 when min == max),
 * kept in double for 32 bits and rounded to float for 16 and 8 bits; the
 * zero point 0; q[u] = quantize_k(c[u]) = ROUND(c / scale) clamped to the
 * width; an exact uint64 sum of q over each row's in-edges, and the update
 * r' = base + d * next with next = (float)(scale * (double)sum) (32 bits)
 * or scale * (float)sum in float (16 and 8 bits).
 *
 * ROUND(x) is (uint32_t)(x + 0.5): x86-64 code converts a double to a u32
 * through a 64-bit signed truncation and keeps the low 32 bits, so a value
 * of 2^32 or more (c / scale exceeds RANGE_MAX_k when min > 0) wraps, and
 * one of 2^63 or more (or NaN) gives the indefinite integer, low bits 0;
 * CLAMP_16/CLAMP_8 then saturate the wrapped u32 at the width's maximum. */
static double quant_scale(float mn, float mx, int bits)
{
    if (mn == mx) return 1.0;
    const double d = fabs((double)mx - (double)mn);
    switch (bits) {
    case 32: return d / 4294967295.0;
    case 16: return (double)(float)(d / 65535.0);
    default: return (double)(float)(d / 255.0);
    }
}

static uint32_t quant_round(double x)
{
    const double y = x + 0.5;
    if (!(y < 9223372036854775808.0)) return 0u;
    return (uint32_t)(uint64_t)(int64_t)y;
}

static uint64_t quant_value(float c, double scale, int bits)
{
    const uint32_t q = quant_round((double)c / scale);
    switch (bits) {
    case 32: return q;
    case 16: return q > 65535u ? 65535u : q;
    default: return q > 255u ? 255u : q;
    }
}

static float quant_update(float base, float damp, double scale, uint64_t s, int bits)
{
    float next;
    if (bits == 32) next = (float)(scale * (double)s);
    else next = (float)scale * (float)s;
    return base + damp * next;
}

int orc_pr_pull_quant(uint32_t N, const uint32_t *outdeg, const uint64_t *row_ptr,
                      const uint32_t *col, int bits, uint32_t max_iters, double epsilon,
                      float *ranks_out, uint32_t *iters_out, double *err_out)
{
    if ((bits != 32 && bits != 16 && bits != 8) || N == 0) return -1;
    float *c = malloc((size_t)N * sizeof(float) + 8);
    uint64_t *q = malloc((size_t)N * sizeof(uint64_t) + 8);
    if (!c || !q) { free(c); free(q); return -1; }
    const float damp = 0.85f, base = 1.0f - damp; /* pageRank.c:55 */
    float *r = ranks_out;
    for (uint32_t v = 0; v < N; v++) r[v] = base;
    uint32_t it = 0;
    double err = 0.0;
    for (it = 0; it < max_iters; it++) {
        float mn = INFINITY, mx = -INFINITY;
        for (uint32_t v = 0; v < N; v++) {
            c[v] = outdeg[v] ? r[v] / (float)outdeg[v] : 0.0f;
            if (c[v] < mn) mn = c[v];
            if (c[v] > mx) mx = c[v];
        }
        const double scale = quant_scale(mn, mx, bits);
#pragma omp parallel for
        for (uint32_t v = 0; v < N; v++) q[v] = quant_value(c[v], scale, bits);
        uint32_t active = 0;
        err = 0.0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : err, active)
        for (uint32_t v = 0; v < N; v++) {
            uint64_t s = 0;
            for (uint64_t j = row_ptr[v]; j < row_ptr[v + 1]; j++) s += q[col[j] & VERTEX_ID_MASK];
            const float prev = r[v];
            const float next = quant_update(base, damp, scale, s, bits);
            r[v] = next;
            const double e = fabs((double)next - (double)prev);
            err += e / N;
            if (e >= epsilon) active++;
        }
        if (active == 0) {
            it++;
            break;
        }
    }
    for (uint32_t v = 0; v < N; v++) r[v] /= N;
    if (iters_out) *iters_out = it < max_iters ? it - 1 : it;
    if (err_out) *err_out = err;
    free(c);
    free(q);
    return 0;
}

/* ---- data-driven pull PageRank (pageRank.c:1711-1870)
 *
 * Every vertex is on the worklist in the first iteration.  Per iteration:
 * c = r / outdeg snapshot (float, 0 for outdeg 0); for each worklist row, a
 * sequential float sum of c over its in-edges, r' = base + d * sum,
 * error = |r' - r| (float difference) added as error / N; only when
 * error >= epsilon is r' stored, the row counted active and its
 * out-neighbours put on the next worklist.  Stop when no row is active. */
int orc_pr_pull_dd(uint32_t N, const uint32_t *outdeg, const uint64_t *row_ptr,
                   const uint32_t *col, const uint64_t *fwd_row_ptr, const uint32_t *fwd_col,
                   uint32_t max_iters, double epsilon, float *ranks_out, uint32_t *iters_out,
                   double *err_out)
{
    float *c = malloc((size_t)N * sizeof(float) + 8);
    uint8_t *cur = malloc((size_t)N + 8), *nxt = malloc((size_t)N + 8);
    if (!c || !cur || !nxt) { free(c); free(cur); free(nxt); return -1; }
    const float damp = 0.85f, base = 1.0f - damp; /* pageRank.c:55 */
    float *r = ranks_out;
    for (uint32_t v = 0; v < N; v++) { r[v] = base; cur[v] = 1; nxt[v] = 0; }
    uint32_t it = 0;
    double err = 0.0;
    for (it = 0; it < max_iters; it++) {
        for (uint32_t v = 0; v < N; v++) c[v] = outdeg[v] ? r[v] / (float)outdeg[v] : 0.0f;
        uint32_t active = 0;
        err = 0.0;
        for (uint32_t v = 0; v < N; v++) {
            if (!cur[v]) continue;
            float s = 0.0f;
            for (uint64_t j = row_ptr[v]; j < row_ptr[v + 1]; j++) s += c[col[j] & VERTEX_ID_MASK];
            const float next = base + damp * s;
            const double e = fabs(next - r[v]);
            err += e / N;
            if (e >= epsilon) {
                r[v] = next;
                for (uint64_t j = fwd_row_ptr[v]; j < fwd_row_ptr[v + 1]; j++) nxt[fwd_col[j] & VERTEX_ID_MASK] = 1;
                active++;
            }
        }
        uint8_t *t = cur; cur = nxt; nxt = t;
        memset(nxt, 0, N);
        if (active == 0) {
            it++;
            break;
        }
    }
    for (uint32_t v = 0; v < N; v++) r[v] /= N;
    if (iters_out) *iters_out = it < max_iters ? it - 1 : it;
    if (err_out) *err_out = err;
    free(c); free(cur); free(nxt);
    return 0;
}

/* ---- data-driven push and pull-push PageRank (pageRank.c:1872-2202)
 *
 * Sequential, in place, in vertex order (one OpenMP thread of the
 * reference).  Residuals start at (1-d) * d * sum over in-edges of
 * 1.0f / outdeg[u] (float, in-edge order); every vertex is on the first
 * worklist.  A worklist vertex v takes its new rank (push: res[v] + r[v];
 * pull-push: base + d * the float sum of r[u] / outdeg[u] over its in-edges,
 * reading ranks already updated in this sweep), adds |r'/N - r/N| to the
 * error, pushes delta = d * (res[v] / outdeg[v]) onto each out-neighbour's
 * residual and then zeroes its own (so a self loop's push is lost).  A push
 * that lifts a residual from <= epsilon to >= epsilon counts as active and
 * puts the neighbour on the next worklist; the pull-push form adds delta a
 * second time on that event.  Stop when nothing became active. */
int orc_pr_dd_push(uint32_t N, const uint32_t *outdeg, const uint64_t *row_ptr,
                   const uint32_t *col, const uint64_t *fwd_row_ptr, const uint32_t *fwd_col,
                   int pull_push, uint32_t max_iters, double epsilon, float *ranks_out,
                   uint32_t *iters_out, double *err_out)
{
    float *res = malloc((size_t)N * sizeof(float) + 8);
    uint8_t *cur = malloc((size_t)N + 8), *nxt = malloc((size_t)N + 8);
    if (!res || !cur || !nxt) { free(res); free(cur); free(nxt); return -1; }
    const float damp = 0.85f, base = 1.0f - damp; /* pageRank.c:55 */
    float *r = ranks_out;
    for (uint32_t v = 0; v < N; v++) {
        r[v] = base;
        cur[v] = 1;
        nxt[v] = 0;
        float s = 0.0f;
        for (uint64_t j = row_ptr[v]; j < row_ptr[v + 1]; j++) {
            const uint32_t u = col[j] & VERTEX_ID_MASK;
            if (outdeg[u]) s += 1.0f / (float)outdeg[u];
        }
        res[v] = (1.0f - damp) * damp * s;
    }
    uint32_t it = 0;
    double err = 0.0;
    for (it = 0; it < max_iters; it++) {
        uint32_t active = 0;
        err = 0.0;
        for (uint32_t v = 0; v < N; v++) {
            if (!cur[v]) continue;
            const float old = r[v];
            float next;
            if (pull_push) {
                float s = 0.0f;
                for (uint64_t j = row_ptr[v]; j < row_ptr[v + 1]; j++) {
                    const uint32_t u = col[j] & VERTEX_ID_MASK;
                    s += r[u] / (float)outdeg[u];
                }
                next = base + damp * s;
            } else {
                next = res[v] + r[v];
            }
            err += fabs(next / (float)N - old / (float)N);
            r[v] = next;
            const float delta = damp * (res[v] / (float)outdeg[v]);
            for (uint64_t j = fwd_row_ptr[v]; j < fwd_row_ptr[v + 1]; j++) {
                const uint32_t u = fwd_col[j] & VERTEX_ID_MASK;
                const float prev = res[u];
                res[u] += delta;
                if (fabs(prev + delta) >= epsilon && prev <= epsilon) {
                    active++;
                    if (pull_push) res[u] += delta;
                    nxt[u] = 1;
                }
            }
            res[v] = 0.0f;
        }
        uint8_t *t = cur; cur = nxt; nxt = t;
        memset(nxt, 0, N);
        if (active == 0) {
            it++;
            break;
        }
    }
    for (uint32_t v = 0; v < N; v++) r[v] /= N;
    if (iters_out) *iters_out = it < max_iters ? it - 1 : it;
    if (err_out) *err_out = err;
    free(res); free(cur); free(nxt);
    return 0;
}

/* ---- R-MAT generator (counter-based; identical integer arithmetic on GPU) */

static uint64_t mix64(uint64_t z)
{
    /* SplitMix64 finaliser (Steele, Lea, Flood 2014). */
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint32_t orc_rmat_permute(uint32_t x, uint32_t scale, uint64_t seed)
{
    if (scale == 0) return 0;
    const uint32_t mask = scale >= 32 ? 0xFFFFFFFFu : ((1u << scale) - 1u);
    const uint32_t sh = scale / 2 + 1;
    for (uint32_t round = 0; round < 4; round++) {
        uint64_t k = mix64(seed ^ (0xA5A5A5A5ull + round * 0x9E3779B97F4A7C15ull));
        uint32_t mul = ((uint32_t)k | 1u);
        uint32_t add = (uint32_t)(k >> 32);
        x = (x * mul + add) & mask;  /* odd multiply + add: bijection mod 2^s */
        x ^= x >> sh;                /* xorshift-right: bijection on s bits */
    }
    return x & mask;
}

int orc_rmat_generate(uint32_t scale, uint32_t edge_factor, double a, double b,
                      double c, uint64_t seed, uint32_t *src, uint32_t *dst)
{
    const uint64_t M = (uint64_t)edge_factor << scale;
    const uint32_t ta = (uint32_t)(a * 4294967296.0);
    const uint32_t tab = (uint32_t)((a + b) * 4294967296.0);
    const uint32_t tabc = (uint32_t)((a + b + c) * 4294967296.0);
    const uint64_t key = mix64(seed);
#pragma omp parallel for schedule(static)
    for (uint64_t e = 0; e < M; e++) {
        uint32_t u = 0, v = 0;
        for (uint32_t l = 0; l < scale; l += 2) {
            uint64_t h = mix64(key + (e << 5) + (l >> 1) + 1);
            for (uint32_t q = 0; q < 2 && l + q < scale; q++) {
                uint32_t r = q ? (uint32_t)(h >> 32) : (uint32_t)h;
                uint32_t bu = r >= tab;                 /* quadrants c,d */
                uint32_t bv = (r >= ta && r < tab) || r >= tabc; /* b,d */
                u = (u << 1) | bu;
                v = (v << 1) | bv;
            }
        }
        src[e] = orc_rmat_permute(u, scale, seed);
        dst[e] = orc_rmat_permute(v, scale, seed);
    }
    return 0;
}

/* LSD radix sort of (64-bit key, 32-bit index) pairs, 16-bit digits. */
static int sort_key_index(uint64_t *key, uint32_t *idx, uint64_t n)
{
    uint64_t *k2 = malloc(n * sizeof(uint64_t) + 8);
    uint32_t *i2 = malloc(n * sizeof(uint32_t) + 8);
    uint64_t *cnt = malloc(65537 * sizeof(uint64_t));
    if (!k2 || !i2 || !cnt) { free(k2); free(i2); free(cnt); return -1; }
    for (int pass = 0; pass < 4; pass++) {
        int sh = 16 * pass;
        memset(cnt, 0, 65537 * sizeof(uint64_t));
        for (uint64_t i = 0; i < n; i++) cnt[((key[i] >> sh) & 0xFFFF) + 1]++;
        for (int d = 0; d < 65536; d++) cnt[d + 1] += cnt[d];
        for (uint64_t i = 0; i < n; i++) {
            uint64_t p = cnt[(key[i] >> sh) & 0xFFFF]++;
            k2[p] = key[i];
            i2[p] = idx[i];
        }
        memcpy(key, k2, n * sizeof(uint64_t));
        memcpy(idx, i2, n * sizeof(uint32_t));
    }
    free(k2); free(i2); free(cnt);
    return 0;
}

uint64_t orc_dedupe_keep_first(uint32_t *src, uint32_t *dst, float *w,
                               uint64_t num_edges)
{
    uint64_t n = num_edges;
    uint64_t *key = malloc(n * sizeof(uint64_t) + 8);
    uint32_t *idx = malloc(n * sizeof(uint32_t) + 8);
    uint8_t *keep = calloc(n + 1, 1);
    if (!key || !idx || !keep) { free(key); free(idx); free(keep); return (uint64_t)-1; }
    for (uint64_t i = 0; i < n; i++) {
        key[i] = ((uint64_t)src[i] << 32) | dst[i];
        idx[i] = (uint32_t)i;
    }
    if (sort_key_index(key, idx, n)) { free(key); free(idx); free(keep); return (uint64_t)-1; }
    for (uint64_t i = 0; i < n; i++) {
        int first = (i == 0) || key[i] != key[i - 1];
        int loop = (uint32_t)(key[i] >> 32) == (uint32_t)key[i];
        if (first && !loop) keep[idx[i]] = 1; /* stable sort: first = lowest index */
    }
    uint64_t j = 0;
    for (uint64_t i = 0; i < n; i++) {
        if (!keep[i]) continue;
        src[j] = src[i];
        dst[j] = dst[i];
        if (w) w[j] = w[i];
        j++;
    }
    free(key); free(idx); free(keep);
    return j;
}
