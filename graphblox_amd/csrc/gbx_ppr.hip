// gbx_ppr.hip -- batched personalised PageRank: k teleport vectors advanced
// together by one multi-vector pull launch per iteration.
//
// No reference counterpart; the single-vector semantics are those of the
// stochastic form in gbx_pr.hip (uniform dangling redistribution, FP64 mass
// and residual reductions), applied to each vector independently:
//   r_q' = (1-d) t_q + d (P^T r_q + D_q/N) / S_q
// with t_q uniform over seed set q, D_q the dangling mass and S_q the rank
// sum of vector q.  Ranks and contributions are stored vertex-major,
// [vertex][V] with V = k rounded up to a power of two, so one gather of a
// source vertex reads all k contributions from one or two cache lines.
// A team of V/VPL lanes (VPL vectors a lane) walks one row, its gathers
// issued a batch at a time with the next batch's columns, and the next row's
// pointers, first columns and update inputs loaded during the current row;
// hub rows are split into edge chunks whose per-team partial rows are
// reduced in LDS and combined in chunk order by the last chunk to finish.  Per-vector FP64 block partials
// are reduced by a second small kernel that also evaluates the stop rule.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "gbx_plan.hpp"

namespace gbx {

constexpr uint32_t kMaxVec = 64;

struct BScal {
    double err, dang, sum;
};

struct BState {
    uint32_t done;    // every vector met the stop rule
    uint32_t iters;   // iterations executed
    uint32_t last;    // scalar slot written by the last executed iteration
    uint32_t pad;
    uint32_t conv_iter[kMaxVec];  // iteration at which vector q first met the rule (0: not yet)
};

struct BPlan {
    const uint32_t *__restrict__ rp;
    const uint32_t *__restrict__ col;
    const uint32_t *__restrict__ outdeg;
    const HubChunk *__restrict__ chunks;
    uint32_t *__restrict__ hub_cnt;
    float *__restrict__ hub_part;  // [n_chunks][V]
    uint32_t N, n_chunks, row_begin, n_tiles;
};

struct BatchView {
    float *__restrict__ R;           // [N][V] ranks, plan order, updated in place
    const float *__restrict__ C_in;  // [N][V] contributions of the previous iteration
    float *__restrict__ C_out;
    const uint64_t *__restrict__ mask;  // [N] bit q: vertex in seed set q (nullable: uniform)
    const float *__restrict__ inv_size; // [V]
    const BScal *__restrict__ scal_in;  // [V]
    double *__restrict__ partials;      // [grid][3][V]
    BState *state;
    float damp;
};

}  // namespace gbx

struct gbx_ppr_state {
    uint32_t k = 0, V = 0, vpl = 1;
    uint32_t hub_threshold = 0, hub_chunk = 0, n_chunks = 0, n_hub_rows = 0, n_tiles = 0;
    uint32_t grid = 0;
    gbx::DevBuf R, C[2], mask, inv_size, scal, partials, state, counter, chunks, hub_cnt, hub_part, out;
    bool has_mask = false;
};

namespace gbx {

// VPL consecutive contributions of one source vertex (one 4/8/16-byte load).
template <int VPL>
__device__ __forceinline__ void load_vec(const float *__restrict__ p, float (&x)[VPL])
{
    if constexpr (VPL == 1) {
        x[0] = p[0];
    } else if constexpr (VPL == 2) {
        const float2 t = *reinterpret_cast<const float2 *>(p);
        x[0] = t.x;
        x[1] = t.y;
    } else {
        const float4 t = *reinterpret_cast<const float4 *>(p);
        x[0] = t.x;
        x[1] = t.y;
        x[2] = t.z;
        x[3] = t.w;
    }
}

template <int VPL>
__device__ __forceinline__ void store_vec(float *__restrict__ p, const float (&x)[VPL])
{
    if constexpr (VPL == 1) {
        p[0] = x[0];
    } else if constexpr (VPL == 2) {
        *reinterpret_cast<float2 *>(p) = make_float2(x[0], x[1]);
    } else {
        *reinterpret_cast<float4 *>(p) = make_float4(x[0], x[1], x[2], x[3]);
    }
}

// Gathers in flight per lane: GBX_UNROLL, half as many with four vectors a
// lane (the same bytes in flight, half the registers).
template <int VPL>
constexpr uint32_t team_batch() { return VPL >= 4 ? GBX_UNROLL / 2 : GBX_UNROLL; }

// Sum of one row's contributions for this lane's VPL vectors, edges
// j, j+step, ... in order, U = team_batch() at a time: each batch's U
// gathers (predicated for a partial batch) overlap the next batch's column
// loads, so a batch costs about one round trip.
template <int V, int VPL>
__device__ __forceinline__ void team_row_sum(uint32_t j, uint32_t end, uint32_t step,
                                             const uint32_t *__restrict__ col,
                                             const float *__restrict__ C, uint32_t q, float (&acc)[VPL])
{
    constexpr uint32_t U = team_batch<VPL>();
#pragma unroll
    for (int v = 0; v < VPL; v++) acc[v] = 0.0f;
    const float *base = C + (size_t)q * VPL;
    uint32_t c[U];
#pragma unroll
    for (uint32_t k = 0; k < U; k++) c[k] = j + k * step < end ? col[j + k * step] : 0u;
    for (; j < end; j += U * step) {
        float x[U][VPL];
#pragma unroll
        for (uint32_t k = 0; k < U; k++) {
            if (j + k * step < end) {
                load_vec<VPL>(base + (size_t)c[k] * V, x[k]);
            } else {
#pragma unroll
                for (int v = 0; v < VPL; v++) x[k][v] = 0.0f;
            }
        }
        const uint32_t jn = j + U * step;
        if (jn < end) {
#pragma unroll
            for (uint32_t k = 0; k < U; k++) c[k] = jn + k * step < end ? col[jn + k * step] : 0u;
        }
#pragma unroll
        for (uint32_t k = 0; k < U; k++)
            if (j + k * step < end)
#pragma unroll
                for (int v = 0; v < VPL; v++) acc[v] += x[k][v];
    }
}

// V vectors; a team of L = V/VPL lanes walks one row, lane q holding
// vectors q*VPL .. q*VPL+VPL-1.
template <int V, int VPL>
__global__ void __launch_bounds__(kBlock) ppr_iter_kernel(BPlan P, BatchView A)
{
    if (A.state->done) return;
    constexpr uint32_t L = V / VPL;       // lanes per team
    constexpr uint32_t TPB = kBlock / L;  // teams per block
    const uint32_t q = threadIdx.x % L, team = threadIdx.x / L;
    const float fa = 1.0f - A.damp;
    const float uni = (float)(1.0 / (double)P.N);
    float fs[VPL], fd[VPL], isz[VPL];
    double perr[VPL], pdang[VPL], psum[VPL];
#pragma unroll
    for (int v = 0; v < VPL; v++) {
        const BScal sc = A.scal_in[q * VPL + v];
        fs[v] = sc.sum > 0.0 ? (float)((double)A.damp / sc.sum) : 0.0f;
        fd[v] = (float)(sc.dang / (double)P.N);
        isz[v] = A.mask ? A.inv_size[q * VPL + v] : uni;
        perr[v] = pdang[v] = psum[v] = 0.0;
    }
    // update inputs of row i: out-degree, seed bits, this lane's old ranks
    struct RowIn {
        uint32_t od;
        uint64_t m;
        float r_old[VPL];
    };
    auto row_in = [&](uint32_t i) {
        RowIn in;
        in.od = P.outdeg[i];
        in.m = A.mask ? A.mask[i] : 0ull;
        load_vec<VPL>(A.R + (size_t)i * V + q * VPL, in.r_old);
        return in;
    };
    auto epi = [&](uint32_t i, const RowIn &in, const float (&acc)[VPL]) {
        float r_new[VPL], c_new[VPL];
#pragma unroll
        for (int v = 0; v < VPL; v++) {
            const uint32_t qv = q * VPL + v;
            const float t = A.mask ? (((in.m >> qv) & 1ull) ? isz[v] : 0.0f) : uni;
            r_new[v] = fa * t + fs[v] * (acc[v] + fd[v]);
            c_new[v] = in.od ? r_new[v] / (float)in.od : 0.0f;
            perr[v] += (double)fabsf(r_new[v] - in.r_old[v]);
            if (!in.od) pdang[v] += (double)r_new[v];
            psum[v] += (double)r_new[v];
        }
        const size_t o = (size_t)i * V + q * VPL;
        store_vec<VPL>(A.R + o, r_new);
        store_vec<VPL>(A.C_out + o, c_new);
    };
    __shared__ float red[kBlock * VPL];
    __shared__ bool last_chunk;
    // hub chunks (the first n_chunks tiles), then rows
    uint32_t tile = blockIdx.x;
    for (; tile < P.n_tiles && tile < P.n_chunks; tile += gridDim.x) {
        const HubChunk hc = P.chunks[tile];
        float acc[VPL];
        team_row_sum<V, VPL>(hc.ebeg + team, hc.eend, TPB, P.col, A.C_in, q, acc);
#pragma unroll
        for (int v = 0; v < VPL; v++) red[threadIdx.x * VPL + v] = acc[v];
        __syncthreads();
        if (team == 0) {
#pragma unroll
            for (int v = 0; v < VPL; v++) {
                float s = red[q * VPL + v];
                for (uint32_t t2 = 1; t2 < TPB; t2++) s += red[(t2 * L + q) * VPL + v];
                handoff_store(P.hub_part + (size_t)tile * V + q * VPL + v, s);
            }
            stores_performed();  // before the block's ticket (after the barrier)
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint32_t prev = atomicAdd(&P.hub_cnt[hc.hub], 1u);
            last_chunk = prev == hc.nch - 1;
        }
        __syncthreads();
        if (last_chunk && team == 0) {
            const float *hp = P.hub_part;
            float tot[VPL];
#pragma unroll
            for (int v = 0; v < VPL; v++) {
                tot[v] = 0.0f;
                for (uint32_t k = 0; k < hc.nch; k++)
                    tot[v] += handoff_load(hp + (size_t)(hc.first + k) * V + q * VPL + v);
            }
            epi(hc.row, row_in(hc.row), tot);
            if (q == 0) P.hub_cnt[hc.hub] = 0;
        }
        __syncthreads();
    }
    // Rows, software-pipelined across a team's successive rows: the next
    // row's pointers are loaded when the current row starts, its first
    // U columns and its update inputs when the current row's gathers
    // are done (in flight during the current row's update), and each batch's
    // gathers overlap the next batch's column loads.  Edges are summed in
    // order, U at a time (the sums are team_row_sum's).
    constexpr uint32_t U = team_batch<VPL>();
    const float *cbase = A.C_in + (size_t)q * VPL;
    auto row_of = [&](uint32_t t) { return P.row_begin + (t - P.n_chunks) * TPB + team; };
    auto load_cols = [&](uint32_t j, uint32_t end, uint32_t (&c)[U]) {
#pragma unroll
        for (uint32_t k = 0; k < U; k++) c[k] = j + k < end ? P.col[j + k] : 0u;
    };
    uint32_t row = tile < P.n_tiles ? row_of(tile) : P.N;
    bool valid = row < P.N;
    uint32_t beg = 0, end = 0, c[U];
    RowIn in{};
    if (valid) {
        beg = P.rp[row];
        end = P.rp[row + 1];
        load_cols(beg, end, c);
        in = row_in(row);
    }
    while (tile < P.n_tiles) {
        const uint32_t ntile = tile + gridDim.x;
        const uint32_t nrow = ntile < P.n_tiles ? row_of(ntile) : P.N;
        const bool nvalid = nrow < P.N;
        uint32_t nbeg = 0, nend = 0;
        if (nvalid) {
            nbeg = P.rp[nrow];
            nend = P.rp[nrow + 1];
        }
        float acc[VPL];
#pragma unroll
        for (int v = 0; v < VPL; v++) acc[v] = 0.0f;
        if (valid) {
            for (uint32_t j = beg; j < end; j += U) {
                float x[U][VPL];
#pragma unroll
                for (uint32_t k = 0; k < U; k++) {
                    if (j + k < end) {
                        load_vec<VPL>(cbase + (size_t)c[k] * V, x[k]);
                    } else {
#pragma unroll
                        for (int v = 0; v < VPL; v++) x[k][v] = 0.0f;
                    }
                }
                if (j + U < end) load_cols(j + U, end, c);
#pragma unroll
                for (uint32_t k = 0; k < U; k++)
                    if (j + k < end)
#pragma unroll
                        for (int v = 0; v < VPL; v++) acc[v] += x[k][v];
            }
        }
        RowIn nin{};
        if (nvalid) {
            load_cols(nbeg, nend, c);
            nin = row_in(nrow);
        }
        if (valid) epi(row, in, acc);
        tile = ntile;
        row = nrow;
        valid = nvalid;
        beg = nbeg;
        end = nend;
        in = nin;
    }
    // per-vector block partials: teams of a wave first (lanes L apart share
    // their vectors), then waves through LDS, in a fixed order
#pragma unroll
    for (int v = 0; v < VPL; v++) {
        for (uint32_t off = L; off < 64; off <<= 1) {
            perr[v] += __shfl_xor(perr[v], off, 64);
            pdang[v] += __shfl_xor(pdang[v], off, 64);
            psum[v] += __shfl_xor(psum[v], off, 64);
        }
    }
    constexpr uint32_t W = kBlock / 64;
    __shared__ double pw[3][W][V];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane < L) {
#pragma unroll
        for (int v = 0; v < VPL; v++) {
            pw[0][wave][lane * VPL + v] = perr[v];
            pw[1][wave][lane * VPL + v] = pdang[v];
            pw[2][wave][lane * VPL + v] = psum[v];
        }
    }
    __syncthreads();
    if (threadIdx.x < V) {
        const uint32_t x = threadIdx.x;
        double e = 0.0, d = 0.0, s = 0.0;
        for (uint32_t w = 0; w < W; w++) {
            e += pw[0][w][x];
            d += pw[1][w][x];
            s += pw[2][w][x];
        }
        double *out = A.partials + (size_t)blockIdx.x * 3 * V;
        out[x] = e;
        out[V + x] = d;
        out[2 * V + x] = s;
    }
}

// Initial ranks (teleport vectors) and contributions; block partials of the
// rank sum and dangling mass.
template <int V>
__global__ void __launch_bounds__(kBlock) ppr_reset_kernel(BPlan P, BatchView A)
{
    constexpr uint32_t TPB = kBlock / V;
    const uint32_t q = threadIdx.x % V, team = threadIdx.x / V;
    const float uni = (float)(1.0 / (double)P.N);
    const float isz = A.mask ? A.inv_size[q] : uni;
    double pdang = 0.0, psum = 0.0;
    for (uint32_t i = blockIdx.x * TPB + team; i < P.N; i += gridDim.x * TPB) {
        const size_t o = (size_t)i * V + q;
        const float t = A.mask ? (((A.mask[i] >> q) & 1ull) ? isz : 0.0f) : uni;
        A.R[o] = t;
        const uint32_t od = P.outdeg[i];
        A.C_out[o] = od ? t / (float)od : 0.0f;
        if (!od) pdang += (double)t;
        psum += (double)t;
    }
    __shared__ double pr[2][kBlock];
    pr[0][threadIdx.x] = pdang;
    pr[1][threadIdx.x] = psum;
    __syncthreads();
    if (team == 0) {
        double d = pr[0][q], s = pr[1][q];
        for (uint32_t t2 = 1; t2 < TPB; t2++) {
            d += pr[0][t2 * V + q];
            s += pr[1][t2 * V + q];
        }
        double *out = A.partials + (size_t)blockIdx.x * 3 * V;
        out[q] = 0.0;
        out[V + q] = d;
        out[2 * V + q] = s;
    }
}

// One block per vector: sum the block partials in block order into
// scal_out[q]; the last block evaluates the stop rule over the k vectors.
__global__ void __launch_bounds__(256)
ppr_reduce_kernel(const double *__restrict__ partials, uint32_t nblocks, uint32_t V, uint32_t k,
                  BScal *__restrict__ scal_out, uint32_t slot, BState *state, uint32_t *counter,
                  int stop, double eps, int is_reset)
{
    if (!is_reset && state->done) return;
    const uint32_t q = blockIdx.x;
    double e = 0.0, d = 0.0, s = 0.0;
    for (uint32_t b = threadIdx.x; b < nblocks; b += 256) {
        const double *p = partials + (size_t)b * 3 * V;
        e += p[q];
        d += p[V + q];
        s += p[2 * V + q];
    }
    __shared__ double red[3][256];
    red[0][threadIdx.x] = e;
    red[1][threadIdx.x] = d;
    red[2][threadIdx.x] = s;
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
            red[2][threadIdx.x] += red[2][threadIdx.x + w];
        }
        __syncthreads();
    }
    __shared__ bool last;
    if (threadIdx.x == 0) {
        handoff_store_words<uint64_t>(scal_out + q, BScal{red[0][0], red[1][0], red[2][0]});
        stores_performed();
        last = atomicAdd(counter, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last || threadIdx.x != 0) return;
    *counter = 0;
    if (is_reset) {
        state->done = 0;
        state->iters = 0;
        state->last = slot;
        for (uint32_t v = 0; v < kMaxVec; v++) state->conv_iter[v] = 0;
        return;
    }
    const uint32_t it = state->iters + 1;
    bool all = true;
    for (uint32_t v = 0; v < k; v++) {
        const bool ok = stop == GBX_STOP_L1 && handoff_load(&scal_out[v].err) < eps;
        if (ok && state->conv_iter[v] == 0) state->conv_iter[v] = it;
        all = all && ok;
    }
    state->iters = it;
    state->last = slot;
    if (stop != GBX_STOP_NONE && all) state->done = 1;
}

__global__ void ppr_finalize_kernel(const float *__restrict__ R, const uint32_t *__restrict__ perm,
                                    uint32_t N, uint32_t V, uint32_t k, const BScal *__restrict__ scal,
                                    float *__restrict__ out)
{
    const size_t total = (size_t)N * k;
    for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < total;
         x += (size_t)gridDim.x * blockDim.x) {
        const uint32_t i = (uint32_t)(x / k), q = (uint32_t)(x % k);
        const double S = scal[q].sum > 0.0 ? scal[q].sum : 1.0;
        out[(size_t)perm[i] * k + q] = R[(size_t)i * V + q] / (float)S;
    }
}

// seed-set membership bits in plan order
__global__ void ppr_mask_kernel(const uint32_t *__restrict__ seeds, const uint32_t *__restrict__ owner,
                                uint32_t nseeds, const uint32_t *__restrict__ inv, uint32_t N,
                                unsigned long long *__restrict__ mask)
{
    for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < nseeds; x += gridDim.x * blockDim.x) {
        const uint32_t v = seeds[x];
        if (v < N) atomicOr(&mask[inv[v]], 1ull << owner[x]);
    }
}

static BPlan bplan(const gbx_pr *pr)
{
    const gbx_plan *p = pr->plan;
    const gbx_ppr_state *b = pr->batch;
    BPlan v;
    v.rp = p->rp.as<uint32_t>();
    v.col = p->col.as<uint32_t>();
    v.outdeg = p->outdeg.as<uint32_t>();
    v.chunks = b->chunks.as<HubChunk>();
    v.hub_cnt = b->hub_cnt.as<uint32_t>();
    v.hub_part = b->hub_part.as<float>();
    v.N = p->N;
    v.n_chunks = b->n_chunks;
    v.row_begin = b->n_hub_rows;
    v.n_tiles = b->n_tiles;
    return v;
}

static BatchView bview(const gbx_pr *pr, uint32_t t)
{
    const gbx_ppr_state *b = pr->batch;
    BatchView a;
    a.R = b->R.as<float>();
    a.C_in = b->C[t & 1].as<float>();
    a.C_out = b->C[(t + 1) & 1].as<float>();
    a.mask = b->has_mask ? b->mask.as<uint64_t>() : nullptr;
    a.inv_size = b->inv_size.as<float>();
    a.scal_in = b->scal.as<BScal>() + (size_t)(t & 1) * kMaxVec;
    a.partials = b->partials.as<double>();
    a.state = b->state.as<BState>();
    a.damp = pr->p.damp;
    return a;
}

template <int V>
static void launch_v(bool reset, uint32_t vpl, unsigned grid, hipStream_t st, const BPlan &P, const BatchView &A)
{
    if (reset) ppr_reset_kernel<V><<<grid, kBlock, 0, st>>>(P, A);
    else if constexpr (V >= 4) {
        if (vpl == 4) ppr_iter_kernel<V, 4><<<grid, kBlock, 0, st>>>(P, A);
        else if (vpl == 2) ppr_iter_kernel<V, 2><<<grid, kBlock, 0, st>>>(P, A);
        else ppr_iter_kernel<V, 1><<<grid, kBlock, 0, st>>>(P, A);
    } else if constexpr (V == 2) {
        if (vpl >= 2) ppr_iter_kernel<V, 2><<<grid, kBlock, 0, st>>>(P, A);
        else ppr_iter_kernel<V, 1><<<grid, kBlock, 0, st>>>(P, A);
    } else {
        ppr_iter_kernel<V, 1><<<grid, kBlock, 0, st>>>(P, A);
    }
}

static int launch_any(gbx_pr *pr, bool reset)
{
    gbx_ctx *ctx = pr->ctx;
    gbx_ppr_state *b = pr->batch;
    const uint32_t t = reset ? 1u : pr->t;  // reset writes C[0] (= C_out of t=1... see below)
    BPlan P = bplan(pr);
    BatchView A = bview(pr, t);
    if (reset) A.C_out = b->C[0].as<float>();
    const unsigned grid = b->grid;
    hipStream_t st = ctx->stream;
    switch (b->V) {
    case 1: launch_v<1>(reset, b->vpl, grid, st, P, A); break;
    case 2: launch_v<2>(reset, b->vpl, grid, st, P, A); break;
    case 4: launch_v<4>(reset, b->vpl, grid, st, P, A); break;
    case 8: launch_v<8>(reset, b->vpl, grid, st, P, A); break;
    case 16: launch_v<16>(reset, b->vpl, grid, st, P, A); break;
    case 32: launch_v<32>(reset, b->vpl, grid, st, P, A); break;
    default: launch_v<64>(reset, b->vpl, grid, st, P, A); break;
    }
    GBX_HIP(ctx, hipGetLastError());
    const uint32_t slot = reset ? 0u : ((pr->t + 1) & 1);
    ppr_reduce_kernel<<<b->V, 256, 0, st>>>(b->partials.as<double>(), grid, b->V, b->k,
                                            b->scal.as<BScal>() + (size_t)slot * kMaxVec, slot,
                                            b->state.as<BState>(), b->counter.as<uint32_t>(), pr->p.stop,
                                            pr->p.epsilon, reset ? 1 : 0);
    GBX_HIP(ctx, hipGetLastError());
    return GBX_OK;
}

int ppr_reset(gbx_pr *pr)
{
    pr->t = 0;
    return launch_any(pr, true);
}

int ppr_launch_iter(gbx_pr *pr)
{
    GBX_TRY(launch_any(pr, false));
    pr->t++;
    return GBX_OK;
}

int ppr_read(gbx_pr *pr, float *ranks_out, gbx_pr_report *rep)
{
    gbx_ctx *ctx = pr->ctx;
    gbx_ppr_state *b = pr->batch;
    BState hs;
    GBX_HIP(ctx, hipMemcpyAsync(&hs, b->state.p, sizeof(BState), hipMemcpyDeviceToHost, ctx->stream));
    GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const BScal *sc = b->scal.as<BScal>() + (size_t)hs.last * kMaxVec;
    const uint32_t N = pr->plan->N;
    ppr_finalize_kernel<<<grid_for((size_t)N * b->k, 256, 16384), 256, 0, ctx->stream>>>(
        b->R.as<float>(), pr->plan->perm.as<uint32_t>(), N, b->V, b->k, sc, b->out.as<float>());
    GBX_HIP(ctx, hipGetLastError());
    std::vector<BScal> hsc(b->k);
    GBX_HIP(ctx, hipMemcpyAsync(hsc.data(), sc, b->k * sizeof(BScal), hipMemcpyDeviceToHost, ctx->stream));
    if (ranks_out) GBX_TRY(copy_to_host(ctx, ranks_out, b->out.p, (size_t)N * b->k * 4));
    GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (rep) {
        // one report per vector
        for (uint32_t q = 0; q < b->k; q++) {
            gbx_pr_report &r = rep[q];
            r.executed = hs.iters;
            r.converged = hs.conv_iter[q] != 0;
            r.iterations = r.converged ? hs.conv_iter[q] - 1 : hs.iters;
            r.active = 0;
            r.error_total = hsc[q].err;
            double s = 0.0;
            if (ranks_out)
                for (uint32_t i = 0; i < N; i++) s += ranks_out[(size_t)i * b->k + q];
            r.rank_sum = s;
            r.time_total_s = pr->elapsed_ms * 1e-3;
        }
    }
    return GBX_OK;
}

void ppr_state_free(gbx_ppr_state *s) { delete s; }

// Vector width, hub threshold and chunk size: GBX_PPR_HUB (edges, default
// 4096) and GBX_PPR_CHUNK (edges per hub chunk, default 8192).
static int build_batch(gbx_pr *pr, uint32_t k, const uint32_t *seed_ptr, const uint32_t *seeds)
{
    gbx_ctx *ctx = pr->ctx;
    gbx_plan *p = pr->plan;
    gbx_ppr_state *b = pr->batch;
    b->k = k;
    uint32_t V = 1;
    while (V < k) V <<= 1;
    b->V = V;
    b->vpl = std::min(ctx->tune.ppr_vectors_per_lane, V);
    b->hub_threshold = ctx->tune.ppr_hub;
    b->hub_chunk = ctx->tune.ppr_chunk;
    GBX_TRY(plan_rows_at_least(ctx, p, b->hub_threshold, &b->n_hub_rows));
    std::vector<uint32_t> hrp;
    GBX_TRY(plan_row_ptr_prefix(ctx, p, b->n_hub_rows, hrp));
    std::vector<HubChunk> chunks;
    for (uint32_t h = 0; h < b->n_hub_rows; h++) {
        const uint32_t e0 = hrp[h], e1 = hrp[h + 1];
        const uint32_t nch = (e1 - e0 + b->hub_chunk - 1) / b->hub_chunk;
        const uint32_t first = (uint32_t)chunks.size();
        for (uint32_t c = 0; c < nch; c++) {
            HubChunk hc;
            hc.row = h;
            hc.ebeg = e0 + c * b->hub_chunk;
            hc.eend = std::min(e1, hc.ebeg + b->hub_chunk);
            hc.hub = h;
            hc.first = first;
            hc.nch = nch;
            chunks.push_back(hc);
        }
    }
    b->n_chunks = (uint32_t)chunks.size();
    const uint32_t TPB = kBlock / (V / b->vpl);
    b->n_tiles = b->n_chunks + (p->N - b->n_hub_rows + TPB - 1) / TPB;
    const unsigned per_cu = ctx->tune.ppr_blocks_per_cu;
    b->grid = std::max(1u, std::min((unsigned)std::max(1, ctx->num_cus) * per_cu, b->n_tiles));
    const size_t nv = (size_t)p->N * V * 4;
    GBX_TRY(dev_alloc(ctx, b->R, nv));
    GBX_TRY(dev_alloc(ctx, b->C[0], nv));
    GBX_TRY(dev_alloc(ctx, b->C[1], nv));
    GBX_TRY(dev_alloc(ctx, b->out, (size_t)p->N * k * 4));
    GBX_TRY(dev_alloc(ctx, b->scal, 2 * kMaxVec * sizeof(BScal)));
    GBX_TRY(dev_alloc(ctx, b->partials, (size_t)std::max(b->grid, 1u) * 3 * V * sizeof(double)));
    GBX_TRY(dev_alloc(ctx, b->state, sizeof(BState)));
    GBX_TRY(dev_alloc(ctx, b->counter, 16));
    GBX_TRY(fill_u32(ctx, b->counter.as<uint32_t>(), 0, 4));
    GBX_TRY(dev_alloc(ctx, b->chunks, chunks.size() * sizeof(HubChunk) + 16));
    GBX_TRY(dev_alloc(ctx, b->hub_cnt, (size_t)b->n_hub_rows * 4 + 4));
    GBX_TRY(dev_alloc(ctx, b->hub_part, chunks.size() * V * 4 + 16));
    GBX_TRY(fill_u32(ctx, b->hub_cnt.as<uint32_t>(), 0, b->n_hub_rows + 1));
    if (!chunks.empty())
        GBX_HIP(ctx, hipMemcpyAsync(b->chunks.p, chunks.data(), chunks.size() * sizeof(HubChunk),
                                    hipMemcpyHostToDevice, ctx->stream));
    // teleport vectors
    std::vector<float> isz(V, 0.0f);
    GBX_TRY(dev_alloc(ctx, b->inv_size, V * 4));
    b->has_mask = seed_ptr != nullptr;
    if (b->has_mask) {
        const uint32_t ns = seed_ptr[k];
        std::vector<uint32_t> owner(ns);
        for (uint32_t q = 0; q < k; q++) {
            if (seed_ptr[q + 1] < seed_ptr[q]) return set_error(ctx, GBX_ERR_ARG, "seed_ptr not monotone");
            const uint32_t n = seed_ptr[q + 1] - seed_ptr[q];
            if (n == 0) return set_error(ctx, GBX_ERR_ARG, "seed set %u is empty", q);
            isz[q] = (float)(1.0 / (double)n);
            for (uint32_t x = seed_ptr[q]; x < seed_ptr[q + 1]; x++) {
                if (seeds[x] >= p->N) return set_error(ctx, GBX_ERR_RANGE, "seed %u >= N", seeds[x]);
                owner[x] = q;
            }
        }
        GBX_TRY(dev_alloc(ctx, b->mask, (size_t)p->N * 8));
        GBX_TRY(fill_u32(ctx, b->mask.as<uint32_t>(), 0, (size_t)p->N * 2));
        DevBuf ds, dow;
        GBX_TRY(dev_alloc(ctx, ds, (size_t)ns * 4 + 4));
        GBX_TRY(dev_alloc(ctx, dow, (size_t)ns * 4 + 4));
        if (ns) {
            GBX_HIP(ctx, hipMemcpyAsync(ds.p, seeds, (size_t)ns * 4, hipMemcpyHostToDevice, ctx->stream));
            GBX_HIP(ctx, hipMemcpyAsync(dow.p, owner.data(), (size_t)ns * 4, hipMemcpyHostToDevice, ctx->stream));
            ppr_mask_kernel<<<grid_for(ns, 256, 4096), 256, 0, ctx->stream>>>(
                ds.as<uint32_t>(), dow.as<uint32_t>(), ns, p->inv.as<uint32_t>(), p->N,
                reinterpret_cast<unsigned long long *>(b->mask.p));
            GBX_HIP(ctx, hipGetLastError());
        }
    }
    GBX_HIP(ctx, hipMemcpyAsync(b->inv_size.p, isz.data(), V * 4, hipMemcpyHostToDevice, ctx->stream));
    GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GBX_OK;
}

}  // namespace gbx

using namespace gbx;

extern "C" {

int gbx_ppr_create(gbx_ctx *ctx, gbx_graph *g, const gbx_pr_params *p, uint32_t num_vectors,
                   const uint32_t *seed_ptr, const uint32_t *seeds, gbx_pr **out)
{
    if (!ctx || !g || !p || !out) return set_error(ctx, GBX_ERR_ARG, "gbx_ppr_create: NULL argument");
    *out = nullptr;
    if (num_vectors == 0 || num_vectors > kMaxVec)
        return set_error(ctx, GBX_ERR_ARG, "gbx_ppr_create: num_vectors %u not in 1..%u", num_vectors, kMaxVec);
    if (seed_ptr && !seeds && seed_ptr[num_vectors] > 0)
        return set_error(ctx, GBX_ERR_ARG, "gbx_ppr_create: seeds is NULL");
    if (p->mode != GBX_PR_STOCHASTIC || p->precision != GBX_F32 || p->weighted || p->stop == GBX_STOP_ACTIVE)
        return set_error(ctx, GBX_ERR_ARG,
                         "gbx_ppr_create: batched form is STOCHASTIC, F32, unweighted, stop L1 or NONE");
    if (!(p->damp >= 0.0f && p->damp < 1.0f))
        return set_error(ctx, GBX_ERR_ARG, "gbx_ppr_create: damp %f not in [0,1)", (double)p->damp);
    GBX_HIP(ctx, hipSetDevice(ctx->device));
    gbx_plan *plan = nullptr;
    GBX_TRY(ensure_plan(ctx, g, p->direction, &plan));
    if (plan->segmented) return set_error(ctx, GBX_ERR_STATE, "gbx_ppr_create: not available for a batch of graphs");
    gbx_pr *pr = new (std::nothrow) gbx_pr();
    if (!pr) return set_error(ctx, GBX_ERR_NOMEM, "gbx_ppr_create: out of host memory");
    pr->ctx = ctx;
    pr->g = g;
    pr->plan = plan;
    pr->plan_ref = g->plan[p->direction];
    pr->p = *p;
    pr->kind = K_PR_STOCH;
    pr->batch = new (std::nothrow) gbx_ppr_state();
    int rc = pr->batch ? GBX_OK : set_error(ctx, GBX_ERR_NOMEM, "gbx_ppr_create: out of host memory");
    if (!rc) rc = build_batch(pr, num_vectors, seed_ptr, seeds);
    if (!rc && (hipEventCreate(&pr->ev_begin) != hipSuccess || hipEventCreate(&pr->ev_end) != hipSuccess))
        rc = set_error(ctx, GBX_ERR_HIP, "gbx_ppr_create: event creation");
    if (rc) {
        gbx_pr_free(pr);
        return rc;
    }
    *out = pr;
    return GBX_OK;
}

int gbx_ppr_info(const gbx_pr *pr, uint32_t *num_vectors, uint32_t *vector_width, uint32_t *hub_rows,
                 uint32_t *hub_chunks)
{
    if (!pr || !pr->batch) return set_error(nullptr, GBX_ERR_ARG, "gbx_ppr_info: not a batched object");
    if (num_vectors) *num_vectors = pr->batch->k;
    if (vector_width) *vector_width = pr->batch->V / pr->batch->vpl;
    if (hub_rows) *hub_rows = pr->batch->n_hub_rows;
    if (hub_chunks) *hub_chunks = pr->batch->n_chunks;
    return GBX_OK;
}

}  // extern "C"
