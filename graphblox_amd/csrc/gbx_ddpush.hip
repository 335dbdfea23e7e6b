// gbx_ddpush.hip -- the reference's data-driven push and pull-push PageRank
// (pageRankDataDrivenPushGraphCSR, pageRank.c:1872-2027; pageRankDataDriven
// PullPushGraphCSR, pageRank.c:2029-2202) as frontier iterations.
//
// The reference sweeps its worklist in vertex order and updates ranks and
// residuals in place, so its result depends on the order (and, with several
// threads, on the schedule).  Here one iteration is one kernel over the
// current frontier (a compacted list of vertex ids): one 64-lane wave per
// frontier vertex takes the vertex's residual (atomic exchange with 0), sets
// its rank (push: residual + rank; pull-push: base + d * sum of r[u]/outdeg[u]
// over its in-edges, reading ranks as other waves write them), adds
// |r'/N - r/N| to the error, and pushes delta = d * residual / outdeg along
// its out-edges with float atomics.  A push that lifts a residual from
// <= epsilon to >= epsilon counts as active (the pull-push form adds delta a
// second time, as the reference does) and appends the neighbour to the next
// frontier: one append per wave through a ballot, each vertex once per
// iteration through a stamp.  A vertex's push to itself is dropped: the
// reference zeroes the residual after its push loop, which discards it.
// The last block to finish reduces the error and the active count into the
// next scalar slot (stop when nothing became active) and clears the consumed
// frontier's counter for the iteration after next.
#include <algorithm>
#include <cmath>
#include <new>

#include "gbx_iter.hpp"

struct gbx_ddp_state {
    gbx::DevBuf res;       // f32[N] residuals
    gbx::DevBuf front[2];  // u32[N] frontiers (ping-pong)
    gbx::DevBuf cnt;       // u32[2] frontier sizes
    gbx::DevBuf mark;      // u32[N] iteration stamp of the last append
    unsigned grid = 0;
};

namespace gbx {

struct DdpView {
    const uint32_t *__restrict__ rdeg;  // pull rows: in-degree (DIRECTED) or out-degree
    const uint32_t *__restrict__ ridx;
    const uint32_t *__restrict__ rcol;
    const uint32_t *__restrict__ od;    // out-degree (the divisor and the push fan-out)
    const uint32_t *__restrict__ fidx;  // forward CSR: first out-edge, out-neighbours
    const uint32_t *__restrict__ fcol;
    float *r;
    float *res;
    uint32_t *front0, *front1;
    uint32_t *cnt;
    uint32_t *mark;
    uint32_t N;
};

// r0 = 1-d; res0[v] = (1-d) * d * sum over in-edges of 1/outdeg[u]
// (pageRank.c:1908-1923); frontier 0 = every vertex; slot 0.
__global__ void __launch_bounds__(kBlock) ddp_reset_kernel(DdpView D, IterView A)
{
    const float damp = A.damp, base = 1.0f - damp;
    for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < D.N; v += gridDim.x * kBlock) {
        const uint32_t beg = D.ridx[v], deg = D.rdeg[v];
        float s = 0.0f;
        for (uint32_t j = 0; j < deg; j++) {
            const uint32_t u = D.rcol[beg + j] & kIdMask;
            const uint32_t du = D.od[u];
            if (du) s += 1.0f / (float)du;
        }
        D.res[v] = (1.0f - damp) * damp * s;
        D.r[v] = base;
        D.front0[v] = v;
        D.mark[v] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        D.cnt[0] = D.N;
        D.cnt[1] = 0;
        Slot s = {};
        s.active = D.N;
        A.slots[0] = s;
        *A.counter = 0;
    }
}

template <bool PULL_PUSH>
__global__ void __launch_bounds__(kBlock) ddp_iter_kernel(DdpView D, IterView A)
{
    const Slot *cur = A.slots + (A.t & 1);
    Slot *nxt = A.slots + ((A.t + 1) & 1);
    if (cur->done) {
        if (blockIdx.x == 0 && threadIdx.x == 0) *nxt = *cur;
        return;
    }
    const uint32_t *fin = (A.t & 1) ? D.front1 : D.front0;
    uint32_t *fout = (A.t & 1) ? D.front0 : D.front1;
    uint32_t *cnt_out = D.cnt + ((A.t + 1) & 1);
    const uint32_t n = D.cnt[A.t & 1];
    const uint32_t stamp = A.t + 1;
    const float damp = A.damp, base = 1.0f - damp;
    const float fN = (float)D.N;
    const double eps = A.eps;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t nwaves = gridDim.x * (kBlock / 64);
    Partial part = {0.0, 0.0, 0.0, 0u, 0u};
    for (uint32_t i = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); i < n; i += nwaves) {
        const uint32_t v = fin[i];
        float resv = 0.0f;
        if (lane == 0) resv = atomicExch(D.res + v, 0.0f);
        resv = __shfl(resv, 0, 64);
        const float r_old = D.r[v];
        float r_new;
        if constexpr (PULL_PUSH) {
            // pageRank.c:2114-2122: ranks of this iteration that are already written
            const uint32_t beg = D.ridx[v], deg = D.rdeg[v];
            float s = 0.0f;
            for (uint32_t j = lane; j < deg; j += 64) {
                const uint32_t u = D.rcol[beg + j] & kIdMask;
                s += D.r[u] / (float)D.od[u];
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
            r_new = base + damp * s;
        } else {
            r_new = resv + r_old;  // pageRank.c:1950
        }
        if (lane == 0) {
            part.err += (double)fabsf(r_new / fN - r_old / fN);  // pageRank.c:1951
            D.r[v] = r_new;
        }
        const uint32_t deg = D.od[v], beg = D.fidx[v];
        const float delta = damp * (resv / (float)deg);  // pageRank.c:1955 (unused when deg is 0)
        for (uint32_t j0 = 0; j0 < deg; j0 += 64) {
            const uint32_t j = j0 + lane;
            uint32_t u = 0;
            bool add = false;
            if (j < deg) {
                u = D.fcol[beg + j] & kIdMask;
                if (u != v) {
                    const float prev = atomicAdd(D.res + u, delta);
                    // pageRank.c:1966-1975
                    if (fabs((double)(prev + delta)) >= eps && (double)prev <= eps) {
                        part.active += 1u;
                        if constexpr (PULL_PUSH) atomicAdd(D.res + u, delta);  // pageRank.c:2140
                        add = atomicExch(D.mark + u, stamp) != stamp;
                    }
                }
            }
            const uint64_t m = __ballot(add);
            if (m) {
                const uint32_t leader = (uint32_t)__ffsll((unsigned long long)m) - 1u;
                uint32_t pos = 0;
                if (lane == leader) pos = atomicAdd(cnt_out, (uint32_t)__popcll(m));
                pos = __shfl(pos, (int)leader, 64);
                if (add) fout[pos + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = u;
            }
        }
    }
    const uint32_t iters = cur->iters;
    const int stop = A.stop;
    uint32_t *cnt_in = D.cnt + (A.t & 1);
    grid_reduce(part, A.partials, A.counter, [&](const Partial &tot) {
        *nxt = next_slot<K_DDP>(tot, iters, eps, stop);
        *cnt_in = 0;  // the consumed frontier's counter is the next iteration's output counter
    });
}

static DdpView ddp_view(gbx_pr *pr)
{
    gbx_graph *g = pr->g;
    const gbx_csr_dev &rows = pr->p.direction == GBX_PULL_IN ? g->inv : g->fwd;
    gbx_ddp_state *st = pr->ddp;
    DdpView D;
    D.rdeg = rows.out_degree.as<uint32_t>();
    D.ridx = rows.edges_idx.as<uint32_t>();
    D.rcol = rows.dst.as<uint32_t>();
    D.od = g->fwd.out_degree.as<uint32_t>();
    D.fidx = g->fwd.edges_idx.as<uint32_t>();
    D.fcol = g->fwd.dst.as<uint32_t>();
    D.r = pr->r.as<float>();
    D.res = st->res.as<float>();
    D.front0 = st->front[0].as<uint32_t>();
    D.front1 = st->front[1].as<uint32_t>();
    D.cnt = st->cnt.as<uint32_t>();
    D.mark = st->mark.as<uint32_t>();
    D.N = g->N;
    return D;
}

static IterView ddp_iter_view(gbx_pr *pr)
{
    IterView a = {};
    a.r = pr->r.as<float>();
    a.slots = pr->slots.as<Slot>();
    a.partials = pr->partials.as<Partial>();
    a.counter = pr->counter.as<uint32_t>();
    a.t = pr->t;
    a.damp = pr->p.damp;
    a.eps = pr->p.epsilon;
    a.stop = pr->p.stop;
    return a;
}

int ddp_create(gbx_pr *pr)
{
    gbx_ctx *ctx = pr->ctx;
    const size_t N = pr->g->N;
    gbx_ddp_state *st = new (std::nothrow) gbx_ddp_state();
    if (!st) return set_error(ctx, GBX_ERR_NOMEM, "data-driven state: host memory");
    pr->ddp = st;
    GBX_TRY(dev_alloc(ctx, st->res, N * 4 + 16));
    GBX_TRY(dev_alloc(ctx, st->front[0], N * 4 + 16));
    GBX_TRY(dev_alloc(ctx, st->front[1], N * 4 + 16));
    GBX_TRY(dev_alloc(ctx, st->mark, N * 4 + 16));
    GBX_TRY(dev_alloc(ctx, st->cnt, 16));
    // a persistent grid of waves; every block's partial is reduced, so at most
    // the partials' capacity
    st->grid = (unsigned)std::min<size_t>((size_t)std::max(1, ctx->num_cus) * 8, pr->partials.bytes / sizeof(Partial));
    return GBX_OK;
}

int ddp_reset(gbx_pr *pr)
{
    gbx_ctx *ctx = pr->ctx;
    ddp_reset_kernel<<<grid_for(pr->g->N, kBlock, (unsigned)std::max(1, ctx->num_cus) * 8), kBlock, 0,
                       ctx->stream>>>(ddp_view(pr), ddp_iter_view(pr));
    GBX_HIP(ctx, hipGetLastError());
    return GBX_OK;
}

int ddp_launch_iter(gbx_pr *pr)
{
    gbx_ctx *ctx = pr->ctx;
    const DdpView D = ddp_view(pr);
    const IterView A = ddp_iter_view(pr);
    if (pr->kind == K_DDPP) ddp_iter_kernel<true><<<pr->ddp->grid, kBlock, 0, ctx->stream>>>(D, A);
    else ddp_iter_kernel<false><<<pr->ddp->grid, kBlock, 0, ctx->stream>>>(D, A);
    GBX_HIP(ctx, hipGetLastError());
    pr->t++;
    return GBX_OK;
}

void ddp_state_free(gbx_ddp_state *s) { delete s; }

}  // namespace gbx
