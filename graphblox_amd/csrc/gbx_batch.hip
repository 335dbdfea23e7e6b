// gbx_batch.hip -- a batch of independent graphs advanced by one launch per
// iteration (no reference counterpart).
//
// The graphs are stored block-diagonally in one gbx_graph (graph k owns
// vertices [seg_off[k], seg_off[k+1])).  Its plan relabels rows by degree
// inside each graph and cuts every graph into graph-pure tiles (hub chunks
// and lanes-per-row tiles, gbx_pr.hip: build_segmented_tiles).  Each graph
// runs the stochastic PageRank of its own: teleport 1/N_k, its own dangling
// mass D_k and rank sum S_k, its own L1 stop rule.  A tile adds nothing to
// any other graph, so one kernel walks all tiles; every wave writes its FP64
// partials (|dr|, dangling mass, rank sum) to a slot of its own, and a
// per-graph kernel sums the graph's slots in slot order (deterministic),
// derives the next coefficients and marks converged graphs, whose tiles are
// skipped from then on.
#include <algorithm>
#include <vector>

#include "gbx_plan.hpp"

namespace gbx {

struct SegView {
    const TileDesc *__restrict__ tiles;
    const SegRange *__restrict__ ranges;
    GCoef *__restrict__ coef;
    GStat *__restrict__ stat;
    double *__restrict__ slots;  // [n_wave_slots + n_hubs][3]
    uint32_t hub_slot_base, n_tiles, nseg;
    float *__restrict__ r;
    const float *__restrict__ c_in;
    float *__restrict__ c_out;
    const uint32_t *__restrict__ outdeg;
    uint32_t *all_done;
    uint32_t *counter;
    float damp;
};

}  // namespace gbx

struct gbx_seg_state {
    gbx::DevBuf r, c[2], coef, stat, slots, flags, out;
    uint32_t grid = 0;
    // on the binned iteration (K_SEG): the units' wave partials and every
    // graph's count of rows without out-edges
    gbx::DevBuf uslots, gdang;
};

namespace gbx {

struct SegEpi {
    float *__restrict__ r;
    float *__restrict__ c_out;
    const uint32_t *__restrict__ outdeg;
    float fa, ti, fs, fd;
    double err, dang, sum;
    uint32_t rows;
    __device__ __forceinline__ float init(uint32_t) const { return 0.0f; }
    __device__ __forceinline__ void operator()(uint32_t i, float acc)
    {
        const float r_old = r[i];
        const float r_new = fa * ti + fs * (acc + fd);
        r[i] = r_new;
        const uint32_t od = outdeg[i];
        c_out[i] = od ? r_new / (float)od : 0.0f;
        err += (double)fabsf(r_new - r_old);
        if (!od) dang += (double)r_new;
        sum += (double)r_new;
        rows++;
    }
};

__device__ __forceinline__ void wave_sum3(double &a, double &b, double &c)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
        c += __shfl_xor(c, off, 64);
    }
}

__device__ __forceinline__ void put_slot(double *slots, uint32_t s, double a, double b, double c)
{
    double *p = slots + (size_t)s * 3;
    p[0] = a;
    p[1] = b;
    p[2] = c;
}

__global__ void __launch_bounds__(kBlock) seg_iter_kernel(PlanView P, SegView S)
{
    if (*S.all_done) return;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    GatherC<float, false> gth{S.c_in, 0};
    for (uint32_t tile = blockIdx.x; tile < S.n_tiles; tile += gridDim.x) {
        const TileDesc td = S.tiles[tile];
        const GCoef cf = S.coef[td.graph];
        if (cf.done) continue;
        SegEpi epi{S.r, S.c_out, P.outdeg, 1.0f - S.damp, cf.ti, cf.fs, cf.fd, 0.0, 0.0, 0.0, 0u};
        if (td.lanes == 0) {
            do_hub<AccF32>(td.row_begin, P, gth, epi);
            if (threadIdx.x == 0 && epi.rows)
                put_slot(S.slots, S.hub_slot_base + P.chunks[td.row_begin].hub, epi.err, epi.dang, epi.sum);
            continue;
        }
        const ClassDesc cd{td.lanes, td.row_begin, td.row_end, 0u, 0u};
        switch (td.lanes) {
        case 64: do_rows<64, AccF32>(cd, 0, P, gth, epi); break;
        case 32: do_rows<32, AccF32>(cd, 0, P, gth, epi); break;
        case 16: do_rows<16, AccF32>(cd, 0, P, gth, epi); break;
        case 8: do_rows<8, AccF32>(cd, 0, P, gth, epi); break;
        case 4: do_rows<4, AccF32>(cd, 0, P, gth, epi); break;
        case 2: do_rows<2, AccF32>(cd, 0, P, gth, epi); break;
        default: do_rows<1, AccF32>(cd, 0, P, gth, epi); break;
        }
        double e = epi.err, d = epi.dang, s = epi.sum;
        wave_sum3(e, d, s);
        if (lane == 0) put_slot(S.slots, td.slot + wave, e, d, s);
    }
}

// r0 = 1/N_k, c0 = r0/outdeg, slots of the rank sum and dangling mass.
__global__ void __launch_bounds__(kBlock) seg_reset_kernel(PlanView P, SegView S, const GCoef *__restrict__ init)
{
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint32_t tile = blockIdx.x; tile < S.n_tiles; tile += gridDim.x) {
        const TileDesc td = S.tiles[tile];
        const float t0 = init[td.graph].ti;
        if (td.lanes == 0) {
            const HubChunk hc = P.chunks[td.row_begin];
            if (threadIdx.x == 0 && td.row_begin == hc.first) {
                S.r[hc.row] = t0;
                const uint32_t od = P.outdeg[hc.row];
                S.c_out[hc.row] = od ? t0 / (float)od : 0.0f;
                put_slot(S.slots, S.hub_slot_base + hc.hub, 0.0, od ? 0.0 : (double)t0, (double)t0);
            }
            continue;
        }
        const uint32_t rpt = kBlock / td.lanes;
        const uint32_t end = min(td.row_end, td.row_begin + rpt);
        const uint32_t i = td.row_begin + threadIdx.x;
        double d = 0.0, s = 0.0, e = 0.0;
        if (i < end) {
            S.r[i] = t0;
            const uint32_t od = P.outdeg[i];
            S.c_out[i] = od ? t0 / (float)od : 0.0f;
            if (!od) d = (double)t0;
            s = (double)t0;
        }
        wave_sum3(e, d, s);
        if (lane == 0) put_slot(S.slots, td.slot + wave, 0.0, d, s);
    }
}

// One block per graph: sum the graph's slots in slot order, derive the next
// coefficients, apply the stop rule; the last block raises all_done.
__global__ void __launch_bounds__(256) seg_reduce_kernel(SegView S, int stop, double eps, int is_reset)
{
    const uint32_t k = blockIdx.x;
    const bool was_done = S.coef[k].done != 0;
    __shared__ double red[3][256];
    __shared__ bool last;
    double e = 0.0, d = 0.0, s = 0.0;
    if (is_reset || !was_done) {
        const SegRange rg = S.ranges[k];
        for (uint32_t x = rg.slot_begin + threadIdx.x; x < rg.slot_end; x += 256) {
            const double *p = S.slots + (size_t)x * 3;
            e += p[0];
            d += p[1];
            s += p[2];
        }
        for (uint32_t x = rg.hub_begin + threadIdx.x; x < rg.hub_end; x += 256) {
            const double *p = S.slots + (size_t)(S.hub_slot_base + x) * 3;
            e += p[0];
            d += p[1];
            s += p[2];
        }
    }
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
    if (threadIdx.x == 0) {
        if (is_reset || !was_done) {
            GStat st = S.stat[k];
            const double n = (double)st.n;
            const double sum = red[2][0] > 0.0 ? red[2][0] : 1.0;
            GCoef cf;
            cf.ti = (float)(1.0 / n);
            cf.fs = (float)((double)S.damp / sum);
            cf.fd = (float)(red[1][0] / n);
            cf.done = 0;
            st.err = is_reset ? 0.0 : red[0][0];
            st.dang = red[1][0];
            st.sum = red[2][0];
            if (is_reset) {
                st.iters = 0;
                st.conv_iter = 0;
            } else {
                st.iters += 1;
                if (stop == GBX_STOP_L1 && st.err < eps) {
                    cf.done = 1;
                    st.conv_iter = st.iters;
                }
            }
            handoff_store_words<uint32_t>(S.coef + k, cf);
            handoff_store_words<uint64_t>(S.stat + k, st);
        }
        stores_performed();
        last = atomicAdd(S.counter, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    // the last block: every graph's flag, read by all of its threads
    int all = 1;
    for (uint32_t j = threadIdx.x; j < S.nseg; j += blockDim.x) all &= handoff_load(&S.coef[j].done) ? 1 : 0;
    all = __syncthreads_and(all);
    if (threadIdx.x == 0) {
        *S.all_done = all ? 1u : 0u;
        *S.counter = 0;
    }
}

// ---------------------------------------------------------------- batch on the binned iteration

// r0 = 1/N_k, c0 = r0/outdeg for every row (graph from the plan's row_gid).
__global__ void segbin_reset_rows(const uint32_t *__restrict__ gid, const uint32_t *__restrict__ outdeg,
                                  const GCoef *__restrict__ coef, uint32_t N, float *__restrict__ r,
                                  float *__restrict__ c)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const float t0 = coef[gid[i]].ti;
        r[i] = t0;
        const uint32_t od = outdeg[i];
        c[i] = od ? t0 / (float)od : 0.0f;
    }
}

// Rows without out-edges per graph (wave-aggregated where a wave's rows are
// of one graph).
__global__ void segbin_dangling_counts(const uint32_t *__restrict__ gid, const uint32_t *__restrict__ outdeg,
                                       uint32_t N, uint32_t *__restrict__ cnt)
{
    for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < N; i0 += gridDim.x * blockDim.x) {
        const uint32_t i = i0 + threadIdx.x;
        const bool in = i < N;
        const uint32_t g = in ? gid[i] : 0xFFFFFFFFu;
        const bool dang = in && outdeg[i] == 0;
        const uint32_t g0 = __builtin_amdgcn_readfirstlane(g);
        const uint64_t same = __ballot(dang && g == g0);
        if ((threadIdx.x & 63u) == 0 && same) atomicAdd(cnt + g0, (uint32_t)__popcll(same));
        if (dang && g != g0) atomicAdd(cnt + g, 1u);
    }
}

// One block per graph: at the reset, the rank sum and dangling mass of
// r0 = 1/N_k (N_k and the dangling count times the float 1/N_k, exact in
// FP64, as a sum of the rows' values is); after an iteration, the sums of
// its units' wave slots in ascending unit order.  Then the next coefficients
// and the stop rule as seg_reduce_kernel; the last block writes the batch's
// slot for the next iteration (done when every graph is, the rank sum over
// all graphs for the accumulate's fixed-point scale).
__global__ void __launch_bounds__(256) segbin_reduce_kernel(
    GCoef *__restrict__ coef, GStat *__restrict__ stat, const double *__restrict__ uslots,
    const uint32_t *__restrict__ uptr, const uint32_t *__restrict__ uidx, const uint32_t *__restrict__ gdang,
    uint32_t nseg, float damp, int stop, double eps, int is_reset, uint32_t *counter, uint32_t *all_done, Slot *nxt)
{
    const uint32_t k = blockIdx.x;
    const bool was_done = coef[k].done != 0;
    __shared__ double red[3][256];
    __shared__ bool last;
    double e = 0.0, d = 0.0, s = 0.0;
    if (is_reset) {
        if (threadIdx.x == 0) {
            const double t0 = (double)coef[k].ti;
            s = (double)stat[k].n * t0;
            d = (double)gdang[k] * t0;
        }
    } else if (!was_done) {
        const uint32_t n = (uptr[k + 1] - uptr[k]) * kSegUnitSlots;
        for (uint32_t x = threadIdx.x; x < n; x += 256) {
            const uint32_t u = uidx[uptr[k] + x / kSegUnitSlots];
            const double *q = uslots + ((size_t)u * kSegUnitSlots + x % kSegUnitSlots) * 3u;
            e += q[0];
            d += q[1];
            s += q[2];
        }
    }
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
    if (threadIdx.x == 0) {
        if (is_reset || !was_done) {
            GStat st = stat[k];
            const double n = (double)st.n;
            const double sum = red[2][0] > 0.0 ? red[2][0] : 1.0;
            GCoef cf;
            cf.ti = (float)(1.0 / n);
            cf.fs = (float)((double)damp / sum);
            cf.fd = (float)(red[1][0] / n);
            cf.done = 0;
            st.err = is_reset ? 0.0 : red[0][0];
            st.dang = red[1][0];
            st.sum = red[2][0];
            if (is_reset) {
                st.iters = 0;
                st.conv_iter = 0;
            } else {
                st.iters += 1;
                if (stop == GBX_STOP_L1 && st.err < eps) {
                    cf.done = 1;
                    st.conv_iter = st.iters;
                }
            }
            handoff_store_words<uint32_t>(coef + k, cf);
            handoff_store_words<uint64_t>(stat + k, st);
        }
        stores_performed();
        last = atomicAdd(counter, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    // the last block: every graph's flag and rank sum, read by all of its
    // threads, the sums added in a fixed tree order (deterministic)
    int all = 1;
    double total = 0.0;
    for (uint32_t j = threadIdx.x; j < nseg; j += 256) {
        all &= handoff_load(&coef[j].done) ? 1 : 0;
        total += handoff_load(&stat[j].sum);
    }
    all = __syncthreads_and(all);
    red[0][threadIdx.x] = total;
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[0][threadIdx.x] += red[0][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        *all_done = all ? 1u : 0u;
        Slot sl{};
        sl.sum = red[0][0];
        sl.done = all ? 1u : 0u;
        *nxt = sl;
        *counter = 0;
    }
}

__global__ void seg_finalize_kernel(const float *__restrict__ r, const uint32_t *__restrict__ perm,
                                    const uint32_t *__restrict__ gid, const GStat *__restrict__ stat, uint32_t N,
                                    float *__restrict__ out)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const double S = stat[gid[i]].sum > 0.0 ? stat[gid[i]].sum : 1.0;
        out[perm[i]] = r[i] / (float)S;
    }
}

// ---------------------------------------------------------------- union build

__global__ void copy_offset(const uint32_t *__restrict__ in, uint32_t *__restrict__ out, uint32_t n, uint32_t add)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = in[i] + add;
}

// edges_idx of the union: shifted by the edge offset, still 0 for empty rows
__global__ void copy_edges_idx(const uint32_t *__restrict__ idx, const uint32_t *__restrict__ deg,
                               uint32_t *__restrict__ out, uint32_t n, uint32_t add)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = deg[i] ? idx[i] + add : 0u;
}

static PlanView seg_plan_view(const gbx_plan *p)
{
    PlanView v;
    v.rp = p->rp.as<uint32_t>();
    v.col = p->col.as<uint32_t>();
    v.wt = nullptr;
    v.outdeg = p->outdeg.as<uint32_t>();
    v.chunks = p->seg_chunks.as<HubChunk>();
    v.hub_cnt = p->seg_hub_cnt.as<uint32_t>();
    v.hub_part = p->seg_hub_part.as<float>();
    v.N = p->N;
    v.n_tiles = 0;
    v.n_chunks = p->n_seg_chunks;
    v.nclass = 0;
    return v;
}

static SegView seg_view(gbx_pr *pr, uint32_t t)
{
    gbx_plan *p = pr->plan;
    gbx_seg_state *st = pr->seg;
    SegView v;
    v.tiles = p->tiles.as<TileDesc>();
    v.ranges = p->seg_ranges.as<SegRange>();
    v.coef = st->coef.as<GCoef>();
    v.stat = st->stat.as<GStat>();
    v.slots = st->slots.as<double>();
    v.hub_slot_base = p->n_wave_slots;
    v.n_tiles = p->n_seg_tiles;
    v.nseg = p->nseg;
    v.r = st->r.as<float>();
    v.c_in = st->c[t & 1].as<float>();
    v.c_out = st->c[(t + 1) & 1].as<float>();
    v.outdeg = p->outdeg.as<uint32_t>();
    v.all_done = st->flags.as<uint32_t>();
    v.counter = st->flags.as<uint32_t>() + 1;
    v.damp = pr->p.damp;
    return v;
}

int seg_create(gbx_pr *pr)
{
    gbx_ctx *ctx = pr->ctx;
    gbx_plan *p = pr->plan;
    gbx_graph *g = pr->g;
    pr->seg = new (std::nothrow) gbx_seg_state();
    if (!pr->seg) return set_error(ctx, GBX_ERR_NOMEM, "batch of graphs: out of host memory");
    gbx_seg_state *st = pr->seg;
    const size_t nb = (size_t)p->N * 4 + 16;
    GBX_TRY(dev_alloc(ctx, st->r, nb));
    GBX_TRY(dev_alloc(ctx, st->c[0], nb));
    GBX_TRY(dev_alloc(ctx, st->c[1], nb));
    GBX_TRY(dev_alloc(ctx, st->out, nb));
    GBX_TRY(dev_alloc(ctx, st->coef, (size_t)p->nseg * sizeof(GCoef) + 16));
    GBX_TRY(dev_alloc(ctx, st->stat, (size_t)p->nseg * sizeof(GStat) + 16));
    GBX_TRY(dev_alloc(ctx, st->slots, ((size_t)p->n_wave_slots + p->n_seg_hubs + 1) * 3 * sizeof(double)));
    GBX_TRY(dev_alloc(ctx, st->flags, 16));
    GBX_TRY(fill_u32(ctx, st->flags.as<uint32_t>(), 0, 4));
    std::vector<GStat> hs(p->nseg);
    std::vector<GCoef> hc(p->nseg);
    for (uint32_t k = 0; k < p->nseg; k++) {
        hs[k] = GStat{0.0, 0.0, 0.0, 0u, 0u, g->seg_off[k + 1] - g->seg_off[k], 0u};
        hc[k] = GCoef{(float)(1.0 / (double)std::max(1u, hs[k].n)), 0.0f, 0.0f, 0u};
    }
    GBX_HIP(ctx, hipMemcpyAsync(st->stat.p, hs.data(), hs.size() * sizeof(GStat), hipMemcpyHostToDevice, ctx->stream));
    GBX_HIP(ctx, hipMemcpyAsync(st->coef.p, hc.data(), hc.size() * sizeof(GCoef), hipMemcpyHostToDevice, ctx->stream));
    GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const unsigned per_cu = ctx->tune.ppr_blocks_per_cu;
    st->grid = std::max(1u, std::min((unsigned)std::max(1, ctx->num_cus) * per_cu, p->n_seg_tiles));
    if (pr->kind == K_SEG) {
        // the binned iteration: its layout (bins cut at every graph's first
        // row), the batch's slot pair and unit counter, the units' partials
        GBX_TRY(dev_alloc(ctx, pr->slots, 2 * sizeof(Slot)));
        GBX_TRY(dev_alloc(ctx, pr->counter, 64));
        GBX_TRY(fill_u32(ctx, pr->counter.as<uint32_t>(), 0, 16));
        GBX_TRY(binned_create(pr));
        GBX_TRY(dev_alloc(ctx, st->uslots, ((size_t)binned_units(pr) * kSegUnitSlots + 1) * 3 * sizeof(double)));
        GBX_TRY(dev_alloc(ctx, st->gdang, (size_t)p->nseg * 4 + 4));
        GBX_TRY(fill_u32(ctx, st->gdang.as<uint32_t>(), 0, p->nseg));
        segbin_dangling_counts<<<grid_for(p->N, 256, 16384), 256, 0, ctx->stream>>>(
            p->row_gid.as<uint32_t>(), p->outdeg.as<uint32_t>(), p->N, st->gdang.as<uint32_t>());
        GBX_HIP(ctx, hipGetLastError());
        GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return GBX_OK;
}

int seg_reset(gbx_pr *pr)
{
    gbx_ctx *ctx = pr->ctx;
    gbx_seg_state *st = pr->seg;
    pr->t = 0;
    if (pr->kind == K_SEG) {
        gbx_plan *p = pr->plan;
        // the coefficient array still holds ti = 1/N_k in every state
        segbin_reset_rows<<<grid_for(p->N, 256, 16384), 256, 0, ctx->stream>>>(
            p->row_gid.as<uint32_t>(), p->outdeg.as<uint32_t>(), st->coef.as<GCoef>(), p->N, st->r.as<float>(),
            st->c[0].as<float>());
        GBX_HIP(ctx, hipGetLastError());
        const uint32_t *uptr = nullptr, *uidx = nullptr;
        binned_graph_units(pr, &uptr, &uidx);
        segbin_reduce_kernel<<<p->nseg, 256, 0, ctx->stream>>>(
            st->coef.as<GCoef>(), st->stat.as<GStat>(), st->uslots.as<double>(), uptr, uidx, st->gdang.as<uint32_t>(),
            p->nseg, pr->p.damp, pr->p.stop, pr->p.epsilon, 1,
            st->flags.as<uint32_t>() + 1, st->flags.as<uint32_t>(), pr->slots.as<Slot>());
        GBX_HIP(ctx, hipGetLastError());
        return GBX_OK;
    }
    PlanView P = seg_plan_view(pr->plan);
    SegView S = seg_view(pr, 1);
    S.c_out = st->c[0].as<float>();
    // the coefficient array still holds ti = 1/N_k in every state
    seg_reset_kernel<<<st->grid, kBlock, 0, ctx->stream>>>(P, S, S.coef);
    GBX_HIP(ctx, hipGetLastError());
    seg_reduce_kernel<<<pr->plan->nseg, 256, 0, ctx->stream>>>(S, pr->p.stop, pr->p.epsilon, 1);
    GBX_HIP(ctx, hipGetLastError());
    return GBX_OK;
}

int seg_launch_iter(gbx_pr *pr)
{
    gbx_ctx *ctx = pr->ctx;
    gbx_seg_state *st = pr->seg;
    if (pr->kind == K_SEG) {
        gbx_plan *p = pr->plan;
        const uint32_t t = pr->t;
        GBX_TRY(binned_seg_iter(pr, st->r.as<float>(), st->c[t & 1].as<float>(), st->c[(t + 1) & 1].as<float>(),
                                st->coef.as<GCoef>(), st->uslots.as<double>()));
        const uint32_t *uptr = nullptr, *uidx = nullptr;
        binned_graph_units(pr, &uptr, &uidx);
        segbin_reduce_kernel<<<p->nseg, 256, 0, ctx->stream>>>(
            st->coef.as<GCoef>(), st->stat.as<GStat>(), st->uslots.as<double>(), uptr, uidx, st->gdang.as<uint32_t>(),
            p->nseg, pr->p.damp, pr->p.stop, pr->p.epsilon, 0,
            st->flags.as<uint32_t>() + 1, st->flags.as<uint32_t>(), pr->slots.as<Slot>() + ((t + 1) & 1));
        GBX_HIP(ctx, hipGetLastError());
        pr->t++;
        return GBX_OK;
    }
    PlanView P = seg_plan_view(pr->plan);
    SegView S = seg_view(pr, pr->t);
    seg_iter_kernel<<<st->grid, kBlock, 0, ctx->stream>>>(P, S);
    GBX_HIP(ctx, hipGetLastError());
    seg_reduce_kernel<<<pr->plan->nseg, 256, 0, ctx->stream>>>(S, pr->p.stop, pr->p.epsilon, 0);
    GBX_HIP(ctx, hipGetLastError());
    pr->t++;
    return GBX_OK;
}

int seg_read(gbx_pr *pr, float *ranks_out, gbx_pr_report *rep)
{
    gbx_ctx *ctx = pr->ctx;
    gbx_plan *p = pr->plan;
    gbx_seg_state *st = pr->seg;
    seg_finalize_kernel<<<grid_for(p->N, 256, 16384), 256, 0, ctx->stream>>>(
        st->r.as<float>(), p->perm.as<uint32_t>(), p->row_gid.as<uint32_t>(), st->stat.as<GStat>(), p->N,
        st->out.as<float>());
    GBX_HIP(ctx, hipGetLastError());
    std::vector<GStat> hs(p->nseg);
    GBX_HIP(ctx, hipMemcpyAsync(hs.data(), st->stat.p, hs.size() * sizeof(GStat), hipMemcpyDeviceToHost, ctx->stream));
    if (ranks_out) GBX_TRY(copy_to_host(ctx, ranks_out, st->out.p, (size_t)p->N * 4));
    GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (rep) {
        const std::vector<uint32_t> &off = pr->g->seg_off;
        for (uint32_t k = 0; k < p->nseg; k++) {
            gbx_pr_report &r = rep[k];
            r.executed = hs[k].iters;
            r.converged = hs[k].conv_iter != 0;
            r.iterations = r.converged ? hs[k].conv_iter - 1 : hs[k].iters;
            r.active = 0;
            r.error_total = hs[k].err;
            double s = 0.0;
            if (ranks_out)
                for (uint32_t v = off[k]; v < off[k + 1]; v++) s += ranks_out[v];
            r.rank_sum = s;
            r.time_total_s = pr->elapsed_ms * 1e-3;
        }
    }
    return GBX_OK;
}

void seg_state_free(gbx_seg_state *s) { delete s; }

}  // namespace gbx

using namespace gbx;

extern "C" {

int gbx_graph_batch_create(gbx_ctx *ctx, uint32_t num_graphs, gbx_graph *const *graphs, gbx_graph **out)
{
    if (!ctx || !out || !graphs || num_graphs == 0)
        return set_error(ctx, GBX_ERR_ARG, "gbx_graph_batch_create: NULL argument or no graphs");
    *out = nullptr;
    uint64_t N = 0, E = 0;
    bool weighted = true;
    for (uint32_t k = 0; k < num_graphs; k++) {
        if (!graphs[k]) return set_error(ctx, GBX_ERR_ARG, "gbx_graph_batch_create: graph %u is NULL", k);
        if (graphs[k]->seg_off.size() >= 2)
            return set_error(ctx, GBX_ERR_ARG, "gbx_graph_batch_create: graph %u is itself a batch", k);
        N += graphs[k]->N;
        E += graphs[k]->E;
        weighted = weighted && graphs[k]->weighted;
    }
    if (N > kMaxVertices || E >= 0x80000000ull)
        return set_error(ctx, GBX_ERR_RANGE, "gbx_graph_batch_create: %llu vertices / %llu edges too many",
                         (unsigned long long)N, (unsigned long long)E);
    GBX_HIP(ctx, hipSetDevice(ctx->device));
    gbx_graph *u = new (std::nothrow) gbx_graph();
    if (!u) return set_error(ctx, GBX_ERR_NOMEM, "gbx_graph_batch_create: out of host memory");
    u->ctx = ctx;
    u->N = (uint32_t)N;
    u->E = (uint32_t)E;
    u->weighted = weighted;
    int rc = GBX_OK;
    const size_t nb = (size_t)N * 4 + 4, eb = (size_t)E * 4 + 4;
    for (gbx_csr_dev *c : {&u->fwd, &u->inv}) {
        if (!rc) rc = dev_alloc(ctx, c->out_degree, nb);
        if (!rc) rc = dev_alloc(ctx, c->in_degree, nb);
        if (!rc) rc = dev_alloc(ctx, c->edges_idx, nb);
        if (!rc) rc = dev_alloc(ctx, c->src, eb);
        if (!rc) rc = dev_alloc(ctx, c->dst, eb);
        if (!rc) rc = dev_alloc(ctx, c->w, eb);
    }
    if (!rc) rc = dev_alloc(ctx, u->seg_gid, nb);
    uint32_t voff = 0, eoff = 0;
    u->seg_off.push_back(0);
    u->seg_edge_off.push_back(0);
    for (uint32_t k = 0; k < num_graphs && !rc; k++) {
        const gbx_graph *gk = graphs[k];
        const uint32_t n = gk->N, e = gk->E;
        for (int side = 0; side < 2 && !rc; side++) {
            const gbx_csr_dev &a = side ? gk->inv : gk->fwd;
            gbx_csr_dev &b = side ? u->inv : u->fwd;
            if (hipMemcpyAsync(b.out_degree.as<uint32_t>() + voff, a.out_degree.p, (size_t)n * 4,
                               hipMemcpyDeviceToDevice, ctx->stream) ||
                hipMemcpyAsync(b.in_degree.as<uint32_t>() + voff, a.in_degree.p, (size_t)n * 4,
                               hipMemcpyDeviceToDevice, ctx->stream) ||
                (e && hipMemcpyAsync(b.w.as<float>() + eoff, a.w.p, (size_t)e * 4, hipMemcpyDeviceToDevice,
                                     ctx->stream))) {
                rc = set_error(ctx, GBX_ERR_HIP, "gbx_graph_batch_create: copy failed");
                break;
            }
            copy_edges_idx<<<grid_for(n, 256, 4096), 256, 0, ctx->stream>>>(
                a.edges_idx.as<uint32_t>(), a.out_degree.as<uint32_t>(), b.edges_idx.as<uint32_t>() + voff, n, eoff);
            if (e) {
                copy_offset<<<grid_for(e, 256, 16384), 256, 0, ctx->stream>>>(a.src.as<uint32_t>(),
                                                                               b.src.as<uint32_t>() + eoff, e, voff);
                copy_offset<<<grid_for(e, 256, 16384), 256, 0, ctx->stream>>>(a.dst.as<uint32_t>(),
                                                                               b.dst.as<uint32_t>() + eoff, e, voff);
            }
            if (hipGetLastError() != hipSuccess) rc = set_error(ctx, GBX_ERR_HIP, "gbx_graph_batch_create: kernel");
        }
        if (!rc) rc = fill_u32(ctx, u->seg_gid.as<uint32_t>() + voff, k, n);
        voff += n;
        eoff += e;
        u->seg_off.push_back(voff);
        u->seg_edge_off.push_back(eoff);
    }
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess)
        rc = set_error(ctx, GBX_ERR_HIP, "gbx_graph_batch_create: sync failed");
    if (rc) {
        delete u;
        return rc;
    }
    *out = u;
    return GBX_OK;
}

int gbx_graph_batch_info(const gbx_graph *g, uint32_t *num_graphs, uint32_t *vertex_offsets)
{
    if (!g) return set_error(nullptr, GBX_ERR_ARG, "gbx_graph_batch_info: NULL graph");
    const uint32_t n = g->seg_off.size() >= 2 ? (uint32_t)g->seg_off.size() - 1 : 1;
    if (num_graphs) *num_graphs = n;
    if (vertex_offsets) {
        if (g->seg_off.size() >= 2)
            for (uint32_t k = 0; k <= n; k++) vertex_offsets[k] = g->seg_off[k];
        else {
            vertex_offsets[0] = 0;
            vertex_offsets[1] = g->N;
        }
    }
    return GBX_OK;
}

}  // extern "C"
