// gbx_pr.hip -- degree-binned pull PageRank and pull SpMV for gfx950.
//
// Reference semantics: pageRankPullGraphCSR (01_host/src/algorithms/openmp/
// pageRank.c:245-368) and SPMVPullGraphCSR (01_host/src/algorithms/openmp/
// SPMV.c:117-215).
//
// Plan (built once per graph and direction, device resident):
//   * rows are relabelled by in-degree, descending (stable by id), so every
//     degree bin is one contiguous row range and the most gathered vertices
//     (R-MAT in/out degrees are correlated) share cache lines;
//   * within a row the edge order of the reference CSR is kept, so the row
//     sum visits sources in the reference's order;
//   * bins: hub rows (degree >= hub threshold) are cut into fixed-size edge
//     chunks, one 256-thread block each, whose partial sums the last chunk
//     to finish combines in chunk order; other rows get G lanes per row
//     (G = 64, 32, ..., 2, 1 by degree), G=1 being the packed thread-per-row
//     tail whose sequential sum is bit-identical to the reference's.
// Iteration: one kernel.  It gathers the previous contributions, applies the
// update, writes the next contributions (ping-pong) and the ranks (in place),
// and reduces |dr|, the active count, the dangling mass and the rank sum in
// FP64 per block; the last block to finish reduces the block partials in a
// fixed order into the next scalar slot and evaluates the stop rule, after
// which later iteration kernels return immediately.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <type_traits>
#include <vector>

#include "gbx_iter.hpp"


namespace gbx {

// One iteration.  T: contribution element type (float or bf16 bits);
// Acc: row accumulator; LDS: hot prefix of the contributions in LDS.
template <int KIND, class T, class Acc, bool LDS>
__global__ void __launch_bounds__(kBlock) pr_iter_kernel(PlanView P, IterView A)
{
    const Slot *cur = A.slots + (A.t & 1);
    Slot *nxt = A.slots + ((A.t + 1) & 1);
    if (cur->done) {
        if (blockIdx.x == 0 && threadIdx.x == 0) *nxt = *cur;
        return;
    }
    const T *c_in = reinterpret_cast<const T *>(A.c_in);
    if constexpr (LDS) load_hot(c_in, A.hot);
    std::conditional_t<KIND == K_DD, DdEpi, PrEpi<KIND, T>> epi;
    if constexpr (KIND == K_DD) epi = make_dd_epi(P.outdeg, P.N, A, cur);
    else epi = make_epi<KIND, T>(P.outdeg, P.N, A, cur);
    if constexpr (KIND == K_PR_STOCH_W || KIND == K_ADJ) {
        GatherCW<T, LDS> g{c_in, reinterpret_cast<const T *>(P.wt), A.hot};
        tile_loop<Acc>(P, g, epi);
    } else {
        GatherC<T, LDS> g{c_in, A.hot};
        tile_loop<Acc>(P, g, epi);
    }
    const uint32_t iters = cur->iters;
    const double eps = A.eps;
    const int stop = A.stop;
    grid_reduce(epi.part, A.partials, A.counter,
                [&](const Partial &tot) { *nxt = next_slot<KIND>(tot, iters, eps, stop); });
}

// r0, c0 and slot 0 (sum and dangling mass of r0).
template <int KIND, class T>
__global__ void __launch_bounds__(kBlock) pr_reset_kernel(PlanView P, IterView A)
{
    Partial part = {0.0, 0.0, 0.0, 0u, 0u};
    const float base = 1.0f - A.damp;
    const float uni = (float)(1.0 / (double)P.N);
    T *c_out = reinterpret_cast<T *>(A.c_out);
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < P.N; i += gridDim.x * kBlock) {
        float r0 = (KIND == K_PR_REF || is_int_kind(KIND)) ? base : (A.tele ? A.tele[i] : uni);
        A.r[i] = r0;
        const uint32_t od = P.outdeg[i];
        if constexpr (is_fixed_kind(KIND)) {
            c_out[i] = od ? (T)fixed_conv<KIND>(r0 / (float)od) : (T)0;  // pageRank.c:543-548
        } else if constexpr (is_quant_kind(KIND)) {
            const float c0 = od ? r0 / (float)od : 0.0f;  // pageRank.c:1205-1211
            c_out[i] = c0;
            note_contribution(part, c0);
        } else {
            float c0;
            if (KIND == K_ADJ) c0 = r0;
            else if (KIND == K_PR_STOCH_W) c0 = od ? r0 : 0.0f;
            else c0 = od ? r0 / (float)od : 0.0f;
            store_val(c_out, i, c0);
        }
        if (KIND == K_ADJ) part.dang += (double)fabsf(r0);
        else if (!od) part.dang += (double)r0;
        part.sum += (double)r0;
    }
    Slot *s0 = A.slots;
    const uint32_t N = P.N;
    grid_reduce(part, A.partials, A.counter, [&](const Partial &tot) {
        Slot s;
        s.err = 0.0;
        s.dang = tot.dang;
        s.sum = tot.sum;
        s.active = N;
        s.iters = 0;
        s.done = 0;
        s.conv = 0;
        s.cmax = tot.cmax;
        s.cnmin = tot.cnmin;
        *s0 = s;
    });
}

// FP64 sum of the returned ranks into dst->sum (fixed grid, fixed order:
// deterministic); the read-back reports it, so the host does not sum N
// floats after the copy.
__global__ void __launch_bounds__(kBlock) pr_out_sum_kernel(const float *__restrict__ out, uint32_t N,
                                                            Partial *partials, uint32_t *counter, Slot *dst)
{
    Partial part = {0.0, 0.0, 0.0, 0u, 0u, 0u, 0u};
    double a = 0.0, b = 0.0;  // two chains a thread
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t step = gridDim.x * kBlock;
    for (; i + step < N; i += 2 * step) {
        a += (double)out[i];
        b += (double)out[i + step];
    }
    if (i < N) a += (double)out[i];
    part.sum = a + b;
    grid_reduce(part, partials, counter, [&](const Partial &tot) { dst->sum = tot.sum; });
}

// ranks in original order: REF divides by N (pageRank.c:344-350); the
// stochastic forms divide by the last FP64 rank sum.
__global__ void pr_finalize_kernel(const float *__restrict__ r, const uint32_t *__restrict__ perm,
                                   uint32_t N, const Slot *slot, int ref, float *__restrict__ out)
{
    const float fN = (float)N;
    const double S = slot->sum > 0.0 ? slot->sum : 1.0;
    const float fS = (float)S;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
        out[perm ? perm[i] : i] = ref == 2 ? r[i] : (ref ? r[i] / fN : r[i] / fS);
}

// ------------------------------------------------------------------ SpMV

struct SpmvEpi {
    float *__restrict__ y;
    __device__ __forceinline__ float init(uint32_t i) const { return y[i]; }
    __device__ __forceinline__ void operator()(uint32_t i, float acc) { y[i] = acc; }
};

template <int KIND>
__global__ void __launch_bounds__(kBlock)
spmv_kernel(PlanView P, const float *__restrict__ x, float *__restrict__ y)
{
    SpmvEpi epi{y};
    if constexpr (KIND == K_SPMV_W) {
        GatherCW<float, false> g{x, P.wt, 0};
        tile_loop<AccF32>(P, g, epi);
    } else {
        GatherConstW g{x, 0.0001f};  // SPMV.c:177
        tile_loop<AccF32>(P, g, epi);
    }
}

// ------------------------------------------------------------------ integer gather-accumulate
// The FPGA overlay's Single.PR program (02_device/ol/GraphBlox/Single/
// topology.json:162-201, host side buffer_mapping.PR.cpp:2-68): for each row
// v, read its edge range, gather the u32 value of every column u and add
// them with ALU_ACC's unsigned 32-bit add (engine_alu_ops_kernel.sv:104-112).
struct GatherU32 {
    const uint32_t *__restrict__ v;
    __device__ __forceinline__ uint32_t operator()(uint32_t, uint32_t c) const { return v[c]; }
};

struct U32Epi {
    uint32_t *__restrict__ out;
    __device__ __forceinline__ uint32_t init(uint32_t) const { return 0u; }
    __device__ __forceinline__ void operator()(uint32_t i, uint32_t acc) { out[i] = acc; }
};

__global__ void __launch_bounds__(kBlock)
gather_acc_u32_kernel(PlanView P, const uint32_t *__restrict__ in, uint32_t *__restrict__ out)
{
    U32Epi epi{out};
    GatherU32 g{in};
    tile_loop<AccU32>(P, g, epi);
}

__global__ void gather_perm_u32(const uint32_t *__restrict__ in, const uint32_t *__restrict__ perm, uint32_t N,
                                uint32_t *__restrict__ out)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
        out[i] = in[perm[i]];
}

__global__ void scatter_perm_u32(const uint32_t *__restrict__ in, const uint32_t *__restrict__ perm, uint32_t N,
                                 uint32_t *__restrict__ out)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
        out[perm[i]] = in[i];
}

// ------------------------------------------------------------------ plan build

__global__ void degree_keys(const uint32_t *__restrict__ deg, uint32_t N, uint32_t maxdeg,
                            uint32_t *__restrict__ key, uint32_t *__restrict__ val)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        key[i] = maxdeg - deg[i];
        val[i] = i;
    }
}

__global__ void plan_vertices(const uint32_t *__restrict__ perm, uint32_t N,
                              const uint32_t *__restrict__ row_deg, const uint32_t *__restrict__ fwd_outdeg,
                              uint32_t *__restrict__ inv, uint32_t *__restrict__ deg_new,
                              uint32_t *__restrict__ outdeg_new)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const uint32_t v = perm[i];
        inv[v] = i;
        deg_new[i] = row_deg[v];
        outdeg_new[i] = fwd_outdeg[v];
    }
}

// key[i] = graph of the vertex now at position i (batches of graphs)
__global__ void gid_keys(const uint32_t *__restrict__ perm, const uint32_t *__restrict__ gid, uint32_t N,
                         uint32_t *__restrict__ key)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
        key[i] = gid[perm[i]];
}

// Copy each edge of the row CSR into the relabelled plan, keeping its
// position inside the row.
__global__ void plan_edges(const uint32_t *__restrict__ rsrc, const uint32_t *__restrict__ rdst,
                           const float *__restrict__ rw, const uint32_t *__restrict__ ridx,
                           const uint32_t *__restrict__ inv, const uint32_t *__restrict__ rp,
                           uint32_t E, uint32_t *__restrict__ col, float *__restrict__ wt)
{
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x) {
        const uint32_t v = rsrc[e] & kIdMask;
        const uint32_t pos = rp[inv[v]] + (e - ridx[v]);
        col[pos] = inv[rdst[e] & kIdMask];
        if (wt) wt[pos] = rw[e];
    }
}

// bound[c] = number of rows with degree >= thr[c] (degrees are descending).
__global__ void class_bounds(const uint32_t *__restrict__ deg, uint32_t N,
                             const uint32_t *__restrict__ thr, uint32_t nthr,
                             uint32_t *__restrict__ bound)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const uint32_t d = deg[i];
        const uint32_t dn = i + 1 < N ? deg[i + 1] : 0u;
        for (uint32_t c = 0; c < nthr; c++)
            if (d >= thr[c] && dn < thr[c]) bound[c] = i + 1;
    }
}

__global__ void plan_degrees(const uint32_t *__restrict__ rp, uint32_t N, uint32_t *__restrict__ deg)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
        deg[i] = rp[i + 1] - rp[i];
}

// Sum of each vertex's out-weights in the forward CSR, one wave per vertex.
__global__ void out_weight_sums(const uint32_t *__restrict__ od, const uint32_t *__restrict__ idx,
                                const float *__restrict__ w, uint32_t N, float *__restrict__ S)
{
    const uint32_t lane = threadIdx.x & 63;
    for (uint32_t v = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; v < N;
         v += (gridDim.x * blockDim.x) >> 6) {
        double s = 0.0;
        const uint32_t b = idx[v], e = b + od[v];
        for (uint32_t j = b + lane; j < e; j += 64) s += (double)w[j];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        if (lane == 0) S[v] = (float)s;
    }
}

// Row-normalised transition weight of each plan edge: w / S[source].
__global__ void normalise_weights(const uint32_t *__restrict__ col, const float *__restrict__ wt,
                                  const uint32_t *__restrict__ perm, const float *__restrict__ S,
                                  uint32_t E, float *__restrict__ wn)
{
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < E; j += gridDim.x * blockDim.x) {
        const float s = S[perm[col[j]]];
        wn[j] = s > 0.0f ? (float)((double)wt[j] / (double)s) : 0.0f;
    }
}

// Per-edge transition weight in forward-CSR order: w[e] / S[src[e]].
__global__ void row_normalise_fwd(const uint32_t *__restrict__ src, const float *__restrict__ w,
                                  const float *__restrict__ S, uint32_t E, float *__restrict__ out)
{
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x) {
        const float s = S[src[e] & kIdMask];
        out[e] = s > 0.0f ? (float)((double)w[e] / (double)s) : 0.0f;
    }
}

// Per-edge values of the row CSR placed at their plan positions.
__global__ void plan_edge_values(const uint32_t *__restrict__ rsrc, const uint32_t *__restrict__ ridx,
                                 const uint32_t *__restrict__ inv, const uint32_t *__restrict__ rp, uint32_t E,
                                 const float *__restrict__ in, float *__restrict__ out)
{
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x) {
        const uint32_t v = rsrc[e] & kIdMask;
        out[rp[inv[v]] + (e - ridx[v])] = in[e];
    }
}

// dL/dW[e] for forward-CSR edge e = (u, v):  d p[u] (lambda[v] - (P lambda)[u]) / S[u]
__global__ void adjoint_edge_grad(const uint32_t *__restrict__ src, const uint32_t *__restrict__ dst, uint32_t E,
                                  const float *__restrict__ p, const float *__restrict__ lam,
                                  const float *__restrict__ plam, const float *__restrict__ S, float damp,
                                  float *__restrict__ grad)
{
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x) {
        const uint32_t u = src[e] & kIdMask, v = dst[e] & kIdMask;
        const double s = (double)S[u];
        grad[e] = s > 0.0 ? (float)((double)damp * (double)p[u] * ((double)lam[v] - (double)plam[u]) / s) : 0.0f;
    }
}

__global__ void to_bf16(const float *__restrict__ in, uint32_t n, uint16_t *__restrict__ out)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = f32_to_bf16(in[i]);
}

__global__ void gather_perm(const float *__restrict__ in, const uint32_t *__restrict__ perm,
                            uint32_t N, float *__restrict__ out)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
        out[i] = in[perm[i]];
}

__global__ void scatter_perm(const float *__restrict__ in, const uint32_t *__restrict__ perm,
                             uint32_t N, float *__restrict__ out)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
        out[perm[i]] = in[i];
}

__global__ void spmv_init_x(const uint32_t *__restrict__ od, uint32_t N, float *__restrict__ x,
                            float *__restrict__ y)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        x[i] = od[i] ? 1.0f / od[i] : 0.001f;  // SPMV.c:156-163
        y[i] = 0.0f;
    }
}

__global__ void fill_f32(float *p, float v, uint32_t n)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        p[i] = v;
}

// Degree thresholds of the bins (hub, then 64, 32, 16, 8, 4, 2 lanes per
// row; rows below the last threshold use 1 lane) and the hub chunk size, from
// the context's tuning (a hub threshold or chunk of 0 chosen by the graph's
// size).
struct BinConfig {
    uint32_t thr[7];
    uint32_t hub_chunk;
    bool operator==(const BinConfig &o) const
    {
        return hub_chunk == o.hub_chunk && memcmp(thr, o.thr, sizeof(thr)) == 0;
    }
};

static BinConfig bin_config(const gbx_ctx *ctx, const gbx_graph *g)
{
    BinConfig b;
    memcpy(b.thr, ctx->tune.bin_thresholds, sizeof(b.thr));
    if (b.thr[0] == 0) b.thr[0] = auto_hub_threshold(cut_edges(g));
    b.hub_chunk = ctx->tune.hub_chunk ? ctx->tune.hub_chunk : 4u * b.thr[0];
    return b;
}

}  // namespace gbx

uint64_t gbx::cut_edges(const gbx_graph *g)
{
    uint64_t m = g->E;
    if (g->seg_edge_off.size() >= 2) {
        m = 0;
        for (size_t k = 0; k + 1 < g->seg_edge_off.size(); k++)
            m = std::max<uint64_t>(m, g->seg_edge_off[k + 1] - g->seg_edge_off[k]);
    }
    return m;
}

namespace gbx {

// Batch of graphs: graph-pure tiles.  Inside graph k (plan rows
// [seg_off[k], seg_off[k+1]), degrees descending) hub rows become chunk tiles
// and the other rows tiles of kBlock/lanes rows; each row tile owns one
// partial slot per wave, each hub row one slot of its own.
static int build_segmented_tiles(gbx_ctx *ctx, gbx_graph *g, gbx_plan *p, const DevBuf &deg_new,
                                 const BinConfig &bc)
{
    const uint32_t N = p->N, nseg = (uint32_t)g->seg_off.size() - 1;
    std::vector<uint32_t> deg(N), rph((size_t)N + 1);
    GBX_HIP(ctx, hipMemcpyAsync(deg.data(), deg_new.p, (size_t)N * 4, hipMemcpyDeviceToHost, ctx->stream));
    GBX_HIP(ctx, hipMemcpyAsync(rph.data(), p->rp.p, ((size_t)N + 1) * 4, hipMemcpyDeviceToHost, ctx->stream));
    GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    static const uint32_t lanes[8] = {0, 64, 32, 16, 8, 4, 2, 1};
    const uint32_t wpb = kBlock / 64;
    std::vector<TileDesc> tiles;
    std::vector<HubChunk> chunks;
    std::vector<SegRange> ranges(nseg);
    uint32_t slot = 0, hub = 0;
    for (uint32_t k = 0; k < nseg; k++) {
        const uint32_t r0 = g->seg_off[k], r1 = g->seg_off[k + 1];
        auto first_below = [&](uint32_t t) {
            uint32_t lo = r0, hi = r1;
            while (lo < hi) {
                const uint32_t mid = lo + (hi - lo) / 2;
                if (deg[mid] >= t) lo = mid + 1;
                else hi = mid;
            }
            return lo;
        };
        uint32_t b[7];
        for (int c = 0; c < 7; c++) b[c] = std::max(first_below(bc.thr[c]), c ? b[c - 1] : r0);
        ranges[k].hub_begin = hub;
        for (uint32_t r = r0; r < b[0]; r++, hub++) {
            const uint32_t e0 = rph[r], e1 = rph[r + 1];
            const uint32_t nch = (e1 - e0 + bc.hub_chunk - 1) / bc.hub_chunk;
            const uint32_t first = (uint32_t)chunks.size();
            for (uint32_t c = 0; c < nch; c++) {
                HubChunk hc;
                hc.row = r;
                hc.ebeg = e0 + c * bc.hub_chunk;
                hc.eend = std::min(e1, hc.ebeg + bc.hub_chunk);
                hc.hub = hub;
                hc.first = first;
                hc.nch = nch;
                hc.graph = k;
                hc.pad = 0;
                tiles.push_back(TileDesc{(uint32_t)chunks.size(), 0u, 0u, k, 0u});
                chunks.push_back(hc);
            }
        }
        ranges[k].hub_end = hub;
        ranges[k].slot_begin = slot;
        for (int c = 1; c < 8; c++) {
            const uint32_t beg = b[c - 1], end = c == 7 ? r1 : b[c];
            const uint32_t rpt = kBlock / lanes[c];
            for (uint32_t r = beg; r < end; r += rpt) {
                tiles.push_back(TileDesc{r, end, lanes[c], k, slot});
                slot += wpb;
            }
        }
        ranges[k].slot_end = slot;
    }
    p->segmented = true;
    p->nseg = nseg;
    p->n_seg_tiles = (uint32_t)tiles.size();
    p->n_wave_slots = slot;
    p->n_seg_hubs = hub;
    p->n_seg_chunks = (uint32_t)chunks.size();
    GBX_TRY(dev_alloc(ctx, p->tiles, tiles.size() * sizeof(TileDesc) + 16));
    GBX_TRY(dev_alloc(ctx, p->seg_ranges, ranges.size() * sizeof(SegRange) + 16));
    GBX_TRY(dev_alloc(ctx, p->seg_chunks, chunks.size() * sizeof(HubChunk) + 16));
    GBX_TRY(dev_alloc(ctx, p->seg_hub_cnt, (size_t)hub * 4 + 4));
    GBX_TRY(dev_alloc(ctx, p->seg_hub_part, chunks.size() * 4 + 4));
    GBX_TRY(dev_alloc(ctx, p->row_gid, (size_t)N * 4 + 4));
    GBX_TRY(fill_u32(ctx, p->seg_hub_cnt.as<uint32_t>(), 0, hub + 1));
    if (!tiles.empty())
        GBX_HIP(ctx, hipMemcpyAsync(p->tiles.p, tiles.data(), tiles.size() * sizeof(TileDesc),
                                    hipMemcpyHostToDevice, ctx->stream));
    GBX_HIP(ctx, hipMemcpyAsync(p->seg_ranges.p, ranges.data(), ranges.size() * sizeof(SegRange),
                                hipMemcpyHostToDevice, ctx->stream));
    if (!chunks.empty())
        GBX_HIP(ctx, hipMemcpyAsync(p->seg_chunks.p, chunks.data(), chunks.size() * sizeof(HubChunk),
                                    hipMemcpyHostToDevice, ctx->stream));
    gid_keys<<<grid_for(N, 256, 16384), 256, 0, ctx->stream>>>(p->perm.as<uint32_t>(), g->seg_gid.as<uint32_t>(),
                                                                N, p->row_gid.as<uint32_t>());
    GBX_HIP(ctx, hipGetLastError());
    GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GBX_OK;
}

static int build_plan(gbx_ctx *ctx, gbx_graph *g, int direction, gbx_plan **out)
{
    if (g->E >= 0x80000000u) return set_error(ctx, GBX_ERR_RANGE, "plan: more than 2^31 edges");
    const gbx_csr_dev &rows = direction == GBX_PULL_IN ? g->inv : g->fwd;
    const uint32_t N = g->N, E = g->E;
    gbx_plan *p = new (std::nothrow) gbx_plan();
    if (!p) return set_error(ctx, GBX_ERR_NOMEM, "plan: out of host memory");
    auto fail = [&](int rc) { delete p; return rc; };
    p->direction = direction;
    p->N = N;
    p->E = E;
    int rc;
    uint32_t maxdeg = 0;
    if ((rc = max_u32(ctx, rows.out_degree.as<uint32_t>(), N, &maxdeg))) return fail(rc);
    DevBuf key, deg_new;
    if ((rc = dev_alloc(ctx, key, (size_t)N * 4))) return fail(rc);
    if ((rc = dev_alloc(ctx, deg_new, (size_t)N * 4 + 4))) return fail(rc);
    if ((rc = dev_alloc(ctx, p->perm, (size_t)N * 4))) return fail(rc);
    if ((rc = dev_alloc(ctx, p->inv, (size_t)N * 4))) return fail(rc);
    if ((rc = dev_alloc(ctx, p->outdeg, (size_t)N * 4))) return fail(rc);
    if ((rc = dev_alloc(ctx, p->rp, ((size_t)N + 1) * 4))) return fail(rc);
    if ((rc = dev_alloc(ctx, p->col, (size_t)E * 4))) return fail(rc);
    const unsigned gN = grid_for(N, 256, 16384), gE = grid_for(E, 256, 16384);
    degree_keys<<<gN, 256, 0, ctx->stream>>>(rows.out_degree.as<uint32_t>(), N, maxdeg,
                                              key.as<uint32_t>(), p->perm.as<uint32_t>());
    if (hipGetLastError() != hipSuccess) return fail(set_error(ctx, GBX_ERR_HIP, "degree_keys"));
    // stable sort by (maxdeg - degree): descending degree, ascending id
    if ((rc = radix_sort_pairs(ctx, key.as<uint32_t>(), p->perm.as<uint32_t>(), nullptr, N,
                               bits_for(maxdeg))))
        return fail(rc);
    const bool seg = g->seg_off.size() >= 2;
    if (seg && g->seg_off.size() > 2) {
        // then stably by graph: graph-major, degree-descending inside each graph
        gid_keys<<<gN, 256, 0, ctx->stream>>>(p->perm.as<uint32_t>(), g->seg_gid.as<uint32_t>(), N,
                                               key.as<uint32_t>());
        if (hipGetLastError() != hipSuccess) return fail(set_error(ctx, GBX_ERR_HIP, "gid_keys"));
        if ((rc = radix_sort_pairs(ctx, key.as<uint32_t>(), p->perm.as<uint32_t>(), nullptr, N,
                                   bits_for((uint32_t)g->seg_off.size() - 2))))
            return fail(rc);
    }
    plan_vertices<<<gN, 256, 0, ctx->stream>>>(p->perm.as<uint32_t>(), N, rows.out_degree.as<uint32_t>(),
                                                g->fwd.out_degree.as<uint32_t>(), p->inv.as<uint32_t>(),
                                                deg_new.as<uint32_t>(), p->outdeg.as<uint32_t>());
    if (hipGetLastError() != hipSuccess) return fail(set_error(ctx, GBX_ERR_HIP, "plan_vertices"));
    if ((rc = exclusive_scan_u32(ctx, deg_new.as<uint32_t>(), p->rp.as<uint32_t>(), N,
                                 p->rp.as<uint32_t>() + N)))
        return fail(rc);
    p->has_wt = g->weighted;
    if (p->has_wt && (rc = dev_alloc(ctx, p->wt, (size_t)E * 4))) return fail(rc);
    if (E) {
        plan_edges<<<gE, 256, 0, ctx->stream>>>(rows.src.as<uint32_t>(), rows.dst.as<uint32_t>(),
                                                 rows.w.as<float>(), rows.edges_idx.as<uint32_t>(),
                                                 p->inv.as<uint32_t>(), p->rp.as<uint32_t>(), E,
                                                 p->col.as<uint32_t>(), p->has_wt ? p->wt.as<float>() : nullptr);
        if (hipGetLastError() != hipSuccess) return fail(set_error(ctx, GBX_ERR_HIP, "plan_edges"));
    }
    // bins
    const BinConfig bc = bin_config(ctx, g);
    p->bins_thr_chunk[7] = bc.hub_chunk;
    memcpy(p->bins_thr_chunk, bc.thr, sizeof(bc.thr));
    DevBuf dthr, dbound;
    if ((rc = dev_alloc(ctx, dthr, 64)) || (rc = dev_alloc(ctx, dbound, 64))) return fail(rc);
    uint32_t bound[7] = {0, 0, 0, 0, 0, 0, 0};
    if (hipMemcpyAsync(dthr.p, bc.thr, sizeof(bc.thr), hipMemcpyHostToDevice, ctx->stream) ||
        hipMemcpyAsync(dbound.p, bound, sizeof(bound), hipMemcpyHostToDevice, ctx->stream))
        return fail(set_error(ctx, GBX_ERR_HIP, "plan: threshold upload"));
    class_bounds<<<gN, 256, 0, ctx->stream>>>(deg_new.as<uint32_t>(), N, dthr.as<uint32_t>(), 7,
                                               dbound.as<uint32_t>());
    if (hipGetLastError() != hipSuccess) return fail(set_error(ctx, GBX_ERR_HIP, "class_bounds"));
    if (hipMemcpyAsync(bound, dbound.p, sizeof(bound), hipMemcpyDeviceToHost, ctx->stream) ||
        hipStreamSynchronize(ctx->stream))
        return fail(set_error(ctx, GBX_ERR_HIP, "plan: bounds download"));
    for (int c = 1; c < 7; c++) bound[c] = std::max(bound[c], bound[c - 1]);
    // hub chunks
    const uint32_t n_hubs = bound[0];
    std::vector<uint32_t> hrp(n_hubs + 1);
    if (n_hubs) {
        if (hipMemcpy(hrp.data(), p->rp.p, (n_hubs + 1) * 4, hipMemcpyDeviceToHost))
            return fail(set_error(ctx, GBX_ERR_HIP, "plan: hub rows download"));
    }
    std::vector<HubChunk> chunks;
    for (uint32_t h = 0; h < n_hubs; h++) {
        const uint32_t b = hrp[h], e = hrp[h + 1];
        const uint32_t nch = (e - b + bc.hub_chunk - 1) / bc.hub_chunk;
        const uint32_t first = (uint32_t)chunks.size();
        for (uint32_t k = 0; k < nch; k++) {
            HubChunk hc;
            hc.row = h;
            hc.ebeg = b + k * bc.hub_chunk;
            hc.eend = std::min(e, hc.ebeg + bc.hub_chunk);
            hc.hub = h;
            hc.first = first;
            hc.nch = nch;
            chunks.push_back(hc);
        }
    }
    p->n_hubs = n_hubs;
    p->n_chunks = (uint32_t)chunks.size();
    if ((rc = dev_alloc(ctx, p->chunks, chunks.size() * sizeof(HubChunk) + 16))) return fail(rc);
    if ((rc = dev_alloc(ctx, p->hub_cnt, (size_t)n_hubs * 4 + 4))) return fail(rc);
    if ((rc = dev_alloc(ctx, p->hub_part, chunks.size() * 4 + 4))) return fail(rc);
    if (!chunks.empty() &&
        hipMemcpy(p->chunks.p, chunks.data(), chunks.size() * sizeof(HubChunk), hipMemcpyHostToDevice))
        return fail(set_error(ctx, GBX_ERR_HIP, "plan: chunk upload"));
    if ((rc = fill_u32(ctx, p->hub_cnt.as<uint32_t>(), 0, n_hubs + 1))) return fail(rc);
    // class table: [0] hubs, then 64..2 lanes, then 1 lane
    static const uint32_t lanes[8] = {256, 64, 32, 16, 8, 4, 2, 1};
    uint32_t tile = 0;
    std::vector<uint64_t> hedges(8, 0);
    p->nclass = 8;
    for (int c = 0; c < 8; c++) {
        ClassDesc &cd = p->cls[c];
        cd.lanes = lanes[c];
        cd.row_begin = c == 0 ? 0 : bound[c - 1];
        cd.row_end = c == 7 ? N : bound[c];
        cd.tile_begin = tile;
        if (c == 0) tile += p->n_chunks;
        else {
            const uint32_t rpt = kBlock / lanes[c];
            tile += (cd.row_end - cd.row_begin + rpt - 1) / rpt;
        }
        cd.tile_end = tile;
    }
    p->n_tiles = tile;
    // edges per class (for reporting)
    std::vector<uint32_t> brp(8, 0);
    for (int c = 0; c < 8; c++) {
        uint32_t r0 = p->cls[c].row_begin, r1 = p->cls[c].row_end;
        uint32_t e0 = 0, e1 = 0;
        if (r0 < r1) {
            if (hipMemcpy(&e0, p->rp.as<uint32_t>() + r0, 4, hipMemcpyDeviceToHost) ||
                hipMemcpy(&e1, p->rp.as<uint32_t>() + r1, 4, hipMemcpyDeviceToHost))
                return fail(set_error(ctx, GBX_ERR_HIP, "plan: class edge download"));
        }
        p->cls_edges[c] = (uint64_t)(e1 - e0);
    }
    if (seg && (rc = build_segmented_tiles(ctx, g, p, deg_new, bc))) return fail(rc);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess)
        return fail(set_error(ctx, GBX_ERR_HIP, "plan build failed"));
    *out = p;
    return GBX_OK;
}

int ensure_plan(gbx_ctx *ctx, gbx_graph *g, int direction, gbx_plan **out)
{
    if (direction != GBX_PULL_IN && direction != GBX_PULL_FWD)
        return set_error(ctx, GBX_ERR_ARG, "direction %d not in {0,1}", direction);
    std::shared_ptr<gbx_plan> &slot = g->plan[direction];
    if (slot) {
        // rebuild when the tuning's bins differ from the ones the plan was cut with
        const BinConfig bc = bin_config(ctx, g);
        BinConfig built;
        memcpy(built.thr, slot->bins_thr_chunk, sizeof(built.thr));
        built.hub_chunk = slot->bins_thr_chunk[7];
        if (!(built == bc)) {
            GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
            slot.reset();
        }
    }
    if (!slot) {
        gbx_plan *raw = nullptr;
        GBX_TRY(build_plan(ctx, g, direction, &raw));
        slot.reset(raw);
    }
    *out = slot.get();
    return GBX_OK;
}

static int ensure_normalised(gbx_ctx *ctx, gbx_graph *g, gbx_plan *p)
{
    if (p->has_wn) return GBX_OK;
    if (!p->has_wt) return set_error(ctx, GBX_ERR_STATE, "weighted PageRank needs a weighted graph");
    DevBuf S;
    GBX_TRY(dev_alloc(ctx, S, (size_t)p->N * 4));
    GBX_TRY(dev_alloc(ctx, p->wn, (size_t)p->E * 4));
    out_weight_sums<<<grid_for((size_t)p->N * 64, 256, 16384), 256, 0, ctx->stream>>>(
        g->fwd.out_degree.as<uint32_t>(), g->fwd.edges_idx.as<uint32_t>(), g->fwd.w.as<float>(), p->N,
        S.as<float>());
    GBX_HIP(ctx, hipGetLastError());
    if (p->E) {
        normalise_weights<<<grid_for(p->E, 256, 16384), 256, 0, ctx->stream>>>(
            p->col.as<uint32_t>(), p->wt.as<float>(), p->perm.as<uint32_t>(), S.as<float>(), p->E,
            p->wn.as<float>());
        GBX_HIP(ctx, hipGetLastError());
    }
    GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    p->has_wn = true;
    return GBX_OK;
}

static int ensure_normalised_bf16(gbx_ctx *ctx, gbx_graph *g, gbx_plan *p)
{
    if (p->has_wn16) return GBX_OK;
    GBX_TRY(ensure_normalised(ctx, g, p));
    GBX_TRY(dev_alloc(ctx, p->wn16, (size_t)p->E * 2 + 16));
    if (p->E) {
        to_bf16<<<grid_for(p->E, 256, 16384), 256, 0, ctx->stream>>>(p->wn.as<float>(), p->E,
                                                                      p->wn16.as<uint16_t>());
        GBX_HIP(ctx, hipGetLastError());
    }
    GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    p->has_wn16 = true;
    return GBX_OK;
}

// Row-normalised weights of a forward-direction plan (rows = sources): the
// transition matrix P itself, for the adjoint's pull P lambda.
static int ensure_row_normalised(gbx_ctx *ctx, gbx_graph *g, gbx_plan *p)
{
    if (p->has_wr) return GBX_OK;
    if (p->direction != GBX_PULL_FWD) return set_error(ctx, GBX_ERR_STATE, "row weights need the forward plan");
    if (!g->weighted) return set_error(ctx, GBX_ERR_STATE, "adjoint needs a weighted graph");
    DevBuf S, wfn;
    GBX_TRY(dev_alloc(ctx, S, (size_t)p->N * 4));
    GBX_TRY(dev_alloc(ctx, wfn, (size_t)p->E * 4 + 4));
    GBX_TRY(dev_alloc(ctx, p->wr, (size_t)p->E * 4 + 4));
    out_weight_sums<<<grid_for((size_t)p->N * 64, 256, 16384), 256, 0, ctx->stream>>>(
        g->fwd.out_degree.as<uint32_t>(), g->fwd.edges_idx.as<uint32_t>(), g->fwd.w.as<float>(), p->N,
        S.as<float>());
    GBX_HIP(ctx, hipGetLastError());
    if (p->E) {
        row_normalise_fwd<<<grid_for(p->E, 256, 16384), 256, 0, ctx->stream>>>(
            g->fwd.src.as<uint32_t>(), g->fwd.w.as<float>(), S.as<float>(), p->E, wfn.as<float>());
        GBX_HIP(ctx, hipGetLastError());
        plan_edge_values<<<grid_for(p->E, 256, 16384), 256, 0, ctx->stream>>>(
            g->fwd.src.as<uint32_t>(), g->fwd.edges_idx.as<uint32_t>(), p->inv.as<uint32_t>(),
            p->rp.as<uint32_t>(), p->E, wfn.as<float>(), p->wr.as<float>());
        GBX_HIP(ctx, hipGetLastError());
    }
    GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    p->has_wr = true;
    return GBX_OK;
}


int plan_rows_at_least(gbx_ctx *ctx, const gbx_plan *p, uint32_t threshold, uint32_t *count)
{
    DevBuf dthr, dbound;
    GBX_TRY(dev_alloc(ctx, dthr, 16));
    GBX_TRY(dev_alloc(ctx, dbound, 16));
    uint32_t thr = threshold, bound = 0;
    GBX_HIP(ctx, hipMemcpyAsync(dthr.p, &thr, 4, hipMemcpyHostToDevice, ctx->stream));
    GBX_TRY(fill_u32(ctx, dbound.as<uint32_t>(), 0, 1));
    // degrees of the plan rows from the row pointers
    DevBuf deg;
    GBX_TRY(dev_alloc(ctx, deg, (size_t)p->N * 4 + 4));
    plan_degrees<<<grid_for(p->N, 256, 16384), 256, 0, ctx->stream>>>(p->rp.as<uint32_t>(), p->N,
                                                                        deg.as<uint32_t>());
    GBX_HIP(ctx, hipGetLastError());
    class_bounds<<<grid_for(p->N, 256, 16384), 256, 0, ctx->stream>>>(deg.as<uint32_t>(), p->N,
                                                                       dthr.as<uint32_t>(), 1,
                                                                       dbound.as<uint32_t>());
    GBX_HIP(ctx, hipGetLastError());
    GBX_HIP(ctx, hipMemcpyAsync(&bound, dbound.p, 4, hipMemcpyDeviceToHost, ctx->stream));
    GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *count = bound;
    return GBX_OK;
}

int plan_row_ptr_prefix(gbx_ctx *ctx, const gbx_plan *p, uint32_t nrows, std::vector<uint32_t> &out)
{
    out.resize((size_t)nrows + 1);
    GBX_HIP(ctx, hipMemcpyAsync(out.data(), p->rp.p, ((size_t)nrows + 1) * 4, hipMemcpyDeviceToHost,
                                ctx->stream));
    GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GBX_OK;
}

PlanView plan_view(const gbx_plan *p, const float *w)
{
    PlanView v;
    v.rp = p->rp.as<uint32_t>();
    v.col = p->col.as<uint32_t>();
    v.wt = w;
    v.outdeg = p->outdeg.as<uint32_t>();
    v.chunks = p->chunks.as<HubChunk>();
    v.hub_cnt = p->hub_cnt.as<uint32_t>();
    v.hub_part = p->hub_part.as<float>();
    v.N = p->N;
    v.n_tiles = p->n_tiles;
    v.n_chunks = p->n_chunks;
    v.nclass = p->nclass;
    for (uint32_t c = 0; c < kMaxClasses; c++) v.cls[c] = p->cls[c];
    return v;
}

unsigned iter_grid(const gbx_ctx *ctx, const gbx_plan *p)
{
    const unsigned per_cu = ctx->tune.blocks_per_cu;
    unsigned g = (unsigned)std::max(1, ctx->num_cus) * per_cu;
    return std::max(1u, std::min(g, p->n_tiles));
}

// ------------------------------------------------------------------ PR object

static IterView iter_view(gbx_pr *pr)
{
    IterView a;
    a.r = pr->r.as<float>();
    a.c_in = pr->c[pr->t & 1].p;
    a.c_out = pr->c[(pr->t + 1) & 1].p;
    a.tele = pr->has_tele ? pr->tele.as<float>() : nullptr;
    a.slots = pr->slots.as<Slot>();
    a.partials = pr->partials.as<Partial>();
    a.counter = pr->counter.as<uint32_t>();
    a.t = pr->t;
    a.hot = pr->hot;
    a.damp = pr->p.damp;
    a.eps = pr->p.epsilon;
    a.stop = pr->p.stop;
    return a;
}

template <int KIND, class T, class Acc>
static void launch_pr_kernel(bool lds, unsigned grid, size_t shm, hipStream_t st, const PlanView &P,
                             const IterView &A)
{
    if (lds) {
        // more than 64 KiB of dynamic LDS needs the per-kernel opt-in
        static size_t granted = 64 * 1024;
        if (shm > granted &&
            hipFuncSetAttribute(reinterpret_cast<const void *>(&pr_iter_kernel<KIND, T, Acc, true>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm) == hipSuccess)
            granted = shm;
        pr_iter_kernel<KIND, T, Acc, true><<<grid, kBlock, shm, st>>>(P, A);
    } else {
        pr_iter_kernel<KIND, T, Acc, false><<<grid, kBlock, 0, st>>>(P, A);
    }
}

static const float *plan_weights(const gbx_pr *pr)
{
    if (pr->kind == K_ADJ) return pr->plan->wr.as<float>();
    if (pr->kind != K_PR_STOCH_W) return nullptr;
    if (pr->elem == 2) return reinterpret_cast<const float *>(pr->plan->wn16.p);
    return pr->plan->wn.as<float>();
}

static int launch_iter(gbx_pr *pr)
{
    if (pr->batch) return ppr_launch_iter(pr);
    if (pr->seg) return seg_launch_iter(pr);
    if (pr->bin) return binned_launch_iter(pr);
    if (pr->ddp) return ddp_launch_iter(pr);
    gbx_ctx *ctx = pr->ctx;
    PlanView P = plan_view(pr->plan, plan_weights(pr));
    IterView A = iter_view(pr);
    const bool lds = pr->hot > 0;
    const size_t shm = (size_t)pr->hot * pr->elem;
    hipStream_t st = ctx->stream;
    const unsigned g = pr->grid;
    if (pr->elem == 4) {
        switch (pr->kind) {
        case K_PR_REF: launch_pr_kernel<K_PR_REF, float, AccF32>(lds, g, shm, st, P, A); break;
        case K_PR_STOCH: launch_pr_kernel<K_PR_STOCH, float, AccF32>(lds, g, shm, st, P, A); break;
        case K_ADJ: launch_pr_kernel<K_ADJ, float, AccF32>(lds, g, shm, st, P, A); break;
        case K_DD: launch_pr_kernel<K_DD, float, AccDD>(lds, g, shm, st, P, A); break;
        default: launch_pr_kernel<K_PR_STOCH_W, float, AccF32>(lds, g, shm, st, P, A); break;
        }
    } else {
        switch (pr->kind) {
        case K_PR_REF: launch_pr_kernel<K_PR_REF, uint16_t, AccKahan>(lds, g, shm, st, P, A); break;
        case K_PR_STOCH: launch_pr_kernel<K_PR_STOCH, uint16_t, AccKahan>(lds, g, shm, st, P, A); break;
        default: launch_pr_kernel<K_PR_STOCH_W, uint16_t, AccKahan>(lds, g, shm, st, P, A); break;
        }
    }
    GBX_HIP(ctx, hipGetLastError());
    pr->t++;
    return GBX_OK;
}

// Contributions kept in LDS per block: GBX_LDS_BYTES (default below), a
// multiple of 16 bytes, never more than the vector itself.
static uint32_t hot_count(const gbx_ctx *ctx, uint32_t N, size_t elem)
{
    size_t bytes = ctx->tune.lds_bytes;
    if (bytes > 160 * 1024) bytes = 160 * 1024;
    size_t n = bytes / elem;
    const size_t align = 16 / elem;
    if (n > N) n = N;
    n -= n % align;
    return (uint32_t)n;
}

// ------------------------------------------------------------------ fixed-point SpMV
// SPMVPullFixedPointGraphCSR (SPMV.c:313-449, 32 bits) and
// SPMVPushFixedPointGraphCSR (SPMV.c:451-590, 64 bits; it pushes the same
// integer terms, so a pull gives its sums): per edge an integer product of
// fixed-point weight and input, added in uint64 (order-free, so bit-exact).
// 32: weight = (float)FloatToFixed32(w), x = FloatToFixed32(x), term =
// ((u64)weight * x) >> 16; 64: DoubleToFixed64 of both, term = ((u128) w * x)
// >> 32 mod 2^64.  Float to integer conversions are x86-64's: a 64-bit
// truncation (low 32 bits for u32).
__device__ __forceinline__ uint32_t fix32_of(float v)
{
    return (uint32_t)(unsigned long long)(long long)(v * 65536.0f + 0.5f);
}

__device__ __forceinline__ uint64_t fix64_of(float v)
{
    const double d = (double)v * 4294967296.0 + 0.5;
    return d >= 9223372036854775808.0 ? (uint64_t)(d - 9223372036854775808.0) + (1ull << 63) : (uint64_t)d;
}

__global__ void spmv_fixed_input(const uint32_t *__restrict__ outdeg, uint32_t N, int bits, float *__restrict__ x,
                                 uint64_t *__restrict__ xq, uint64_t *__restrict__ y)
{
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        const uint32_t od = outdeg[v];
        const float xv = od ? 1.0f / (float)od : 0.001f;  // SPMV.c:156-163
        x[v] = xv;
        xq[v] = bits == 32 ? (uint64_t)fix32_of(xv) : fix64_of(xv);
        y[v] = 0;
    }
}

// Edge-parallel over a CSR whose edges are sorted by row: each wave takes 64
// consecutive edges, adds the terms of equal rows with a segmented scan, and
// the last lane of each row run adds its sum with one atomic.
template <int BITS>
__global__ void spmv_fixed_kernel(const uint32_t *__restrict__ row, const uint32_t *__restrict__ col,
                                  const float *__restrict__ w, uint32_t E, const uint64_t *__restrict__ xq,
                                  uint64_t wconst, uint64_t *__restrict__ y)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t nwaves = gridDim.x * (blockDim.x / 64);
    for (uint32_t base = (blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6)) * 64u; base < E;
         base += nwaves * 64u) {
        const uint32_t j = base + lane;
        const bool valid = j < E;
        const uint32_t r = valid ? row[j] : 0xFFFFFFFFu;
        uint64_t t = 0;
        if (valid) {
            const uint64_t x = xq[col[j] & kIdMask];
            if constexpr (BITS == 32) {
                const float wf = w ? (float)fix32_of(w[j]) : (float)wconst;
                t = ((uint64_t)wf * x) >> 16;
            } else {
                const uint64_t wq = w ? fix64_of(w[j]) : wconst;
                t = (__umul64hi(wq, x) << 32) | ((wq * x) >> 32);
            }
        }
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint64_t o = __shfl_up(t, off, 64);
            const uint32_t ro = __shfl_up(r, off, 64);
            if ((int)lane >= off && ro == r) t += o;
        }
        const uint32_t rn = __shfl_down(r, 1, 64);
        if (valid && (lane == 63 || rn != r))
            atomicAdd(reinterpret_cast<unsigned long long *>(y + r), (unsigned long long)t);
    }
}

__global__ void spmv_fixed_output(const uint64_t *__restrict__ y, uint32_t N, int bits, float *__restrict__ out)
{
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x)
        out[v] = bits == 32 ? (float)y[v] / 65536.0f : (float)((double)y[v] / 4294967296.0);
}

}  // namespace gbx

using namespace gbx;

extern "C" {

void gbx_pr_params_default(gbx_pr_params *p)
{
    if (!p) return;
    memset(p, 0, sizeof(*p));
    p->damp = 0.85f;     // pageRank.h:13
    p->max_iters = 20;   // graphConfig.c argumentsNew
    p->epsilon = 1e-4;   // graphConfig.c argumentsNew
    p->mode = GBX_PR_REF;
    p->stop = GBX_STOP_ACTIVE;
    p->precision = GBX_F32;
    p->direction = GBX_PULL_IN;
    p->weighted = 0;
}

static int pr_create_kind(gbx_ctx *ctx, gbx_graph *g, const gbx_pr_params *p, int kind, gbx_pr **out);

int gbx_pr_create(gbx_ctx *ctx, gbx_graph *g, const gbx_pr_params *p, gbx_pr **out)
{
    return pr_create_kind(ctx, g, p, -1, out);
}

static int pr_create_kind(gbx_ctx *ctx, gbx_graph *g, const gbx_pr_params *p, int kind_override, gbx_pr **out)
{
    if (!ctx || !g || !p || !out) return set_error(ctx, GBX_ERR_ARG, "gbx_pr_create: NULL argument");
    *out = nullptr;
    if (p->mode < GBX_PR_REF || p->mode > GBX_PR_DD_PULL_PUSH)
        return set_error(ctx, GBX_ERR_ARG, "gbx_pr_create: mode %d", p->mode);
    const bool dd = p->mode == GBX_PR_DATA_DRIVEN;
    const bool ddp = p->mode == GBX_PR_DD_PUSH || p->mode == GBX_PR_DD_PULL_PUSH;
    if ((dd || ddp) && (p->weighted || p->precision != GBX_F32 || p->method == GBX_METHOD_BINNED))
        return set_error(ctx, GBX_ERR_ARG, "gbx_pr_create: the data-driven modes are unweighted FP32 and do not "
                                           "run on the binned method");
    if (ddp && g->N == 0) return set_error(ctx, GBX_ERR_ARG, "gbx_pr_create: empty graph");
    const bool fixed = p->mode >= GBX_PR_FIXED64 && p->mode <= GBX_PR_FIXED8;
    const bool quant = p->mode >= GBX_PR_QUANT32 && p->mode <= GBX_PR_QUANT8;
    if ((fixed || quant) && (p->weighted || p->method == GBX_METHOD_PULL))
        return set_error(ctx, GBX_ERR_ARG, "gbx_pr_create: the fixed-point and quantized modes are unweighted "
                                           "and run on the binned method");
    if (p->stop < 0 || p->stop > 2) return set_error(ctx, GBX_ERR_ARG, "gbx_pr_create: stop %d", p->stop);
    if (p->precision != GBX_F32 && p->precision != GBX_BF16)
        return set_error(ctx, GBX_ERR_ARG, "gbx_pr_create: precision %d not supported", p->precision);
    if (p->weighted && p->mode != GBX_PR_STOCHASTIC)
        return set_error(ctx, GBX_ERR_ARG, "gbx_pr_create: weighted PageRank needs STOCHASTIC mode");
    if (!(p->damp >= 0.0f && p->damp < 1.0f))
        return set_error(ctx, GBX_ERR_ARG, "gbx_pr_create: damp %f not in [0,1)", (double)p->damp);
    GBX_HIP(ctx, hipSetDevice(ctx->device));
    gbx_plan *plan = nullptr;
    GBX_TRY(ensure_plan(ctx, g, p->direction, &plan));
    if (plan->segmented) {
        if (p->mode != GBX_PR_STOCHASTIC || p->precision != GBX_F32 || p->weighted || p->stop == GBX_STOP_ACTIVE)
            return set_error(ctx, GBX_ERR_ARG,
                             "gbx_pr_create: a batch of graphs runs STOCHASTIC, F32, unweighted, stop L1 or NONE");
        gbx_pr *pr = new (std::nothrow) gbx_pr();
        if (!pr) return set_error(ctx, GBX_ERR_NOMEM, "gbx_pr_create: out of host memory");
        pr->ctx = ctx;
        pr->g = g;
        pr->plan = plan;
        pr->plan_ref = g->plan[p->direction];
        pr->p = *p;
        // the binned iteration from binned_min_vertices (all graphs together), else graph-pure pull tiles
        if (p->method != GBX_METHOD_AUTO && p->method != GBX_METHOD_PULL && p->method != GBX_METHOD_BINNED) {
            delete pr;
            return set_error(ctx, GBX_ERR_ARG, "gbx_pr_create: method %d", p->method);
        }
        const bool binned = p->method == GBX_METHOD_BINNED ||
                            (p->method == GBX_METHOD_AUTO && g->N >= ctx->tune.binned_min_vertices);
        pr->kind = binned ? K_SEG : K_PR_STOCH;
        pr->elem = 4;
        int rc = seg_create(pr);
        if (!rc && (hipEventCreate(&pr->ev_begin) != hipSuccess || hipEventCreate(&pr->ev_end) != hipSuccess))
            rc = set_error(ctx, GBX_ERR_HIP, "gbx_pr_create: event creation");
        if (rc) {
            gbx_pr_free(pr);
            return rc;
        }
        *out = pr;
        return GBX_OK;
    }
    if (kind_override == K_SPMV_W) {
        if (!plan->has_wt) return set_error(ctx, GBX_ERR_STATE, "weighted SpMV needs a weighted graph");
    } else if (kind_override == K_ADJ) {
        GBX_TRY(ensure_row_normalised(ctx, g, plan));
    } else if (p->weighted) {
        if (p->precision == GBX_BF16) GBX_TRY(ensure_normalised_bf16(ctx, g, plan));
        else GBX_TRY(ensure_normalised(ctx, g, plan));
    }
    gbx_pr *pr = new (std::nothrow) gbx_pr();
    if (!pr) return set_error(ctx, GBX_ERR_NOMEM, "gbx_pr_create: out of host memory");
    pr->ctx = ctx;
    pr->g = g;
    pr->plan = plan;
    pr->plan_ref = g->plan[p->direction];
    pr->p = *p;
    pr->kind = p->mode == GBX_PR_REF ? K_PR_REF : (p->weighted ? K_PR_STOCH_W : K_PR_STOCH);
    if (fixed) pr->kind = K_FIX64 + (p->mode - GBX_PR_FIXED64);
    if (quant) pr->kind = K_Q32 + (p->mode - GBX_PR_QUANT32);
    if (dd) pr->kind = K_DD;
    if (ddp) pr->kind = p->mode == GBX_PR_DD_PUSH ? K_DDP : K_DDPP;
    if (kind_override >= 0) pr->kind = kind_override;
    pr->grid = iter_grid(ctx, plan);
    pr->elem = p->precision == GBX_BF16 ? 2 : 4;
    if (fixed) pr->elem = (size_t)8 >> (p->mode - GBX_PR_FIXED64);  // 8, 4, 2, 1 bytes
    if (quant) pr->elem = 4;  // FP32 contributions; quantized when loaded
    pr->hot = hot_count(ctx, g->N, pr->elem);
    const size_t nb = (size_t)g->N * 4;
    const size_t cb = (size_t)g->N * pr->elem + 16;
    int rc = GBX_OK;
    if (!rc) rc = dev_alloc(ctx, pr->r, nb);
    if (!rc) rc = dev_alloc(ctx, pr->c[0], cb);
    if (!rc) rc = dev_alloc(ctx, pr->c[1], cb);
    if (!rc) rc = dev_alloc(ctx, pr->out, nb);
    if (!rc) rc = dev_alloc(ctx, pr->slots, 3 * sizeof(Slot));  // ping-pong, and the read-back's rank sum
    if (!rc) rc = dev_alloc(ctx, pr->partials, (size_t)std::max(pr->grid, 16384u) * sizeof(Partial));
    if (!rc) rc = dev_alloc(ctx, pr->counter, 64);
    if (!rc) rc = fill_u32(ctx, pr->counter.as<uint32_t>(), 0, 16);
    if (!rc && (hipEventCreate(&pr->ev_begin) != hipSuccess || hipEventCreate(&pr->ev_end) != hipSuccess))
        rc = set_error(ctx, GBX_ERR_HIP, "gbx_pr_create: event creation");
    if (!rc && hipHostMalloc((void **)&pr->h_slot, 2 * sizeof(Slot), 0) != hipSuccess)
        rc = set_error(ctx, GBX_ERR_NOMEM, "gbx_pr_create: pinned slot");
    if (!rc) {
        bool binned = false;
        if (p->method == GBX_METHOD_BINNED) {
            if (!binned_supported(pr))
                rc = set_error(ctx, GBX_ERR_ARG, "gbx_pr_create: the binned method runs REF, STOCHASTIC (FP32 or "
                                                 "BF16) and weighted STOCHASTIC (FP32) PageRank on one graph");
            binned = true;
        } else if (p->method == GBX_METHOD_AUTO) {
            binned = (kind_override < 0 || kind_override == K_ADJ) && binned_supported(pr) &&
                     (fixed || quant || g->N >= ctx->tune.binned_min_vertices);
        } else if (p->method != GBX_METHOD_PULL) {
            rc = set_error(ctx, GBX_ERR_ARG, "gbx_pr_create: method %d", p->method);
        }
        if (!rc && !binned && (fixed || quant))
            rc = set_error(ctx, GBX_ERR_ARG, "gbx_pr_create: the fixed-point and quantized modes run on the binned "
                                             "method, which this graph does not support");
        if (!rc && binned) rc = binned_create(pr);
        if (!rc && ddp) rc = ddp_create(pr);
    }
    if (rc) {
        gbx_pr_free(pr);
        return rc;
    }
    *out = pr;
    return GBX_OK;
}

int gbx_pr_reset(gbx_pr *pr, const float *teleport)
{
    if (!pr) return set_error(nullptr, GBX_ERR_ARG, "gbx_pr_reset: NULL");
    gbx_ctx *ctx = pr->ctx;
    GBX_HIP(ctx, hipSetDevice(ctx->device));
    const uint32_t N = pr->g->N;
    pr->has_tele = false;
    if (pr->batch || pr->seg) {
        if (teleport) return set_error(ctx, GBX_ERR_ARG, "gbx_pr_reset: batched teleports are fixed at creation");
        pr->timing_started = false;
        pr->elapsed_ms = 0.0;
        return pr->batch ? ppr_reset(pr) : seg_reset(pr);
    }
    if (teleport) {
        if (pr->p.mode != GBX_PR_STOCHASTIC && pr->kind != K_ADJ)
            return set_error(ctx, GBX_ERR_ARG, "gbx_pr_reset: teleport needs STOCHASTIC mode");
        DevBuf tmp;
        GBX_TRY(dev_alloc(ctx, tmp, (size_t)N * 4));
        if (!pr->tele.p) GBX_TRY(dev_alloc(ctx, pr->tele, (size_t)N * 4));
        GBX_HIP(ctx, hipMemcpyAsync(tmp.p, teleport, (size_t)N * 4, hipMemcpyHostToDevice, ctx->stream));
        gather_perm<<<grid_for(N, 256, 16384), 256, 0, ctx->stream>>>(
            tmp.as<float>(), pr->plan->perm.as<uint32_t>(), N, pr->tele.as<float>());
        GBX_HIP(ctx, hipGetLastError());
        GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        pr->has_tele = true;
    }
    pr->t = 0;
    pr->timing_started = false;
    pr->elapsed_ms = 0.0;
    if (pr->ddp) return ddp_reset(pr);
    PlanView P = plan_view(pr->plan, plan_weights(pr));
    IterView A = iter_view(pr);
    A.c_out = pr->c[0].p;
    const unsigned grid = grid_for(N, kBlock, (unsigned)std::max(1, ctx->num_cus) * 8);
    hipStream_t st = ctx->stream;
    if (is_fixed_kind(pr->kind)) {
        switch (pr->kind) {
        case K_FIX64: pr_reset_kernel<K_FIX64, uint64_t><<<grid, kBlock, 0, st>>>(P, A); break;
        case K_FIX32: pr_reset_kernel<K_FIX32, uint32_t><<<grid, kBlock, 0, st>>>(P, A); break;
        case K_FIX16: pr_reset_kernel<K_FIX16, uint16_t><<<grid, kBlock, 0, st>>>(P, A); break;
        default: pr_reset_kernel<K_FIX8, uint8_t><<<grid, kBlock, 0, st>>>(P, A); break;
        }
    } else if (is_quant_kind(pr->kind)) {
        // contributions and their min/max do not depend on the width
        pr_reset_kernel<K_Q32, float><<<grid, kBlock, 0, st>>>(P, A);
    } else if (pr->elem == 4) {
        switch (pr->kind) {
        case K_PR_REF:
        case K_DD: pr_reset_kernel<K_PR_REF, float><<<grid, kBlock, 0, st>>>(P, A); break;
        case K_PR_STOCH: pr_reset_kernel<K_PR_STOCH, float><<<grid, kBlock, 0, st>>>(P, A); break;
        case K_ADJ: pr_reset_kernel<K_ADJ, float><<<grid, kBlock, 0, st>>>(P, A); break;
        default: pr_reset_kernel<K_PR_STOCH_W, float><<<grid, kBlock, 0, st>>>(P, A); break;
        }
    } else {
        switch (pr->kind) {
        case K_PR_REF: pr_reset_kernel<K_PR_REF, uint16_t><<<grid, kBlock, 0, st>>>(P, A); break;
        case K_PR_STOCH: pr_reset_kernel<K_PR_STOCH, uint16_t><<<grid, kBlock, 0, st>>>(P, A); break;
        default: pr_reset_kernel<K_PR_STOCH_W, uint16_t><<<grid, kBlock, 0, st>>>(P, A); break;
        }
    }
    GBX_HIP(ctx, hipGetLastError());
    return GBX_OK;
}

int gbx_pr_launch(gbx_pr *pr, uint32_t n_iters)
{
    if (!pr) return set_error(nullptr, GBX_ERR_ARG, "gbx_pr_launch: NULL");
    gbx_ctx *ctx = pr->ctx;
    GBX_HIP(ctx, hipSetDevice(ctx->device));
    if (!pr->timing_started && n_iters) {
        GBX_HIP(ctx, hipEventRecord(pr->ev_begin, ctx->stream));
        pr->timing_started = true;
    }
    for (uint32_t k = 0; k < n_iters; k++) GBX_TRY(launch_iter(pr));
    if (n_iters) GBX_HIP(ctx, hipEventRecord(pr->ev_end, ctx->stream));
    return GBX_OK;
}

int gbx_pr_launch_graph(gbx_pr *pr, uint32_t n_iters)
{
    if (!pr) return set_error(nullptr, GBX_ERR_ARG, "gbx_pr_launch_graph: NULL");
    gbx_ctx *ctx = pr->ctx;
    GBX_HIP(ctx, hipSetDevice(ctx->device));
    if (n_iters == 0) return GBX_OK;
    // the captured launches hold the teleport pointer (or none) and the
    // ping-pong parity they start from
    const uint64_t key = ((uint64_t)n_iters << 2) | ((pr->has_tele ? 1u : 0u) << 1) | (pr->t & 1u);
    hipGraphExec_t exec = nullptr;
    for (auto &e : pr->graph_cache)
        if (e.first == key) exec = e.second;
    if (!exec) {
        // capture n iterations (kernel launches only: no memsets, copies or syncs)
        const uint32_t t0 = pr->t;
        GBX_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
        int rc = GBX_OK;
        for (uint32_t k = 0; k < n_iters && !rc; k++) rc = launch_iter(pr);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
        pr->t = t0;
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess) return set_error(ctx, GBX_ERR_HIP, "capture failed: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return set_error(ctx, GBX_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
        pr->graph_cache.emplace_back(key, exec);
    }
    if (!pr->timing_started) {
        GBX_HIP(ctx, hipEventRecord(pr->ev_begin, ctx->stream));
        pr->timing_started = true;
    }
    GBX_HIP(ctx, hipGraphLaunch(exec, ctx->stream));
    GBX_HIP(ctx, hipEventRecord(pr->ev_end, ctx->stream));
    pr->t += n_iters;
    return GBX_OK;
}

int gbx_pr_launch_timed(gbx_pr *pr, uint32_t n_iters, double *kernel_ms_avg)
{
    if (!pr) return set_error(nullptr, GBX_ERR_ARG, "gbx_pr_launch_timed: NULL");
    gbx_ctx *ctx = pr->ctx;
    GBX_HIP(ctx, hipSetDevice(ctx->device));
    std::vector<hipEvent_t> ev(2 * (size_t)n_iters, nullptr);
    int rc = GBX_OK;
    for (auto &e : ev)
        if (hipEventCreate(&e) != hipSuccess) rc = set_error(ctx, GBX_ERR_HIP, "event creation");
    if (!rc && !pr->timing_started && n_iters) {
        if (hipEventRecord(pr->ev_begin, ctx->stream) != hipSuccess) rc = set_error(ctx, GBX_ERR_HIP, "event");
        pr->timing_started = true;
    }
    for (uint32_t k = 0; k < n_iters && !rc; k++) {
        if (hipEventRecord(ev[2 * k], ctx->stream) != hipSuccess) rc = set_error(ctx, GBX_ERR_HIP, "event");
        if (!rc) rc = launch_iter(pr);
        if (!rc && hipEventRecord(ev[2 * k + 1], ctx->stream) != hipSuccess)
            rc = set_error(ctx, GBX_ERR_HIP, "event");
    }
    if (!rc && n_iters && hipEventRecord(pr->ev_end, ctx->stream) != hipSuccess)
        rc = set_error(ctx, GBX_ERR_HIP, "event");
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess)
        rc = set_error(ctx, GBX_ERR_HIP, "sync: %s", hipGetErrorString(hipGetLastError()));
    double tot = 0.0;
    for (uint32_t k = 0; k < n_iters && !rc; k++) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]) != hipSuccess)
            rc = set_error(ctx, GBX_ERR_HIP, "event elapsed");
        tot += ms;
    }
    for (auto &e : ev)
        if (e) (void)hipEventDestroy(e);
    if (!rc && kernel_ms_avg) *kernel_ms_avg = n_iters ? tot / n_iters : 0.0;
    return rc;
}

int gbx_pr_read(gbx_pr *pr, float *ranks_out, gbx_pr_report *rep)
{
    if (!pr) return set_error(nullptr, GBX_ERR_ARG, "gbx_pr_read: NULL");
    gbx_ctx *ctx = pr->ctx;
    GBX_HIP(ctx, hipSetDevice(ctx->device));
    if (pr->batch || pr->seg) {
        GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (pr->timing_started) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, pr->ev_begin, pr->ev_end) == hipSuccess) pr->elapsed_ms = ms;
        }
        return pr->batch ? ppr_read(pr, ranks_out, rep) : seg_read(pr, ranks_out, rep);
    }
    const uint32_t N = pr->g->N;
    const Slot *slot = pr->slots.as<Slot>() + (pr->t & 1);
    // REF-update kinds divide by N, the adjoint is raw, the stochastic forms divide by the rank sum
    const bool ref_update =
        pr->kind == K_PR_REF || pr->kind == K_DD || pr->kind == K_DDP || pr->kind == K_DDPP || is_int_kind(pr->kind);
    // the data-driven push ranks are kept in vertex order, the others in plan order
    const uint32_t *perm = pr->ddp ? nullptr : pr->plan->perm.as<uint32_t>();
    pr_finalize_kernel<<<grid_for(N, 256, 16384), 256, 0, ctx->stream>>>(
        pr->r.as<float>(), perm, N, slot, pr->kind == K_ADJ ? 2 : (ref_update ? 1 : 0), pr->out.as<float>());
    GBX_HIP(ctx, hipGetLastError());
    Slot *sum_slot = pr->slots.as<Slot>() + 2;
    pr_out_sum_kernel<<<std::max(1u, std::min((N + kBlock - 1) / kBlock, 4u * (unsigned)std::max(1, ctx->num_cus))),
                        kBlock, 0, ctx->stream>>>(pr->out.as<float>(), N, pr->partials.as<Partial>(),
                                                  pr->counter.as<uint32_t>(), sum_slot);
    GBX_HIP(ctx, hipGetLastError());
    GBX_HIP(ctx, hipMemcpyAsync(pr->h_slot, slot, sizeof(Slot), hipMemcpyDeviceToHost, ctx->stream));
    GBX_HIP(ctx, hipMemcpyAsync(pr->h_slot + 1, sum_slot, sizeof(Slot), hipMemcpyDeviceToHost, ctx->stream));
    if (ranks_out) GBX_TRY(copy_to_host(ctx, ranks_out, pr->out.p, (size_t)N * 4));
    GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (pr->timing_started) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, pr->ev_begin, pr->ev_end) == hipSuccess) pr->elapsed_ms = ms;
    }
    if (rep) {
        const Slot s = *pr->h_slot;
        rep->executed = s.iters;
        rep->converged = s.conv;
        rep->iterations = s.conv ? (s.iters ? s.iters - 1 : 0) : s.iters;
        rep->active = s.active;
        rep->error_total = s.err;
        rep->rank_sum = pr->h_slot[1].sum;  // FP64 sum of the returned ranks, summed on the GPU
        rep->time_total_s = pr->elapsed_ms * 1e-3;
    }
    return GBX_OK;
}

int gbx_pr_free(gbx_pr *pr)
{
    if (!pr) return GBX_OK;
    (void)hipSetDevice(pr->ctx->device);
    (void)hipStreamSynchronize(pr->ctx->stream);
    if (pr->ev_begin) (void)hipEventDestroy(pr->ev_begin);
    if (pr->ev_end) (void)hipEventDestroy(pr->ev_end);
    if (pr->h_slot) (void)hipHostFree(pr->h_slot);
    if (pr->batch) ppr_state_free(pr->batch);
    if (pr->seg) seg_state_free(pr->seg);
    if (pr->bin) binned_state_free(pr->bin);
    if (pr->ddp) ddp_state_free(pr->ddp);
    for (auto &e : pr->graph_cache) (void)hipGraphExecDestroy(e.second);
    delete pr;
    return GBX_OK;
}

int gbx_pr_pull(gbx_ctx *ctx, gbx_graph *g, const gbx_pr_params *p, const float *teleport,
                float *ranks_out, gbx_pr_report *rep)
{
    gbx_pr *pr = nullptr;
    GBX_TRY(gbx_pr_create(ctx, g, p, &pr));
    int rc = gbx_pr_reset(pr, teleport);
    // launch in chunks; after each, read the latest slot and stop once done
    const uint32_t chunk = 8;
    uint32_t launched = 0;
    while (!rc && launched < p->max_iters) {
        const uint32_t n = std::min(chunk, p->max_iters - launched);
        rc = gbx_pr_launch(pr, n);
        launched += n;
        if (rc) break;
        const Slot *slot = pr->slots.as<Slot>() + (pr->t & 1);
        if (hipMemcpyAsync(pr->h_slot, slot, sizeof(Slot), hipMemcpyDeviceToHost, ctx->stream) ||
            hipStreamSynchronize(ctx->stream))
            rc = set_error(ctx, GBX_ERR_HIP, "gbx_pr_pull: %s", hipGetErrorString(hipGetLastError()));
        else if (pr->h_slot->done)
            break;
    }
    if (!rc) rc = gbx_pr_read(pr, ranks_out, rep);
    gbx_pr_free(pr);
    return rc;
}

// Launch in chunks of 8 iterations until the stop rule fires or max_iters.
static int run_to_stop(gbx_pr *pr)
{
    gbx_ctx *ctx = pr->ctx;
    uint32_t launched = 0;
    while (launched < pr->p.max_iters) {
        const uint32_t n = std::min(8u, pr->p.max_iters - launched);
        GBX_TRY(gbx_pr_launch(pr, n));
        launched += n;
        const Slot *slot = pr->slots.as<Slot>() + (pr->t & 1);
        GBX_HIP(ctx, hipMemcpyAsync(pr->h_slot, slot, sizeof(Slot), hipMemcpyDeviceToHost, ctx->stream));
        GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (pr->h_slot->done) break;
    }
    return GBX_OK;
}

int gbx_pr_adjoint(gbx_ctx *ctx, gbx_graph *g, const gbx_pr_params *p, const float *teleport,
                   const float *loss_weights, float *ranks_out, float *lambda_out, float *grad_w,
                   float *grad_teleport, gbx_pr_report *fwd_rep, gbx_pr_report *adj_rep)
{
    if (!ctx || !g || !p || !loss_weights) return set_error(ctx, GBX_ERR_ARG, "gbx_pr_adjoint: NULL argument");
    if (!g->weighted) return set_error(ctx, GBX_ERR_STATE, "gbx_pr_adjoint: graph has no weights");
    if (g->seg_off.size() >= 2) return set_error(ctx, GBX_ERR_STATE, "gbx_pr_adjoint: not for a batch of graphs");
    if (p->precision != GBX_F32 || p->stop == GBX_STOP_ACTIVE)
        return set_error(ctx, GBX_ERR_ARG, "gbx_pr_adjoint: F32 with stop L1 or NONE");
    GBX_HIP(ctx, hipSetDevice(ctx->device));
    const uint32_t N = g->N, E = g->E;
    gbx_pr_params q = *p;
    q.mode = GBX_PR_STOCHASTIC;
    q.weighted = 1;
    q.direction = GBX_PULL_IN;
    gbx_pr *fwd = nullptr, *adj = nullptr;
    int rc = gbx_pr_create(ctx, g, &q, &fwd);
    if (!rc) rc = gbx_pr_reset(fwd, teleport);
    if (!rc) rc = run_to_stop(fwd);
    if (!rc) rc = gbx_pr_read(fwd, ranks_out, fwd_rep);
    // adjoint fixed point on the forward CSR: lambda = g + d (P lambda + delta Lambda/N)
    gbx_pr_params qa = q;
    qa.direction = GBX_PULL_FWD;
    if (!rc) rc = pr_create_kind(ctx, g, &qa, K_ADJ, &adj);
    if (!rc) rc = gbx_pr_reset(adj, loss_weights);
    if (!rc) rc = run_to_stop(adj);
    if (!rc) rc = gbx_pr_read(adj, lambda_out, adj_rep);
    if (!rc && grad_w && E) {
        gbx_plan *pf = adj->plan;
        DevBuf y, plam, S;
        rc = dev_alloc(ctx, y, (size_t)N * 4);
        if (!rc) rc = dev_alloc(ctx, plam, (size_t)N * 4);
        if (!rc) rc = dev_alloc(ctx, S, (size_t)N * 4);
        DevBuf gw;
        if (!rc) rc = dev_alloc(ctx, gw, (size_t)E * 4);
        if (!rc) {
            const unsigned gN = grid_for(N, 256, 16384);
            fill_f32<<<gN, 256, 0, ctx->stream>>>(y.as<float>(), 0.0f, N);
            // (P lambda)[u] in plan order, then original order
            PlanView P = plan_view(pf, pf->wr.as<float>());
            spmv_kernel<K_SPMV_W><<<iter_grid(ctx, pf), kBlock, 0, ctx->stream>>>(P, adj->r.as<float>(), y.as<float>());
            scatter_perm<<<gN, 256, 0, ctx->stream>>>(y.as<float>(), pf->perm.as<uint32_t>(), N, plam.as<float>());
            out_weight_sums<<<grid_for((size_t)N * 64, 256, 16384), 256, 0, ctx->stream>>>(
                g->fwd.out_degree.as<uint32_t>(), g->fwd.edges_idx.as<uint32_t>(), g->fwd.w.as<float>(), N,
                S.as<float>());
            adjoint_edge_grad<<<grid_for(E, 256, 16384), 256, 0, ctx->stream>>>(
                g->fwd.src.as<uint32_t>(), g->fwd.dst.as<uint32_t>(), E, fwd->out.as<float>(), adj->out.as<float>(),
                plam.as<float>(), S.as<float>(), q.damp, gw.as<float>());
            if (hipGetLastError() != hipSuccess) rc = set_error(ctx, GBX_ERR_HIP, "gbx_pr_adjoint: gradient kernels");
        }
        if (!rc && (hipMemcpyAsync(grad_w, gw.p, (size_t)E * 4, hipMemcpyDeviceToHost, ctx->stream) ||
                    hipStreamSynchronize(ctx->stream)))
            rc = set_error(ctx, GBX_ERR_HIP, "gbx_pr_adjoint: gradient copy");
    }
    if (!rc && grad_teleport) {
        std::vector<float> lam(N);
        if (hipMemcpyAsync(lam.data(), adj->out.p, (size_t)N * 4, hipMemcpyDeviceToHost, ctx->stream) ||
            hipStreamSynchronize(ctx->stream))
            rc = set_error(ctx, GBX_ERR_HIP, "gbx_pr_adjoint: lambda copy");
        else
            for (uint32_t v = 0; v < N; v++) grad_teleport[v] = (float)((1.0 - (double)q.damp) * lam[v]);
    }
    gbx_pr_free(fwd);
    gbx_pr_free(adj);
    return rc;
}

int gbx_pr_plan_info(const gbx_pr *pr, uint32_t *n_classes, uint32_t *class_lanes,
                     uint32_t *class_rows, uint64_t *class_edges)
{
    if (!pr || !n_classes) return set_error(nullptr, GBX_ERR_ARG, "gbx_pr_plan_info: NULL");
    const gbx_plan *p = pr->plan;
    *n_classes = p->nclass;
    for (uint32_t c = 0; c < p->nclass; c++) {
        if (class_lanes) class_lanes[c] = p->cls[c].lanes;
        if (class_rows) class_rows[c] = p->cls[c].row_end - p->cls[c].row_begin;
        if (class_edges) class_edges[c] = p->cls_edges[c];
    }
    return GBX_OK;
}

// Sum over the plan's edges of |w[j] x[col[j]]| in FP64, in a fixed order:
// per-block partials, then one block.  Bounds every row sum of a weighted
// SpMV pass (the scale of the binned iteration's fixed-point row sums).
__global__ void __launch_bounds__(256) spmv_bound_partial(const uint32_t *__restrict__ col,
                                                          const float *__restrict__ w, uint32_t E,
                                                          const float *__restrict__ x, double *__restrict__ part)
{
    double s = 0.0;
    for (uint32_t j = blockIdx.x * 256 + threadIdx.x; j < E; j += gridDim.x * 256)
        s += fabs((double)w[j] * (double)x[col[j]]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    __shared__ double ws[4];
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (ws[0] + ws[1]) + (ws[2] + ws[3]);
}

__global__ void __launch_bounds__(256) spmv_bound_final(const double *__restrict__ part, uint32_t n, Slot *slot)
{
    double s = 0.0;
    for (uint32_t b = threadIdx.x; b < n; b += 256) s += part[b];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    __shared__ double ws[4];
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        Slot z = {};
        z.sum = (ws[0] + ws[1]) + (ws[2] + ws[3]);
        *slot = z;
    }
}

// Weighted SpMV on the binned (propagation-blocking) iteration: a gbx_pr of
// kind K_SPMV_W whose input vector stays in c[0] and whose rank array is y.
static int spmv_binned(gbx_ctx *ctx, gbx_graph *g, uint32_t iterations, int direction, const float *x_in,
                       float *x_out, float *y_out, double *seconds)
{
    gbx_pr_params q;
    gbx_pr_params_default(&q);
    q.mode = GBX_PR_STOCHASTIC;
    q.weighted = 1;
    q.stop = GBX_STOP_NONE;
    q.direction = direction;
    q.method = GBX_METHOD_BINNED;
    q.max_iters = iterations;
    gbx_pr *pr = nullptr;
    GBX_TRY(pr_create_kind(ctx, g, &q, K_SPMV_W, &pr));
    const gbx_plan *plan = pr->plan;
    const uint32_t N = g->N, E = g->E;
    const unsigned gN = grid_for(N, 256, 16384);
    hipStream_t st = ctx->stream;
    int rc = GBX_OK;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    DevBuf part;
    const unsigned gb = std::max(1u, std::min(grid_for(E, 256, 4096), 4096u));
    rc = dev_alloc(ctx, part, (size_t)gb * 8);
    if (!rc) {
        spmv_init_x<<<gN, 256, 0, st>>>(plan->outdeg.as<uint32_t>(), N, pr->c[0].as<float>(), pr->r.as<float>());
        if (x_in && (hipMemcpyAsync(pr->out.p, x_in, (size_t)N * 4, hipMemcpyHostToDevice, st) != hipSuccess))
            rc = set_error(ctx, GBX_ERR_HIP, "gbx_spmv_pull: input copy");
        if (!rc && x_in)
            gather_perm<<<gN, 256, 0, st>>>(pr->out.as<float>(), plan->perm.as<uint32_t>(), N, pr->c[0].as<float>());
        spmv_bound_partial<<<gb, 256, 0, st>>>(plan->col.as<uint32_t>(), plan->wt.as<float>(), E,
                                               pr->c[0].as<float>(), part.as<double>());
        spmv_bound_final<<<1, 256, 0, st>>>(part.as<double>(), gb, pr->slots.as<Slot>());
        if (hipGetLastError() != hipSuccess) rc = set_error(ctx, GBX_ERR_HIP, "gbx_spmv_pull: set-up kernels");
    }
    pr->t = 0;
    if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = set_error(ctx, GBX_ERR_HIP, "gbx_spmv_pull: events");
    if (!rc && hipEventRecord(e0, st) != hipSuccess) rc = set_error(ctx, GBX_ERR_HIP, "gbx_spmv_pull: event");
    for (uint32_t it = 0; it < iterations && !rc; it++) rc = binned_launch_iter(pr);
    if (!rc && hipEventRecord(e1, st) != hipSuccess) rc = set_error(ctx, GBX_ERR_HIP, "gbx_spmv_pull: event");
    if (!rc && y_out) {
        scatter_perm<<<gN, 256, 0, st>>>(pr->r.as<float>(), plan->perm.as<uint32_t>(), N, pr->out.as<float>());
        if (hipMemcpyAsync(y_out, pr->out.p, (size_t)N * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            rc = set_error(ctx, GBX_ERR_HIP, "gbx_spmv_pull: output copy");
    }
    if (!rc && x_out) {
        scatter_perm<<<gN, 256, 0, st>>>(pr->c[0].as<float>(), plan->perm.as<uint32_t>(), N, pr->out.as<float>());
        if (hipMemcpyAsync(x_out, pr->out.p, (size_t)N * 4, hipMemcpyDeviceToHost, st) != hipSuccess)
            rc = set_error(ctx, GBX_ERR_HIP, "gbx_spmv_pull: output copy");
    }
    if (!rc && hipStreamSynchronize(st) != hipSuccess)
        rc = set_error(ctx, GBX_ERR_HIP, "gbx_spmv_pull: %s", hipGetErrorString(hipGetLastError()));
    float ms = 0.0f;
    if (!rc && seconds && hipEventElapsedTime(&ms, e0, e1) == hipSuccess) *seconds = ms * 1e-3;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    gbx_pr_free(pr);
    return rc;
}

int gbx_spmv_pull(gbx_ctx *ctx, gbx_graph *g, uint32_t iterations, int weighted, int direction,
                  const float *x_in, float *x_out, float *y_out, double *seconds)
{
    if (!ctx || !g) return set_error(ctx, GBX_ERR_ARG, "gbx_spmv_pull: NULL argument");
    GBX_HIP(ctx, hipSetDevice(ctx->device));
    gbx_plan *plan = nullptr;
    GBX_TRY(ensure_plan(ctx, g, direction, &plan));
    if (plan->segmented) return set_error(ctx, GBX_ERR_STATE, "gbx_spmv_pull: not available for a batch of graphs");
    if (weighted && !plan->has_wt)
        return set_error(ctx, GBX_ERR_STATE, "gbx_spmv_pull: weighted SpMV needs a weighted graph");
    // large weighted graphs: the binned iteration, as GBX_METHOD_AUTO picks for PageRank
    if (weighted && g->N >= ctx->tune.binned_min_vertices && g->E)
        return spmv_binned(ctx, g, iterations, direction, x_in, x_out, y_out, seconds);
    const uint32_t N = g->N;
    DevBuf x, y, tmp;
    GBX_TRY(dev_alloc(ctx, x, (size_t)N * 4));
    GBX_TRY(dev_alloc(ctx, y, (size_t)N * 4));
    GBX_TRY(dev_alloc(ctx, tmp, (size_t)N * 4));
    const unsigned gN = grid_for(N, 256, 16384);
    spmv_init_x<<<gN, 256, 0, ctx->stream>>>(plan->outdeg.as<uint32_t>(), N, x.as<float>(), y.as<float>());
    GBX_HIP(ctx, hipGetLastError());
    if (x_in) {
        GBX_HIP(ctx, hipMemcpyAsync(tmp.p, x_in, (size_t)N * 4, hipMemcpyHostToDevice, ctx->stream));
        gather_perm<<<gN, 256, 0, ctx->stream>>>(tmp.as<float>(), plan->perm.as<uint32_t>(), N, x.as<float>());
        GBX_HIP(ctx, hipGetLastError());
    }
    PlanView P = plan_view(plan, weighted ? plan->wt.as<float>() : nullptr);
    const unsigned grid = iter_grid(ctx, plan);
    hipEvent_t e0, e1;
    GBX_HIP(ctx, hipEventCreate(&e0));
    GBX_HIP(ctx, hipEventCreate(&e1));
    GBX_HIP(ctx, hipEventRecord(e0, ctx->stream));
    for (uint32_t it = 0; it < iterations; it++) {
        if (weighted)
            spmv_kernel<K_SPMV_W><<<grid, kBlock, 0, ctx->stream>>>(P, x.as<float>(), y.as<float>());
        else
            spmv_kernel<K_SPMV><<<grid, kBlock, 0, ctx->stream>>>(P, x.as<float>(), y.as<float>());
        GBX_HIP(ctx, hipGetLastError());
    }
    GBX_HIP(ctx, hipEventRecord(e1, ctx->stream));
    if (x_out) {
        scatter_perm<<<gN, 256, 0, ctx->stream>>>(x.as<float>(), plan->perm.as<uint32_t>(), N, tmp.as<float>());
        GBX_HIP(ctx, hipGetLastError());
        GBX_HIP(ctx, hipMemcpyAsync(x_out, tmp.p, (size_t)N * 4, hipMemcpyDeviceToHost, ctx->stream));
        GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    if (y_out) {
        scatter_perm<<<gN, 256, 0, ctx->stream>>>(y.as<float>(), plan->perm.as<uint32_t>(), N, tmp.as<float>());
        GBX_HIP(ctx, hipGetLastError());
        GBX_HIP(ctx, hipMemcpyAsync(y_out, tmp.p, (size_t)N * 4, hipMemcpyDeviceToHost, ctx->stream));
    }
    GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (seconds) *seconds = ms * 1e-3;
    return GBX_OK;
}

int gbx_gather_accumulate_u32(gbx_ctx *ctx, gbx_graph *g, int direction, const uint32_t *aux1, uint32_t *aux2)
{
    if (!ctx || !g || !aux1 || !aux2) return set_error(ctx, GBX_ERR_ARG, "gbx_gather_accumulate_u32: NULL argument");
    GBX_HIP(ctx, hipSetDevice(ctx->device));
    gbx_plan *plan = nullptr;
    GBX_TRY(ensure_plan(ctx, g, direction, &plan));
    if (plan->segmented)
        return set_error(ctx, GBX_ERR_STATE, "gbx_gather_accumulate_u32: not available for a batch of graphs");
    const uint32_t N = g->N;
    const size_t nb = (size_t)N * 4;
    DevBuf in, in_plan, out_plan, out;
    GBX_TRY(dev_alloc(ctx, in, nb));
    GBX_TRY(dev_alloc(ctx, in_plan, nb));
    GBX_TRY(dev_alloc(ctx, out_plan, nb));
    GBX_TRY(dev_alloc(ctx, out, nb));
    // the second halves of the two 2N-word buffers hold the vertex values
    // (the READ_WRITE engines' Index_Start = N, buffer_mapping.PR.cpp:30,58)
    GBX_HIP(ctx, hipMemcpyAsync(in.p, aux1 + N, nb, hipMemcpyHostToDevice, ctx->stream));
    const unsigned gN = grid_for(N, 256, 16384);
    const uint32_t *perm = plan->perm.as<uint32_t>();
    gather_perm_u32<<<gN, 256, 0, ctx->stream>>>(in.as<uint32_t>(), perm, N, in_plan.as<uint32_t>());
    GBX_HIP(ctx, hipGetLastError());
    PlanView P = plan_view(plan, nullptr);
    gather_acc_u32_kernel<<<iter_grid(ctx, plan), kBlock, 0, ctx->stream>>>(P, in_plan.as<uint32_t>(),
                                                                           out_plan.as<uint32_t>());
    GBX_HIP(ctx, hipGetLastError());
    scatter_perm_u32<<<gN, 256, 0, ctx->stream>>>(out_plan.as<uint32_t>(), perm, N, out.as<uint32_t>());
    GBX_HIP(ctx, hipGetLastError());
    GBX_HIP(ctx, hipMemcpyAsync(aux2 + N, out.p, nb, hipMemcpyDeviceToHost, ctx->stream));
    GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GBX_OK;
}

int gbx_spmv_fixed(gbx_ctx *ctx, gbx_graph *g, uint32_t iterations, int weighted, int direction, int bits,
                   float *x_out, float *y_out, double *seconds)
{
    if (!ctx || !g) return set_error(ctx, GBX_ERR_ARG, "gbx_spmv_fixed: NULL argument");
    if (bits != 32 && bits != 64) return set_error(ctx, GBX_ERR_ARG, "gbx_spmv_fixed: bits %d (32 or 64)", bits);
    if (direction != GBX_PULL_IN && direction != GBX_PULL_FWD)
        return set_error(ctx, GBX_ERR_ARG, "gbx_spmv_fixed: direction %d", direction);
    if (!g->seg_off.empty()) return set_error(ctx, GBX_ERR_STATE, "gbx_spmv_fixed: not available for a batch of graphs");
    if (weighted && !g->weighted)
        return set_error(ctx, GBX_ERR_STATE, "gbx_spmv_fixed: weighted SpMV needs a weighted graph");
    GBX_HIP(ctx, hipSetDevice(ctx->device));
    const uint32_t N = g->N, E = g->E;
    // the pull runs over the transpose rows (DIRECTED) or the forward rows;
    // the push form always sums over each destination's in-edges
    const gbx_csr_dev &rows = (bits == 64 || direction == GBX_PULL_IN) ? g->inv : g->fwd;
    DevBuf x, xq, y, out;
    GBX_TRY(dev_alloc(ctx, x, (size_t)N * 4 + 4));
    GBX_TRY(dev_alloc(ctx, xq, (size_t)N * 8 + 8));
    GBX_TRY(dev_alloc(ctx, y, (size_t)N * 8 + 8));
    GBX_TRY(dev_alloc(ctx, out, (size_t)N * 4 + 4));
    const unsigned gN = grid_for(N, 256, 16384);
    spmv_fixed_input<<<gN, 256, 0, ctx->stream>>>(g->fwd.out_degree.as<uint32_t>(), N, bits, x.as<float>(),
                                                 xq.as<uint64_t>(), y.as<uint64_t>());
    GBX_HIP(ctx, hipGetLastError());
    const float *w = weighted ? rows.w.as<float>() : nullptr;
    // SPMV.c:354 and :497: the constant weight 0.0001f, converted
    const float wc_f = 0.0001f;
    const uint64_t wconst = bits == 32 ? (uint64_t)(uint32_t)(long long)(wc_f * 65536.0f + 0.5f)
                                       : (uint64_t)((double)wc_f * 4294967296.0 + 0.5);
    const unsigned gE = grid_for(((size_t)E + 63) / 64 * 64, 256, (unsigned)std::max(1, ctx->num_cus) * 16);
    hipEvent_t e0, e1;
    GBX_HIP(ctx, hipEventCreate(&e0));
    GBX_HIP(ctx, hipEventCreate(&e1));
    GBX_HIP(ctx, hipEventRecord(e0, ctx->stream));
    for (uint32_t it = 0; it < iterations && E; it++) {
        if (bits == 32)
            spmv_fixed_kernel<32><<<gE, 256, 0, ctx->stream>>>(rows.src.as<uint32_t>(), rows.dst.as<uint32_t>(), w, E,
                                                               xq.as<uint64_t>(), wconst, y.as<uint64_t>());
        else
            spmv_fixed_kernel<64><<<gE, 256, 0, ctx->stream>>>(rows.src.as<uint32_t>(), rows.dst.as<uint32_t>(), w, E,
                                                               xq.as<uint64_t>(), wconst, y.as<uint64_t>());
        GBX_HIP(ctx, hipGetLastError());
    }
    GBX_HIP(ctx, hipEventRecord(e1, ctx->stream));
    spmv_fixed_output<<<gN, 256, 0, ctx->stream>>>(y.as<uint64_t>(), N, bits, out.as<float>());
    GBX_HIP(ctx, hipGetLastError());
    if (x_out) GBX_HIP(ctx, hipMemcpyAsync(x_out, x.p, (size_t)N * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (y_out) GBX_HIP(ctx, hipMemcpyAsync(y_out, out.p, (size_t)N * 4, hipMemcpyDeviceToHost, ctx->stream));
    GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (seconds) *seconds = ms * 1e-3;
    return GBX_OK;
}

}  // extern "C"
