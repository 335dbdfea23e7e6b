// gbx_sort.hip -- GPU CSR build, bit-exact with the reference's count sort.
//
// The reference builds CSR with a stable counting sort keyed on the source id
// (01_host/src/preprocess/countsort.c:28-140), optionally preceded by a
// stable pass on the destination id (countsort.c:262-270), then derives
// out_degree / edges_idx from the sorted source array
// (01_host/src/structures/vertex.c:204-328).  A stable sort by key is unique,
// so any stable sort reproduces those arrays exactly.  On the GPU the stable
// sort is an LSD radix sort with 8-bit digits; the stable in-tile rank of
// each element comes from a wavefront multisplit (64-bit ballots over the
// digit bits), so no atomics decide the order.
#include <algorithm>

#include "gbx_internal.hpp"

namespace gbx {

constexpr int kSortBlock = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortBlock * kSortItems;  // 4096 keys per tile
constexpr int kRadix = 256;
constexpr int kWaves = kSortBlock / 64;

// Per-tile digit histogram, written digit-major: counts[d * tiles + tile].
__global__ void __launch_bounds__(kSortBlock)
radix_upsweep(const uint32_t *__restrict__ key, size_t n, uint32_t shift, uint32_t mask,
              uint32_t *__restrict__ counts, uint32_t tiles)
{
    __shared__ uint32_t hist[kRadix];
    hist[threadIdx.x] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * kSortTile;
    const uint64_t lt = (1ull << (threadIdx.x & 63)) - 1ull;
#pragma unroll 4
    for (int k = 0; k < kSortItems; k++) {
        size_t i = base + (size_t)k * kSortBlock + threadIdx.x;
        bool valid = i < n;
        uint32_t d = valid ? (key[i] >> shift) & mask : 0u;
        // wavefront multisplit: peers with the same digit
        uint64_t match = __ballot(valid);
        for (uint32_t b = 0; (mask >> b) != 0; b++) {
            uint64_t bal = __ballot((d >> b) & 1u);
            match &= ((d >> b) & 1u) ? bal : ~bal;
        }
        if (valid && (match & lt) == 0) atomicAdd(&hist[d], (uint32_t)__popcll(match));
    }
    __syncthreads();
    counts[(size_t)threadIdx.x * tiles + blockIdx.x] = hist[threadIdx.x];
}

// Stable scatter of one tile: element order inside a tile is (item, wave,
// lane) = index order, and digit offsets are exclusive over (item, wave).
template <bool HAS_W>
__global__ void __launch_bounds__(kSortBlock)
radix_downsweep(const uint32_t *__restrict__ key_in, const uint32_t *__restrict__ pay_in,
                const uint32_t *__restrict__ w_in, uint32_t *__restrict__ key_out,
                uint32_t *__restrict__ pay_out, uint32_t *__restrict__ w_out, size_t n,
                uint32_t shift, uint32_t mask, const uint32_t *__restrict__ offsets,
                uint32_t tiles)
{
    __shared__ uint16_t wcnt[kSortItems][kWaves][kRadix];  // 32 KiB
    __shared__ uint32_t dbase[kRadix];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    {
        uint32_t *z = reinterpret_cast<uint32_t *>(&wcnt[0][0][0]);
        constexpr int words = kSortItems * kWaves * kRadix / 2;
        for (int i = tid; i < words; i += kSortBlock) z[i] = 0;
    }
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * kSortTile;
    const uint64_t lt = (1ull << lane) - 1ull;
    uint32_t kv[kSortItems], pv[kSortItems], wv[kSortItems], dr[kSortItems];
#pragma unroll
    for (int k = 0; k < kSortItems; k++) {
        size_t i = base + (size_t)k * kSortBlock + tid;
        bool valid = i < n;
        kv[k] = valid ? key_in[i] : 0u;
        pv[k] = valid ? pay_in[i] : 0u;
        if (HAS_W) wv[k] = valid ? w_in[i] : 0u;
        uint32_t d = (kv[k] >> shift) & mask;
        uint64_t match = __ballot(valid);
        for (uint32_t b = 0; (mask >> b) != 0; b++) {
            uint64_t bal = __ballot((d >> b) & 1u);
            match &= ((d >> b) & 1u) ? bal : ~bal;
        }
        uint32_t rank = (uint32_t)__popcll(match & lt);
        if (valid && rank == 0) wcnt[k][wid][d] = (uint16_t)__popcll(match);
        dr[k] = valid ? (d | (rank << 8)) : 0xFFFFFFFFu;
    }
    __syncthreads();
    {
        // thread d: exclusive prefix over (item, wave) for digit d
        const int d = tid;
        uint32_t run = 0;
        for (int k = 0; k < kSortItems; k++)
#pragma unroll
            for (int w = 0; w < kWaves; w++) {
                uint32_t c = wcnt[k][w][d];
                wcnt[k][w][d] = (uint16_t)run;
                run += c;
            }
        dbase[d] = offsets[(size_t)d * tiles + blockIdx.x];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kSortItems; k++) {
        if (dr[k] == 0xFFFFFFFFu) continue;
        uint32_t d = dr[k] & 0xFFu, rank = dr[k] >> 8;
        uint32_t pos = dbase[d] + wcnt[k][wid][d] + rank;
        key_out[pos] = kv[k];
        pay_out[pos] = pv[k];
        if (HAS_W) w_out[pos] = wv[k];
    }
}

int radix_sort_pairs(gbx_ctx *ctx, uint32_t *key, uint32_t *pay, float *w, size_t n,
                     uint32_t bits)
{
    if (n <= 1 || bits == 0) return GBX_OK;
    if (n > 0xFFFFFFFFull) return set_error(ctx, GBX_ERR_ARG, "radix sort: n too large");
    const uint32_t tiles = (uint32_t)((n + kSortTile - 1) / kSortTile);
    const size_t ncount = (size_t)kRadix * tiles;
    GBX_TRY(dev_reserve(ctx, ctx->sort_hist, ncount * sizeof(uint32_t)));
    GBX_TRY(dev_reserve(ctx, ctx->sort_tmp[0], n * sizeof(uint32_t)));
    GBX_TRY(dev_reserve(ctx, ctx->sort_tmp[1], n * sizeof(uint32_t)));
    if (w) GBX_TRY(dev_reserve(ctx, ctx->sort_tmp[2], n * sizeof(uint32_t)));
    uint32_t *counts = ctx->sort_hist.as<uint32_t>();
    uint32_t *k_a = key, *p_a = pay, *w_a = reinterpret_cast<uint32_t *>(w);
    uint32_t *k_b = ctx->sort_tmp[0].as<uint32_t>(), *p_b = ctx->sort_tmp[1].as<uint32_t>();
    uint32_t *w_b = w ? ctx->sort_tmp[2].as<uint32_t>() : nullptr;
    uint32_t passes = 0;
    for (uint32_t shift = 0; shift < bits; shift += 8, passes++) {
        uint32_t nb = bits - shift < 8 ? bits - shift : 8;
        uint32_t mask = (1u << nb) - 1u;
        radix_upsweep<<<tiles, kSortBlock, 0, ctx->stream>>>(k_a, n, shift, mask, counts, tiles);
        GBX_HIP(ctx, hipGetLastError());
        GBX_TRY(exclusive_scan_u32(ctx, counts, counts, ncount, nullptr));
        if (w)
            radix_downsweep<true><<<tiles, kSortBlock, 0, ctx->stream>>>(
                k_a, p_a, w_a, k_b, p_b, w_b, n, shift, mask, counts, tiles);
        else
            radix_downsweep<false><<<tiles, kSortBlock, 0, ctx->stream>>>(
                k_a, p_a, nullptr, k_b, p_b, nullptr, n, shift, mask, counts, tiles);
        GBX_HIP(ctx, hipGetLastError());
        std::swap(k_a, k_b);
        std::swap(p_a, p_b);
        std::swap(w_a, w_b);
    }
    if (passes & 1) {
        // results are in the scratch buffers: copy back
        GBX_HIP(ctx, hipMemcpyAsync(key, k_a, n * 4, hipMemcpyDeviceToDevice, ctx->stream));
        GBX_HIP(ctx, hipMemcpyAsync(pay, p_a, n * 4, hipMemcpyDeviceToDevice, ctx->stream));
        if (w)
            GBX_HIP(ctx, hipMemcpyAsync(w, w_a, n * 4, hipMemcpyDeviceToDevice, ctx->stream));
    }
    return GBX_OK;
}

// ---------------------------------------------------------------- vertex map

// Run boundaries of the sorted source array: edges_idx[v] = first index of
// v's run, out_degree[v] temporarily = one past its last index
// (mapVerticesWithInOutDegree_base, vertex.c:267-288).
__global__ void vertex_runs(const uint32_t *__restrict__ src, uint32_t E,
                            uint32_t *__restrict__ out_degree, uint32_t *__restrict__ edges_idx)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < E; i += gridDim.x * blockDim.x) {
        uint32_t s = src[i];
        uint32_t v = s & kIdMask;
        if (i == 0 || src[i - 1] != s) edges_idx[v] = i;
        if (i + 1 == E || src[i + 1] != s) out_degree[v] = i + 1;
    }
}

__global__ void vertex_degrees(uint32_t N, uint32_t *__restrict__ out_degree,
                               const uint32_t *__restrict__ edges_idx)
{
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        uint32_t e = out_degree[v];
        if (e) out_degree[v] = e - edges_idx[v];
    }
}

int build_csr_device(gbx_ctx *ctx, uint32_t *src, uint32_t *dst, float *w, uint32_t E,
                     uint32_t N, int sort_mode, bool move_weights, uint32_t *out_degree,
                     uint32_t *edges_idx)
{
    const uint32_t bits = N > 1 ? bits_for(N - 1) : 0;
    float *wm = move_weights ? w : nullptr;
    if (sort_mode == GBX_SORT_RADIX_SRC_DST || sort_mode == GBX_SORT_COUNT_SRC_DST)
        GBX_TRY(radix_sort_pairs(ctx, dst, src, wm, E, bits));  // by destination
    GBX_TRY(radix_sort_pairs(ctx, src, dst, wm, E, bits));      // then by source
    GBX_TRY(fill_u32(ctx, out_degree, 0, N));
    GBX_TRY(fill_u32(ctx, edges_idx, 0, N));
    if (E) {
        vertex_runs<<<grid_for(E, 256, 16384), 256, 0, ctx->stream>>>(src, E, out_degree, edges_idx);
        GBX_HIP(ctx, hipGetLastError());
    }
    vertex_degrees<<<grid_for(N, 256, 16384), 256, 0, ctx->stream>>>(N, out_degree, edges_idx);
    GBX_HIP(ctx, hipGetLastError());
    return GBX_OK;
}

// ---------------------------------------------------------------- dedupe

// keep[i] for a (src,dst)-sorted list: not a self loop and not equal to the
// previous pair (removeDulpicatesSelfLoopEdges, edgeList.c:255-299).
__global__ void dedupe_flags(const uint32_t *__restrict__ src, const uint32_t *__restrict__ dst,
                             uint32_t E, uint32_t *__restrict__ keep)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < E; i += gridDim.x * blockDim.x) {
        uint32_t s = src[i], d = dst[i];
        bool k = s != d && (i == 0 || src[i - 1] != s || dst[i - 1] != d);
        keep[i] = k ? 1u : 0u;
    }
}

// Compact kept edges; weights pass through a uint32_t (edgeList.c:261,270).
__global__ void dedupe_compact(const uint32_t *__restrict__ src, const uint32_t *__restrict__ dst,
                               const float *__restrict__ w, const uint32_t *__restrict__ keep,
                               const uint32_t *__restrict__ pos, uint32_t E,
                               uint32_t *__restrict__ so, uint32_t *__restrict__ dout,
                               float *__restrict__ wo)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < E; i += gridDim.x * blockDim.x) {
        if (!keep[i]) continue;
        uint32_t p = pos[i];
        so[p] = src[i];
        dout[p] = dst[i];
        wo[p] = (float)(uint32_t)w[i];
    }
}

// flag[0] |= 1 if src is not sorted, |= 2 if an id is >= N
__global__ void check_sorted(const uint32_t *__restrict__ src, const uint32_t *__restrict__ dst,
                             uint32_t E, uint32_t N, uint32_t *__restrict__ flag)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < E; i += gridDim.x * blockDim.x) {
        uint32_t f = 0;
        if (i > 0 && src[i - 1] > src[i]) f |= 1u;
        if ((src[i] & kIdMask) >= N || (dst[i] & kIdMask) >= N) f |= 2u;
        if (f) atomicOr(flag, f);
    }
}

// The vertex arrays of a CSR against its sorted edges: flag[0] |= 4 if an
// edge e of source v = src[e] lies outside [edges_idx[v], edges_idx[v] +
// out_degree[v]); *degsum += sum of out_degree.  With src sorted, every
// edge inside its source's range and the degrees summing to E, each range
// holds exactly its vertex's edges (no range can hold more edges than its
// length), which is what the plan build indexes with.
__global__ void check_csr_index(const uint32_t *__restrict__ src, uint32_t E, const uint32_t *__restrict__ od,
                                const uint32_t *__restrict__ idx, uint32_t N, uint32_t *__restrict__ flag,
                                unsigned long long *__restrict__ degsum)
{
    const uint32_t stride = gridDim.x * blockDim.x;
    uint32_t bad = 0;
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += stride) {
        const uint32_t v = src[e] & kIdMask;
        if (v >= N) continue;  // reported by check_sorted
        const uint64_t b = idx[v];
        if ((uint64_t)e < b || (uint64_t)e >= b + od[v]) bad = 4u;
    }
    if (bad) atomicOr(flag, bad);
    unsigned long long s = 0;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += stride) s += od[v];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(degsum, s);
}

__global__ void fill_f32_kernel(float *p, float v, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void symmetric_append(uint32_t *src, uint32_t *dst, float *w, uint32_t E)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < E; i += gridDim.x * blockDim.x) {
        src[i + E] = dst[i];
        dst[i + E] = src[i];
        w[i + E] = w[i];
    }
}

int graph_from_device_edges(gbx_ctx *ctx, DevBuf &src, DevBuf &dst, DevBuf &w, uint32_t E,
                            uint32_t N, int sort_mode, unsigned flags, gbx_graph **out,
                            const gbx_relabel *relabel, uint32_t *labels_out)
{
    const bool weighted = (flags & GBX_GRAPH_WEIGHTED) != 0;
    if (flags & GBX_GRAPH_DEDUPE) sort_mode = GBX_SORT_RADIX_SRC_DST;  // GraphBlox.c:358-359
    gbx_graph *g = new (std::nothrow) gbx_graph();
    if (!g) return set_error(ctx, GBX_ERR_NOMEM, "graph: out of host memory");
    g->ctx = ctx;
    g->N = N;
    g->weighted = weighted;
    auto fail = [&](int rc) { delete g; return rc; };
    int rc;
    // forward CSR
    if ((rc = dev_alloc(ctx, g->fwd.out_degree, (size_t)N * 4))) return fail(rc);
    if ((rc = dev_alloc(ctx, g->fwd.edges_idx, (size_t)N * 4))) return fail(rc);
    if ((rc = build_csr_device(ctx, src.as<uint32_t>(), dst.as<uint32_t>(), w.as<float>(), E, N,
                               sort_mode, weighted, g->fwd.out_degree.as<uint32_t>(),
                               g->fwd.edges_idx.as<uint32_t>())))
        return fail(rc);
    if (flags & GBX_GRAPH_DEDUPE) {
        DevBuf keep, pos, s2, d2, w2;
        if ((rc = dev_alloc(ctx, keep, (size_t)E * 4 + 4))) return fail(rc);
        if ((rc = dev_alloc(ctx, pos, (size_t)E * 4 + 4))) return fail(rc);
        if (E) {
            dedupe_flags<<<grid_for(E, 256, 16384), 256, 0, ctx->stream>>>(
                src.as<uint32_t>(), dst.as<uint32_t>(), E, keep.as<uint32_t>());
            if (hipGetLastError() != hipSuccess) return fail(set_error(ctx, GBX_ERR_HIP, "dedupe_flags"));
        }
        if ((rc = exclusive_scan_u32(ctx, keep.as<uint32_t>(), pos.as<uint32_t>(), E,
                                     pos.as<uint32_t>() + E)))
            return fail(rc);
        uint32_t E2 = 0;
        if (hipMemcpyAsync(&E2, pos.as<uint32_t>() + E, 4, hipMemcpyDeviceToHost, ctx->stream) !=
                hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
            return fail(set_error(ctx, GBX_ERR_HIP, "dedupe: count copy failed"));
        if ((rc = dev_alloc(ctx, s2, (size_t)E2 * 4))) return fail(rc);
        if ((rc = dev_alloc(ctx, d2, (size_t)E2 * 4))) return fail(rc);
        if ((rc = dev_alloc(ctx, w2, (size_t)E2 * 4))) return fail(rc);
        if (E) {
            dedupe_compact<<<grid_for(E, 256, 16384), 256, 0, ctx->stream>>>(
                src.as<uint32_t>(), dst.as<uint32_t>(), w.as<float>(), keep.as<uint32_t>(),
                pos.as<uint32_t>(), E, s2.as<uint32_t>(), d2.as<uint32_t>(), w2.as<float>());
            if (hipGetLastError() != hipSuccess) return fail(set_error(ctx, GBX_ERR_HIP, "dedupe_compact"));
        }
        src = std::move(s2);
        dst = std::move(d2);
        w = std::move(w2);
        E = E2;
        // degrees of the deduplicated list (graphCSR.c:260-266 then :291)
        if ((rc = fill_u32(ctx, g->fwd.out_degree.as<uint32_t>(), 0, N))) return fail(rc);
        if ((rc = fill_u32(ctx, g->fwd.edges_idx.as<uint32_t>(), 0, N))) return fail(rc);
        if (E) {
            vertex_runs<<<grid_for(E, 256, 16384), 256, 0, ctx->stream>>>(
                src.as<uint32_t>(), E, g->fwd.out_degree.as<uint32_t>(), g->fwd.edges_idx.as<uint32_t>());
            if (hipGetLastError() != hipSuccess) return fail(set_error(ctx, GBX_ERR_HIP, "vertex_runs"));
        }
        vertex_degrees<<<grid_for(N, 256, 16384), 256, 0, ctx->stream>>>(
            N, g->fwd.out_degree.as<uint32_t>(), g->fwd.edges_idx.as<uint32_t>());
        if (hipGetLastError() != hipSuccess) return fail(set_error(ctx, GBX_ERR_HIP, "vertex_degrees"));
    }
    if (relabel_requested(relabel)) {
        // EdgeList.avg_degree: E / N as read, recomputed by the dedupe (edgeList.c:305,681)
        const uint32_t avg = N ? E / N : 0;
        DevBuf total;
        if ((rc = relabel_edges(ctx, src, dst, w, E, N, avg, sort_mode, weighted, g->fwd.out_degree.as<uint32_t>(),
                                g->fwd.edges_idx.as<uint32_t>(), relabel, total)))
            return fail(rc);
        if (labels_out &&
            (hipMemcpyAsync(labels_out, total.p, (size_t)N * 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
             hipStreamSynchronize(ctx->stream) != hipSuccess))
            return fail(set_error(ctx, GBX_ERR_HIP, "relabel: label copy failed"));
    } else if (labels_out) {
        for (uint32_t v = 0; v < N; v++) labels_out[v] = v;
    }
    g->E = E;
    g->fwd.src = std::move(src);
    g->fwd.dst = std::move(dst);
    g->fwd.w = std::move(w);
    // transpose: swap src/dst (edgeList.c:799-801), same sort (graphCSR.c:301-316)
    if ((rc = dev_alloc(ctx, g->inv.src, (size_t)E * 4))) return fail(rc);
    if ((rc = dev_alloc(ctx, g->inv.dst, (size_t)E * 4))) return fail(rc);
    if ((rc = dev_alloc(ctx, g->inv.w, (size_t)E * 4))) return fail(rc);
    if ((rc = dev_alloc(ctx, g->inv.out_degree, (size_t)N * 4))) return fail(rc);
    if ((rc = dev_alloc(ctx, g->inv.edges_idx, (size_t)N * 4))) return fail(rc);
    if ((rc = dev_alloc(ctx, g->inv.in_degree, (size_t)N * 4))) return fail(rc);
    if ((rc = dev_alloc(ctx, g->fwd.in_degree, (size_t)N * 4))) return fail(rc);
    if (E) {
        if (hipMemcpyAsync(g->inv.src.p, g->fwd.dst.p, (size_t)E * 4, hipMemcpyDeviceToDevice, ctx->stream) ||
            hipMemcpyAsync(g->inv.dst.p, g->fwd.src.p, (size_t)E * 4, hipMemcpyDeviceToDevice, ctx->stream) ||
            hipMemcpyAsync(g->inv.w.p, g->fwd.w.p, (size_t)E * 4, hipMemcpyDeviceToDevice, ctx->stream))
            return fail(set_error(ctx, GBX_ERR_HIP, "transpose copy failed"));
    }
    if ((rc = build_csr_device(ctx, g->inv.src.as<uint32_t>(), g->inv.dst.as<uint32_t>(),
                               g->inv.w.as<float>(), E, N, sort_mode, weighted,
                               g->inv.out_degree.as<uint32_t>(), g->inv.edges_idx.as<uint32_t>())))
        return fail(rc);
    // in_degree of each side = out_degree of the other (vertex.c:291-311)
    if (N) {
        if (hipMemcpyAsync(g->fwd.in_degree.p, g->inv.out_degree.p, (size_t)N * 4, hipMemcpyDeviceToDevice, ctx->stream) ||
            hipMemcpyAsync(g->inv.in_degree.p, g->fwd.out_degree.p, (size_t)N * 4, hipMemcpyDeviceToDevice, ctx->stream))
            return fail(set_error(ctx, GBX_ERR_HIP, "in_degree copy failed"));
    }
    if (hipStreamSynchronize(ctx->stream) != hipSuccess)
        return fail(set_error(ctx, GBX_ERR_HIP, "graph build failed: %s", hipGetErrorString(hipGetLastError())));
    *out = g;
    return GBX_OK;
}

static int upload_edges(gbx_ctx *ctx, const uint32_t *src, const uint32_t *dst, const float *w,
                        uint32_t E, uint32_t cap, DevBuf &ds, DevBuf &dd, DevBuf &dw)
{
    GBX_TRY(dev_alloc(ctx, ds, (size_t)cap * 4));
    GBX_TRY(dev_alloc(ctx, dd, (size_t)cap * 4));
    GBX_TRY(dev_alloc(ctx, dw, (size_t)cap * 4));
    if (E) {
        GBX_HIP(ctx, hipMemcpyAsync(ds.p, src, (size_t)E * 4, hipMemcpyHostToDevice, ctx->stream));
        GBX_HIP(ctx, hipMemcpyAsync(dd.p, dst, (size_t)E * 4, hipMemcpyHostToDevice, ctx->stream));
        if (w) {
            GBX_HIP(ctx, hipMemcpyAsync(dw.p, w, (size_t)E * 4, hipMemcpyHostToDevice, ctx->stream));
        } else {
            fill_f32_kernel<<<grid_for(E, 256, 8192), 256, 0, ctx->stream>>>(dw.as<float>(), 1.0f, E);
            GBX_HIP(ctx, hipGetLastError());
        }
    }
    return GBX_OK;
}

int check_ids(gbx_ctx *ctx, const DevBuf &ds, const DevBuf &dd, uint32_t E, uint32_t *N)
{
    uint32_t ms = 0, md = 0;
    GBX_TRY(max_u32(ctx, ds.as<uint32_t>(), E, &ms));
    GBX_TRY(max_u32(ctx, dd.as<uint32_t>(), E, &md));
    uint32_t m = ms > md ? ms : md;
    if (E && m >= kMaxVertices)
        return set_error(ctx, GBX_ERR_RANGE, "vertex id %u >= 2^30 (ids are masked to 30 bits)", m);
    if (*N == 0) *N = E ? m + 1 : 1;  // edgeList.c:679
    else if (E && m >= *N)
        return set_error(ctx, GBX_ERR_RANGE, "vertex id %u >= num_vertices %u", m, *N);
    if (*N > kMaxVertices) return set_error(ctx, GBX_ERR_RANGE, "num_vertices %u > 2^30", *N);
    return GBX_OK;
}

}  // namespace gbx

using namespace gbx;

// FloatToFixed32SORT (fixedPoint.h:38): (u32)(r * 2^20 + 0.5f) in float,
// -0.5f for negative r; the u32 conversion is x86-64's (64-bit signed
// truncation keeping the low 32 bits; 2^63 and beyond, or NaN, give 0).
__global__ void rank_sort_keys(const float *__restrict__ ranks, uint32_t n, uint32_t *__restrict__ key)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float r = ranks[i];
        const float x = r * 1048576.0f + (r >= 0.0f ? 0.5f : -0.5f);
        key[i] = (x < 9223372036854775808.0f && x >= -9223372036854775808.0f)
                     ? (uint32_t)(unsigned long long)(long long)x
                     : 0u;
    }
}

extern "C" {

int gbx_rank_sort(gbx_ctx *ctx, const float *ranks, uint32_t num_vertices, uint32_t *labels)
{
    if (!ctx) return set_error(nullptr, GBX_ERR_ARG, "gbx_rank_sort: ctx is NULL");
    if (num_vertices && (!ranks || !labels)) return set_error(ctx, GBX_ERR_ARG, "gbx_rank_sort: NULL array");
    if (num_vertices <= 1) return GBX_OK;
    GBX_HIP(ctx, hipSetDevice(ctx->device));
    const size_t nb = (size_t)num_vertices * 4;
    DevBuf r, key, lab;
    GBX_TRY(dev_alloc(ctx, r, nb));
    GBX_TRY(dev_alloc(ctx, key, nb));
    GBX_TRY(dev_alloc(ctx, lab, nb));
    GBX_HIP(ctx, hipMemcpyAsync(r.p, ranks, nb, hipMemcpyHostToDevice, ctx->stream));
    GBX_HIP(ctx, hipMemcpyAsync(lab.p, labels, nb, hipMemcpyHostToDevice, ctx->stream));
    rank_sort_keys<<<grid_for(num_vertices, 256, 16384), 256, 0, ctx->stream>>>(r.as<float>(), num_vertices,
                                                                                 key.as<uint32_t>());
    GBX_HIP(ctx, hipGetLastError());
    GBX_TRY(radix_sort_pairs(ctx, key.as<uint32_t>(), lab.as<uint32_t>(), nullptr, num_vertices, 32));
    GBX_HIP(ctx, hipMemcpyAsync(labels, lab.p, nb, hipMemcpyDeviceToHost, ctx->stream));
    GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GBX_OK;
}

int gbx_csr_build(gbx_ctx *ctx, const uint32_t *src, const uint32_t *dst, const float *w,
                  uint32_t num_edges, uint32_t num_vertices, int sort_mode, int move_weights,
                  uint32_t *sorted_src, uint32_t *sorted_dst, float *sorted_w,
                  uint32_t *out_degree, uint32_t *edges_idx)
{
    if (!ctx) return set_error(nullptr, GBX_ERR_ARG, "gbx_csr_build: ctx is NULL");
    if (num_edges && (!src || !dst))
        return set_error(ctx, GBX_ERR_ARG, "gbx_csr_build: src/dst is NULL");
    if (sort_mode < 0 || sort_mode > 3)
        return set_error(ctx, GBX_ERR_ARG, "gbx_csr_build: sort_mode %d not in 0..3", sort_mode);
    if (num_vertices == 0) return set_error(ctx, GBX_ERR_ARG, "gbx_csr_build: num_vertices is 0");
    GBX_HIP(ctx, hipSetDevice(ctx->device));
    DevBuf ds, dd, dw, od, idx;
    GBX_TRY(upload_edges(ctx, src, dst, w, num_edges, num_edges, ds, dd, dw));
    uint32_t N = num_vertices;
    GBX_TRY(check_ids(ctx, ds, dd, num_edges, &N));
    GBX_TRY(dev_alloc(ctx, od, (size_t)N * 4));
    GBX_TRY(dev_alloc(ctx, idx, (size_t)N * 4));
    GBX_TRY(build_csr_device(ctx, ds.as<uint32_t>(), dd.as<uint32_t>(), dw.as<float>(), num_edges,
                             N, sort_mode, move_weights != 0, od.as<uint32_t>(), idx.as<uint32_t>()));
    const size_t eb = (size_t)num_edges * 4, nb = (size_t)N * 4;
    if (sorted_src && eb) GBX_HIP(ctx, hipMemcpyAsync(sorted_src, ds.p, eb, hipMemcpyDeviceToHost, ctx->stream));
    if (sorted_dst && eb) GBX_HIP(ctx, hipMemcpyAsync(sorted_dst, dd.p, eb, hipMemcpyDeviceToHost, ctx->stream));
    if (sorted_w && eb) GBX_HIP(ctx, hipMemcpyAsync(sorted_w, dw.p, eb, hipMemcpyDeviceToHost, ctx->stream));
    if (out_degree) GBX_HIP(ctx, hipMemcpyAsync(out_degree, od.p, nb, hipMemcpyDeviceToHost, ctx->stream));
    if (edges_idx) GBX_HIP(ctx, hipMemcpyAsync(edges_idx, idx.p, nb, hipMemcpyDeviceToHost, ctx->stream));
    GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GBX_OK;
}

int gbx_graph_create(gbx_ctx *ctx, const uint32_t *src, const uint32_t *dst, const float *w,
                     uint32_t num_edges, uint32_t num_vertices, int sort_mode, unsigned flags,
                     gbx_graph **out)
{
    return gbx_graph_create_relabel(ctx, src, dst, w, num_edges, num_vertices, sort_mode, flags, nullptr, nullptr,
                                    out);
}

int gbx_graph_create_relabel(gbx_ctx *ctx, const uint32_t *src, const uint32_t *dst, const float *w,
                             uint32_t num_edges, uint32_t num_vertices, int sort_mode, unsigned flags,
                             const gbx_relabel *relabel, uint32_t *labels_out, gbx_graph **out)
{
    if (!ctx || !out) return set_error(ctx, GBX_ERR_ARG, "gbx_graph_create: NULL argument");
    GBX_TRY(relabel_validate(ctx, relabel));
    *out = nullptr;
    if (num_edges && (!src || !dst))
        return set_error(ctx, GBX_ERR_ARG, "gbx_graph_create: src/dst is NULL");
    if (sort_mode < 0 || sort_mode > 3)
        return set_error(ctx, GBX_ERR_ARG, "gbx_graph_create: sort_mode %d not in 0..3", sort_mode);
    const bool sym = (flags & GBX_GRAPH_SYMMETRIC) != 0;
    if (sym && num_edges > 0x7FFFFFFFu)
        return set_error(ctx, GBX_ERR_ARG, "gbx_graph_create: too many edges to symmetrise");
    GBX_HIP(ctx, hipSetDevice(ctx->device));
    const uint32_t E = sym ? 2 * num_edges : num_edges;
    DevBuf ds, dd, dw;
    GBX_TRY(upload_edges(ctx, src, dst, w, num_edges, E, ds, dd, dw));
    if (sym && num_edges) {
        symmetric_append<<<grid_for(num_edges, 256, 16384), 256, 0, ctx->stream>>>(
            ds.as<uint32_t>(), dd.as<uint32_t>(), dw.as<float>(), num_edges);
        GBX_HIP(ctx, hipGetLastError());
    }
    uint32_t N = num_vertices;
    GBX_TRY(check_ids(ctx, ds, dd, E, &N));
    return graph_from_device_edges(ctx, ds, dd, dw, E, N, sort_mode, flags, out, relabel, labels_out);
}

int gbx_graph_from_csr(gbx_ctx *ctx, uint32_t num_vertices, uint32_t num_edges,
                       const uint32_t *fwd_out_degree, const uint32_t *fwd_edges_idx,
                       const uint32_t *fwd_src, const uint32_t *fwd_dst, const float *fwd_w,
                       const uint32_t *inv_out_degree, const uint32_t *inv_edges_idx,
                       const uint32_t *inv_src, const uint32_t *inv_dst, const float *inv_w,
                       gbx_graph **out)
{
    if (!ctx || !out) return set_error(ctx, GBX_ERR_ARG, "gbx_graph_from_csr: NULL argument");
    *out = nullptr;
    if (num_vertices == 0 || num_vertices > kMaxVertices)
        return set_error(ctx, GBX_ERR_ARG, "gbx_graph_from_csr: num_vertices %u", num_vertices);
    const bool fwd_only = !inv_out_degree && !inv_edges_idx && !inv_src && !inv_dst && !inv_w;
    if (!fwd_out_degree || !fwd_edges_idx || (num_edges && (!fwd_src || !fwd_dst)) ||
        (!fwd_only && (!inv_out_degree || !inv_edges_idx || (num_edges && (!inv_src || !inv_dst)))))
        return set_error(ctx, GBX_ERR_ARG, "gbx_graph_from_csr: NULL array");
    GBX_HIP(ctx, hipSetDevice(ctx->device));
    gbx_graph *g = new (std::nothrow) gbx_graph();
    if (!g) return set_error(ctx, GBX_ERR_NOMEM, "gbx_graph_from_csr: out of host memory");
    g->ctx = ctx;
    g->N = num_vertices;
    g->E = num_edges;
    g->weighted = fwd_w && (fwd_only || inv_w);
    const size_t nb = (size_t)num_vertices * 4, eb = (size_t)num_edges * 4;
    struct Side {
        gbx_csr_dev *d;
        const uint32_t *od, *idx, *src, *dst;
        const float *w;
    } sides[2] = {{&g->fwd, fwd_out_degree, fwd_edges_idx, fwd_src, fwd_dst, fwd_w},
                  {&g->inv, inv_out_degree, inv_edges_idx, inv_src, inv_dst, inv_w}};
    int rc = GBX_OK;
    for (int k = 0; k < 2 && !rc; k++) {
        Side &sd = sides[k];
        if (!rc) rc = dev_alloc(ctx, sd.d->out_degree, nb);
        if (!rc) rc = dev_alloc(ctx, sd.d->in_degree, nb);
        if (!rc) rc = dev_alloc(ctx, sd.d->edges_idx, nb);
        if (!rc) rc = dev_alloc(ctx, sd.d->src, eb);
        if (!rc) rc = dev_alloc(ctx, sd.d->dst, eb);
        if (!rc) rc = dev_alloc(ctx, sd.d->w, eb);
        if (rc) break;
        if (k == 1 && fwd_only) {
            // the transpose of the forward CSR: swapped edges, stable sort by the
            // new source (edgeList.c:799-801, then countsort.c:28-140)
            if (eb && (hipMemcpyAsync(sd.d->src.p, g->fwd.dst.p, eb, hipMemcpyDeviceToDevice, ctx->stream) ||
                       hipMemcpyAsync(sd.d->dst.p, g->fwd.src.p, eb, hipMemcpyDeviceToDevice, ctx->stream) ||
                       hipMemcpyAsync(sd.d->w.p, g->fwd.w.p, eb, hipMemcpyDeviceToDevice, ctx->stream))) {
                rc = set_error(ctx, GBX_ERR_HIP, "gbx_graph_from_csr: transpose copy failed");
                break;
            }
            continue;  // sorted after the forward side has been checked
        }
        if (hipMemcpyAsync(sd.d->out_degree.p, sd.od, nb, hipMemcpyHostToDevice, ctx->stream) ||
            hipMemcpyAsync(sd.d->edges_idx.p, sd.idx, nb, hipMemcpyHostToDevice, ctx->stream) ||
            (eb && hipMemcpyAsync(sd.d->src.p, sd.src, eb, hipMemcpyHostToDevice, ctx->stream)) ||
            (eb && hipMemcpyAsync(sd.d->dst.p, sd.dst, eb, hipMemcpyHostToDevice, ctx->stream))) {
            rc = set_error(ctx, GBX_ERR_HIP, "gbx_graph_from_csr: upload failed");
            break;
        }
        if (eb && sd.w) {
            if (hipMemcpyAsync(sd.d->w.p, sd.w, eb, hipMemcpyHostToDevice, ctx->stream))
                rc = set_error(ctx, GBX_ERR_HIP, "gbx_graph_from_csr: upload failed");
        } else if (eb) {
            fill_f32_kernel<<<grid_for(num_edges, 256, 8192), 256, 0, ctx->stream>>>(sd.d->w.as<float>(), 1.0f,
                                                                                   num_edges);
        }
    }
    // every uploaded side: sorted sources, ids below N, vertex arrays that
    // agree with the edges (checked before any kernel indexes with them)
    const int nchecked = fwd_only ? 1 : 2;
    if (!rc) {
        DevBuf flag;
        uint32_t hflag[4] = {0, 0, 0, 0};
        rc = dev_alloc(ctx, flag, 32);
        if (!rc) rc = fill_u32(ctx, flag.as<uint32_t>(), 0, 8);
        for (int k = 0; k < nchecked && !rc; k++) {
            const Side &sd = sides[k];
            if (num_edges)
                check_sorted<<<grid_for(num_edges, 256, 16384), 256, 0, ctx->stream>>>(
                    sd.d->src.as<uint32_t>(), sd.d->dst.as<uint32_t>(), num_edges, num_vertices,
                    flag.as<uint32_t>());
            check_csr_index<<<grid_for(std::max(num_edges, num_vertices), 256, 16384), 256, 0, ctx->stream>>>(
                sd.d->src.as<uint32_t>(), num_edges, sd.d->out_degree.as<uint32_t>(), sd.d->edges_idx.as<uint32_t>(),
                num_vertices, flag.as<uint32_t>(), reinterpret_cast<unsigned long long *>(flag.as<uint32_t>() + 2 + 2 * k));
            if (hipGetLastError() != hipSuccess) rc = set_error(ctx, GBX_ERR_HIP, "gbx_graph_from_csr: check launch");
        }
        uint64_t degsum[2] = {0, 0};
        if (!rc && (hipMemcpyAsync(hflag, flag.p, 16, hipMemcpyDeviceToHost, ctx->stream) ||
                    hipMemcpyAsync(degsum, flag.as<uint32_t>() + 2, 16, hipMemcpyDeviceToHost, ctx->stream) ||
                    hipStreamSynchronize(ctx->stream)))
            rc = set_error(ctx, GBX_ERR_HIP, "gbx_graph_from_csr: check failed");
        for (int k = 0; k < nchecked; k++)
            if (degsum[k] != num_edges) hflag[0] |= 4u;
        if (!rc && (hflag[0] & 2)) rc = set_error(ctx, GBX_ERR_RANGE, "gbx_graph_from_csr: vertex id >= num_vertices");
        if (!rc && (hflag[0] & 1)) rc = set_error(ctx, GBX_ERR_ARG, "gbx_graph_from_csr: edges not sorted by source");
        if (!rc && (hflag[0] & 4))
            rc = set_error(ctx, GBX_ERR_ARG, "gbx_graph_from_csr: out_degree / edges_idx do not match the sorted "
                                             "edges (an edge outside its source's range, or degrees not summing "
                                             "to num_edges)");
    }
    if (!rc && fwd_only)
        rc = build_csr_device(ctx, g->inv.src.as<uint32_t>(), g->inv.dst.as<uint32_t>(), g->inv.w.as<float>(),
                              num_edges, num_vertices, GBX_SORT_COUNT_SRC, g->weighted,
                              g->inv.out_degree.as<uint32_t>(), g->inv.edges_idx.as<uint32_t>());
    if (!rc) {
        // in_degree of each side = out_degree of the other (vertex.c:291-311)
        if (hipMemcpyAsync(g->fwd.in_degree.p, g->inv.out_degree.p, nb, hipMemcpyDeviceToDevice, ctx->stream) ||
            hipMemcpyAsync(g->inv.in_degree.p, g->fwd.out_degree.p, nb, hipMemcpyDeviceToDevice, ctx->stream))
            rc = set_error(ctx, GBX_ERR_HIP, "gbx_graph_from_csr: copy failed");
    }
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess)
        rc = set_error(ctx, GBX_ERR_HIP, "gbx_graph_from_csr: sync failed");
    if (rc) {
        delete g;
        return rc;
    }
    *out = g;
    return GBX_OK;
}

int gbx_graph_info(const gbx_graph *g, uint32_t *num_vertices, uint32_t *num_edges)
{
    if (!g) return set_error(nullptr, GBX_ERR_ARG, "gbx_graph_info: graph is NULL");
    if (num_vertices) *num_vertices = g->N;
    if (num_edges) *num_edges = g->E;
    return GBX_OK;
}

int gbx_graph_export(const gbx_graph *g, int inverse, uint32_t *out_degree, uint32_t *in_degree,
                     uint32_t *edges_idx, uint32_t *src, uint32_t *dst, float *w)
{
    if (!g) return set_error(nullptr, GBX_ERR_ARG, "gbx_graph_export: graph is NULL");
    gbx_ctx *ctx = g->ctx;
    const gbx_csr_dev &c = inverse ? g->inv : g->fwd;
    const size_t nb = (size_t)g->N * 4, eb = (size_t)g->E * 4;
    GBX_HIP(ctx, hipSetDevice(ctx->device));
    if (out_degree) GBX_HIP(ctx, hipMemcpyAsync(out_degree, c.out_degree.p, nb, hipMemcpyDeviceToHost, ctx->stream));
    if (in_degree) GBX_HIP(ctx, hipMemcpyAsync(in_degree, c.in_degree.p, nb, hipMemcpyDeviceToHost, ctx->stream));
    if (edges_idx) GBX_HIP(ctx, hipMemcpyAsync(edges_idx, c.edges_idx.p, nb, hipMemcpyDeviceToHost, ctx->stream));
    if (eb) {
        if (src) GBX_HIP(ctx, hipMemcpyAsync(src, c.src.p, eb, hipMemcpyDeviceToHost, ctx->stream));
        if (dst) GBX_HIP(ctx, hipMemcpyAsync(dst, c.dst.p, eb, hipMemcpyDeviceToHost, ctx->stream));
        if (w) GBX_HIP(ctx, hipMemcpyAsync(w, c.w.p, eb, hipMemcpyDeviceToHost, ctx->stream));
    }
    GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GBX_OK;
}

int gbx_graph_free(gbx_graph *g)
{
    if (!g) return GBX_OK;
    (void)hipSetDevice(g->ctx->device);
    (void)hipStreamSynchronize(g->ctx->stream);
    delete g;
    return GBX_OK;
}

}  // extern "C"
