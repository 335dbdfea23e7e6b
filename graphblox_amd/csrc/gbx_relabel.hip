// gbx_relabel.hip -- the relabel levels of graphCSRPreProcessingStep
// (graphCSR.c:268-290): reorderGraphProcess (reorder.c:1556-1616) computes a
// new id for every vertex from its degree, the edges are renamed
// (relabelEdgeList, reorder.c:1618-1637) and sorted again.
//
// Every order below is a stable sort of vertex ids by a 32-bit key on the
// GPU (radix_sort_pairs), the ids entering in ascending or descending order so
// that equal keys keep the reference's tie order:
//   degree (lmode 1, 2, 3, 10 and codes without a method of their own,
//     reorder.c:296-385): degree descending, equal degrees by id descending
//     -> key ~degree over descending ids;
//   DBG and Corder (4, 5, 12, 13; Corder runs the DBG list function,
//     reorder.c:574-624): 11 degree buckets, thresholds t0 = avg / 2 (1 when
//     avg <= 1) doubling, the last unbounded, taken from high to low with ids
//     ascending -> key 10 - bucket over ascending ids;
//   HUBCluster (8, 9): the same with 2 buckets split at the average degree;
//   HUBSort (6, 7): hubs (degree > t0) by degree descending, equal degrees by
//     id descending, then the other vertices by id ascending -> the hubs,
//     compacted in descending id order, sorted by ~degree; the others ranked
//     by a scan;
//   file (11): the caller's labels.
// Degrees (reorder.c:1483-1552): out (1, 4, 6, 8, 12, 13 and other codes),
// in (2, 5, 7, 9), in + out (3), or MT19937 outputs modulo N (10; one
// generator stream, as one reference thread draws them).
#include <algorithm>
#include <cstring>
#include <vector>

#include "gbx_internal.hpp"

namespace gbx {

namespace {

__global__ void degree_count(const uint32_t *__restrict__ src, const uint32_t *__restrict__ dst, uint32_t E,
                             int which, uint32_t *__restrict__ deg)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < E; i += gridDim.x * blockDim.x) {
        if (which != 1) atomicAdd(deg + src[i], 1u);  // out (and in + out)
        if (which != 0) atomicAdd(deg + dst[i], 1u);  // in (and in + out)
    }
}

// ids N-1-i (descending) or i, with the key of each id
__global__ void order_keys(const uint32_t *__restrict__ deg, uint32_t N, int method, int descending,
                           const uint32_t *__restrict__ th, uint32_t nth, uint32_t *__restrict__ key,
                           uint32_t *__restrict__ ids)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const uint32_t v = descending ? N - 1 - i : i;
        const uint32_t d = deg[v];
        uint32_t k;
        if (method == 0) {
            k = ~d;
        } else {
            uint32_t b = 0;
            while (b + 1 < nth && d > th[b]) b++;  // first bucket with d <= threshold
            k = nth - 1 - b;
        }
        key[i] = k;
        ids[i] = v;
    }
}

__global__ void scatter_ranks(const uint32_t *__restrict__ ids, uint32_t n, uint32_t base,
                              uint32_t *__restrict__ labels)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        labels[ids[i]] = base + i;
}

// HUBSort: flags of hubs in descending id order and of the others in id order
__global__ void hub_flags(const uint32_t *__restrict__ deg, uint32_t N, uint32_t t0, uint32_t *__restrict__ hub_desc,
                          uint32_t *__restrict__ cold)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        hub_desc[i] = deg[N - 1 - i] > t0 ? 1u : 0u;
        cold[i] = deg[i] > t0 ? 0u : 1u;
    }
}

__global__ void hub_compact(const uint32_t *__restrict__ deg, uint32_t N, const uint32_t *__restrict__ hub_desc,
                            const uint32_t *__restrict__ pos, uint32_t *__restrict__ key, uint32_t *__restrict__ ids)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
        if (hub_desc[i]) {
            const uint32_t v = N - 1 - i;
            key[pos[i]] = ~deg[v];
            ids[pos[i]] = v;
        }
}

__global__ void cold_labels(const uint32_t *__restrict__ cold, const uint32_t *__restrict__ pos, uint32_t N,
                            uint32_t H, uint32_t *__restrict__ labels)
{
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x)
        if (cold[v]) labels[v] = H + pos[v];
}

__global__ void rename_edges(uint32_t *__restrict__ src, uint32_t *__restrict__ dst, uint32_t E,
                             const uint32_t *__restrict__ labels)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < E; i += gridDim.x * blockDim.x) {
        src[i] = labels[src[i]];
        dst[i] = labels[dst[i]];
    }
}

__global__ void compose_labels(uint32_t *__restrict__ total, uint32_t N, const uint32_t *__restrict__ labels)
{
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x)
        total[v] = labels[total[v]];
}

__global__ void iota_u32(uint32_t *__restrict__ a, uint32_t N)
{
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) a[v] = v;
}

// MT19937 (Matsumoto & Nishimura 1998) seeded as the 1999 reference code
// (mt19937.c:46-57): each state word from two steps of x <- 69069 x + 1,
// upper 16 bits each; tempered 32-bit outputs.
void mt19937_fill(uint32_t seed, uint32_t *out, size_t count)
{
    constexpr int n = 624, m = 397;
    uint32_t mt[n];
    uint32_t x = seed;
    for (int i = 0; i < n; i++) {
        const uint32_t hi = x & 0xFFFF0000u;
        x = 69069u * x + 1u;
        mt[i] = hi | ((x & 0xFFFF0000u) >> 16);
        x = 69069u * x + 1u;
    }
    int idx = n;
    for (size_t c = 0; c < count; c++) {
        if (idx >= n) {
            for (int k = 0; k < n; k++) {
                const uint32_t y = (mt[k] & 0x80000000u) | (mt[(k + 1) % n] & 0x7FFFFFFFu);
                mt[k] = mt[(k + m) % n] ^ (y >> 1) ^ ((y & 1u) ? 0x9908B0DFu : 0u);
            }
            idx = 0;
        }
        uint32_t y = mt[idx++];
        y ^= y >> 11;
        y ^= (y << 7) & 0x9D2C5680u;
        y ^= (y << 15) & 0xEFC60000u;
        y ^= y >> 18;
        out[c] = y;
    }
}

unsigned blocks_for(uint32_t n) { return grid_for(n, 256, 16384); }

// New ids of one level into labels (device, N).
int level_labels(gbx_ctx *ctx, const uint32_t *src, const uint32_t *dst, uint32_t E, uint32_t N, uint32_t avg,
                 uint32_t lmode, const gbx_relabel *rl, DevBuf &labels)
{
    hipStream_t st = ctx->stream;
    GBX_TRY(dev_alloc(ctx, labels, (size_t)N * 4 + 4));
    if (lmode == 11) {
        if (!rl->file_labels || rl->num_file_labels < N)
            return set_error(ctx, GBX_ERR_ARG, "relabel from file: %u labels for %u vertices", rl->num_file_labels,
                             N);
        for (uint32_t v = 0; v < N; v++)
            if (rl->file_labels[v] >= N)
                return set_error(ctx, GBX_ERR_RANGE, "relabel from file: label %u of vertex %u is not below %u",
                                 rl->file_labels[v], v, N);
        GBX_HIP(ctx, hipMemcpyAsync(labels.p, rl->file_labels, (size_t)N * 4, hipMemcpyHostToDevice, st));
        return GBX_OK;
    }
    DevBuf deg, key, ids, th;
    GBX_TRY(dev_alloc(ctx, deg, (size_t)N * 4 + 4));
    GBX_TRY(dev_alloc(ctx, key, (size_t)N * 4 + 4));
    GBX_TRY(dev_alloc(ctx, ids, (size_t)N * 4 + 4));
    if (lmode == 10) {
        std::vector<uint32_t> h(N);
        mt19937_fill(27491095u, h.data(), N);  // reorder.c:1541-1546
        for (uint32_t v = 0; v < N; v++) h[v] %= N;
        GBX_HIP(ctx, hipMemcpyAsync(deg.p, h.data(), (size_t)N * 4, hipMemcpyHostToDevice, st));
        GBX_HIP(ctx, hipStreamSynchronize(st));
    } else {
        GBX_TRY(fill_u32(ctx, deg.as<uint32_t>(), 0, N));
        const int which = (lmode == 2 || lmode == 5 || lmode == 7 || lmode == 9) ? 1 : (lmode == 3 ? 2 : 0);
        if (E) degree_count<<<blocks_for(E), 256, 0, st>>>(src, dst, E, which, deg.as<uint32_t>());
        GBX_HIP(ctx, hipGetLastError());
    }
    const uint32_t t0 = avg <= 1 ? 1u : avg / 2;
    if (lmode == 6 || lmode == 7) {
        DevBuf hub, cold, hpos, cpos, total;
        GBX_TRY(dev_alloc(ctx, hub, (size_t)N * 4 + 4));
        GBX_TRY(dev_alloc(ctx, cold, (size_t)N * 4 + 4));
        GBX_TRY(dev_alloc(ctx, hpos, (size_t)N * 4 + 4));
        GBX_TRY(dev_alloc(ctx, cpos, (size_t)N * 4 + 4));
        GBX_TRY(dev_alloc(ctx, total, 16));
        hub_flags<<<blocks_for(N), 256, 0, st>>>(deg.as<uint32_t>(), N, t0, hub.as<uint32_t>(), cold.as<uint32_t>());
        GBX_HIP(ctx, hipGetLastError());
        GBX_TRY(exclusive_scan_u32(ctx, hub.as<uint32_t>(), hpos.as<uint32_t>(), N, total.as<uint32_t>()));
        GBX_TRY(exclusive_scan_u32(ctx, cold.as<uint32_t>(), cpos.as<uint32_t>(), N, nullptr));
        uint32_t H = 0;
        GBX_HIP(ctx, hipMemcpyAsync(&H, total.p, 4, hipMemcpyDeviceToHost, st));
        GBX_HIP(ctx, hipStreamSynchronize(st));
        hub_compact<<<blocks_for(N), 256, 0, st>>>(deg.as<uint32_t>(), N, hub.as<uint32_t>(), hpos.as<uint32_t>(),
                                                   key.as<uint32_t>(), ids.as<uint32_t>());
        GBX_HIP(ctx, hipGetLastError());
        GBX_TRY(radix_sort_pairs(ctx, key.as<uint32_t>(), ids.as<uint32_t>(), nullptr, H, 32));
        if (H) scatter_ranks<<<blocks_for(H), 256, 0, st>>>(ids.as<uint32_t>(), H, 0, labels.as<uint32_t>());
        cold_labels<<<blocks_for(N), 256, 0, st>>>(cold.as<uint32_t>(), cpos.as<uint32_t>(), N, H,
                                                   labels.as<uint32_t>());
        GBX_HIP(ctx, hipGetLastError());
        GBX_HIP(ctx, hipStreamSynchronize(st));
        return GBX_OK;
    }
    int method = 0;  // degree order
    uint32_t nth = 0;
    uint32_t h_th[11];
    if (lmode == 4 || lmode == 5 || lmode == 12 || lmode == 13) {
        method = 1;
        nth = 11;
        h_th[0] = t0;
        for (int i = 1; i < 10; i++) h_th[i] = h_th[i - 1] * 2u;  // reorder.c:409-413
        h_th[10] = 0xFFFFFFFFu;
    } else if (lmode == 8 || lmode == 9) {
        method = 1;
        nth = 2;
        h_th[0] = avg;  // reorder.c:1003
        h_th[1] = 0xFFFFFFFFu;
    }
    GBX_TRY(dev_alloc(ctx, th, sizeof(h_th)));
    GBX_HIP(ctx, hipMemcpyAsync(th.p, h_th, sizeof(h_th), hipMemcpyHostToDevice, st));
    order_keys<<<blocks_for(N), 256, 0, st>>>(deg.as<uint32_t>(), N, method, method == 0, th.as<uint32_t>(), nth,
                                              key.as<uint32_t>(), ids.as<uint32_t>());
    GBX_HIP(ctx, hipGetLastError());
    GBX_TRY(radix_sort_pairs(ctx, key.as<uint32_t>(), ids.as<uint32_t>(), nullptr, N, method == 0 ? 32 : 8));
    scatter_ranks<<<blocks_for(N), 256, 0, st>>>(ids.as<uint32_t>(), N, 0, labels.as<uint32_t>());
    GBX_HIP(ctx, hipGetLastError());
    GBX_HIP(ctx, hipStreamSynchronize(st));
    return GBX_OK;
}

}  // namespace

bool relabel_requested(const gbx_relabel *rl)
{
    return rl && (rl->lmode[0] || rl->lmode[1] || rl->lmode[2]);
}

int relabel_validate(gbx_ctx *ctx, const gbx_relabel *rl)
{
    if (!rl) return GBX_OK;
    for (int l = 0; l < 3; l++)
        if (rl->lmode[l] == 11 && !rl->file_labels)
            return set_error(ctx, GBX_ERR_ARG, "relabel: lmode 11 needs file labels");
    return GBX_OK;
}

int relabel_edges(gbx_ctx *ctx, DevBuf &src, DevBuf &dst, DevBuf &w, uint32_t E, uint32_t N, uint32_t avg,
                  int sort_mode, bool weighted, uint32_t *out_degree, uint32_t *edges_idx, const gbx_relabel *rl,
                  DevBuf &total)
{
    hipStream_t st = ctx->stream;
    GBX_TRY(dev_alloc(ctx, total, (size_t)N * 4 + 4));
    iota_u32<<<blocks_for(N), 256, 0, st>>>(total.as<uint32_t>(), N);
    GBX_HIP(ctx, hipGetLastError());
    for (int l = 0; l < 3; l++) {
        if (!rl->lmode[l]) continue;
        DevBuf labels;
        GBX_TRY(level_labels(ctx, src.as<uint32_t>(), dst.as<uint32_t>(), E, N, avg, rl->lmode[l], rl, labels));
        if (E) rename_edges<<<blocks_for(E), 256, 0, st>>>(src.as<uint32_t>(), dst.as<uint32_t>(), E,
                                                           labels.as<uint32_t>());
        compose_labels<<<blocks_for(N), 256, 0, st>>>(total.as<uint32_t>(), N, labels.as<uint32_t>());
        GBX_HIP(ctx, hipGetLastError());
        // sortRunAlgorithms again on the new ids (graphCSR.c:270-271)
        GBX_TRY(build_csr_device(ctx, src.as<uint32_t>(), dst.as<uint32_t>(), w.as<float>(), E, N, sort_mode,
                                 weighted, out_degree, edges_idx));
        GBX_HIP(ctx, hipStreamSynchronize(st));
    }
    return GBX_OK;
}

}  // namespace gbx
