// gbx_rmat.hip -- seeded counter-based R-MAT generator on the GPU.
//
// The reference ships no generator; benchmark graphs are produced here.
// Every edge is a pure function of (seed, edge index): for each pair of
// recursion levels one SplitMix64 finaliser output supplies two 32-bit
// draws compared against integer thresholds floor(p * 2^32), so the CPU
// oracle's generator reproduces the same edges bit for bit.  Vertex ids are
// relabelled by a seeded bijection of [0, 2^scale) (odd multiply + add, then
// xorshift, four rounds).  Self loops and duplicate pairs are removed keeping
// the first occurrence, in generation order.
#include <cmath>

#include "gbx_internal.hpp"

namespace gbx {

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct PermKeys {
    uint32_t mul[4], add[4];
    uint32_t mask, sh;
};

static PermKeys make_perm_keys(uint32_t scale, uint64_t seed)
{
    PermKeys k;
    k.mask = scale >= 32 ? 0xFFFFFFFFu : ((1u << scale) - 1u);
    k.sh = scale / 2 + 1;
    for (uint32_t r = 0; r < 4; r++) {
        uint64_t h = mix64(seed ^ (0xA5A5A5A5ull + r * 0x9E3779B97F4A7C15ull));
        k.mul[r] = (uint32_t)h | 1u;
        k.add[r] = (uint32_t)(h >> 32);
    }
    return k;
}

__device__ __forceinline__ uint32_t permute(uint32_t x, const PermKeys &k)
{
#pragma unroll
    for (int r = 0; r < 4; r++) {
        x = (x * k.mul[r] + k.add[r]) & k.mask;
        x ^= x >> k.sh;
    }
    return x & k.mask;
}

__global__ void rmat_kernel(uint32_t scale, uint64_t M, uint32_t ta, uint32_t tab, uint32_t tabc,
                            uint64_t key, PermKeys pk, uint32_t *__restrict__ src,
                            uint32_t *__restrict__ dst)
{
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < M;
         e += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t u = 0, v = 0;
        for (uint32_t l = 0; l < scale; l += 2) {
            uint64_t h = mix64(key + (e << 5) + (l >> 1) + 1);
            uint32_t r = (uint32_t)h;
            uint32_t bu = r >= tab;
            uint32_t bv = (r >= ta && r < tab) || r >= tabc;
            u = (u << 1) | bu;
            v = (v << 1) | bv;
            if (l + 1 < scale) {
                r = (uint32_t)(h >> 32);
                bu = r >= tab;
                bv = (r >= ta && r < tab) || r >= tabc;
                u = (u << 1) | bu;
                v = (v << 1) | bv;
            }
        }
        src[e] = permute(u, pk);
        dst[e] = permute(v, pk);
    }
}

__global__ void iota_kernel(uint32_t *p, uint32_t n)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        p[i] = i;
}

// keep[orig] = first occurrence of a non-loop pair in a (src,dst)-stable-sorted list
__global__ void keep_first_flags(const uint32_t *__restrict__ s, const uint32_t *__restrict__ d,
                                 const uint32_t *__restrict__ orig, uint32_t n,
                                 uint32_t *__restrict__ keep)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        bool first = i == 0 || s[i - 1] != s[i] || d[i - 1] != d[i];
        keep[orig[i]] = (first && s[i] != d[i]) ? 1u : 0u;
    }
}

__global__ void compact_pairs(const uint32_t *__restrict__ s, const uint32_t *__restrict__ d,
                              const uint32_t *__restrict__ keep, const uint32_t *__restrict__ pos,
                              uint32_t n, uint32_t *__restrict__ so, uint32_t *__restrict__ dout)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (!keep[i]) continue;
        uint32_t p = pos[i];
        so[p] = s[i];
        dout[p] = d[i];
    }
}

// weight_kind 1: uniform [0,1) with 24-bit resolution; 2: lognormal(0,1)
// by Box-Muller on two 24-bit uniforms in (0,1].
__global__ void weights_kernel(float *w, uint32_t n, int kind, uint64_t key)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint64_t h = mix64(key + (uint64_t)i * 0x9E3779B97F4A7C15ull);
        if (kind == 1) {
            w[i] = (float)(h >> 40) * (1.0f / 16777216.0f);
        } else if (kind == 2) {
            float u1 = ((float)(h >> 40) + 1.0f) * (1.0f / 16777216.0f);
            float u2 = (float)((h >> 8) & 0xFFFFFFull) * (1.0f / 16777216.0f);
            float z = sqrtf(-2.0f * logf(u1)) * cosf(6.283185307179586f * u2);
            w[i] = expf(z);
        } else {
            w[i] = 1.0f;
        }
    }
}

// Generates into ds/dd (length M) then, when dedupe, compacts in place;
// returns the kept count in *E_out.
static int rmat_device(gbx_ctx *ctx, uint32_t scale, uint32_t edge_factor, double a, double b,
                       double c, uint64_t seed, bool dedupe, int weight_kind, uint64_t weight_seed,
                       DevBuf &ds, DevBuf &dd, DevBuf &dw, uint32_t *E_out)
{
    if (scale == 0 || scale > 30)
        return set_error(ctx, GBX_ERR_ARG, "rmat: scale %u not in 1..30", scale);
    if (!(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0))
        return set_error(ctx, GBX_ERR_ARG, "rmat: bad quadrant probabilities");
    const uint64_t M = (uint64_t)edge_factor << scale;
    if (M == 0 || M > 0xFFFFFFF0ull)
        return set_error(ctx, GBX_ERR_ARG, "rmat: edge count %llu out of range",
                         (unsigned long long)M);
    const uint32_t ta = (uint32_t)(a * 4294967296.0);
    const uint32_t tab = (uint32_t)((a + b) * 4294967296.0);
    const uint32_t tabc = (uint32_t)((a + b + c) * 4294967296.0);
    GBX_TRY(dev_alloc(ctx, ds, M * 4));
    GBX_TRY(dev_alloc(ctx, dd, M * 4));
    rmat_kernel<<<grid_for(M, 256, 65536), 256, 0, ctx->stream>>>(
        scale, M, ta, tab, tabc, mix64(seed), make_perm_keys(scale, seed), ds.as<uint32_t>(),
        dd.as<uint32_t>());
    GBX_HIP(ctx, hipGetLastError());
    uint32_t n = (uint32_t)M;
    if (dedupe) {
        DevBuf s2, d2, idx, keep, pos;
        GBX_TRY(dev_alloc(ctx, s2, M * 4));
        GBX_TRY(dev_alloc(ctx, d2, M * 4));
        GBX_TRY(dev_alloc(ctx, idx, M * 4));
        GBX_HIP(ctx, hipMemcpyAsync(s2.p, ds.p, M * 4, hipMemcpyDeviceToDevice, ctx->stream));
        GBX_HIP(ctx, hipMemcpyAsync(d2.p, dd.p, M * 4, hipMemcpyDeviceToDevice, ctx->stream));
        iota_kernel<<<grid_for(M, 256, 16384), 256, 0, ctx->stream>>>(idx.as<uint32_t>(), n);
        GBX_HIP(ctx, hipGetLastError());
        // stable (src,dst) order carrying the original index as third payload
        GBX_TRY(radix_sort_pairs(ctx, d2.as<uint32_t>(), s2.as<uint32_t>(), idx.as<float>(), n, scale));
        GBX_TRY(radix_sort_pairs(ctx, s2.as<uint32_t>(), d2.as<uint32_t>(), idx.as<float>(), n, scale));
        GBX_TRY(dev_alloc(ctx, keep, M * 4 + 4));
        GBX_TRY(dev_alloc(ctx, pos, M * 4 + 4));
        keep_first_flags<<<grid_for(M, 256, 16384), 256, 0, ctx->stream>>>(
            s2.as<uint32_t>(), d2.as<uint32_t>(), idx.as<uint32_t>(), n, keep.as<uint32_t>());
        GBX_HIP(ctx, hipGetLastError());
        GBX_TRY(exclusive_scan_u32(ctx, keep.as<uint32_t>(), pos.as<uint32_t>(), n,
                                   pos.as<uint32_t>() + n));
        uint32_t kept = 0;
        GBX_HIP(ctx, hipMemcpyAsync(&kept, pos.as<uint32_t>() + n, 4, hipMemcpyDeviceToHost, ctx->stream));
        GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        // compact from the generation-order arrays into the sorted scratch
        compact_pairs<<<grid_for(M, 256, 16384), 256, 0, ctx->stream>>>(
            ds.as<uint32_t>(), dd.as<uint32_t>(), keep.as<uint32_t>(), pos.as<uint32_t>(), n,
            s2.as<uint32_t>(), d2.as<uint32_t>());
        GBX_HIP(ctx, hipGetLastError());
        ds = std::move(s2);
        dd = std::move(d2);
        n = kept;
    }
    GBX_TRY(dev_alloc(ctx, dw, (size_t)n * 4));
    if (n) {
        weights_kernel<<<grid_for(n, 256, 16384), 256, 0, ctx->stream>>>(
            dw.as<float>(), n, weight_kind, mix64(weight_seed ^ 0x5EEDull));
        GBX_HIP(ctx, hipGetLastError());
    }
    GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *E_out = n;
    return GBX_OK;
}

}  // namespace gbx

using namespace gbx;

extern "C" {

int gbx_rmat_edges(gbx_ctx *ctx, uint32_t scale, uint32_t edge_factor, double a, double b,
                   double c, uint64_t seed, int dedupe, int weight_kind, uint64_t weight_seed,
                   uint32_t *src, uint32_t *dst, float *w, uint64_t *num_edges)
{
    if (!ctx || !num_edges) return set_error(ctx, GBX_ERR_ARG, "gbx_rmat_edges: NULL argument");
    GBX_HIP(ctx, hipSetDevice(ctx->device));
    DevBuf ds, dd, dw;
    uint32_t E = 0;
    GBX_TRY(rmat_device(ctx, scale, edge_factor, a, b, c, seed, dedupe != 0, weight_kind,
                        weight_seed, ds, dd, dw, &E));
    const size_t eb = (size_t)E * 4;
    if (eb) {
        if (src) GBX_HIP(ctx, hipMemcpyAsync(src, ds.p, eb, hipMemcpyDeviceToHost, ctx->stream));
        if (dst) GBX_HIP(ctx, hipMemcpyAsync(dst, dd.p, eb, hipMemcpyDeviceToHost, ctx->stream));
        if (w) GBX_HIP(ctx, hipMemcpyAsync(w, dw.p, eb, hipMemcpyDeviceToHost, ctx->stream));
    }
    GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *num_edges = E;
    return GBX_OK;
}

int gbx_graph_create_rmat(gbx_ctx *ctx, uint32_t scale, uint32_t edge_factor, double a, double b,
                          double c, uint64_t seed, int weight_kind, uint64_t weight_seed,
                          int sort_mode, unsigned flags, gbx_graph **out)
{
    if (!ctx || !out) return set_error(ctx, GBX_ERR_ARG, "gbx_graph_create_rmat: NULL argument");
    *out = nullptr;
    if (sort_mode < 0 || sort_mode > 3)
        return set_error(ctx, GBX_ERR_ARG, "gbx_graph_create_rmat: sort_mode %d", sort_mode);
    GBX_HIP(ctx, hipSetDevice(ctx->device));
    DevBuf ds, dd, dw;
    uint32_t E = 0;
    GBX_TRY(rmat_device(ctx, scale, edge_factor, a, b, c, seed, true, weight_kind, weight_seed,
                        ds, dd, dw, &E));
    uint32_t ms = 0, md = 0;
    GBX_TRY(max_u32(ctx, ds.as<uint32_t>(), E, &ms));
    GBX_TRY(max_u32(ctx, dd.as<uint32_t>(), E, &md));
    const uint32_t N = E ? (ms > md ? ms : md) + 1 : 1;  // edgeList.c:679
    // the generated list is already free of loops and duplicates
    return graph_from_device_edges(ctx, ds, dd, dw, E, N, sort_mode, flags & ~GBX_GRAPH_DEDUPE, out);
}

}  // extern "C"
