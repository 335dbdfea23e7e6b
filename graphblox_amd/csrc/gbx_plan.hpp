// gbx_plan.hpp -- the degree-binned pull plan shared by the single-vector
// (gbx_pr.hip) and multi-vector (gbx_ppr.hip) kernels, and device helpers.
#pragma once

#include "gbx_internal.hpp"

namespace gbx {

// 1024-thread blocks, one per CU, each holding the hottest contributions in
// (almost) all of the CU's 160 KiB of LDS.
#ifndef GBX_BLOCK
#define GBX_BLOCK 1024
#endif
constexpr int kBlock = GBX_BLOCK;
#ifndef GBX_LDS_DEFAULT
#define GBX_LDS_DEFAULT 159744
#endif
constexpr int kMaxClasses = 9;

enum Kind : int {
    K_PR_REF = 0,      // pageRank.c: r' = (1-d) + d * sum(r/outdeg)
    K_PR_STOCH = 1,    // stochastic, unweighted
    K_PR_STOCH_W = 2,  // stochastic, row-normalised edge weights
    K_SPMV = 3,        // y += 0.0001f * x  (SPMV.c:177)
    K_SPMV_W = 4,      // y += w * x
};

struct ClassDesc {
    uint32_t lanes;       // lanes per row (1..64); 256 = hub chunks
    uint32_t row_begin, row_end;
    uint32_t tile_begin, tile_end;
};

struct HubChunk {
    uint32_t row, ebeg, eend;
    uint32_t hub;    // hub index (counter slot)
    uint32_t first;  // first chunk of this hub
    uint32_t nch;    // chunks of this hub
};

struct Slot {
    double err, dang, sum;
    uint32_t active, done, iters, conv;
};

struct Partial {
    double err, dang, sum;
    uint32_t active, pad;
};

}  // namespace gbx

struct gbx_plan {
    int direction = 0;
    uint32_t N = 0, E = 0;
    gbx::DevBuf perm, inv, rp, col, wt, wn, wn16, outdeg, chunks, hub_cnt, hub_part;
    uint32_t nclass = 0;
    gbx::ClassDesc cls[gbx::kMaxClasses];
    uint64_t cls_edges[gbx::kMaxClasses];
    uint32_t n_chunks = 0, n_hubs = 0, n_tiles = 0;
    bool has_wt = false, has_wn = false, has_wn16 = false;
};

struct gbx_ppr_state;

struct gbx_pr {
    gbx_ctx *ctx = nullptr;
    gbx_graph *g = nullptr;
    gbx_plan *plan = nullptr;
    gbx_pr_params p;
    int kind = 0;
    size_t elem = 4;      // bytes per contribution (4: f32, 2: bf16)
    uint32_t hot = 0;     // contributions held in LDS per block
    gbx::DevBuf r, c[2], tele, slots, partials, counter, out;
    bool has_tele = false;
    uint32_t t = 0, grid = 0;
    hipEvent_t ev_begin = nullptr, ev_end = nullptr;
    bool timing_started = false;
    double elapsed_ms = 0.0;
    gbx::Slot *h_slot = nullptr;  // pinned
    gbx_ppr_state *batch = nullptr;  // multi-vector state (gbx_ppr.hip), else null
};

namespace gbx {

// Streaming (read-once) data is loaded and stored non-temporally so that the
// L2 keeps the gathered contribution vector.
#ifndef GBX_NT
#define GBX_NT 0
#endif
#ifndef GBX_UNROLL
#define GBX_UNROLL 8
#endif

template <class T>
__device__ __forceinline__ T ld_stream(const T *p)
{
#if GBX_NT
    return __builtin_nontemporal_load(p);
#else
    return *p;
#endif
}

template <class T>
__device__ __forceinline__ void st_stream(T *p, T v)
{
#if GBX_NT
    __builtin_nontemporal_store(v, p);
#else
    *p = v;
#endif
}

// ------------------------------------------------------------------ accumulators

// Plain FP32 running sum: `acc += x` in visit order, as pageRank.c:313-314.
struct AccF32 {
    float s;
    __device__ __forceinline__ void init(float v) { s = v; }
    __device__ __forceinline__ void add(float x) { s += x; }
    __device__ __forceinline__ void merge(const AccF32 &o) { s += o.s; }
    __device__ __forceinline__ void shfl_merge(int off) { s += __shfl_xor(s, off, 64); }
    __device__ __forceinline__ float value() const { return s; }
};

// Compensated FP32 sum (Kahan 1965); partial sums are merged with Knuth's
// TwoSum so the rounding error of every merge is carried as well.
// Represented value: s - c.
struct AccKahan {
    float s, c;
    __device__ __forceinline__ void init(float v) { s = v; c = 0.0f; }
    __device__ __forceinline__ void add(float x)
    {
        const float y = x - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ __forceinline__ void merge(const AccKahan &o)
    {
        const float t = s + o.s;
        const float bp = t - s;
        const float e = (s - (t - bp)) + (o.s - bp);  // s + o.s == t + e exactly
        c = c + o.c - e;
        s = t;
    }
    __device__ __forceinline__ void shfl_merge(int off)
    {
        AccKahan o;
        o.s = __shfl_xor(s, off, 64);
        o.c = __shfl_xor(c, off, 64);
        merge(o);
    }
    __device__ __forceinline__ float value() const { return s - c; }
};

// ------------------------------------------------------------------ element types

__device__ __forceinline__ float load_val(const float *p, uint32_t i) { return p[i]; }
__device__ __forceinline__ float load_val(const uint16_t *p, uint32_t i)
{
    return __uint_as_float((uint32_t)p[i] << 16);  // bf16 -> f32 is exact
}
__device__ __forceinline__ uint16_t f32_to_bf16(float v)
{
    // round to nearest even; values here are finite
    uint32_t u = __float_as_uint(v);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ void store_val(float *p, uint32_t i, float v) { st_stream(p + i, v); }
__device__ __forceinline__ void store_val(uint16_t *p, uint32_t i, float v) { p[i] = f32_to_bf16(v); }

// Hot prefix of the contribution vector held in LDS for the whole kernel.
extern __shared__ __align__(16) unsigned char gbx_dyn_lds[];

// Multi-vector hooks implemented in gbx_ppr.hip.
int ppr_reset(gbx_pr *pr);
int ppr_launch_iter(gbx_pr *pr);
int ppr_read(gbx_pr *pr, float *ranks_out, gbx_pr_report *rep);
void ppr_state_free(gbx_ppr_state *s);

// Host helpers implemented in gbx_pr.hip.
int ensure_plan(gbx_ctx *ctx, gbx_graph *g, int direction, gbx_plan **out);
// Number of leading plan rows with degree >= threshold (rows are sorted by
// degree, descending).
int plan_rows_at_least(gbx_ctx *ctx, const gbx_plan *p, uint32_t threshold, uint32_t *count);
// Row pointers rp[0..nrows] of the plan, copied to host.
int plan_row_ptr_prefix(gbx_ctx *ctx, const gbx_plan *p, uint32_t nrows, std::vector<uint32_t> &out);

}  // namespace gbx
