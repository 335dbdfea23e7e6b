// gbx_plan.hpp -- the degree-binned pull plan shared by the single-vector
// (gbx_pr.hip) and multi-vector (gbx_ppr.hip) kernels, and device helpers.
#pragma once

#include "gbx_internal.hpp"

#include <memory>
#include <type_traits>
#include <utility>
#include <vector>

namespace gbx {

// 1024-thread blocks, one per CU, each holding the hottest contributions in
// (almost) all of the CU's 160 KiB of LDS.
#ifndef GBX_BLOCK
#define GBX_BLOCK 1024
#endif
constexpr int kBlock = GBX_BLOCK;
constexpr int kMaxClasses = 9;

enum Kind : int {
    K_PR_REF = 0,      // pageRank.c: r' = (1-d) + d * sum(r/outdeg)
    K_PR_STOCH = 1,    // stochastic, unweighted
    K_PR_STOCH_W = 2,  // stochastic, row-normalised edge weights
    K_SPMV = 3,        // y += 0.0001f * x  (SPMV.c:177)
    K_SPMV_W = 4,      // y += w * x
    K_ADJ = 5,         // adjoint of the weighted stochastic fixed point (forward-CSR rows)
    // the reference's fixed-point pull PageRank (pageRank.c:496-1030): integer
    // contributions of 64/32/16/8 bits, exact integer row sums
    K_FIX64 = 6,
    K_FIX32 = 7,
    K_FIX16 = 8,
    K_FIX8 = 9,
    // the reference's quantized pull PageRank (pageRank.c:1159-1577): FP32
    // contributions quantized per iteration to 32/16/8 bits with a min/max
    // scale, exact integer row sums
    K_Q32 = 10,
    K_Q16 = 11,
    K_Q8 = 12,
    // the reference's data-driven pull PageRank (pageRank.c:1711-1870):
    // worklist rows only, conditional update (pull iteration)
    K_DD = 13,
    // the reference's data-driven push and pull-push PageRank
    // (pageRank.c:1872-2202): residual pushes over a frontier (gbx_ddpush.hip)
    K_DDP = 14,
    K_DDPP = 15,
    // stochastic PageRank of a batch of graphs on the binned iteration: each
    // graph its own teleport 1/N_k, dangling mass and rank sum (gbx_batch.hip)
    K_SEG = 16,
};

constexpr bool is_fixed_kind(int k) { return k >= K_FIX64 && k <= K_FIX8; }
constexpr bool is_quant_kind(int k) { return k >= K_Q32 && k <= K_Q8; }
// kinds whose binned row sums are the reference's own uint64 integer sums
constexpr bool is_int_kind(int k) { return is_fixed_kind(k) || is_quant_kind(k); }
// kinds whose gathered values are edge weight times source value
constexpr bool is_weighted_kind(int k) { return k == K_PR_STOCH_W || k == K_ADJ || k == K_SPMV_W; }

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
    uint32_t graph;  // graph of the row (batched graphs), else 0
    uint32_t pad;
};

// One unit of work of a batch of graphs: rows [row_begin, row_end) of one
// graph with `lanes` lanes per row, or (lanes == 0) hub chunk row_begin.
struct TileDesc {
    uint32_t row_begin, row_end;
    uint32_t lanes;
    uint32_t graph;
    uint32_t slot;  // first per-wave partial slot of this tile
};

// Per-graph slot ranges of a batch of graphs.
struct SegRange {
    uint32_t slot_begin, slot_end;  // per-wave slots of the graph's row tiles
    uint32_t hub_begin, hub_end;    // hub-row slots (offset by hub_slot_base)
};

// cmax / cnmin: the largest contribution and the complement of the smallest
// as float bits (contributions are >= 0, so unsigned order is float order and
// both reduce with max from 0); the quantized kinds take their scale from them.
struct Slot {
    double err, dang, sum;
    uint32_t active, done, iters, conv;
    uint32_t cmax, cnmin;
};

struct Partial {
    double err, dang, sum;
    uint32_t active, pad;
    uint32_t cmax, cnmin;
};

// Batch of graphs: the update coefficients and statistics of graph k.
struct GCoef {
    float ti;  // 1/N_k
    float fs;  // d / S_k
    float fd;  // D_k / N_k
    uint32_t done;
};

struct GStat {
    double err, dang, sum;
    uint32_t iters, conv_iter, n, pad;
};

// Accumulate units of a batch of graphs on the binned iteration: every unit
// belongs to one graph; each of its waves leaves (err, dang, sum) of the
// rows it updated in its slot, kSegUnitSlots per unit.
constexpr uint32_t kSegUnitSlots = 16;

}  // namespace gbx

struct gbx_bplan;

struct gbx_plan {
    int direction = 0;
    uint32_t N = 0, E = 0;
    gbx::DevBuf perm, inv, rp, col, wt, wn, wn16, wr, outdeg, chunks, hub_cnt, hub_part;
    uint32_t nclass = 0;
    gbx::ClassDesc cls[gbx::kMaxClasses];
    uint64_t cls_edges[gbx::kMaxClasses];
    uint32_t n_chunks = 0, n_hubs = 0, n_tiles = 0;
    bool has_wt = false, has_wn = false, has_wn16 = false, has_wr = false;
    // batch of graphs: explicit graph-pure tiles and per-graph slot ranges
    uint32_t bins_thr_chunk[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // bins the plan was cut with
    bool segmented = false;
    uint32_t nseg = 0, n_seg_tiles = 0, n_wave_slots = 0, n_seg_hubs = 0, n_seg_chunks = 0;
    gbx::DevBuf tiles, seg_ranges, row_gid, seg_chunks, seg_hub_cnt, seg_hub_part;
    // binned (propagation-blocking) layouts, built on first use per cut;
    // shared with the PageRank objects that use them
    std::vector<std::shared_ptr<gbx_bplan>> bplans;
    ~gbx_plan();
};

struct gbx_ppr_state;
struct gbx_seg_state;
struct gbx_binned_state;
struct gbx_ddp_state;

struct gbx_pr {
    gbx_ctx *ctx = nullptr;
    gbx_graph *g = nullptr;
    gbx_plan *plan = nullptr;
    std::shared_ptr<gbx_plan> plan_ref;  // keeps the plan alive for this object
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
    gbx_seg_state *seg = nullptr;    // batch-of-graphs state (gbx_batch.hip), else null
    gbx_binned_state *bin = nullptr; // binned-method state (gbx_binned.hip), else null
    gbx_ddp_state *ddp = nullptr;    // data-driven push state (gbx_ddpush.hip), else null
    // captured iteration sequences: key = (iterations << 2) | (teleport set << 1) | starting parity
    std::vector<std::pair<uint64_t, hipGraphExec_t>> graph_cache;
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

// ------------------------------------------------------------------ cross-block hand-off
// Values that one block of a launch hands to another (a split row's chunk
// sums, the blocks' partials of a grid reduction, per-graph flags) are
// written and read with device-scope atomic stores and loads.  gfx950 keeps
// those coherent across the L2s of its XCDs (they are performed at the
// device's coherence point), so neither side needs a device-scope fence:
// such a release writes back the whole L2 of the writer's XCD (every block
// of an iteration fences, so the write-backs queue behind each other) and
// the acquire invalidates the reader's L2, together about 7 us of a 21 us
// iteration at R-MAT s16.  The protocol:
//   writer: handoff_store(...) of every value, stores_performed(), then the
//           ticket (a device-scope atomicAdd);
//   reader: the ticket's result says every writer has taken its ticket,
//           then handoff_load(...) of the values.
// A wave's stores are performed when its vmcnt has drained (gfx9 counts
// stores there too), so every value is at the coherence point before its
// writer's ticket is; the reader's loads issue after its own ticket
// returned and read the coherence point.  Values read back by later
// launches need nothing: a launch ends with its writes visible.
template <class W>
__device__ __forceinline__ void handoff_store(W *p, W v)
{
    static_assert(sizeof(W) == 4 || sizeof(W) == 8, "one atomic store per value");
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <class W>
__device__ __forceinline__ W handoff_load(const W *p)
{
    static_assert(sizeof(W) == 4 || sizeof(W) == 8, "one atomic load per value");
    return __hip_atomic_load(const_cast<W *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Every store this wave issued has been performed (waits for the vector
// memory counter; the memory clobber keeps the compiler from moving memory
// accesses across it).
__device__ __forceinline__ void stores_performed() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// A struct of 8-byte words (or 4-byte ones), word by word.
template <class W, class S>
__device__ __forceinline__ void handoff_store_words(S *p, const S &v)
{
    static_assert(sizeof(S) % sizeof(W) == 0, "whole words");
    W w[sizeof(S) / sizeof(W)];
    __builtin_memcpy(w, &v, sizeof(S));
#pragma unroll
    for (size_t k = 0; k < sizeof(S) / sizeof(W); k++) handoff_store(reinterpret_cast<W *>(p) + k, w[k]);
}

template <class W, class S>
__device__ __forceinline__ S handoff_load_words(const S *p)
{
    static_assert(sizeof(S) % sizeof(W) == 0, "whole words");
    W w[sizeof(S) / sizeof(W)];
#pragma unroll
    for (size_t k = 0; k < sizeof(S) / sizeof(W); k++) w[k] = handoff_load(reinterpret_cast<const W *>(p) + k);
    S v;
    __builtin_memcpy(&v, w, sizeof(S));
    return v;
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

// Unsigned 32-bit sum with wrap-around: the FPGA overlay's ALU_ACC
// (engine_alu_ops_kernel.sv:104-112), exact in any order.
struct AccU32 {
    uint32_t s;
    __device__ __forceinline__ void init(uint32_t v) { s = v; }
    __device__ __forceinline__ void add(uint32_t x) { s += x; }
    __device__ __forceinline__ void merge(const AccU32 &o) { s += o.s; }
    __device__ __forceinline__ void shfl_merge(int off) { s += (uint32_t)__shfl_xor((int)s, off, 64); }
    __device__ __forceinline__ uint32_t value() const { return s; }
};

// Data-driven pull: contributions carry "updated in the previous iteration"
// in their sign bit (they are >= 0 otherwise).  Sums the magnitudes in the
// same order as AccF32 and ORs the sign bits; value() returns the sum with
// the OR as its sign (a flagged row's sum is > 0: its flagged in-neighbour
// has an out-edge, so a positive contribution).
struct AccDD {
    float s;
    uint32_t f;
    __device__ __forceinline__ void init(float v)
    {
        s = fabsf(v);
        f = __float_as_uint(v) & 0x80000000u;
    }
    __device__ __forceinline__ void add(float x)
    {
        s += fabsf(x);
        f |= __float_as_uint(x) & 0x80000000u;
    }
    __device__ __forceinline__ void merge(const AccDD &o)
    {
        s += o.s;
        f |= o.f;
    }
    __device__ __forceinline__ void shfl_merge(int off)
    {
        s += __shfl_xor(s, off, 64);
        f |= (uint32_t)__shfl_xor((int)f, off, 64);
    }
    __device__ __forceinline__ float value() const { return __uint_as_float(__float_as_uint(s) | f); }
};

// ------------------------------------------------------------------ element types

__device__ __forceinline__ float load_val(const float *p, uint32_t i) { return p[i]; }
__device__ __forceinline__ float load_val(const uint16_t *p, uint32_t i)
{
    return __uint_as_float((uint32_t)p[i] << 16);  // bf16 -> f32 is exact
}
__device__ __forceinline__ uint16_t f32_to_bf16(float v)
{
    // round to nearest even: one v_cvt_pk_bf16_f32 on gfx950
    return __builtin_bit_cast(uint16_t, (__bf16)v);
}
__device__ __forceinline__ void store_val(float *p, uint32_t i, float v) { st_stream(p + i, v); }
__device__ __forceinline__ void store_val(uint16_t *p, uint32_t i, float v) { p[i] = f32_to_bf16(v); }

// Hot prefix of the contribution vector held in LDS for the whole kernel.
extern __shared__ __align__(16) unsigned char gbx_dyn_lds[];

struct PlanView {
    const uint32_t *__restrict__ rp;
    const uint32_t *__restrict__ col;
    const float *__restrict__ wt;  // raw (SpMV) or row-normalised (PR) weights
    const uint32_t *__restrict__ outdeg;
    const HubChunk *__restrict__ chunks;
    uint32_t *__restrict__ hub_cnt;
    float *__restrict__ hub_part;
    uint32_t N, n_tiles, n_chunks, nclass;
    ClassDesc cls[kMaxClasses];
};

// ------------------------------------------------------------------ gathers

// Contribution gather; with LDS the first `hot` (most gathered) vertices
// come from the block's LDS copy.
template <class T, bool LDS>
struct GatherC {
    const T *__restrict__ v;
    uint32_t hot;
    __device__ __forceinline__ float operator()(uint32_t, uint32_t c) const
    {
        if constexpr (LDS) {
            if (c < hot) return load_val(reinterpret_cast<const T *>(gbx_dyn_lds), c);
        }
        return load_val(v, c);
    }
};

// Weighted contribution gather: w[j] * v[c], weights of type T.
template <class T, bool LDS>
struct GatherCW {
    const T *__restrict__ v;
    const T *__restrict__ w;
    uint32_t hot;
    __device__ __forceinline__ float operator()(uint32_t j, uint32_t c) const
    {
        float x;
        if constexpr (LDS) {
            if (c < hot) x = load_val(reinterpret_cast<const T *>(gbx_dyn_lds), c);
            else x = load_val(v, c);
        } else {
            x = load_val(v, c);
        }
        float wj;
        if constexpr (sizeof(T) == 4) wj = ld_stream(reinterpret_cast<const float *>(w) + j);
        else wj = load_val(w, j);
        return wj * x;
    }
};

struct GatherConstW {
    const float *__restrict__ v;
    float w;
    __device__ __forceinline__ float operator()(uint32_t, uint32_t c) const { return w * v[c]; }
};

// Lane-strided, in-order partial sum of one row: GBX_UNROLL column loads,
// then GBX_UNROLL gathers in flight; the tail is one predicated batch.  The
// gathered values are float, or uint32_t for the integer accumulator.
template <int G, class Acc, class Gather>
__device__ __forceinline__ void lane_sum(uint32_t j, uint32_t end, const uint32_t *__restrict__ col,
                                         const Gather &g, Acc &acc)
{
    constexpr uint32_t U = GBX_UNROLL;
    using Val = decltype(g(0u, 0u));
    for (; j + (U - 1) * G < end; j += U * G) {
        uint32_t c[U];
        Val t[U];
#pragma unroll
        for (uint32_t k = 0; k < U; k++) c[k] = ld_stream(col + j + k * G);
#pragma unroll
        for (uint32_t k = 0; k < U; k++) t[k] = g(j + k * G, c[k]);
#pragma unroll
        for (uint32_t k = 0; k < U; k++) acc.add(t[k]);
    }
    if (j < end) {
        uint32_t c[U];
        Val t[U];
#pragma unroll
        for (uint32_t k = 0; k < U; k++) c[k] = (j + k * G < end) ? ld_stream(col + j + k * G) : 0u;
#pragma unroll
        for (uint32_t k = 0; k < U; k++) t[k] = (j + k * G < end) ? g(j + k * G, c[k]) : Val(0);
#pragma unroll
        for (uint32_t k = 0; k < U; k++)
            if (j + k * G < end) acc.add(t[k]);
    }
}

template <int G, class Acc, class Gather, class Epi>
__device__ __forceinline__ void do_rows(const ClassDesc &cd, uint32_t local, const PlanView &P,
                                        const Gather &g, Epi &epi)
{
    constexpr uint32_t RPT = kBlock / G;
    const uint32_t grp = threadIdx.x / G, lig = threadIdx.x % G;
    const uint32_t row = cd.row_begin + local * RPT + grp;
    const bool valid = row < cd.row_end;
    using Val = decltype(epi.init(0u));
    Acc acc;
    acc.init((G == 1 && valid) ? epi.init(row) : Val(0));
    if (valid) {
        const uint32_t beg = ld_stream(P.rp + row), end = ld_stream(P.rp + row + 1);
        lane_sum<G>(beg + lig, end, P.col, g, acc);
    }
    if constexpr (G > 1) {
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1) acc.shfl_merge(off);
    }
    if (valid && lig == 0) epi(row, G == 1 ? acc.value() : epi.init(row) + acc.value());
}

template <class Acc, class Gather, class Epi>
__device__ __forceinline__ void do_hub(uint32_t tile, const PlanView &P, const Gather &g, Epi &epi)
{
    __shared__ Acc red[kBlock / 64];
    using Val = decltype(epi.init(0u));
    // chunk partials: float bits, or the integer accumulator's own value
    using Part = std::conditional_t<std::is_same_v<Val, uint32_t>, uint32_t, float>;
    const HubChunk hc = P.chunks[tile];
    Acc acc;
    acc.init(Val(0));
    lane_sum<kBlock>(hc.ebeg + threadIdx.x, hc.eend, P.col, g, acc);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc.shfl_merge(off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        Acc s = red[0];
#pragma unroll
        for (int w = 1; w < kBlock / 64; w++) s.merge(red[w]);
        if (hc.nch == 1) {
            // the whole row in one chunk: no cross-block combine, no fences
            epi(hc.row, epi.init(hc.row) + s.value());
        } else {
            Part *hp = reinterpret_cast<Part *>(P.hub_part);
            handoff_store(hp + tile, (Part)s.value());
            stores_performed();
            uint32_t prev = atomicAdd(&P.hub_cnt[hc.hub], 1u);
            if (prev == hc.nch - 1) {
                Acc tot;
                tot.init(Val(0));
                for (uint32_t k = 0; k < hc.nch; k++) tot.add(handoff_load(hp + hc.first + k));
                P.hub_cnt[hc.hub] = 0;
                epi(hc.row, epi.init(hc.row) + tot.value());
            }
        }
    }
    __syncthreads();
}

template <class Acc, class Gather, class Epi>
__device__ __forceinline__ void tile_loop(const PlanView &P, const Gather &g, Epi &epi)
{
    for (uint32_t tile = blockIdx.x; tile < P.n_tiles; tile += gridDim.x) {
        if (tile < P.n_chunks) {
            do_hub<Acc>(tile, P, g, epi);
            continue;
        }
        uint32_t c = 1;
        while (c + 1 < P.nclass && tile >= P.cls[c].tile_end) c++;
        const ClassDesc &cd = P.cls[c];
        const uint32_t local = tile - cd.tile_begin;
        switch (cd.lanes) {
        case 64: do_rows<64, Acc>(cd, local, P, g, epi); break;
        case 32: do_rows<32, Acc>(cd, local, P, g, epi); break;
        case 16: do_rows<16, Acc>(cd, local, P, g, epi); break;
        case 8: do_rows<8, Acc>(cd, local, P, g, epi); break;
        case 4: do_rows<4, Acc>(cd, local, P, g, epi); break;
        case 2: do_rows<2, Acc>(cd, local, P, g, epi); break;
        default: do_rows<1, Acc>(cd, local, P, g, epi); break;
        }
    }
}

// Copy the hot prefix of the contribution vector into LDS: each thread's
// loads (up to 10 of 16 bytes) issued before its first LDS write, so a full
// 156 KiB copy (9,984 words over 1,024 threads) takes one round trip.
template <class T>
__device__ __forceinline__ void load_hot(const T *__restrict__ c_in, uint32_t hot)
{
    constexpr uint32_t K = 10;
    T *h = reinterpret_cast<T *>(gbx_dyn_lds);
    const uint32_t words = (uint32_t)((hot * sizeof(T) + 15) / 16);
    const uint4 *src = reinterpret_cast<const uint4 *>(c_in);
    uint4 *dst = reinterpret_cast<uint4 *>(h);
    for (uint32_t i0 = threadIdx.x; i0 < words; i0 += kBlock * K) {
        // indexes past the last word repeat it (same value, same address)
        uint4 w[K];
#pragma unroll
        for (uint32_t k = 0; k < K; k++) w[k] = src[min(i0 + k * kBlock, words - 1u)];
#pragma unroll
        for (uint32_t k = 0; k < K; k++) dst[min(i0 + k * kBlock, words - 1u)] = w[k];
    }
    __syncthreads();
}

// Multi-vector hooks implemented in gbx_ppr.hip.
int ppr_reset(gbx_pr *pr);
int ppr_launch_iter(gbx_pr *pr);
int ppr_read(gbx_pr *pr, float *ranks_out, gbx_pr_report *rep);
void ppr_state_free(gbx_ppr_state *s);

// Batch-of-graphs hooks implemented in gbx_batch.hip.
int seg_create(gbx_pr *pr);
int seg_reset(gbx_pr *pr);
int seg_launch_iter(gbx_pr *pr);
int seg_read(gbx_pr *pr, float *ranks_out, gbx_pr_report *rep);
void seg_state_free(gbx_seg_state *s);

// Binned-method hooks implemented in gbx_binned.hip.
bool binned_supported(const gbx_pr *pr);
int binned_create(gbx_pr *pr);       // builds the plan's binned layout on first use
int binned_launch_iter(gbx_pr *pr);  // scatter + accumulate: one iteration
// A batch of graphs (K_SEG): one iteration from c_in to c_out, the update
// with each unit's graph coefficients, every wave's partials into uslots.
int binned_seg_iter(gbx_pr *pr, float *r, const float *c_in, float *c_out, const GCoef *coef, double *uslots);
// The layout's accumulate units, and per graph the units it owns
// (ascending unit index): uptr [nseg+1], uidx [units].
uint32_t binned_units(const gbx_pr *pr);
void binned_graph_units(const gbx_pr *pr, const uint32_t **uptr, const uint32_t **uidx);
void binned_drop_unused(gbx_plan *p);  // drops the plan's layouts no object holds
void binned_state_free(gbx_binned_state *s);

// Data-driven push hooks implemented in gbx_ddpush.hip.
int ddp_create(gbx_pr *pr);
int ddp_reset(gbx_pr *pr);
int ddp_launch_iter(gbx_pr *pr);
void ddp_state_free(gbx_ddp_state *s);

// Host helpers implemented in gbx_pr.hip.
int ensure_plan(gbx_ctx *ctx, gbx_graph *g, int direction, gbx_plan **out);
// Number of leading plan rows with degree >= threshold (rows are sorted by
// degree, descending).
int plan_rows_at_least(gbx_ctx *ctx, const gbx_plan *p, uint32_t threshold, uint32_t *count);
// Row pointers rp[0..nrows] of the plan, copied to host.
int plan_row_ptr_prefix(gbx_ctx *ctx, const gbx_plan *p, uint32_t nrows, std::vector<uint32_t> &out);
// Kernel view of a plan with per-edge weights w (nullable), and the grid of
// its pull kernels on this context.
PlanView plan_view(const gbx_plan *p, const float *w);
unsigned iter_grid(const gbx_ctx *ctx, const gbx_plan *p);

// Per-device set-up of the allocation-free entry point (gbx_noalloc.hip),
// run by gbx_init with the context's device current.
int noalloc_setup(gbx_ctx *ctx);

}  // namespace gbx
