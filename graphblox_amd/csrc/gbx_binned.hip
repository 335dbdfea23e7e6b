// gbx_binned.hip -- the binned (propagation-blocking) PageRank iteration.
//
// The pull iteration gathers one 4-byte contribution per edge from a vector
// far larger than the L2; on MI355X every miss moves a 128-byte line, so at
// R-MAT scale 24 the gathers move about three times the algorithmic bytes.
// The binned method trades the random gather for two streaming passes over
// a value array V with one slot per edge:
//
//   scatter     one block per source chunk (a contiguous range of at most
//               32768 sources): the chunk's contributions are loaded into
//               LDS, and for every destination bin the chunk's edges into
//               that bin are written as one contiguous segment of V.  The
//               chunk's source indices (src_local, u16) are stored segment
//               by segment, each segment padded to a multiple of 32, with the
//               V position of every block of 32 (bdst, u32): the pass streams
//               src_local and bdst and writes V in runs, with no search.
//   accumulate  one block per unit (a destination bin of at most 20000 rows,
//               or a piece of a single giant row): V and the static
//               dst_local array are read sequentially and summed per row in
//               LDS as 64-bit fixed-point integers, then the PageRank update
//               (the pull iteration's epilogue) writes the ranks and the
//               next contributions.
//
// Hot sources: the first H sources of the plan order (the most gathered;
// about a quarter of all edges at R-MAT scale 24 for H = 16384) keep their
// contributions in the accumulate block's LDS instead; their edges are
// stored per bin as static (row, source) u16 pairs, 4 bytes per edge instead
// of 12.5 through V, and never pass through the scatter (with the run
// layout packed into 16 bits per edge, see hot_fill16).
//
// The fixed-point scale is a power of two chosen per iteration from the
// previous FP64 rank sum, which bounds every row sum, so no sum can
// overflow; integer adds are associative, so the result is independent of
// the order of the LDS atomics (deterministic), and the per-edge rounding
// (2^-37 at R-MAT scale 24 in REF mode, 2^-60 in stochastic mode) is below
// FP32 accumulation error.
//
// Layout (built once per plan, device resident): V is bin-major, within a
// bin chunk-major, within a (bin, chunk) segment in the plan's edge order;
// every segment is padded to a multiple of 32 values (pads: value 0, row 0,
// written by the scatter pass, so V is written in whole lines) and the
// accumulate pass reads 8 values per lane.  src_local is in chunk-major
// order, segments padded the same way with the marker 0xFFFF.
//
// Run layout (the FP32/BF16 kinds, unless tuned off): the edges of one
// chunk into one row are consecutive in their segment, and the scatter sums
// each such run (cut at every group of 8 src_local entries) into one V
// value, so V, dst_local and both passes over V shrink to the runs (0.64 of
// the cold edges at R-MAT scale 24).  src_local flags each run's last entry
// (bit 15; 15-bit sources, chunks of at most 32736); a record per block of
// 64 groups and each segment's V offset give the runs' V positions; V
// segments are padded to 8 values (zeros, written once).  The run sums are FP32, so the result
// depends on the cut (not on the launch): deterministic for a plan, within
// FP32 rounding of the one-value-per-edge layout.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <type_traits>
#include <vector>

#include "gbx_iter.hpp"

namespace gbx {

// Streaming (read-once) arrays of the binned iteration -- src_local and the
// block records' groups in the scatter, V, dst_local and the hot pairs in the
// accumulate, and their weights -- are loaded non-temporally, so that the
// caches keep what is read again: the rank, degree and contribution arrays of
// the accumulate's update and V itself between the scatter that writes it and
// the accumulate that reads it.  Measured at R-MAT s24: iteration -4.5 %;
// non-temporal stores of V made the scatter 16 % slower and are not used.
typedef unsigned int nt_u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint4 ld_nt16(const void *p)
{
    const nt_u32x4 v = __builtin_nontemporal_load(reinterpret_cast<const nt_u32x4 *>(p));
    return make_uint4(v.x, v.y, v.z, v.w);
}

}  // namespace gbx

namespace gbx {

constexpr uint32_t kMaxBinRows = 20000;     // LDS int64 row sums of one bin
constexpr uint32_t kNoEdge = 0xFFFFFFFFu;   // no plan edge behind a layout entry
constexpr uint32_t kMaxChunkRows = 32768;   // LDS contributions of one chunk; u16 local ids
constexpr uint32_t kNoMulti = 0xFFFFFFFFu;

// Sources per chunk: a multiple of 8 (chunks start at multiples of 8) in [8, 32768],
// and at most 128 KiB of contributions of `elem` bytes in the scatter's LDS.
inline uint32_t chunk_rows_of(const gbx_tuning &tu, uint32_t elem)
{
    return std::max(8u, std::min(std::min(tu.binned_chunk_rows, kMaxChunkRows), (128u * 1024u) / elem) & ~7u);
}

// The run layout keeps 15 bits of source index (the top bit flags a run's
// end) and one index past the chunk for the pads, and its scatter's LDS
// holds the staging next to the contributions: at most 32736 sources.
constexpr uint32_t kMaxRunChunkRows = 32736;

inline uint32_t chunk_sources_of(const gbx_tuning &tu, uint32_t elem, bool runs)
{
    const uint32_t sc = chunk_rows_of(tu, elem);
    return runs ? std::min(sc, kMaxRunChunkRows) : sc;
}

// Edges per chunk and per bin: the tuning value, or (0) automatic, measured
// on MI355X.  One value per edge (one scatter block per chunk): E / CUs
// clamped to [64K, 512K] (256K best at R-MAT s22, 512K at s24).  Run
// layout: chunks are cut by sources only (the scatter's blocks take equal
// shares of all groups, whatever the chunks: the larger a chunk, the more
// of its edges into one row merge into one run -- 0.386 V values per cold
// edge at R-MAT s24 against 0.578 with 900K edges per chunk); bins 7/8 of
// E / CUs clamped to [64K, 1M].
inline uint32_t cut_edges_of(uint32_t v, const gbx_ctx *ctx, uint32_t E, bool runs, bool chunk)
{
    if (v) return v;
    if (runs && chunk) return 0xFFFFFFFFu;
    const uint32_t per_cu = E / (uint32_t)std::max(1, ctx->num_cus);
    if (runs) return std::min(1u << 20, std::max(1u << 16, (uint32_t)((uint64_t)per_cu * 7 / 8)));
    return std::min(1u << 19, std::max(1u << 16, per_cu));
}

// Run layout: one scatter block per piece -- a chunk, or an equal part of a
// chunk of more than a piece's groups: 1/(2 CUs) of all groups, at least
// kPieceGroups, cut at multiples of 64 groups (a step's start) -- launched
// the largest first.
// Pieces of at least this many groups: the LDS copy of a chunk's
// contributions (128 KiB) stays a small share of a piece's work (measured:
// 23040-30720 best at R-MAT s22, 46208 = 1/(2 CUs) at s24).
constexpr uint32_t kPieceGroups = 24576;

struct BinUnit {
    uint32_t row_begin, row_end;  // rows of the bin (plan order)
    uint32_t vbeg, vend;          // V range of this unit (multiples of 8)
    uint32_t multi;               // accumulator slot of a bin split over units, else kNoMulti
    uint32_t nunits;              // units of the bin
    uint32_t hbeg, hend;          // hot (row, source) pair range of this unit (multiples of 8)
    uint32_t graph, pad;          // a batch of graphs: the graph of the bin's rows (bins never cross graphs)
};

// LDS of an accumulate block: int64 row sums, then the hot contributions.
constexpr uint32_t kAccumLds = 160 * 1024 - 1024;

}  // namespace gbx

struct gbx_bplan {
    uint32_t NC = 0, NB = 0, NU = 0, nseg = 0, n_multi = 0;
    uint32_t Vlen = 0;                // padded length of V
    uint32_t H = 0;                   // hot sources [0, H): their edges bypass V
    uint32_t Hlen = 0;                // padded length of the hot pair array
    uint32_t SB = 0;                  // rows per bin (LDS int64 sums)
    uint32_t cut[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // chunk edges, bin edges, bin rows, chunk rows, pad, hot,
                                                  // hot bytes, run layout
    uint32_t Plen = 0;                // padded length of src_local
    uint32_t SC = 0;                  // sources per chunk (at most)
    bool runs = false;                // run layout: one V value per run of equal rows of a segment
    bool hot16 = false;               // packed hot pairs (with the run layout)
    gbx::DevBuf chunk_lo;             // u32[NC+1] first source of each chunk
    gbx::DevBuf chunk_grp;            // u32[NC+1] first group of 8 src_local entries of each chunk
    uint32_t NP = 0;                  // run layout: scatter blocks (pieces)
    gbx::DevBuf pieces;               // run layout: uint4[NP] first and end source, first and end group
    gbx::DevBuf src_local;            // u16[Plen] source - chunk_lo, chunk-major, segments padded (0xFFFF)
    gbx::DevBuf bdst;                 // u32[Plen/32] V position of each block of 32 src_local entries
    gbx::DevBuf blk_rec;              // run layout: uint4 per block of 64 groups (V position of its first run,
                                      // segments before it, segment-start bits)
    gbx::DevBuf seg_v;                // run layout: u32[nseg] V offset of each segment (chunk-major order)
    gbx::DevBuf dst_local;            // u16[Vlen] row - bin first row, V order
    gbx::DevBuf hot_pairs;            // u32[Hlen] (row - bin first row) | source << 16, per bin (pad: src 0xFFFF);
                                      // packed (hot16): u16[Hlen] source | row step << 14
    gbx::DevBuf hot_base;             // packed: u16[Hlen/8] row - bin first row of each group's first pair
    gbx::DevBuf units;                // BinUnit[NU]
    gbx::DevBuf graph_uptr, graph_uidx;  // a batch of graphs: the units of each graph, ascending
    // run layout: the plan edge behind every src_local entry and every packed
    // hot pair (kNoEdge for pads and fillers), from which a per-edge weight
    // array of the plan is laid out beside them (weighted kinds)
    gbx::DevBuf cold_eidx;            // u32[Plen]
    gbx::DevBuf hot_eidx;             // u32[Hlen]
    struct EdgeWeights {
        const void *src = nullptr;    // the plan's weight array (plan edge order) they were made from
        gbx::DevBuf cold;             // f32[Plen] weight of every src_local entry (0 for pads)
        gbx::DevBuf hot;              // f32[Hlen] weight of every hot pair (0 for fillers)
    };
    std::vector<std::unique_ptr<EdgeWeights>> weights;
};

struct gbx_binned_state {
    std::shared_ptr<gbx_bplan> bp;  // the layout (shared with the plan)
    gbx::DevBuf V;    // f32 or bf16 [Vlen]
    gbx::DevBuf acc;  // i64[n_multi] sums of split bins (zero between iterations)
    gbx::DevBuf cnt;  // u32[n_multi] arrival counters (zero between iterations)
    gbx::DevBuf work;  // u32 next accumulate unit (zero between iterations)
    unsigned accum_grid = 0;
    uint32_t dummy_off = 0;  // run layout: byte offset of the scatter's dummy areas in V
    uint32_t stage_off = 0;  // run layout: byte offset of the scatter's staging in LDS
    size_t shm_scatter = 0, shm_accum = 0;
    const float *wcold = nullptr, *whot = nullptr;  // weighted kinds: the layout's edge weights
};

namespace gbx {

// ------------------------------------------------------------------ build kernels

// rowid[j] = the row of plan edge j: rows [0, nlong) by a wave each (the
// plan order puts the long rows first), the others by a thread each.
__global__ void edge_rows_long(const uint32_t *__restrict__ rp, uint32_t nlong, uint32_t *__restrict__ rowid)
{
    const uint32_t lane = threadIdx.x & 63u, waves = gridDim.x * (blockDim.x / 64u);
    for (uint32_t v = blockIdx.x * (blockDim.x / 64u) + (threadIdx.x >> 6); v < nlong; v += waves)
        for (uint32_t j = rp[v] + lane; j < rp[v + 1]; j += 64u) rowid[j] = v;
}

__global__ void edge_rows(const uint32_t *__restrict__ rp, uint32_t v0, uint32_t N, uint32_t *__restrict__ rowid)
{
    for (uint32_t v = v0 + blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x)
        for (uint32_t j = rp[v]; j < rp[v + 1]; j++) rowid[j] = v;
}

// cnt[u] += occurrences of u in col: ids below kColLds counted in an LDS
// histogram per block first (the plan order puts the sources of most edges
// there), the rest with global atomics.
constexpr uint32_t kColLds = 16384;
__global__ void __launch_bounds__(1024) column_counts(const uint32_t *__restrict__ col, uint32_t E,
                                                      uint32_t *__restrict__ cnt)
{
    __shared__ uint32_t h[kColLds];
    for (uint32_t i = threadIdx.x; i < kColLds; i += blockDim.x) h[i] = 0;
    __syncthreads();
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x) {
        const uint32_t u = col[e];
        if (u < kColLds) atomicAdd(h + u, 1u);
        else atomicAdd(cnt + u, 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kColLds; i += blockDim.x)
        if (h[i]) atomicAdd(cnt + i, h[i]);
}

// grp[2g] = pre[8g], grp[2g + 1] = pre[8g + 1] for the G groups of 8 of n
// items (pre: n + 1 entries, 8g < n).
__global__ void group_prefix(const uint32_t *__restrict__ pre, uint32_t n, uint32_t G, uint32_t *__restrict__ grp)
{
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < G; g += gridDim.x * blockDim.x) {
        grp[2 * g] = pre[8 * g];
        grp[2 * g + 1] = pre[8 * g + 1];
    }
}

// Cold source chunks: from a chunk starting at source 8g (g < G) the next
// one starts at group nxt[g] (>= G: none), the first group g' > g whose
// first source u = 8g' would take the chunk past SC sources with its group
// (u + 8 - 8g > SC) or past EC edges with u itself (pre[u + 1] - pre[8g]:
// grp[2g' + 1] - grp[2g], a binary search: pre ascends); nxt[G] is the same
// from the first cold source H, which need not start a group.
__global__ void chunk_next(const uint32_t *__restrict__ grp, uint32_t G, uint32_t H, uint32_t SC, uint64_t EC,
                           uint32_t *__restrict__ nxt)
{
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g <= G; g += gridDim.x * blockDim.x) {
        const uint32_t start = g < G ? 8u * g : H;
        const uint32_t pstart = g < G ? grp[2 * g] : 0u;  // pre[H] = 0: no cold source below H
        const uint32_t first = start / 8u + 1u;            // the first group past `start`
        const uint32_t gsc = (uint32_t)min(((uint64_t)start + SC - 8u) / 8u + 1u, (uint64_t)G);
        uint32_t lo = first, hi = gsc;  // the first group in [first, gsc) past EC edges, else gsc
        while (lo < hi) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if ((uint64_t)(grp[2 * mid + 1] - pstart) > EC) hi = mid;
            else lo = mid + 1;
        }
        nxt[g] = lo;
    }
}

// clo[0..nc] = the chunks' first sources (H first, N last) and cstart[0..nc]
// their first cold edges (pre at each, Ecold last); *nc_out = nc.  One lane
// walks the chain (a step of at least one group keeps it finite).
__global__ void chunk_chase(const uint32_t *__restrict__ nxt, const uint32_t *__restrict__ grp,
                            const uint32_t *__restrict__ pre, uint32_t G, uint32_t H, uint32_t N,
                            uint32_t *__restrict__ clo, uint32_t *__restrict__ cstart, uint32_t *__restrict__ nc_out)
{
    if (threadIdx.x != 0) return;
    clo[0] = H;
    cstart[0] = 0;
    uint32_t k = 1;
    if (N > H) {
        uint32_t g = nxt[G];
        while (g < G) {
            clo[k] = 8u * g;
            cstart[k] = grp[2 * g];
            k++;
            const uint32_t t = nxt[g];
            g = t > g ? t : g + 1u;
        }
        clo[k] = N;
        cstart[k] = pre[N];
        k++;
    }
    *nc_out = k - 1;
}

// Destination bins: from a bin's first row s the next bin starts at
// nxt[s] = the first of s + SB, the first row v > s with
// rp[v + 1] - rp[s] > EB (binary search: rp ascends), the first graph start
// above s (goff, ngoff ascending offsets; none for one graph) and N.
// Greedy from row 0, this cuts every bin at <= SB rows and <= EB in-edges
// (a longer row is a bin of its own), never across a graph.
__global__ void bin_next(const uint32_t *__restrict__ rp, uint32_t N, uint32_t SB, uint64_t EB,
                         const uint32_t *__restrict__ goff, uint32_t ngoff, uint32_t *__restrict__ nxt)
{
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < N; s += gridDim.x * blockDim.x) {
        uint32_t cut = (uint32_t)min((uint64_t)s + SB, (uint64_t)N);
        const uint64_t T = (uint64_t)rp[s] + EB;
        // the first j in [s + 2, cut] with rp[j] > T makes v = j - 1 < cut
        uint32_t lo = s + 2, hi = cut + 1;
        while (lo < hi) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if ((uint64_t)rp[mid] > T) hi = mid;
            else lo = mid + 1;
        }
        if (lo <= cut) cut = lo - 1;
        lo = 0;
        hi = ngoff;
        while (lo < hi) {  // the first offset above s
            const uint32_t mid = lo + (hi - lo) / 2;
            if (goff[mid] > s) hi = mid;
            else lo = mid + 1;
        }
        if (lo < ngoff) cut = min(cut, goff[lo]);
        nxt[s] = cut;
    }
}

// vb[0..nb] = the bins' first rows from row 0 through nxt, then N; *nb_out =
// nb.  One lane walks the chain; a bin always ends past its first row (a
// step of at least one row keeps the walk finite whatever nxt holds).
__global__ void bin_chase(const uint32_t *__restrict__ nxt, uint32_t N, uint32_t *__restrict__ vb,
                          uint32_t *__restrict__ nb_out)
{
    if (threadIdx.x != 0) return;
    uint32_t s = 0, k = 0;
    while (s < N) {
        vb[k++] = s;
        const uint32_t t = nxt[s];
        s = t > s ? t : s + 1u;
    }
    vb[k] = N;
    *nb_out = k;
}

// *first = min(*first, the first row with fewer than `len` edges).
__global__ void first_short_row(const uint32_t *__restrict__ rp, uint32_t N, uint32_t len, uint32_t *first)
{
    uint32_t m = N;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x)
        if (rp[v + 1] - rp[v] < len) {
            m = v;
            break;
        }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = min(m, (uint32_t)__shfl_xor((int)m, off, 64));
    if ((threadIdx.x & 63u) == 0 && m < N) atomicMin(first, m);
}

// out[i] = the k < nr with bounds[k] <= i < bounds[k + 1] (0 below bounds[0]
// and when nr = 0), i < n; bounds[0..nr] ascending.
__global__ void range_index(const uint32_t *__restrict__ bounds, uint32_t nr, uint32_t n, uint32_t *__restrict__ out)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = nr;  // the largest k in [lo, hi) with bounds[k] <= i, or 0
        while (lo + 1 < hi) {
            const uint32_t mid = (lo + hi) / 2;
            if (bounds[mid] <= i) lo = mid;
            else hi = mid;
        }
        out[i] = lo;
    }
}

// key = bin * NC + chunk (bin-major) or chunk * NB + bin; payload = plan edge index
__global__ void bin_keys(const uint32_t *__restrict__ rowid, const uint32_t *__restrict__ col, uint32_t E,
                         const uint32_t *__restrict__ bin_of, const uint32_t *__restrict__ chunk_of,
                         uint32_t NB, uint32_t NC, int chunk_major, uint32_t *__restrict__ key,
                         uint32_t *__restrict__ pay)
{
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x) {
        const uint32_t b = bin_of[rowid[e]], k = chunk_of[col[e]];
        key[e] = chunk_major ? k * NB + b : b * NC + k;
        pay[e] = e;
    }
}

// The run layout's sort keys (chunk-major), with the payloads that the
// layout needs once the edges are sorted, so nothing is gathered after the
// sort: each edge's local ids (row in its bin << 16 | source in its chunk)
// and its plan edge.
__global__ void run_keys(const uint32_t *__restrict__ rowid, const uint32_t *__restrict__ col,
                         const uint32_t *__restrict__ eidx, uint32_t E, const uint32_t *__restrict__ bin_of,
                         const uint32_t *__restrict__ chunk_of, const uint32_t *__restrict__ bin_row0,
                         const uint32_t *__restrict__ chunk_lo, uint32_t NB, uint32_t *__restrict__ key,
                         uint32_t *__restrict__ local, uint32_t *__restrict__ pedge)
{
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x) {
        const uint32_t row = rowid[e], c = col[e];
        const uint32_t b = bin_of[row], k = chunk_of[c];
        key[e] = k * NB + b;
        local[e] = ((row - bin_row0[b]) << 16) | (c - chunk_lo[k]);
        pedge[e] = eidx[e];
    }
}

// Runs of equal keys in a sorted key array: flag[i] = 1 at a run start.
__global__ void run_flags(const uint32_t *__restrict__ key, uint32_t n, uint32_t *__restrict__ flag)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        flag[i] = (i == 0 || key[i - 1] != key[i]) ? 1u : 0u;
}

// runstart[r] = first index of run r (runidx = exclusive scan of flag)
__global__ void run_starts(const uint32_t *__restrict__ flag, const uint32_t *__restrict__ runidx, uint32_t n,
                           uint32_t *__restrict__ runstart)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (flag[i]) runstart[runidx[i]] = i;
}

// Run length rounded up to a multiple of pad.
__global__ void run_padded_lengths(const uint32_t *__restrict__ runstart, uint32_t nrun, uint32_t n, uint32_t pad,
                                   uint32_t *__restrict__ plen)
{
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < nrun; r += gridDim.x * blockDim.x) {
        const uint32_t len = (r + 1 < nrun ? runstart[r + 1] : n) - runstart[r];
        plen[r] = (len + pad - 1) / pad * pad;
    }
}

// Bin-major sorted edges -> dst_local at their padded V positions, the V
// start of every (chunk, bin) segment in a dense table, and the padded V
// length of every bin.
__global__ void bin_major_fill(const uint32_t *__restrict__ key, const uint32_t *__restrict__ pay, uint32_t E,
                               const uint32_t *__restrict__ flag, const uint32_t *__restrict__ runidx,
                               const uint32_t *__restrict__ runstart, const uint32_t *__restrict__ vstart,
                               const uint32_t *__restrict__ plen, const uint32_t *__restrict__ rowid,
                               const uint32_t *__restrict__ bin_row0, uint32_t NB, uint32_t NC,
                               uint16_t *__restrict__ dst_local, uint32_t *__restrict__ voff_table,
                               uint32_t *__restrict__ bin_len)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < E; i += gridDim.x * blockDim.x) {
        const uint32_t kk = key[i], b = kk / NC, k = kk % NC;
        const uint32_t r = runidx[i] + flag[i] - 1u;
        dst_local[vstart[r] + (i - runstart[r])] = (uint16_t)(rowid[pay[i]] - bin_row0[b]);
        if (flag[i]) {
            voff_table[(size_t)k * NB + b] = vstart[r];
            atomicAdd(bin_len + b, plen[r]);
        }
    }
}

// Chunk-major sorted edges -> src_local at padded positions (pads keep the
// 0xFFFF prefill) and the V position of every group of 8.
__global__ void chunk_major_fill(const uint32_t *__restrict__ key, const uint32_t *__restrict__ pay, uint32_t E,
                                 const uint32_t *__restrict__ flag, const uint32_t *__restrict__ runidx,
                                 const uint32_t *__restrict__ runstart, const uint32_t *__restrict__ pstart,
                                 const uint32_t *__restrict__ col, const uint32_t *__restrict__ chunk_lo,
                                 uint32_t NB, uint16_t *__restrict__ src_local)
{
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < E; j += gridDim.x * blockDim.x) {
        const uint32_t k = key[j] / NB;
        const uint32_t r = runidx[j] + flag[j] - 1u;
        src_local[pstart[r] + (j - runstart[r])] = (uint16_t)(col[pay[j]] - chunk_lo[k]);
    }
}

// V position of every block of 32 src_local entries (segments are padded to 32).
__global__ void block_destinations(const uint32_t *__restrict__ key, const uint32_t *__restrict__ runstart,
                                   const uint32_t *__restrict__ pstart, const uint32_t *__restrict__ plen,
                                   uint32_t nrun, const uint32_t *__restrict__ voff_table,
                                   uint32_t *__restrict__ bdst)
{
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < nrun; r += gridDim.x * blockDim.x) {
        const uint32_t b0 = pstart[r] >> 5, nb = plen[r] >> 5, v0 = voff_table[key[runstart[r]]];
        for (uint32_t q = 0; q < nb; q++) bdst[b0 + q] = v0 + 32u * q;
    }
}

// chunk_grp[k] = first group (of 2^shift entries) of the chunk's first segment.
__global__ void chunk_groups(const uint32_t *__restrict__ cstart, uint32_t NC, const uint32_t *__restrict__ runidx,
                             uint32_t E, uint32_t nrun, const uint32_t *__restrict__ pstart, uint32_t shift,
                             uint32_t *__restrict__ chunk_grp)
{
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k <= NC; k += gridDim.x * blockDim.x) {
        const uint32_t r = cstart[k] >= E ? nrun : runidx[cstart[k]];
        chunk_grp[k] = pstart[r] >> shift;
    }
}

__global__ void hot_flags(const uint32_t *__restrict__ col, uint32_t E, uint32_t H, uint32_t *__restrict__ flag)
{
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x)
        flag[e] = col[e] < H ? 1u : 0u;
}

__global__ void gather_u32(const uint32_t *__restrict__ a, const uint32_t *__restrict__ idx, uint32_t n,
                           uint32_t *__restrict__ out)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = a[idx[i]];
}

// Hot edges -> (row - bin first row) | source << 16 at their per-bin position;
// cold edges -> compacted row and source arrays (plan edge order kept).
__global__ void split_hot_cold(const uint32_t *__restrict__ col, const uint32_t *__restrict__ rowid, uint32_t E,
                               const uint32_t *__restrict__ hflag, const uint32_t *__restrict__ hrank,
                               const uint32_t *__restrict__ bin_of, const uint32_t *__restrict__ bin_row0,
                               const uint32_t *__restrict__ hot_before, const uint32_t *__restrict__ hot_start,
                               uint32_t *__restrict__ hot_pairs, uint32_t *__restrict__ rowid_c,
                               uint32_t *__restrict__ col_c, uint32_t *__restrict__ eidx_c)
{
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x) {
        const uint32_t v = rowid[e], u = col[e];
        if (hflag[e]) {
            if (hot_pairs) {
                const uint32_t b = bin_of[v];
                hot_pairs[hot_start[b] + (hrank[e] - hot_before[b])] = (v - bin_row0[b]) | (u << 16);
            }
        } else {
            const uint32_t c = e - hrank[e];
            rowid_c[c] = v;
            col_c[c] = u;
            eidx_c[c] = e;
        }
    }
}

// ---- run layout: one V value per run of equal rows in a (chunk, bin) segment

// A lane of the scatter takes a group of 8 src_local entries; runs never
// cross a group (segments are padded to 8 entries and a run is cut at every
// group boundary), so every lane sums its own runs.
constexpr uint32_t kRunGroup = 8;
// A wave's step: 64 groups, aligned to blocks of 64 groups of the whole
// chunk-major src_local; each block has a record (blk_rec) and every
// segment its V offset (seg_v), which give the V positions of a step's runs.
constexpr uint32_t kRunBlockGroups = 64;
constexpr uint32_t kRunEnd = 0x8000u;  // src_local flag: the entry ends its run

// head[j] = 1 where a run starts: at a new row and at every group boundary
// (segment starts are group boundaries).  j indexes the chunk-major sorted
// cold edges; segment r = runidx + flag - 1 sits at src_local position
// pstart[r].
__global__ void run_heads(const uint32_t *__restrict__ local, uint32_t n, const uint32_t *__restrict__ flag,
                          const uint32_t *__restrict__ runidx, const uint32_t *__restrict__ runstart,
                          const uint32_t *__restrict__ pstart, uint32_t *__restrict__ head)
{
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const uint32_t r = runidx[j] + flag[j] - 1u;
        const uint32_t p = pstart[r] + (j - runstart[r]);
        // a segment starts at a multiple of 8 (p % 8 == 0), so the rows
        // compared below are in one bin: equal local rows, equal rows
        head[j] = (p % kRunGroup == 0 || (local[j] >> 16) != (local[j - 1] >> 16)) ? 1u : 0u;
    }
}

// Per segment: V length (runs, padded to 8), the bin-major sort key
// (bin * NC + chunk) with the segment as payload, and the bin's V length.
__global__ void run_segments(const uint32_t *__restrict__ key, const uint32_t *__restrict__ runstart, uint32_t nseg,
                             uint32_t n, const uint32_t *__restrict__ hx, uint32_t NB, uint32_t NC,
                             uint32_t *__restrict__ vlen, uint32_t *__restrict__ skey, uint32_t *__restrict__ spay,
                             uint32_t *__restrict__ bin_len)
{
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < nseg; r += gridDim.x * blockDim.x) {
        const uint32_t s0 = runstart[r], s1 = r + 1 < nseg ? runstart[r + 1] : n;
        const uint32_t len = (hx[s1] - hx[s0] + 7u) & ~7u;
        const uint32_t kk = key[s0], k = kk / NB, b = kk % NB;
        vlen[r] = len;
        skey[r] = b * NC + k;
        spay[r] = r;
        atomicAdd(bin_len + b, len);
    }
}

__global__ void scatter_u32(const uint32_t *__restrict__ a, const uint32_t *__restrict__ idx, uint32_t n,
                            uint32_t *__restrict__ out)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[idx[i]] = a[i];
}

// Every cold edge: its source (| kRunEnd on a run's last entry) in
// src_local, its run's row in dst_local, and for the first entry of each
// block of 64 groups the V position of its run (the block's first run).  A
// segment's last run takes the pads of its group (zeros): its end flag sits
// on the group's last entry, so every group ends with a run end.
__global__ void run_fill(const uint32_t *__restrict__ local, const uint32_t *__restrict__ pedge, uint32_t n,
                         const uint32_t *__restrict__ flag, const uint32_t *__restrict__ runidx,
                         const uint32_t *__restrict__ runstart, const uint32_t *__restrict__ pstart,
                         const uint32_t *__restrict__ head, const uint32_t *__restrict__ hx,
                         const uint32_t *__restrict__ voff, uint32_t pad, uint16_t *__restrict__ src_local,
                         uint16_t *__restrict__ dst_local, uint32_t *__restrict__ blk_v,
                         uint32_t *__restrict__ cold_eidx)
{
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const uint32_t r = runidx[j] + flag[j] - 1u, s0 = runstart[r];
        const uint32_t p = pstart[r] + (j - s0);
        const uint32_t vpos = voff[r] + (hx[j] + head[j] - 1u - hx[s0]);
        const bool end = j + 1 == n || head[j + 1];
        const bool seg_end = j + 1 == n || flag[j + 1];
        const uint32_t loc = local[j];
        bool flag_here = end;
        if (seg_end && p % kRunGroup != kRunGroup - 1) {
            flag_here = false;
            src_local[p | (kRunGroup - 1)] = (uint16_t)(pad | kRunEnd);
        }
        src_local[p] = (uint16_t)((loc & 0xFFFFu) | (flag_here ? kRunEnd : 0u));
        cold_eidx[p] = pedge[j];
        if (end) dst_local[vpos] = (uint16_t)(loc >> 16);
        if (p % (kRunGroup * kRunBlockGroups) == 0) blk_v[p / (kRunGroup * kRunBlockGroups)] = vpos;
    }
}

// Bit g of `bits` (u32 words) for every group g that starts a segment.
__global__ void run_segment_bits(const uint32_t *__restrict__ pstart, uint32_t nseg, uint32_t *__restrict__ bits)
{
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < nseg; r += gridDim.x * blockDim.x) {
        const uint32_t g = pstart[r] / kRunGroup;
        atomicOr(bits + g / 32, 1u << (g % 32));
    }
}

// Segment starts per block of 64 groups.
__global__ void run_block_counts(const uint32_t *__restrict__ bits, uint32_t nblk, uint32_t *__restrict__ cnt)
{
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < nblk; b += gridDim.x * blockDim.x)
        cnt[b] = (uint32_t)(__popc(bits[2 * b]) + __popc(bits[2 * b + 1]));
}

// The record of every block of 64 groups: V position of its first run,
// segments starting before it, its segment-start bits.
__global__ void run_block_records(const uint32_t *__restrict__ blk_v, const uint32_t *__restrict__ blk_seg,
                                  const uint32_t *__restrict__ bits, uint32_t nblk, uint4 *__restrict__ rec)
{
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < nblk; b += gridDim.x * blockDim.x)
        rec[b] = make_uint4(blk_v[b], blk_seg[b], bits[2 * b], bits[2 * b + 1]);
}

// ---- packed hot pairs (the run layout's plans): 16 bits per pair

// A hot pair is source (14 bits) | row step (2 bits) << 14: the row of a
// pair is the row of the previous pair of its group of 8 plus the step, the
// first pair of a group takes its row from hot_base (u16 per group).  A gap
// of more than 3 rows between consecutive pairs of a bin gets filler pairs
// (source kHotPad, a zero in LDS, step 3), about 0.4 % of the pairs at
// R-MAT scale 24.
constexpr uint32_t kHotPad = 0x3FFFu;
constexpr uint32_t kHot16Max = 16376;   // hot sources of the packed form
constexpr uint32_t kHot16Slots = 16384; // LDS hot slots (kHotPad included)

// Hot edges (plan order) -> compacted rows and sources.
__global__ void hot_compact(const uint32_t *__restrict__ col, const uint32_t *__restrict__ rowid, uint32_t E,
                            const uint32_t *__restrict__ hflag, const uint32_t *__restrict__ hrank,
                            uint32_t *__restrict__ rowid_h, uint32_t *__restrict__ col_h,
                            uint32_t *__restrict__ eidx_h)
{
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x)
        if (hflag[e]) {
            rowid_h[hrank[e]] = rowid[e];
            col_h[hrank[e]] = col[e];
            eidx_h[hrank[e]] = e;
        }
}

// Filler pairs before hot pair i: a gap of d > 3 rows to the previous pair
// of the bin takes (d - 1) / 3 fillers of step 3; n[i] = fillers + 1.
__global__ void hot_counts(const uint32_t *__restrict__ rowid_h, uint32_t Eh, const uint32_t *__restrict__ bin_of,
                           const uint32_t *__restrict__ hot_before, uint32_t *__restrict__ n)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < Eh; i += gridDim.x * blockDim.x) {
        const uint32_t b = bin_of[rowid_h[i]];
        const uint32_t d = i == hot_before[b] ? 0u : rowid_h[i] - rowid_h[i - 1];
        n[i] = 1u + (d > 3u ? (d - 1u) / 3u : 0u);
    }
}

// Every hot pair and its fillers at their packed positions (bin b starts at
// hot_start[b]; hpos = exclusive scan of n), and the row of every group's
// first pair.
__global__ void hot_fill16(const uint32_t *__restrict__ rowid_h, const uint32_t *__restrict__ col_h, uint32_t Eh,
                           const uint32_t *__restrict__ n, const uint32_t *__restrict__ hpos,
                           const uint32_t *__restrict__ bin_of, const uint32_t *__restrict__ hot_before,
                           const uint32_t *__restrict__ hot_start, const uint32_t *__restrict__ bin_row0,
                           const uint32_t *__restrict__ eidx_h, uint16_t *__restrict__ pairs,
                           uint16_t *__restrict__ hbase, uint32_t *__restrict__ hot_eidx)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < Eh; i += gridDim.x * blockDim.x) {
        const uint32_t row = rowid_h[i], b = bin_of[row], fill = n[i] - 1u;
        const uint32_t base = hot_start[b] + (hpos[i] - hpos[hot_before[b]]);
        const uint32_t d = i == hot_before[b] ? 0u : row - rowid_h[i - 1];
        for (uint32_t k = 1; k <= fill; k++) {
            const uint32_t pos = base + k - 1u;
            pairs[pos] = (uint16_t)(kHotPad | (3u << 14));
            if ((pos & 7u) == 0) hbase[pos >> 3] = (uint16_t)(row - d + 3u * k - bin_row0[b]);
        }
        const uint32_t pos = base + fill;
        pairs[pos] = (uint16_t)(col_h[i] | ((d - 3u * fill) << 14));
        hot_eidx[pos] = eidx_h[i];
        if ((pos & 7u) == 0) hbase[pos >> 3] = (uint16_t)(row - bin_row0[b]);
    }
}

// ------------------------------------------------------------------ iteration kernels

struct ScatterView {
    const uint32_t *__restrict__ chunk_lo;
    const uint32_t *__restrict__ chunk_grp;
    const uint16_t *__restrict__ src_local;
    const uint32_t *__restrict__ bdst;
    uint32_t NC;
};

constexpr int kScatterUnroll = 4;  // groups of 8 per lane in flight

// Eight consecutive values moved as one wide access (8, 16, 32 or 64 bytes).
template <class T>
struct alignas(8 * sizeof(T)) Vec8 {
    T v[8];
};

template <class T>
__device__ __forceinline__ void store8(T *V, uint32_t pos, const T (&v)[8])
{
    Vec8<T> w;
#pragma unroll
    for (int i = 0; i < 8; i++) w.v[i] = v[i];
    *reinterpret_cast<Vec8<T> *>(V + pos) = w;
}

// n values (whole groups of 8 at 8-value aligned addresses, then a tail)
// from global memory into LDS.  A thread's loads of kCopyAhead groups are
// issued before any of them is written, so a fill of a chunk's
// contributions (4 groups a thread) costs one memory round trip, not four.
constexpr int kCopyAhead = 4;
template <class T, int NT = kBlock>
__device__ __forceinline__ void copy_to_lds(T *dst, const T *src, uint32_t n)
{
    const uint32_t nv = n / 8;
    for (uint32_t i0 = threadIdx.x; i0 < nv; i0 += NT * kCopyAhead) {
        Vec8<T> w[kCopyAhead];
        // no branches: indexes past the last group are clamped to it, so a
        // lane past the end loads and writes the last group again (the same
        // value to the same address)
#pragma unroll
        for (int u = 0; u < kCopyAhead; u++) w[u] = reinterpret_cast<const Vec8<T> *>(src)[min(i0 + u * NT, nv - 1u)];
#pragma unroll
        for (int u = 0; u < kCopyAhead; u++) reinterpret_cast<Vec8<T> *>(dst)[min(i0 + u * NT, nv - 1u)] = w[u];
    }
    for (uint32_t i = nv * 8 + threadIdx.x; i < n; i += NT) dst[i] = src[i];
}

// n FP32 contributions quantized into LDS (the quantized kinds): whole
// groups of 8 at 8-value aligned addresses, then a tail.
template <int KIND, class U>
__device__ __forceinline__ void quant_to_lds(U *dst, const float *src, uint32_t n, double scale)
{
    const uint32_t nv = n / 8;
    for (uint32_t i = threadIdx.x; i < nv; i += kBlock) {
        const Vec8<float> w = reinterpret_cast<const Vec8<float> *>(src)[i];
        Vec8<U> q;
#pragma unroll
        for (int k = 0; k < 8; k++) q.v[k] = (U)quantize<KIND>(w.v[k], scale);
        reinterpret_cast<Vec8<U> *>(dst)[i] = q;
    }
    for (uint32_t i = nv * 8 + threadIdx.x; i < n; i += kBlock) dst[i] = (U)quantize<KIND>(src[i], scale);
}

// One block per source chunk.  Each lane takes groups of 8 entries: one
// 16-byte load of source indices, the V position of its block of 32, eight LDS reads, one
// 32-byte (FP32) or 16-byte (BF16) store; pads (0xFFFF) store zeros, so every
// segment is written in whole 128-byte lines.  QK: a quantized kind, whose
// FP32 contributions are quantized on the way into LDS (0: values as stored).
template <class T, int QK>
__global__ void __launch_bounds__(kBlock) bin_scatter_kernel(ScatterView S, const Slot *slots, uint32_t t,
                                                             const void *__restrict__ c_in, T *__restrict__ V)
{
    if (slots[t & 1].done) return;
    const uint32_t k = blockIdx.x;
    T *cl = reinterpret_cast<T *>(gbx_dyn_lds);
    const uint32_t lo = S.chunk_lo[k], n = S.chunk_lo[k + 1] - lo;
    // chunk starts are multiples of 8 sources
    if constexpr (QK != 0) quant_to_lds<QK, T>(cl, reinterpret_cast<const float *>(c_in) + lo, n,
                                               quant_scale<QK>(slots[t & 1]));
    else copy_to_lds(cl, reinterpret_cast<const T *>(c_in) + lo, n);
    __syncthreads();
    const uint32_t g0 = S.chunk_grp[k], g1 = S.chunk_grp[k + 1];
    for (uint32_t base = g0 + threadIdx.x; base < g1; base += kBlock * kScatterUnroll) {
        uint4 src[kScatterUnroll];
        uint32_t dst[kScatterUnroll];
#pragma unroll
        for (int q = 0; q < kScatterUnroll; q++) {
            const uint32_t g = base + (uint32_t)q * kBlock;
            if (g < g1) {
                src[q] = reinterpret_cast<const uint4 *>(S.src_local)[g];
                dst[q] = S.bdst[g >> 2] + ((g & 3u) << 3);
            }
        }
#pragma unroll
        for (int q = 0; q < kScatterUnroll; q++) {
            const uint32_t g = base + (uint32_t)q * kBlock;
            if (g < g1) {
                const uint32_t w[4] = {src[q].x, src[q].y, src[q].z, src[q].w};
                T v[8];
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    const uint32_t sidx = (i & 1) ? (w[i >> 1] >> 16) : (w[i >> 1] & 0xFFFFu);
                    v[i] = sidx == 0xFFFFu ? T(0) : cl[sidx];
                }
                store8(V, dst[q], v);
            }
        }
    }
}

struct RunsView {
    const uint32_t *__restrict__ chunk_lo;
    const uint32_t *__restrict__ chunk_grp;
    const uint4 *__restrict__ pieces;  // per block: sources [x, y) of its chunk, groups [z, w)
    const uint16_t *__restrict__ src_local;  // local source | kRunEnd; pads: `pad`
    const uint4 *__restrict__ blk_rec;       // per block of 64 groups: first run's V position, segments
                                             // before it, segment-start bits (lo, hi)
    const uint32_t *__restrict__ seg_v;      // V offset of each segment
    uint32_t nseg;
    uint32_t pad;        // LDS index of a zero (the plan's chunk size)
    uint32_t stage_off;  // byte offset of the per-wave staging in LDS
    uint32_t dummy;      // byte offset in V of the dummy areas past the last run (written by idle store
                         // lanes, never read): kRunDummyAreas areas of 64 values
    const float *__restrict__ wcold;  // weighted kinds: the weight of every src_local entry (0 for pads)
};

// Values summed in FP32 (BF16 widened; a run's sum rounded to BF16 once).
__device__ __forceinline__ float f32_of(float x) { return x; }
__device__ __forceinline__ float f32_of(uint16_t x) { return __uint_as_float((uint32_t)x << 16); }
__device__ __forceinline__ void put_f32(float *V, uint32_t i, float v) { V[i] = v; }
__device__ __forceinline__ void put_f32(uint16_t *V, uint32_t i, float v) { V[i] = f32_to_bf16(v); }

// LDS written by one lane of a wave and read by another lane of the same
// wave: the LDS performs one wavefront's operations in issue order, so only
// the compiler has to keep the order (no memory access moves across this).
__device__ __forceinline__ void wave_lds_order() { asm volatile("" ::: "memory"); }

constexpr int kRunsBlock = 1024;  // 16 waves: 128 KiB of contributions + 16 x 2 KiB of staging fit the LDS
constexpr uint32_t kRunStage = 64 * kRunGroup;  // staged run sums of a wave
constexpr int kRunsAhead = 4;     // steps per wave whose loads are in flight
// Stores every step issues, the rest on demand (a step has about 200 runs
// at R-MAT s24, 0.39 V values per cold edge).  Measured at s24: 2 always
// beats 6 by 4 % (FP32) and 5 % (BF16), 3-5 by 1-3 %.
constexpr uint32_t kRunStores = 2;
constexpr uint32_t kRunEarly = 4;  // stores whose staged sums are read ahead
// Idle store lanes write a dummy area of their own wave (64 values, 256
// bytes apart): one common dummy address would funnel a large share of
// all stores into one cache line of one L2 channel.
constexpr uint32_t kRunDummyAreas = 4096;
constexpr uint32_t kRunDummyStride = 256;

// Cross-lane moves of the wave (DPP, no LDS): row_shr:n within rows of 16
// lanes, the last lane of a row broadcast to the next row(s), the whole
// wave shifted up one lane.  Lanes without a source read 0.
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t x)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, ROW_MASK, 0xF, true);
}

// Inclusive sum over lanes 0..lane.
__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t v)
{
    v += dpp_u32<0x111>(v);  // row_shr:1
    v += dpp_u32<0x112>(v);  // row_shr:2
    v += dpp_u32<0x114>(v);  // row_shr:4
    v += dpp_u32<0x118>(v);  // row_shr:8
    v += dpp_u32<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp_u32<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
    return v;
}

// One step of the run layout, in three parts so that consecutive steps
// overlap (sum step k + 1, write out step k, stage step k + 1): each lane
// sums the runs of its group of 8 src_local entries (runs never cross a
// group, and every group ends with a run end) and stages each run sum in
// LDS at the run's rank in the wave; the wave then writes the staged sums
// in rank order, 64 consecutive ranks per store.  The runs of the lanes of
// one (chunk, bin) segment have consecutive ranks and consecutive V
// positions, so V position = rank + delta with delta constant over the
// segment; the lanes that start a segment lead it (about 3 per step; their
// V offsets come from seg_v), lane 0 leads the segment the step starts in
// (its V position from the block record), and the write-out takes each
// rank's delta from the last leader at or below it.  The write-out always issues at least
// kRunStores stores (ranks past the wave's runs store to the wave's dummy
// area), so the loads of later steps stay in flight across them.
//
// Staging writes every entry, without branches: an entry inside a run goes
// to the slot of the run's end, which overwrites it later (a lane's writes
// are in entry order, and its last entry ends a run), so a lane only ever
// writes its own ranks; a pad lane (no runs) writes the slot past the wave's
// runs, which the write-out never takes (it has at most 8 x 63 runs then).
struct RunStep {
    float v[8];         // running sum after each entry (a run's sum at its end)
    uint32_t slot[8];   // stage slot of each entry
    uint32_t pre;       // rank of the lane's first run
    uint32_t delta;     // V index - rank, constant over a segment
    uint32_t total;     // runs of the wave
    uint64_t lead;      // lanes that lead a segment
};

// Weights of a group of 8 entries (weighted kinds), two 16-byte loads.
struct W8 {
    uint4 a, b;
};

template <class T, bool W>
__device__ __forceinline__ RunStep runs_sum(const float *cl, uint4 w, const W8 &wt, uint32_t vbase, bool leader)
{
    RunStep st;
    const uint32_t ew[4] = {w.x, w.y, w.z, w.w};
    float x[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const uint32_t h = (i & 1) ? __builtin_amdgcn_ubfe(ew[i >> 1], 16, 15) : (ew[i >> 1] & (kRunEnd - 1u));
        x[i] = cl[h];
    }
    if constexpr (W) {
        const uint32_t wb[8] = {wt.a.x, wt.a.y, wt.a.z, wt.a.w, wt.b.x, wt.b.y, wt.b.z, wt.b.w};
#pragma unroll
        for (int i = 0; i < 8; i++) x[i] = __uint_as_float(wb[i]) * x[i];  // the edge's weight times its source's value
    }
    // an entry ends its run when the top bit of its 16-bit half is set
    bool end[8];
#pragma unroll
    for (int i = 0; i < 8; i++)
        end[i] = (i & 1) ? (int32_t)ew[i >> 1] < 0 : (int16_t)(uint16_t)ew[i >> 1] < 0;
    uint32_t cnt = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) cnt += (uint32_t)__popc(ew[k] & (kRunEnd | (kRunEnd << 16)));
    const uint32_t inc = wave_inclusive_sum(cnt);
    st.pre = inc - cnt;
    st.total = __builtin_amdgcn_readlane(inc, 63);
    st.delta = vbase - st.pre;  // meaningful on leaders
    st.lead = __ballot(leader);
    uint32_t o = cnt ? st.pre : st.total;
    float s = 0.f;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        s += x[i];
        st.v[i] = s;
        st.slot[i] = o;
        o += end[i] ? 1u : 0u;
        s = end[i] ? 0.f : s;
    }
    return st;
}

__device__ __forceinline__ void runs_stage(float *stage, const RunStep &st)
{
#pragma unroll
    for (int i = 0; i < 8; i++) stage[st.slot[i]] = st.v[i];
}

template <class T>
__device__ __forceinline__ void runs_out(const float *stage, const RunStep &st, uint32_t dummy_off, uint32_t lane,
                                         T *__restrict__ V)
{
    char *Vb = reinterpret_cast<char *>(V);
    // the first kRunStores stores always (their count is what the loads'
    // waits assume), more only while ranks are left; the staged sums of the
    // first kRunEarly stores (all of a step's at R-MAT s24, about 200 runs)
    // are read before the first store, the others just before theirs
    const uint32_t nst = max(kRunStores, (st.total + 63u) / 64u);
    float vals[kRunEarly];
#pragma unroll
    for (uint32_t r = 0; r < kRunEarly; r++) vals[r] = stage[lane + 64u * r];
    uint64_t lead = st.lead;  // the lowest leader has rank 0 when the step has runs
    uint32_t cur = 0;
    uint32_t nl = lead ? (uint32_t)__builtin_ctzll(lead) : 0u;
    uint32_t np = lead ? __builtin_amdgcn_readlane(st.pre, nl) : 0xFFFFFFFFu;
#pragma unroll
    for (uint32_t r = 0; r < kRunGroup; r++) {
        if (r >= nst) break;
        const uint32_t q = lane + 64u * r;
        uint32_t d = cur;
        while (np < 64u * (r + 1)) {  // segments whose first rank falls in this store
            const uint32_t dl = __builtin_amdgcn_readlane(st.delta, nl);
            d = q >= np ? dl : d;
            cur = dl;
            lead &= lead - 1;
            nl = lead ? (uint32_t)__builtin_ctzll(lead) : 0u;
            np = lead ? __builtin_amdgcn_readlane(st.pre, nl) : 0xFFFFFFFFu;
        }
        const float v = r < kRunEarly ? vals[r] : stage[q];
        put_f32(reinterpret_cast<T *>(Vb + (q < st.total ? (q + d) * (uint32_t)sizeof(T) : dummy_off)), 0, v);
    }
}

// n contributions into the scatter's FP32 LDS copy: FP32 as they are, BF16
// widened (exact) 8 at a time, so the gathers of the run layout read FP32.
template <class T>
__device__ __forceinline__ void copy_to_lds_f32(float *dst, const T *src, uint32_t n)
{
    if constexpr (sizeof(T) == 4) {
        copy_to_lds<float, kRunsBlock>(dst, reinterpret_cast<const float *>(src), n);
    } else {
        const uint32_t nv = n / 8;
        const uint4 *s8 = reinterpret_cast<const uint4 *>(src);
        for (uint32_t i0 = threadIdx.x; i0 < nv; i0 += kRunsBlock * kCopyAhead) {
            // all loads first (indexes past the last group repeat it: same value, same address)
            uint4 w[kCopyAhead];
#pragma unroll
            for (int k = 0; k < kCopyAhead; k++) w[k] = s8[min(i0 + k * kRunsBlock, nv - 1u)];
#pragma unroll
            for (int k = 0; k < kCopyAhead; k++) {
                const uint32_t u[4] = {w[k].x, w[k].y, w[k].z, w[k].w};
                float4 *d4 = reinterpret_cast<float4 *>(dst + 8u * min(i0 + k * kRunsBlock, nv - 1u));
                d4[0] = make_float4(__uint_as_float(u[0] << 16), __uint_as_float(u[0] & 0xFFFF0000u),
                                    __uint_as_float(u[1] << 16), __uint_as_float(u[1] & 0xFFFF0000u));
                d4[1] = make_float4(__uint_as_float(u[2] << 16), __uint_as_float(u[2] & 0xFFFF0000u),
                                    __uint_as_float(u[3] << 16), __uint_as_float(u[3] & 0xFFFF0000u));
            }
        }
        for (uint32_t i = nv * 8 + threadIdx.x; i < n; i += kRunsBlock) dst[i] = f32_of(src[i]);
    }
}

// Run layout (float kinds): the groups of `src_local` of one chunk, from g0
// to g1, by the whole block: the chunk's contributions (sources c0..c1) are
// copied to LDS, then each wave takes steps of 64 consecutive groups.  The
// loads of the next kRunsAhead steps are in flight while a step is summed;
// a wave always runs whole rounds of kRunsAhead steps (steps past g1 load a
// valid group and stage nothing).
template <class T, bool W>
__device__ __forceinline__ void scatter_runs_piece(const RunsView &S, const T *__restrict__ c_in, T *__restrict__ V,
                                                   float *cl, float *stage, uint32_t dummy_off, uint32_t c0, uint32_t c1,
                                                   uint32_t g0, uint32_t g1)
{
    copy_to_lds_f32<T>(cl, c_in + c0, c1 - c0);
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t padw = S.pad | (S.pad << 16);
    const uint4 padv = make_uint4(padw, padw, padw, padw);
    const uint4 *src = reinterpret_cast<const uint4 *>(S.src_local);
    const uint4 *wsrc = reinterpret_cast<const uint4 *>(S.wcold);
    const uint32_t blk_last = (g1 - 1u) / kRunBlockGroups;
    uint4 w[kRunsAhead], rec[kRunsAhead];
    W8 wt[kRunsAhead];
    uint32_t sv[2];  // the V offset of the segment each lane would start, two steps ahead
    // a step's group data and block record; loads are always issued
    // (clamped into [g0, g1)), lanes outside it take pads when summed
    auto load = [&](int d, uint32_t first) {
        const uint32_t g = min(max(first + lane, g0), g1 - 1u);
        w[d] = ld_nt16(src + g);
        if constexpr (W) {
            wt[d].a = ld_nt16(wsrc + 2u * g);
            wt[d].b = ld_nt16(wsrc + 2u * g + 1u);
        }
        rec[d] = S.blk_rec[min(first / kRunBlockGroups, blk_last)];
    };
    // seg_v of the segment starting at each lane's group (used by segment
    // starts only), from the step's block record in slot d
    auto load_sv = [&](int d, int par) {
        const uint64_t bits = ((uint64_t)rec[d].w << 32) | rec[d].z;
        const uint32_t idx = rec[d].y + (uint32_t)__popcll(bits & ((1ull << lane) - 1ull));
        sv[par] = S.seg_v[min(idx, S.nseg - 1u)];
    };
    // sum the step that starts at group `first` from slot d (and sv[par]),
    // then refill d with the step kRunsAhead later and sv[par] with the
    // step two later
    auto sum = [&](int d, int par, uint32_t first) -> RunStep {
        __builtin_amdgcn_sched_barrier(0);
        const uint32_t g = first + lane;
        const bool act = g >= g0 && g < g1;
        const uint4 cw = act ? w[d] : padv;
        W8 cwt{};  // pads read the LDS zero: any finite weight gives 0
        if constexpr (W) cwt = wt[d];
        const uint64_t bits = ((uint64_t)rec[d].w << 32) | rec[d].z;
        const bool start = act && ((bits >> lane) & 1u);
        // the first active lane of a step leads the segment it is in (a
        // step starts at a multiple of 64 groups, or at g0 = a chunk's
        // first group, a segment start)
        const bool leader = start || (act && lane == 0);
        const uint32_t vbase = start ? sv[par] : rec[d].x;  // the slots are free for the next loads
        __builtin_amdgcn_sched_barrier(0);
        load(d, first + kRunsAhead * kRunsBlock);
        load_sv((d + 2) % kRunsAhead, par);
        return runs_sum<T, W>(cl, cw, cwt, vbase, leader);
    };
    // steps start at multiples of 64 groups of the whole src_local
    uint32_t r = (g0 & ~(kRunBlockGroups - 1u)) + (threadIdx.x & ~63u);
    if (r >= g1) return;
#pragma unroll
    for (int d = 0; d < kRunsAhead; d++) {
        load(d, r + (uint32_t)d * kRunsBlock);  // in slot order: slot 0 is waited for first
        __builtin_amdgcn_sched_barrier(0);
    }
    load_sv(0, 0);
    load_sv(1, 1);
    RunStep cur = sum(0, 0, r);
    runs_stage(stage, cur);
    // a round: for each of its steps, sum the next step, write out this
    // one, stage the next
    auto round = [&](uint32_t rr) {
#pragma unroll
        for (int d = 0; d < kRunsAhead; d++) {
            const RunStep nxt = sum((d + 1) % kRunsAhead, (d + 1) & 1, rr + (uint32_t)(d + 1) * kRunsBlock);
            wave_lds_order();
            runs_out<T>(stage, cur, dummy_off, lane, V);
            wave_lds_order();
            runs_stage(stage, nxt);
            cur = nxt;
        }
    };
    // the first round apart: every entry to the loop then follows a round of
    // the same loads and stores, and the waits inside it stay as short
    round(r);
    for (r += kRunsAhead * kRunsBlock; r < g1; r += kRunsAhead * kRunsBlock) round(r);
}

// Run layout (float kinds): one block per piece (a chunk or a part of one);
// the pieces come the largest first.
template <class T, bool W = false>
__global__ void __launch_bounds__(kRunsBlock) bin_scatter_runs_kernel(RunsView S, const Slot *slots, uint32_t t,
                                                                      const T *__restrict__ c_in, T *__restrict__ V)
{
    // the stop flag and the piece are loaded together (one round trip
    // before the contributions' loads instead of two); the empty asm keeps
    // the piece's loads ahead of the flag's branch
    const auto done = slots[t & 1].done;
    const uint4 pc = S.pieces[blockIdx.x];
    asm volatile("" ::"s"(pc.x), "s"(pc.y), "s"(pc.z), "s"(pc.w));
    if (done) return;
    float *cl = reinterpret_cast<float *>(gbx_dyn_lds);  // the chunk's contributions in FP32
    float *stage = reinterpret_cast<float *>(gbx_dyn_lds + S.stage_off) + (threadIdx.x >> 6) * kRunStage;
    const uint32_t dummy_off = S.dummy + ((blockIdx.x * (kRunsBlock / 64) + (threadIdx.x >> 6)) % kRunDummyAreas) *
                                             kRunDummyStride + (threadIdx.x & 63u) * (uint32_t)sizeof(T);
    if (threadIdx.x == 0) cl[S.pad] = 0.0f;  // the pads' zero, past the chunk's contributions
    scatter_runs_piece<T, W>(S, c_in, V, cl, stage, dummy_off, pc.x, pc.y, pc.z, pc.w);
}

struct AccumView {
    const BinUnit *__restrict__ units;
    const uint16_t *__restrict__ dst_local;
    const uint32_t *__restrict__ hot_pairs;
    const uint16_t *__restrict__ hot_base;  // packed pairs: row of each group's first pair (else null)
    long long *acc;  // split-bin sums
    uint32_t *cnt;   // split-bin arrival counters
    uint32_t H;      // hot sources
    uint32_t hot_off;  // byte offset of the hot contributions in LDS
    uint32_t *work;  // next unit (zero between iterations)
    uint32_t NU;     // units
    const float *__restrict__ whot;  // weighted kinds: the weight of every hot pair (0 for fillers)
    const GCoef *__restrict__ coef;  // a batch of graphs: every graph's coefficients
    double *__restrict__ uslots;     // a batch of graphs: kSegUnitSlots (err, dang, sum) per unit
};

// Values as loaded (raw) and as 64-bit fixed point.  Float kinds keep FP32
// (BF16 widened) and scale by 2^scale; fixed-point and quantized kinds are
// integers already.
template <int KIND>
using RawVal = std::conditional_t<is_int_kind(KIND), uint64_t, float>;

template <int KIND, class T>
__device__ __forceinline__ RawVal<KIND> raw_of(T x)
{
    if constexpr (is_int_kind(KIND)) return (uint64_t)x;
    else if constexpr (sizeof(T) == 2) return __uint_as_float((uint32_t)x << 16);  // bf16
    else return x;
}

__device__ __forceinline__ long long to_fixed(float v, double scale)
{
    return __double2ll_rn((double)v * scale);
}

template <int KIND>
__device__ __forceinline__ long long fixed_of(RawVal<KIND> v, double scale)
{
    if constexpr (is_int_kind(KIND)) return (long long)v;
    else return to_fixed(v, scale);
}

template <int KIND, class T>
__device__ __forceinline__ void load8(const T *V, uint32_t e, RawVal<KIND> (&v)[8])
{
    Vec8<T> w;
    if constexpr (sizeof(T) == 4) {
        const uint4 a = ld_nt16(V + e), b = ld_nt16(V + e + 4);
        const uint32_t u[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
        for (int i = 0; i < 8; i++) w.v[i] = __builtin_bit_cast(T, u[i]);
    } else if constexpr (sizeof(T) == 2) {
        const uint4 a = ld_nt16(V + e);
        const uint32_t u[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
        for (int i = 0; i < 8; i++) w.v[i] = (T)((u[i >> 1] >> ((i & 1) * 16)) & 0xFFFFu);
    } else {
        w = *reinterpret_cast<const Vec8<T> *>(V + e);
    }
#pragma unroll
    for (int i = 0; i < 8; i++) v[i] = raw_of<KIND, T>(w.v[i]);
}

// Eight hot (row, source) pairs: rows and values from the LDS contributions.
template <int KIND, class T>
__device__ __forceinline__ void load_hot8(const uint32_t *pairs, uint32_t e, const T *ch, uint32_t (&d)[8],
                                          RawVal<KIND> (&v)[8])
{
    const uint4 a = ld_nt16(pairs + e);
    const uint4 b = ld_nt16(pairs + e + 4);
    const uint32_t w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
    for (int i = 0; i < 8; i++) {
        d[i] = w[i] & 0xFFFFu;
        const uint32_t src = w[i] >> 16;
        v[i] = src == 0xFFFFu ? RawVal<KIND>(0) : raw_of<KIND, T>(ch[src]);
    }
}

// Eight packed hot pairs: rows from the group's base row and the 2-bit
// steps, values from the LDS contributions (kHotPad holds 0).
template <int KIND, class T>
__device__ __forceinline__ void load_hot8_packed(const uint16_t *pairs, const uint16_t *base, uint32_t e,
                                                 const T *ch, uint32_t (&d)[8], RawVal<KIND> (&v)[8])
{
    const uint4 a = ld_nt16(pairs + e);
    const uint32_t w[4] = {a.x, a.y, a.z, a.w};
    uint32_t r = base[e >> 3];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const uint32_t p = (i & 1) ? (w[i >> 1] >> 16) : (w[i >> 1] & 0xFFFFu);
        if (i) r += p >> 14;
        d[i] = r;
        v[i] = raw_of<KIND, T>(ch[p & kHotPad]);
    }
}

template <int KIND, class T>
__device__ __forceinline__ void load_hot8_any(const AccumView &B, uint32_t e, const T *ch, uint32_t (&d)[8],
                                              RawVal<KIND> (&v)[8])
{
    if (B.hot_base)
        load_hot8_packed<KIND, T>(reinterpret_cast<const uint16_t *>(B.hot_pairs), B.hot_base, e, ch, d, v);
    else
        load_hot8<KIND, T>(B.hot_pairs, e, ch, d, v);
    if constexpr (is_weighted_kind(KIND)) {
        // the pair's edge weight times its source's value (weights of fillers are 0)
        const uint4 a = ld_nt16(B.whot + e), b = ld_nt16(B.whot + e + 4);
        const uint32_t u[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
        for (int i = 0; i < 8; i++) v[i] = __uint_as_float(u[i]) * v[i];
    }
}

// A batch of graphs: the stochastic update of the batch's pull iteration
// (r' = (1-d)/N_k + d/S_k (sum + D_k/N_k), the same float operations) with
// the coefficients of the unit's graph, loaded when the unit starts; every
// wave's (err, dang, sum) of the unit goes to the unit's slots.
struct SegBinEpi {
    float *__restrict__ r;
    float *__restrict__ c_out;
    const uint32_t *__restrict__ outdeg;
    const GCoef *__restrict__ coef;
    double *__restrict__ uslots;
    float fa, ti, fs, fd;
    double err, dang, sum;
    Partial part;  // unused: the partials of a batch are per unit
    struct In {
        float r_old;
        uint32_t od;
    };
    __device__ __forceinline__ float init(uint32_t) const { return 0.0f; }
    // false when the unit's graph has converged (its rows stay as they are)
    __device__ __forceinline__ bool begin(uint32_t graph)
    {
        const GCoef cf = coef[graph];
        ti = cf.ti;
        fs = cf.fs;
        fd = cf.fd;
        err = dang = sum = 0.0;
        return cf.done == 0;
    }
    __device__ __forceinline__ In load(uint32_t i) const { return In{r[i], outdeg[i]}; }
    __device__ __forceinline__ void apply(uint32_t i, float acc, const In &in)
    {
        const float r_new = fa * ti + fs * (acc + fd);
        r[i] = r_new;
        c_out[i] = in.od ? r_new / (float)in.od : 0.0f;
        err += (double)fabsf(r_new - in.r_old);
        if (!in.od) dang += (double)r_new;
        sum += (double)r_new;
    }
    __device__ __forceinline__ void operator()(uint32_t i, float acc) { apply(i, acc, load(i)); }
    __device__ __forceinline__ void put(uint32_t unit, uint32_t slot, double e, double d, double s2) const
    {
        double *q = uslots + ((size_t)unit * kSegUnitSlots + slot) * 3u;
        q[0] = e;
        q[1] = d;
        q[2] = s2;
    }
    // the wave's sums into its slot of the unit
    __device__ __forceinline__ void wave_slot(uint32_t unit, uint32_t wave)
    {
        double e = err, d = dang, s2 = sum;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            e += __shfl_xor(e, off, 64);
            d += __shfl_xor(d, off, 64);
            s2 += __shfl_xor(s2, off, 64);
        }
        if ((threadIdx.x & 63u) == 0) put(unit, wave, e, d, s2);
    }
};

__device__ __forceinline__ SegBinEpi make_segbin_epi(const AccumView &B, const uint32_t *outdeg, const IterView &A)
{
    SegBinEpi e;
    e.r = A.r;
    e.c_out = reinterpret_cast<float *>(A.c_out);
    e.outdeg = outdeg;
    e.coef = B.coef;
    e.uslots = B.uslots;
    e.fa = 1.0f - A.damp;
    e.ti = e.fs = e.fd = 0.0f;
    e.err = e.dang = e.sum = 0.0;
    e.part = Partial{0.0, 0.0, 0.0, 0u, 0u, 0u, 0u};
    return e;
}

template <int KIND, class T>
using EpiOf = std::conditional_t<
    KIND == K_SEG, SegBinEpi,
    std::conditional_t<
    KIND == K_SPMV_W, SpmvBinEpi,
    std::conditional_t<is_fixed_kind(KIND), FixEpi<KIND, T>,
                       std::conditional_t<is_quant_kind(KIND), QuantEpi<KIND>, PrEpi<KIND, T>>>>>;

constexpr int kAccumUnroll = 4;  // groups of 8 per lane in flight
constexpr uint32_t kAccumSlab = 64u * 8u * kAccumUnroll;  // V entries a wave takes at a time
// groups of 8 hot pairs per lane in flight (weighted kinds: one, for the
// registers their weights take)
template <int KIND>
constexpr int hot_unroll() { return is_weighted_kind(KIND) ? 1 : 2; }
template <int KIND>
constexpr uint32_t hot_slab() { return 64u * 8u * hot_unroll<KIND>(); }  // hot pairs a wave takes at a time
// Rows per thread whose update inputs load together: 13 for reference
// semantics, so that a unit of up to 12,480 rows is one batch, loaded before
// its sums stream in (the other kinds hold more per row in their update and
// stay at 8: at 13 the stochastic one spills).
template <int KIND>
constexpr int epi_batch() { return KIND == K_PR_REF ? 13 : 8; }

// Thread 0 takes the next unit from the counter and loads its record.
__device__ __forceinline__ void claim_unit(const AccumView &B, uint32_t *s_w, BinUnit *s_u)
{
    const uint32_t w = atomicAdd(B.work, 1u);
    *s_w = w;
    if (w < B.NU) *s_u = B.units[w];
}

// One accumulate unit: row sums of the unit's V range and hot pairs, then the
// update of its rows (or, for a split bin, the combine), while thread 0
// claims the next unit into s_w, s_u.  The hot contributions are already in
// LDS, the row sums are zero on entry and the update leaves them zero;
// s_slab holds the two slab counters (zero on entry, zero again when the
// update starts).
template <int KIND, class T>
__device__ __forceinline__ void accum_unit(const BinUnit &u, uint32_t uw, const AccumView &B, const T *__restrict__ V,
                                           EpiOf<KIND, T> &epi, double scale, double inv, uint32_t *s_slab,
                                           uint32_t *s_w, BinUnit *s_u)
{
    // a batch of graphs: the units of a converged graph sum and update nothing
    bool live = true;
    if constexpr (KIND == K_SEG) live = epi.begin(u.graph);
    const uint32_t nrows = live ? u.row_end - u.row_begin : 0u;
    const uint32_t stride = kBlock * 8u;
    const T *ch = reinterpret_cast<const T *>(gbx_dyn_lds + B.hot_off);
    const bool has_hot = live && u.hend > u.hbeg;
    if (u.row_end - u.row_begin == 1) {
        // one row: lane sums, block sum, then the split-bin combine if any
        __shared__ long long wsum[kBlock / 64];
        long long s = 0;
        if (has_hot) {
            for (uint32_t e = u.hbeg + threadIdx.x * 8u; e < u.hend; e += stride) {
                uint32_t d[8];
                RawVal<KIND> v[8];
                load_hot8_any<KIND, T>(B, e, ch, d, v);
#pragma unroll
                for (int i = 0; i < 8; i++) s += fixed_of<KIND>(v[i], scale);
            }
        }
        const uint32_t vend = live ? u.vend : u.vbeg;
        for (uint32_t e0 = u.vbeg + threadIdx.x * 8u; e0 < vend; e0 += stride * kAccumUnroll) {
            RawVal<KIND> v[kAccumUnroll][8];
#pragma unroll
            for (int q = 0; q < kAccumUnroll; q++)
                if (e0 + q * stride < vend) load8<KIND, T>(V, e0 + q * stride, v[q]);
#pragma unroll
            for (int q = 0; q < kAccumUnroll; q++)
                if (e0 + q * stride < vend) {
#pragma unroll
                    for (int i = 0; i < 8; i++) s += fixed_of<KIND>(v[q][i], scale);
                }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        if ((threadIdx.x & 63u) == 0) wsum[threadIdx.x >> 6] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            long long tot = 0;
            for (int w = 0; w < kBlock / 64; w++) tot += wsum[w];
            bool apply = live;
            if (live && u.multi != kNoMulti) {
                // both the sum and the ticket are device-scope atomics: the
                // sum is performed before the ticket is taken, and the last
                // unit reads it back with an atomic (no fences)
                atomicAdd(reinterpret_cast<unsigned long long *>(B.acc + u.multi), (unsigned long long)tot);
                stores_performed();
                const uint32_t ticket = atomicAdd(B.cnt + u.multi, 1u);
                apply = ticket == u.nunits - 1;
                if (apply) {
                    tot = (long long)atomicAdd(reinterpret_cast<unsigned long long *>(B.acc + u.multi), 0ull);
                    B.acc[u.multi] = 0;
                    B.cnt[u.multi] = 0;
                }
            }
            if (apply) {
                if constexpr (is_int_kind(KIND)) epi(u.row_begin, (uint64_t)tot);
                else epi(u.row_begin, (float)((double)tot * inv));
            }
            if constexpr (KIND == K_SEG) {
                // the row's partials in slot 0 (zeros for the units of a split bin that do not apply)
                if (live)
                    for (uint32_t w = 0; w < kSegUnitSlots; w++)
                        epi.put(uw, w, w ? 0.0 : epi.err, w ? 0.0 : epi.dang, w ? 0.0 : epi.sum);
            }
            claim_unit(B, s_w, s_u);
        }
    } else {
        long long *y = reinterpret_cast<long long *>(gbx_dyn_lds);
        // The update inputs of this thread's first batch of rows load now and
        // arrive while the unit's sums stream in (the update runs on waves
        // 1.. after the barrier; its first batch then waits for nothing).
        constexpr uint32_t kEpiThreads = kBlock - 64;
        constexpr int kEpiBatch = epi_batch<KIND>();
        typename EpiOf<KIND, T>::In first[kEpiBatch];
        if (threadIdx.x >= 64) {
#pragma unroll
            for (int q = 0; q < kEpiBatch; q++) {
                const uint32_t i = threadIdx.x - 64 + q * kEpiThreads;
                if (i < nrows) first[q] = epi.load(u.row_begin + i);
            }
        }
        // The waves take slabs (kHotSlab hot pairs first, then kAccumSlab V
        // entries) from an LDS counter: waves that run ahead take more, so the
        // block reaches the update barrier together.
        const uint32_t lane = threadIdx.x & 63u;
        constexpr int kHotUnroll = hot_unroll<KIND>();
        constexpr uint32_t kHotSlab = hot_slab<KIND>();
        const uint32_t nh = live ? (u.hend - u.hbeg + kHotSlab - 1) / kHotSlab : 0u;
        const uint32_t nv = live ? (u.vend - u.vbeg + kAccumSlab - 1) / kAccumSlab : 0u;
        constexpr uint32_t kq = 64u * 8u;  // entries of one group of 8 per lane
        for (;;) {
            uint32_t k = 0;
            if (lane == 0) k = atomicAdd(&s_slab[0], 1u);
            k = __builtin_amdgcn_readfirstlane(k);
            if (k >= nh) break;
            {
                const uint32_t e0 = u.hbeg + k * kHotSlab + lane * 8u;
                // groups past the range load the range's last group and add zeros
                uint32_t d[kHotUnroll][8];
                RawVal<KIND> v[kHotUnroll][8];
#pragma unroll
                for (int q = 0; q < kHotUnroll; q++) {
                    const uint32_t e = e0 + q * kq;
                    const bool in = e < u.hend;
                    load_hot8_any<KIND, T>(B, in ? e : u.hend - 8u, ch, d[q], v[q]);
                    if (!in)
#pragma unroll
                        for (int i = 0; i < 8; i++) v[q][i] = RawVal<KIND>(0);
                }
#pragma unroll
                for (int q = 0; q < kHotUnroll; q++) {
                    // hot pairs are in row order: runs of equal rows summed in registers
                    uint32_t rd = d[q][0];
                    long long run = 0;
#pragma unroll
                    for (int i = 0; i < 8; i++) {
                        const long long x = fixed_of<KIND>(v[q][i], scale);
                        if (d[q][i] != rd) {
                            if (run) atomicAdd(reinterpret_cast<unsigned long long *>(y + rd), (unsigned long long)run);
                            rd = d[q][i];
                            run = x;
                        } else {
                            run += x;
                        }
                    }
                    if (run) atomicAdd(reinterpret_cast<unsigned long long *>(y + rd), (unsigned long long)run);
                }
            }
        }
        for (;;) {
            uint32_t k = 0;
            if (lane == 0) k = atomicAdd(&s_slab[1], 1u);
            k = __builtin_amdgcn_readfirstlane(k);
            if (k >= nv) break;
            const uint32_t e0 = u.vbeg + k * kAccumSlab + lane * 8u;
            // all loads of the unrolled groups first, then the sums
            uint4 dw[kAccumUnroll];
            RawVal<KIND> v[kAccumUnroll][8];
#pragma unroll
            for (int q = 0; q < kAccumUnroll; q++)
                if (e0 + q * kq < u.vend) {
                    dw[q] = ld_nt16(B.dst_local + e0 + q * kq);
                    load8<KIND, T>(V, e0 + q * kq, v[q]);
                }
#pragma unroll
            for (int q = 0; q < kAccumUnroll; q++) {
                if (e0 + q * kq >= u.vend) break;
                const uint32_t w[4] = {dw[q].x, dw[q].y, dw[q].z, dw[q].w};
                // runs of equal rows (segments are in row order) are summed
                // in registers before one LDS atomic per run
                uint32_t rd = w[0] & 0xFFFFu;
                long long run = 0;
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    const uint32_t d = (i & 1) ? (w[i >> 1] >> 16) : (w[i >> 1] & 0xFFFFu);
                    const long long x = fixed_of<KIND>(v[q][i], scale);
                    if (d != rd) {
                        if (run) atomicAdd(reinterpret_cast<unsigned long long *>(y + rd), (unsigned long long)run);
                        rd = d;
                        run = x;
                    } else {
                        run += x;
                    }
                }
                if (run) atomicAdd(reinterpret_cast<unsigned long long *>(y + rd), (unsigned long long)run);
            }
        }
        __syncthreads();
        if (threadIdx.x < 64) {
            // wave 0 resets the slab counters (every wave has taken its last
            // slab) and claims the next unit while the other waves update
            if (threadIdx.x < 2) s_slab[threadIdx.x] = 0;
            if constexpr (KIND == K_SEG) {
                if (live && threadIdx.x == 0) epi.put(uw, 0, 0.0, 0.0, 0.0);  // wave 0 updates no row
            }
            if (threadIdx.x == 0) claim_unit(B, s_w, s_u);
            return;
        }
        // the update on waves 1..: inputs of kEpiBatch rows per thread loaded
        // together (the first batch's before the sums); each row sum is
        // cleared once read
        auto update = [&](uint32_t i0, const typename EpiOf<KIND, T>::In (&in)[kEpiBatch]) {
#pragma unroll
            for (int q = 0; q < kEpiBatch; q++) {
                const uint32_t i = i0 + q * kEpiThreads;
                if (i < nrows) {
                    const long long yi = y[i];
                    y[i] = 0;
                    if constexpr (is_int_kind(KIND)) epi.apply(u.row_begin + i, (uint64_t)yi, in[q]);
                    else epi.apply(u.row_begin + i, (float)((double)yi * inv), in[q]);
                }
            }
        };
        if (threadIdx.x - 64 < nrows) update(threadIdx.x - 64, first);
        for (uint32_t i0 = threadIdx.x - 64 + kEpiThreads * kEpiBatch; i0 < nrows; i0 += kEpiThreads * kEpiBatch) {
            typename EpiOf<KIND, T>::In in[kEpiBatch];
#pragma unroll
            for (int q = 0; q < kEpiBatch; q++)
                if (i0 + q * kEpiThreads < nrows) in[q] = epi.load(u.row_begin + i0 + q * kEpiThreads);
            update(i0, in);
        }
        if constexpr (KIND == K_SEG) {
            if (live) epi.wave_slot(uw, threadIdx.x >> 6);
        }
    }
}

// Persistent: one block per CU takes units from a counter until none are
// left, so the hot contributions are loaded once per block and one unit's
// update overlaps the next unit's loads in other blocks.
template <int KIND, class T>
__global__ void __launch_bounds__(kBlock) bin_accum_kernel(AccumView B, const uint32_t *__restrict__ outdeg,
                                                           uint32_t N, const T *__restrict__ V, IterView A)
{
    const Slot *cur = A.slots + (A.t & 1);
    Slot *nxt = A.slots + ((A.t + 1) & 1);
    if (cur->done) {
        if (blockIdx.x == 0 && threadIdx.x == 0) *nxt = *cur;
        return;
    }
    EpiOf<KIND, T> epi;
    if constexpr (is_fixed_kind(KIND)) epi = make_fix_epi<KIND, T>(outdeg, N, A);
    else if constexpr (is_quant_kind(KIND)) epi = make_quant_epi<KIND>(outdeg, N, A, cur);
    else if constexpr (KIND == K_SPMV_W) epi = make_spmv_epi(A);
    else if constexpr (KIND == K_SEG) epi = make_segbin_epi(B, outdeg, A);
    else epi = make_epi<KIND, T>(outdeg, N, A, cur);
    // float kinds: every row sum is at most the sum of the magnitudes of its
    // terms, and those of all rows together are at most: the previous rank
    // sum (each source's transition weights sum to at most 1; bf16 rounding
    // adds < 0.4 %), for the adjoint the previous sum of |lambda| (each row
    // of P sums to 1, so a row sum is at most max |lambda|), for the SpMV the
    // sum over all edges of |w x| that its reset stores in the slot;
    // fixed-point and quantized kinds sum their integers as they are (uint64
    // wrap-around, as the reference's uint64 sums)
    const double bound = (KIND == K_ADJ ? cur->dang : cur->sum) * 1.01 + 1e-30;
    int ex = 0;
    (void)frexp(bound, &ex);
    const double scale = ldexp(1.0, 61 - ex), inv = ldexp(1.0, ex - 61);
    if (B.H) {
        T *hot = reinterpret_cast<T *>(gbx_dyn_lds + B.hot_off);
        if constexpr (is_quant_kind(KIND))
            quant_to_lds<KIND, T>(hot, reinterpret_cast<const float *>(A.c_in), B.H, epi.scale);
        else copy_to_lds(hot, reinterpret_cast<const T *>(A.c_in), B.H);
        if (B.hot_base && threadIdx.x == 0) hot[kHotPad] = T(0);  // packed pairs' pads and fillers
    }
    // Thread 0 claims the next unit while the other waves update the rows
    // of the current one, so neither the claim nor the record load waits at
    // a barrier of its own.
    __shared__ uint32_t s_w, s_slab[2];
    __shared__ BinUnit s_u;
    {
        long long *y = reinterpret_cast<long long *>(gbx_dyn_lds);
        for (uint32_t i = threadIdx.x; i < B.hot_off / 8u; i += kBlock) y[i] = 0;
        if (threadIdx.x < 2) s_slab[threadIdx.x] = 0;
    }
    if (threadIdx.x == 0) claim_unit(B, &s_w, &s_u);
    for (;;) {
        __syncthreads();  // next unit claimed; the previous unit done with LDS (and the hot load finished)
        if (s_w >= B.NU) break;
        // thread 0 claims again only after the unit's sums barrier, when
        // every thread has its copy
        const BinUnit u = s_u;
        const uint32_t uw = s_w;
        accum_unit<KIND, T>(u, uw, B, V, epi, scale, inv, s_slab, &s_w, &s_u);
    }
    if constexpr (KIND == K_SEG) {
        // the partials are per unit (reduced per graph by the next kernel); the
        // last block to leave its unit loop clears the unit counter
        __syncthreads();
        if (threadIdx.x == 0) {
            if (atomicAdd(A.counter, 1u) == gridDim.x - 1) {
                *B.work = 0;
                *A.counter = 0;
            }
        }
        return;
    }
    const uint32_t iters = cur->iters;
    const double eps = A.eps;
    const int stop = A.stop;
    uint32_t *work = B.work;
    grid_reduce(epi.part, A.partials, A.counter, [&](const Partial &tot) {
        if constexpr (KIND == K_SPMV_W) {
            Slot s = *cur;  // the bound of the (unchanging) input vector carries over
            s.iters = iters + 1;
            *nxt = s;
        } else {
            *nxt = next_slot<KIND>(tot, iters, eps, stop);
        }
        *work = 0;  // every block has left its unit loop
    });
}

// ------------------------------------------------------------------ host

// Runs of a sorted key array: flags, run index, starts and padded lengths,
// padded starts (exclusive scan) and their total.
struct Runs {
    DevBuf flag, idx, start, plen, pstart, total;
    uint32_t n = 0, padded = 0;
};

static int find_runs(gbx_ctx *ctx, const uint32_t *key, uint32_t E, uint32_t pad, Runs &R)
{
    GBX_TRY(dev_alloc(ctx, R.flag, (size_t)E * 4 + 4));
    GBX_TRY(dev_alloc(ctx, R.idx, (size_t)E * 4 + 4));
    run_flags<<<grid_for(E, 256, 16384), 256, 0, ctx->stream>>>(key, E, R.flag.as<uint32_t>());
    GBX_HIP(ctx, hipGetLastError());
    GBX_TRY(exclusive_scan_u32(ctx, R.flag.as<uint32_t>(), R.idx.as<uint32_t>(), E, nullptr));
    uint32_t last_idx = 0, last_flag = 0;
    GBX_HIP(ctx, hipMemcpyAsync(&last_idx, R.idx.as<uint32_t>() + (E - 1), 4, hipMemcpyDeviceToHost, ctx->stream));
    GBX_HIP(ctx, hipMemcpyAsync(&last_flag, R.flag.as<uint32_t>() + (E - 1), 4, hipMemcpyDeviceToHost, ctx->stream));
    GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    R.n = last_idx + last_flag;
    GBX_TRY(dev_alloc(ctx, R.start, (size_t)R.n * 4 + 4));
    GBX_TRY(dev_alloc(ctx, R.plen, (size_t)R.n * 4 + 4));
    GBX_TRY(dev_alloc(ctx, R.pstart, (size_t)(R.n + 1) * 4));
    GBX_TRY(dev_alloc(ctx, R.total, 16));
    run_starts<<<grid_for(E, 256, 16384), 256, 0, ctx->stream>>>(R.flag.as<uint32_t>(), R.idx.as<uint32_t>(), E,
                                                                   R.start.as<uint32_t>());
    run_padded_lengths<<<grid_for(R.n, 256, 16384), 256, 0, ctx->stream>>>(R.start.as<uint32_t>(), R.n, E, pad,
                                                                             R.plen.as<uint32_t>());
    GBX_HIP(ctx, hipGetLastError());
    GBX_TRY(exclusive_scan_u32(ctx, R.plen.as<uint32_t>(), R.pstart.as<uint32_t>(), R.n, R.total.as<uint32_t>()));
    GBX_HIP(ctx, hipMemcpyAsync(&R.padded, R.total.p, 4, hipMemcpyDeviceToHost, ctx->stream));
    GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    // a padded total that wrapped around 32 bits is smaller than E
    if ((uint64_t)R.padded < E || R.padded > 0xFFFFFF00u)
        return set_error(ctx, GBX_ERR_RANGE, "binned layout: padded length %u for %u edges", R.padded, E);
    GBX_HIP(ctx, hipMemcpyAsync(R.pstart.as<uint32_t>() + R.n, &R.padded, 4, hipMemcpyHostToDevice, ctx->stream));
    GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GBX_OK;
}

// Run layout of the cold edges (rowid_c, col_c: plan edge order, so the
// edges of one row are consecutive in every segment): src_local chunk-major
// with segments padded to 8 entries (pads: the LDS index bp->SC), V
// bin-major, within a bin chunk-major, one value per run, segments padded
// to 8 values (zeros, never written by the scatter).
static int build_runs(gbx_ctx *ctx, gbx_bplan *bp, uint32_t Ecold, const uint32_t *rowid_c, const uint32_t *col_c,
                      const uint32_t *eidx_c, const uint32_t *bin_of, const uint32_t *chunk_of,
                      const uint32_t *bin_row0, const uint32_t *cstart, uint32_t *bin_len)
{
    const uint32_t NB = bp->NB, NC = bp->NC;
    const uint32_t kbits = bits_for(NB * NC - 1);
    const unsigned gE = grid_for(Ecold, 256, 16384);
    DevBuf key, local, pedge;
    GBX_TRY(dev_alloc(ctx, key, (size_t)Ecold * 4));
    GBX_TRY(dev_alloc(ctx, local, (size_t)Ecold * 4));
    GBX_TRY(dev_alloc(ctx, pedge, (size_t)Ecold * 4));
    run_keys<<<gE, 256, 0, ctx->stream>>>(rowid_c, col_c, eidx_c, Ecold, bin_of, chunk_of, bin_row0,
                                          bp->chunk_lo.as<uint32_t>(), NB, key.as<uint32_t>(), local.as<uint32_t>(),
                                          pedge.as<uint32_t>());
    GBX_HIP(ctx, hipGetLastError());
    // the plan edges ride along as the sort's third payload (moved as bits)
    GBX_TRY(radix_sort_pairs(ctx, key.as<uint32_t>(), local.as<uint32_t>(), reinterpret_cast<float *>(pedge.p),
                             Ecold, kbits));
    Runs R;
    GBX_TRY(find_runs(ctx, key.as<uint32_t>(), Ecold, kRunGroup, R));
    bp->nseg = R.n;
    bp->Plen = R.padded;
    chunk_groups<<<grid_for(NC + 1, 256, 1024), 256, 0, ctx->stream>>>(
        cstart, NC, R.idx.as<uint32_t>(), Ecold, R.n, R.pstart.as<uint32_t>(), 3u, bp->chunk_grp.as<uint32_t>());
    GBX_HIP(ctx, hipGetLastError());
    // the scatter's pieces: every chunk, the large ones cut into equal parts
    // (at step starts), the largest first
    {
        std::vector<uint32_t> cg(NC + 1), lo(NC + 1);
        GBX_HIP(ctx, hipMemcpyAsync(cg.data(), bp->chunk_grp.p, (size_t)(NC + 1) * 4, hipMemcpyDeviceToHost,
                                    ctx->stream));
        GBX_HIP(ctx, hipMemcpyAsync(lo.data(), bp->chunk_lo.p, (size_t)(NC + 1) * 4, hipMemcpyDeviceToHost,
                                    ctx->stream));
        GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        const uint64_t total = cg[NC] - cg[0];
        const uint64_t target = std::max<uint64_t>(total / (2u * (uint64_t)std::max(1, ctx->num_cus)), kPieceGroups);
        const uint32_t S = (uint32_t)((target + kRunBlockGroups - 1) &
                                      ~(uint64_t)(kRunBlockGroups - 1));
        std::vector<uint4> pc;
        for (uint32_t k = 0; k < NC; k++) {
            const uint32_t g0 = cg[k], g1 = cg[k + 1];
            if (g0 >= g1) continue;
            const uint32_t parts = (uint32_t)(((uint64_t)(g1 - g0) + S - 1) / S);
            const uint64_t step = ((uint64_t)(g1 - g0) + parts - 1) / parts;
            uint32_t a = g0;
            for (uint32_t q = 1; q <= parts; q++) {
                const uint32_t b = q == parts ? g1
                                              : std::max(a, (uint32_t)((g0 + step * q) & ~(uint64_t)(kRunBlockGroups - 1)));
                if (b > a) pc.push_back(make_uint4(lo[k], lo[k + 1], a, b));
                a = b;
            }
        }
        std::stable_sort(pc.begin(), pc.end(), [](const uint4 &x, const uint4 &y) { return x.w - x.z > y.w - y.z; });
        if (getenv("GBX_VERBOSE"))
            fprintf(stderr, "[gbx] scatter pieces: %zu of at most %u groups (%llu groups)\n", pc.size(), S,
                    (unsigned long long)total);
        GBX_TRY(dev_alloc(ctx, bp->pieces, pc.size() * sizeof(uint4) + 16));
        if (!pc.empty())
            GBX_HIP(ctx, hipMemcpyAsync(bp->pieces.p, pc.data(), pc.size() * sizeof(uint4), hipMemcpyHostToDevice,
                                        ctx->stream));
        GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        bp->NP = (uint32_t)pc.size();
    }
    // run starts and their running count
    DevBuf head, hx, htot;
    GBX_TRY(dev_alloc(ctx, head, (size_t)Ecold * 4 + 4));
    GBX_TRY(dev_alloc(ctx, hx, (size_t)(Ecold + 1) * 4));
    GBX_TRY(dev_alloc(ctx, htot, 16));
    run_heads<<<gE, 256, 0, ctx->stream>>>(local.as<uint32_t>(), Ecold, R.flag.as<uint32_t>(), R.idx.as<uint32_t>(),
                                           R.start.as<uint32_t>(), R.pstart.as<uint32_t>(), head.as<uint32_t>());
    GBX_HIP(ctx, hipGetLastError());
    GBX_TRY(exclusive_scan_u32(ctx, head.as<uint32_t>(), hx.as<uint32_t>(), Ecold, htot.as<uint32_t>()));
    GBX_HIP(ctx, hipMemcpyAsync(hx.as<uint32_t>() + Ecold, htot.p, 4, hipMemcpyDeviceToDevice, ctx->stream));
    // segments: V lengths, then V offsets in bin-major order
    const uint32_t ns = R.n;
    DevBuf vlen, skey, spay, vsorted, osorted, voff, vtot;
    for (DevBuf *d : {&vlen, &skey, &spay, &vsorted, &osorted, &voff}) GBX_TRY(dev_alloc(ctx, *d, (size_t)ns * 4 + 4));
    GBX_TRY(dev_alloc(ctx, vtot, 16));
    const unsigned gS = grid_for(ns, 256, 16384);
    run_segments<<<gS, 256, 0, ctx->stream>>>(key.as<uint32_t>(), R.start.as<uint32_t>(), ns, Ecold,
                                              hx.as<uint32_t>(), NB, NC, vlen.as<uint32_t>(), skey.as<uint32_t>(),
                                              spay.as<uint32_t>(), bin_len);
    GBX_HIP(ctx, hipGetLastError());
    GBX_TRY(radix_sort_pairs(ctx, skey.as<uint32_t>(), spay.as<uint32_t>(), nullptr, ns, kbits));
    gather_u32<<<gS, 256, 0, ctx->stream>>>(vlen.as<uint32_t>(), spay.as<uint32_t>(), ns, vsorted.as<uint32_t>());
    GBX_HIP(ctx, hipGetLastError());
    GBX_TRY(exclusive_scan_u32(ctx, vsorted.as<uint32_t>(), osorted.as<uint32_t>(), ns, vtot.as<uint32_t>()));
    scatter_u32<<<gS, 256, 0, ctx->stream>>>(osorted.as<uint32_t>(), spay.as<uint32_t>(), ns, voff.as<uint32_t>());
    GBX_HIP(ctx, hipGetLastError());
    uint32_t Vlen = 0;
    GBX_HIP(ctx, hipMemcpyAsync(&Vlen, vtot.p, 4, hipMemcpyDeviceToHost, ctx->stream));
    GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    // runs <= entries in every segment, both padded to 8
    if (Vlen > bp->Plen) return set_error(ctx, GBX_ERR_STATE, "binned layout: %u runs for %u entries", Vlen, bp->Plen);
    bp->Vlen = Vlen;
    GBX_TRY(dev_alloc(ctx, bp->src_local, (size_t)bp->Plen * 2 + 16));
    GBX_TRY(dev_alloc(ctx, bp->dst_local, (size_t)Vlen * 2 + 16));
    const uint32_t ngrp = bp->Plen / kRunGroup, nblk = (ngrp + kRunBlockGroups - 1) / kRunBlockGroups;
    DevBuf blk_v, bits, bcnt, bseg, btot;
    GBX_TRY(dev_alloc(ctx, blk_v, (size_t)nblk * 4 + 4));
    GBX_TRY(dev_alloc(ctx, bits, (size_t)nblk * 8 + 8));
    GBX_TRY(dev_alloc(ctx, bcnt, (size_t)nblk * 4 + 4));
    GBX_TRY(dev_alloc(ctx, bseg, (size_t)nblk * 4 + 4));
    GBX_TRY(dev_alloc(ctx, btot, 16));
    GBX_TRY(dev_alloc(ctx, bp->blk_rec, (size_t)nblk * 16 + 16));
    GBX_TRY(fill_u32(ctx, bits.as<uint32_t>(), 0, (size_t)nblk * 2 + 2));
    GBX_TRY(fill_u32(ctx, blk_v.as<uint32_t>(), 0, nblk + 1));
    GBX_TRY(dev_alloc(ctx, bp->bdst, 16));
    GBX_TRY(fill_u32(ctx, bp->src_local.as<uint32_t>(), bp->SC | (bp->SC << 16), bp->Plen / 2 + 4));
    GBX_HIP(ctx, hipMemsetAsync(bp->dst_local.p, 0, (size_t)Vlen * 2 + 16, ctx->stream));
    GBX_TRY(dev_alloc(ctx, bp->cold_eidx, (size_t)bp->Plen * 4 + 16));
    GBX_TRY(fill_u32(ctx, bp->cold_eidx.as<uint32_t>(), kNoEdge, bp->Plen + 4));

    run_fill<<<gE, 256, 0, ctx->stream>>>(local.as<uint32_t>(), pedge.as<uint32_t>(), Ecold, R.flag.as<uint32_t>(),
                                          R.idx.as<uint32_t>(), R.start.as<uint32_t>(), R.pstart.as<uint32_t>(),
                                          head.as<uint32_t>(), hx.as<uint32_t>(), voff.as<uint32_t>(), bp->SC,
                                          bp->src_local.as<uint16_t>(), bp->dst_local.as<uint16_t>(),
                                          blk_v.as<uint32_t>(), bp->cold_eidx.as<uint32_t>());
    GBX_HIP(ctx, hipGetLastError());
    // block records: segment-start bits, their running count, first V positions
    run_segment_bits<<<gS, 256, 0, ctx->stream>>>(R.pstart.as<uint32_t>(), ns, bits.as<uint32_t>());
    run_block_counts<<<grid_for(nblk, 256, 16384), 256, 0, ctx->stream>>>(bits.as<uint32_t>(), nblk,
                                                                          bcnt.as<uint32_t>());
    GBX_HIP(ctx, hipGetLastError());
    GBX_TRY(exclusive_scan_u32(ctx, bcnt.as<uint32_t>(), bseg.as<uint32_t>(), nblk, btot.as<uint32_t>()));
    run_block_records<<<grid_for(nblk, 256, 16384), 256, 0, ctx->stream>>>(
        blk_v.as<uint32_t>(), bseg.as<uint32_t>(), bits.as<uint32_t>(), nblk, bp->blk_rec.as<uint4>());
    GBX_HIP(ctx, hipGetLastError());
    bp->seg_v = std::move(voff);
    GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GBX_OK;
}

static int build_bplan_into(gbx_ctx *ctx, gbx_plan *p, uint32_t pad, uint32_t H, uint32_t hot_elem, bool runs,
                            const std::vector<uint32_t> *goff, gbx_bplan *bp)
{
    const gbx_tuning &tu = ctx->tune;
    const uint32_t N = p->N, E = p->E;
    const uint32_t EC = cut_edges_of(tu.binned_chunk_edges, ctx, E, runs, true), SC = chunk_sources_of(tu, hot_elem, runs);
    // rows per bin: what the LDS holds next to H hot contributions of hot_elem
    // bytes (a multiple of 8)
    const bool hot16 = runs && H > 0;
    const uint32_t hot_slots = hot16 ? kHot16Slots : H;
    const uint32_t SB =
        std::min(std::min(tu.binned_bin_rows, kMaxBinRows), (kAccumLds - hot_slots * hot_elem) / 8u) & ~7u;
    const uint32_t EB = (std::max(cut_edges_of(tu.binned_bin_edges, ctx, E, runs, false), 8u) + 7u) & ~7u;
    if (SB < 8) return set_error(ctx, GBX_ERR_ARG, "binned layout: %u hot sources leave no room for bins", H);
    bp->H = H;
    bp->SB = SB;
    bp->SC = SC;
    bp->runs = runs;
    bp->hot16 = hot16;
    if (hot16 && H > kHot16Max) return set_error(ctx, GBX_ERR_ARG, "binned layout: %u hot sources, packed pairs take %u", H, kHot16Max);
    // in-edge counts of the cold sources (u >= H; the hot ones count 0), their
    // exclusive prefix pre[0..N] and, per group of 8 sources, pre at its first
    // source and one past it: what the chunk cut reads
    const uint32_t G = (N + 7) / 8;
    DevBuf dcnt, dpre, dgrp;
    GBX_TRY(dev_alloc(ctx, dpre, (size_t)(N + 1) * 4));
    GBX_TRY(dev_alloc(ctx, dgrp, (size_t)G * 8 + 8));
    GBX_TRY(fill_u32(ctx, dpre.as<uint32_t>(), 0, (size_t)std::min(H, N) + 1));
    if (N > H) {
        const uint32_t *cnt = nullptr;
        if (p->direction == GBX_PULL_IN) {
            // the columns are the rows' sources: a column's count is its
            // source's out-degree, which the plan holds
            cnt = p->outdeg.as<uint32_t>();
        } else {
            GBX_TRY(dev_alloc(ctx, dcnt, (size_t)N * 4 + 4));
            GBX_TRY(fill_u32(ctx, dcnt.as<uint32_t>(), 0, N));
            if (E) {
                column_counts<<<(unsigned)std::max(1, ctx->num_cus) * 2u, 1024, 0, ctx->stream>>>(
                    p->col.as<uint32_t>(), E, dcnt.as<uint32_t>());
                GBX_HIP(ctx, hipGetLastError());
            }
            cnt = dcnt.as<uint32_t>();
        }
        GBX_TRY(exclusive_scan_u32(ctx, cnt + H, dpre.as<uint32_t>() + H, N - H, dpre.as<uint32_t>() + N));
    }
    if (G) {
        group_prefix<<<grid_for(G, 256, 16384), 256, 0, ctx->stream>>>(dpre.as<uint32_t>(), N, G, dgrp.as<uint32_t>());
        GBX_HIP(ctx, hipGetLastError());
    }
    // destination bins, on the device (bin_next, bin_chase): <= SB rows and
    // <= EB in-edges each (a longer row is a bin of its own); a batch of
    // graphs also cuts at every graph's first row
    DevBuf d_nxt, d_vb, d_goff, d_nb;
    GBX_TRY(dev_alloc(ctx, d_nxt, (size_t)N * 4 + 4));
    GBX_TRY(dev_alloc(ctx, d_vb, ((size_t)N + 2) * 4));
    GBX_TRY(dev_alloc(ctx, d_nb, 16));
    const uint32_t ngoff = goff ? (uint32_t)goff->size() : 0u;
    if (ngoff) {
        GBX_TRY(dev_alloc(ctx, d_goff, (size_t)ngoff * 4));
        GBX_HIP(ctx, hipMemcpyAsync(d_goff.p, goff->data(), (size_t)ngoff * 4, hipMemcpyHostToDevice, ctx->stream));
    }
    if (N) {
        bin_next<<<grid_for(N, 256, 16384), 256, 0, ctx->stream>>>(p->rp.as<uint32_t>(), N, SB, (uint64_t)EB,
                                                                    ngoff ? d_goff.as<uint32_t>() : nullptr, ngoff,
                                                                    d_nxt.as<uint32_t>());
        GBX_HIP(ctx, hipGetLastError());
    }
    bin_chase<<<1, 64, 0, ctx->stream>>>(d_nxt.as<uint32_t>(), N, d_vb.as<uint32_t>(), d_nb.as<uint32_t>());
    GBX_HIP(ctx, hipGetLastError());
    // cold source chunks, on the device (chunk_next, chunk_chase): <= SC
    // sources and about EC out-edges each, from H on, cut only at multiples
    // of 8 sources (16-byte aligned contributions)
    DevBuf d_cnxt, d_cstart;
    GBX_TRY(dev_alloc(ctx, d_cnxt, ((size_t)G + 1) * 4));
    GBX_TRY(dev_alloc(ctx, bp->chunk_lo, ((size_t)G + 3) * 4));
    GBX_TRY(dev_alloc(ctx, d_cstart, ((size_t)G + 3) * 4));
    chunk_next<<<grid_for((size_t)G + 1, 256, 16384), 256, 0, ctx->stream>>>(dgrp.as<uint32_t>(), G, H, SC, (uint64_t)EC,
                                                                              d_cnxt.as<uint32_t>());
    chunk_chase<<<1, 64, 0, ctx->stream>>>(d_cnxt.as<uint32_t>(), dgrp.as<uint32_t>(), dpre.as<uint32_t>(), G, H, N,
                                           bp->chunk_lo.as<uint32_t>(), d_cstart.as<uint32_t>(),
                                           d_nb.as<uint32_t>() + 1);
    GBX_HIP(ctx, hipGetLastError());
    uint32_t Ecold = 0, counts[2] = {0, 0};  // bins, chunks
    GBX_HIP(ctx, hipMemcpyAsync(&Ecold, dpre.as<uint32_t>() + N, 4, hipMemcpyDeviceToHost, ctx->stream));
    GBX_HIP(ctx, hipMemcpyAsync(counts, d_nb.p, 8, hipMemcpyDeviceToHost, ctx->stream));
    GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const uint32_t NB = counts[0], NC = counts[1];
    if (NB > N + 1u || NC > G + 1u)
        return set_error(ctx, GBX_ERR_STATE, "binned layout: %u bins for %u rows, %u chunks", NB, N, NC);
    std::vector<uint32_t> vb((size_t)NB + 1);
    GBX_HIP(ctx, hipMemcpyAsync(vb.data(), d_vb.p, ((size_t)NB + 1) * 4, hipMemcpyDeviceToHost, ctx->stream));
    const uint64_t Ehot = (uint64_t)E - Ecold;
    if ((uint64_t)NB * std::max(NC, 1u) > (1ull << 28))
        return set_error(ctx, GBX_ERR_RANGE, "binned layout: %u bins x %u chunks too many", NB, NC);
    bp->NC = NC;
    bp->NB = NB;
    int rc = GBX_OK;
    DevBuf d_bin_of, d_chunk_of, d_bin_len, d_rp_bins, d_hot_start;
    auto up = [&](DevBuf &d, const void *h, size_t bytes) {
        if (rc) return;
        rc = dev_alloc(ctx, d, bytes + 16);
        if (!rc && bytes && hipMemcpyAsync(d.p, h, bytes, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
            rc = set_error(ctx, GBX_ERR_HIP, "binned layout: upload");
    };
    // the bin of every row and the chunk of every source (0 for the hot
    // ones), from the cuts; each bin's first edge
    GBX_TRY(dev_alloc(ctx, d_bin_of, (size_t)N * 4 + 16));
    GBX_TRY(dev_alloc(ctx, d_chunk_of, (size_t)N * 4 + 16));
    if (N) {
        range_index<<<grid_for(N, 256, 16384), 256, 0, ctx->stream>>>(d_vb.as<uint32_t>(), NB, N,
                                                                       d_bin_of.as<uint32_t>());
        range_index<<<grid_for(N, 256, 16384), 256, 0, ctx->stream>>>(bp->chunk_lo.as<uint32_t>(), NC, N,
                                                                       d_chunk_of.as<uint32_t>());
        GBX_HIP(ctx, hipGetLastError());
    }
    GBX_TRY(dev_alloc(ctx, d_rp_bins, ((size_t)NB + 1) * 4 + 16));
    gather_u32<<<grid_for(NB + 1, 256, 1024), 256, 0, ctx->stream>>>(p->rp.as<uint32_t>(), d_vb.as<uint32_t>(), NB + 1,
                                                                      d_rp_bins.as<uint32_t>());
    GBX_HIP(ctx, hipGetLastError());
    // the bins' first rows: the first NB entries of vb
    const DevBuf &d_bin_row0 = d_vb;
    GBX_TRY(dev_alloc(ctx, d_bin_len, (size_t)NB * 4 + 4));
    GBX_TRY(fill_u32(ctx, d_bin_len.as<uint32_t>(), 0, NB));
    GBX_TRY(dev_alloc(ctx, bp->chunk_grp, (size_t)(NC + 1) * 4));
    std::vector<uint32_t> hot_start(NB + 1, 0);
    DevBuf rowid;
    if (E) {
        GBX_TRY(dev_alloc(ctx, rowid, (size_t)E * 4));
        // rows of at least 64 edges at the front of the plan order
        uint32_t nlong = N;
        {
            DevBuf d_first;
            GBX_TRY(dev_alloc(ctx, d_first, 16));
            GBX_TRY(fill_u32(ctx, d_first.as<uint32_t>(), N, 1));
            first_short_row<<<grid_for(N, 256, 4096), 256, 0, ctx->stream>>>(p->rp.as<uint32_t>(), N, 64u,
                                                                             d_first.as<uint32_t>());
            GBX_HIP(ctx, hipGetLastError());
            GBX_HIP(ctx, hipMemcpyAsync(&nlong, d_first.p, 4, hipMemcpyDeviceToHost, ctx->stream));
            GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        }
        if (nlong)
            edge_rows_long<<<grid_for(nlong, 4, 16384), 256, 0, ctx->stream>>>(p->rp.as<uint32_t>(), nlong,
                                                                            rowid.as<uint32_t>());
        edge_rows<<<grid_for(N - nlong, 256, 16384), 256, 0, ctx->stream>>>(p->rp.as<uint32_t>(), nlong, N,
                                                                            rowid.as<uint32_t>());
        GBX_HIP(ctx, hipGetLastError());
    }
    // hot edges: per bin, in plan edge order (rows ascending), padded to 8
    DevBuf hflag, hrank, rowid_c, col_c, eidx_c;
    if (E) {
        GBX_TRY(dev_alloc(ctx, hflag, (size_t)E * 4 + 4));
        GBX_TRY(dev_alloc(ctx, hrank, (size_t)(E + 1) * 4));
        DevBuf htot;
        GBX_TRY(dev_alloc(ctx, htot, 16));
        hot_flags<<<grid_for(E, 256, 16384), 256, 0, ctx->stream>>>(p->col.as<uint32_t>(), E, H, hflag.as<uint32_t>());
        GBX_HIP(ctx, hipGetLastError());
        GBX_TRY(exclusive_scan_u32(ctx, hflag.as<uint32_t>(), hrank.as<uint32_t>(), E, htot.as<uint32_t>()));
        GBX_HIP(ctx, hipMemcpyAsync(hrank.as<uint32_t>() + E, htot.p, 4, hipMemcpyDeviceToDevice, ctx->stream));
        // hot edges before each bin's first edge
        DevBuf d_hot_before;
        GBX_TRY(dev_alloc(ctx, d_hot_before, (size_t)(NB + 1) * 4));
        gather_u32<<<grid_for(NB + 1, 256, 1024), 256, 0, ctx->stream>>>(hrank.as<uint32_t>(), d_rp_bins.as<uint32_t>(),
                                                                          NB + 1, d_hot_before.as<uint32_t>());
        GBX_HIP(ctx, hipGetLastError());
        std::vector<uint32_t> hot_before(NB + 1);
        GBX_HIP(ctx, hipMemcpyAsync(hot_before.data(), d_hot_before.p, (size_t)(NB + 1) * 4, hipMemcpyDeviceToHost,
                                    ctx->stream));
        GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if ((uint64_t)hot_before[NB] != Ehot) return set_error(ctx, GBX_ERR_STATE, "binned layout: hot edge count");
        // packed pairs: fillers per pair, their running count and each bin's packed length
        DevBuf rowid_h, col_h, eidx_h, hn, hp;
        std::vector<uint32_t> packed_before(NB + 1);
        const uint32_t Eh = (uint32_t)Ehot;
        if (hot16) {
            DevBuf hptot, d_pb;
            GBX_TRY(dev_alloc(ctx, rowid_h, (size_t)Eh * 4 + 4));
            GBX_TRY(dev_alloc(ctx, col_h, (size_t)Eh * 4 + 4));
            GBX_TRY(dev_alloc(ctx, eidx_h, (size_t)Eh * 4 + 4));
            GBX_TRY(dev_alloc(ctx, hn, (size_t)Eh * 4 + 4));
            GBX_TRY(dev_alloc(ctx, hp, (size_t)(Eh + 1) * 4));
            GBX_TRY(dev_alloc(ctx, hptot, 16));
            GBX_TRY(dev_alloc(ctx, d_pb, (size_t)(NB + 1) * 4));
            hot_compact<<<grid_for(E, 256, 16384), 256, 0, ctx->stream>>>(
                p->col.as<uint32_t>(), rowid.as<uint32_t>(), E, hflag.as<uint32_t>(), hrank.as<uint32_t>(),
                rowid_h.as<uint32_t>(), col_h.as<uint32_t>(), eidx_h.as<uint32_t>());
            hot_counts<<<grid_for(Eh, 256, 16384), 256, 0, ctx->stream>>>(
                rowid_h.as<uint32_t>(), Eh, d_bin_of.as<uint32_t>(), d_hot_before.as<uint32_t>(), hn.as<uint32_t>());
            GBX_HIP(ctx, hipGetLastError());
            GBX_TRY(exclusive_scan_u32(ctx, hn.as<uint32_t>(), hp.as<uint32_t>(), Eh, hptot.as<uint32_t>()));
            GBX_HIP(ctx, hipMemcpyAsync(hp.as<uint32_t>() + Eh, hptot.p, 4, hipMemcpyDeviceToDevice, ctx->stream));
            gather_u32<<<grid_for(NB + 1, 256, 1024), 256, 0, ctx->stream>>>(hp.as<uint32_t>(), d_hot_before.as<uint32_t>(),
                                                                              NB + 1, d_pb.as<uint32_t>());
            GBX_HIP(ctx, hipGetLastError());
            GBX_HIP(ctx, hipMemcpyAsync(packed_before.data(), d_pb.p, (size_t)(NB + 1) * 4, hipMemcpyDeviceToHost,
                                        ctx->stream));
            GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        } else {
            packed_before = hot_before;
        }
        uint64_t hpos = 0;
        for (uint32_t b = 0; b < NB; b++) {
            hot_start[b] = (uint32_t)hpos;
            hpos += ((uint64_t)(packed_before[b + 1] - packed_before[b]) + 7u) & ~7ull;
        }
        hot_start[NB] = (uint32_t)hpos;
        if (hpos >= (1ull << 32)) return set_error(ctx, GBX_ERR_RANGE, "binned layout: hot pairs too many");
        bp->Hlen = (uint32_t)hpos;
        if (hot16) {
            GBX_TRY(dev_alloc(ctx, bp->hot_pairs, (size_t)bp->Hlen * 2 + 16));
            GBX_TRY(fill_u32(ctx, bp->hot_pairs.as<uint32_t>(), kHotPad | (kHotPad << 16), bp->Hlen / 2 + 4));
            GBX_TRY(dev_alloc(ctx, bp->hot_base, (size_t)(bp->Hlen / 8) * 2 + 16));
            GBX_HIP(ctx, hipMemsetAsync(bp->hot_base.p, 0, (size_t)(bp->Hlen / 8) * 2 + 16, ctx->stream));
            GBX_TRY(dev_alloc(ctx, bp->hot_eidx, (size_t)bp->Hlen * 4 + 16));
            GBX_TRY(fill_u32(ctx, bp->hot_eidx.as<uint32_t>(), kNoEdge, bp->Hlen + 4));
        } else {
            GBX_TRY(dev_alloc(ctx, bp->hot_pairs, (size_t)bp->Hlen * 4 + 16));
            GBX_TRY(fill_u32(ctx, bp->hot_pairs.as<uint32_t>(), 0xFFFF0000u, bp->Hlen));
            GBX_TRY(dev_alloc(ctx, bp->hot_base, 16));
        }
        up(d_hot_start, hot_start.data(), (size_t)(NB + 1) * 4);
        GBX_TRY(rc);
        if (hot16 && Eh) {
            hot_fill16<<<grid_for(Eh, 256, 16384), 256, 0, ctx->stream>>>(
                rowid_h.as<uint32_t>(), col_h.as<uint32_t>(), Eh, hn.as<uint32_t>(), hp.as<uint32_t>(),
                d_bin_of.as<uint32_t>(), d_hot_before.as<uint32_t>(), d_hot_start.as<uint32_t>(),
                d_bin_row0.as<uint32_t>(), eidx_h.as<uint32_t>(), bp->hot_pairs.as<uint16_t>(),
                bp->hot_base.as<uint16_t>(), bp->hot_eidx.as<uint32_t>());
            GBX_HIP(ctx, hipGetLastError());
        }
        GBX_TRY(dev_alloc(ctx, rowid_c, (size_t)Ecold * 4 + 4));
        GBX_TRY(dev_alloc(ctx, col_c, (size_t)Ecold * 4 + 4));
        GBX_TRY(dev_alloc(ctx, eidx_c, (size_t)Ecold * 4 + 4));
        split_hot_cold<<<grid_for(E, 256, 16384), 256, 0, ctx->stream>>>(
            p->col.as<uint32_t>(), rowid.as<uint32_t>(), E, hflag.as<uint32_t>(), hrank.as<uint32_t>(),
            d_bin_of.as<uint32_t>(), d_bin_row0.as<uint32_t>(), d_hot_before.as<uint32_t>(),
            d_hot_start.as<uint32_t>(), hot16 ? nullptr : bp->hot_pairs.as<uint32_t>(), rowid_c.as<uint32_t>(),
            col_c.as<uint32_t>(), eidx_c.as<uint32_t>());
        GBX_HIP(ctx, hipGetLastError());
        GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    } else {
        GBX_TRY(dev_alloc(ctx, bp->hot_pairs, 16));
        GBX_TRY(dev_alloc(ctx, bp->hot_base, 16));
    }
    if (Ecold && runs) {
        GBX_TRY(build_runs(ctx, bp, Ecold, rowid_c.as<uint32_t>(), col_c.as<uint32_t>(), eidx_c.as<uint32_t>(),
                           d_bin_of.as<uint32_t>(), d_chunk_of.as<uint32_t>(), d_bin_row0.as<uint32_t>(),
                           d_cstart.as<uint32_t>(), d_bin_len.as<uint32_t>()));
    } else if (Ecold) {
        const uint32_t kbits = bits_for(NB * NC - 1);
        DevBuf key, pay, voff_table;
        GBX_TRY(dev_alloc(ctx, key, (size_t)Ecold * 4));
        GBX_TRY(dev_alloc(ctx, pay, (size_t)Ecold * 4));
        GBX_TRY(dev_alloc(ctx, voff_table, (size_t)NB * NC * 4));
        bin_keys<<<grid_for(Ecold, 256, 16384), 256, 0, ctx->stream>>>(
            rowid_c.as<uint32_t>(), col_c.as<uint32_t>(), Ecold, d_bin_of.as<uint32_t>(), d_chunk_of.as<uint32_t>(), NB,
            NC, 0, key.as<uint32_t>(), pay.as<uint32_t>());
        GBX_HIP(ctx, hipGetLastError());
        GBX_TRY(radix_sort_pairs(ctx, key.as<uint32_t>(), pay.as<uint32_t>(), nullptr, Ecold, kbits));
        // bin-major: V positions (segments padded to `pad` values)
        {
            Runs R;
            GBX_TRY(find_runs(ctx, key.as<uint32_t>(), Ecold, pad, R));
            bp->nseg = R.n;
            bp->Vlen = R.padded;
            GBX_TRY(dev_alloc(ctx, bp->dst_local, (size_t)bp->Vlen * 2 + 16));
            GBX_HIP(ctx, hipMemsetAsync(bp->dst_local.p, 0, (size_t)bp->Vlen * 2 + 16, ctx->stream));
            bin_major_fill<<<grid_for(Ecold, 256, 16384), 256, 0, ctx->stream>>>(
                key.as<uint32_t>(), pay.as<uint32_t>(), Ecold, R.flag.as<uint32_t>(), R.idx.as<uint32_t>(),
                R.start.as<uint32_t>(), R.pstart.as<uint32_t>(), R.plen.as<uint32_t>(), rowid_c.as<uint32_t>(),
                d_bin_row0.as<uint32_t>(), NB, NC, bp->dst_local.as<uint16_t>(), voff_table.as<uint32_t>(),
                d_bin_len.as<uint32_t>());
            GBX_HIP(ctx, hipGetLastError());
        }
        bin_keys<<<grid_for(Ecold, 256, 16384), 256, 0, ctx->stream>>>(
            rowid_c.as<uint32_t>(), col_c.as<uint32_t>(), Ecold, d_bin_of.as<uint32_t>(), d_chunk_of.as<uint32_t>(), NB,
            NC, 1, key.as<uint32_t>(), pay.as<uint32_t>());
        GBX_HIP(ctx, hipGetLastError());
        GBX_TRY(radix_sort_pairs(ctx, key.as<uint32_t>(), pay.as<uint32_t>(), nullptr, Ecold, kbits));
        // chunk-major: source indices and group destinations (same padding)
        {
            Runs R;
            GBX_TRY(find_runs(ctx, key.as<uint32_t>(), Ecold, pad, R));
            if (R.n != bp->nseg) return set_error(ctx, GBX_ERR_STATE, "binned layout: segment mismatch");
            bp->Plen = R.padded;
            GBX_TRY(dev_alloc(ctx, bp->src_local, (size_t)bp->Plen * 2 + 16));
            GBX_TRY(dev_alloc(ctx, bp->bdst, (size_t)(bp->Plen / 32) * 4 + 16));
            GBX_HIP(ctx, hipMemsetAsync(bp->src_local.p, 0xFF, (size_t)bp->Plen * 2 + 16, ctx->stream));
            chunk_major_fill<<<grid_for(Ecold, 256, 16384), 256, 0, ctx->stream>>>(
                key.as<uint32_t>(), pay.as<uint32_t>(), Ecold, R.flag.as<uint32_t>(), R.idx.as<uint32_t>(),
                R.start.as<uint32_t>(), R.pstart.as<uint32_t>(), col_c.as<uint32_t>(), bp->chunk_lo.as<uint32_t>(),
                NB, bp->src_local.as<uint16_t>());
            block_destinations<<<grid_for(R.n, 256, 16384), 256, 0, ctx->stream>>>(
                key.as<uint32_t>(), R.start.as<uint32_t>(), R.pstart.as<uint32_t>(), R.plen.as<uint32_t>(), R.n,
                voff_table.as<uint32_t>(), bp->bdst.as<uint32_t>());
            chunk_groups<<<grid_for(NC + 1, 256, 1024), 256, 0, ctx->stream>>>(
                d_cstart.as<uint32_t>(), NC, R.idx.as<uint32_t>(), Ecold, R.n, R.pstart.as<uint32_t>(), 3u,
                bp->chunk_grp.as<uint32_t>());
            GBX_HIP(ctx, hipGetLastError());
            GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        }
    } else {
        GBX_TRY(dev_alloc(ctx, bp->dst_local, 16));
        GBX_TRY(dev_alloc(ctx, bp->src_local, 16));
        GBX_TRY(dev_alloc(ctx, bp->bdst, 16));
        GBX_TRY(dev_alloc(ctx, bp->blk_rec, 16));
        GBX_TRY(dev_alloc(ctx, bp->seg_v, 16));
        GBX_TRY(fill_u32(ctx, bp->chunk_grp.as<uint32_t>(), 0, NC + 1));
    }
    // units: V range and hot-pair range of each bin (a single row longer than EB is split)
    std::vector<uint32_t> bin_len(NB);
    GBX_HIP(ctx, hipMemcpyAsync(bin_len.data(), d_bin_len.p, (size_t)NB * 4, hipMemcpyDeviceToHost, ctx->stream));
    GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<BinUnit> units;
    uint64_t vpos = 0;
    uint32_t n_multi = 0;
    size_t bg = 0;  // graph of the current bin (a batch of graphs)
    for (uint32_t b = 0; b < NB; b++) {
        if (goff)
            while (bg + 2 < goff->size() && (*goff)[bg + 1] <= vb[b]) bg++;
        const uint32_t len = bin_len[b], rows = vb[b + 1] - vb[b];
        const uint32_t h0 = hot_start[b], hlen = hot_start[b + 1] - hot_start[b];
        uint32_t nu = 1;
        if (rows == 1 && (uint64_t)len + hlen > EB) nu = (uint32_t)(((uint64_t)len + hlen + EB - 1) / EB);
        const uint32_t multi = nu > 1 ? n_multi++ : kNoMulti;
        // each unit takes an equal share (multiple of 8) of both ranges
        const uint32_t vstep = ((len + nu - 1) / nu + 7u) & ~7u, hstep = ((hlen + nu - 1) / nu + 7u) & ~7u;
        for (uint32_t q = 0; q < nu; q++) {
            BinUnit un{};
            un.row_begin = vb[b];
            un.row_end = vb[b + 1];
            un.vbeg = (uint32_t)(vpos + std::min<uint64_t>((uint64_t)q * vstep, len));
            un.vend = (uint32_t)(vpos + std::min<uint64_t>((uint64_t)(q + 1) * vstep, len));
            un.hbeg = h0 + (uint32_t)std::min<uint64_t>((uint64_t)q * hstep, hlen);
            un.hend = h0 + (uint32_t)std::min<uint64_t>((uint64_t)(q + 1) * hstep, hlen);
            un.multi = multi;
            un.nunits = nu;
            un.graph = (uint32_t)bg;
            units.push_back(un);
        }
        vpos += len;
    }
    if (vpos != bp->Vlen) return set_error(ctx, GBX_ERR_STATE, "binned layout: bin lengths do not add up");
    // the accumulate blocks take units in this order: the largest (bytes
    // moved) first, so the last units taken are the smallest
    const uint32_t vel = std::max(hot_elem, 1u);
    auto unit_bytes = [&](const BinUnit &un) {
        return (uint64_t)(un.vend - un.vbeg) * (vel + 2u) + (uint64_t)(un.hend - un.hbeg) * 3u +
               (uint64_t)(un.row_end - un.row_begin) * 16u;
    };
    std::stable_sort(units.begin(), units.end(),
                     [&](const BinUnit &a, const BinUnit &b) { return unit_bytes(a) > unit_bytes(b); });
    bp->NU = (uint32_t)units.size();
    bp->n_multi = n_multi;
    if (goff) {
        // the units of each graph, in ascending unit index (the order its partials are summed in)
        const uint32_t ng = (uint32_t)goff->size() - 1;
        std::vector<uint32_t> uptr(ng + 1, 0), uidx(units.size());
        for (const BinUnit &un : units) uptr[un.graph + 1]++;
        for (uint32_t k = 0; k < ng; k++) uptr[k + 1] += uptr[k];
        std::vector<uint32_t> fill(uptr.begin(), uptr.end() - 1);
        for (uint32_t w = 0; w < (uint32_t)units.size(); w++) uidx[fill[units[w].graph]++] = w;
        GBX_TRY(dev_alloc(ctx, bp->graph_uptr, (size_t)(ng + 1) * 4));
        GBX_TRY(dev_alloc(ctx, bp->graph_uidx, units.size() * 4 + 4));
        GBX_HIP(ctx, hipMemcpyAsync(bp->graph_uptr.p, uptr.data(), (size_t)(ng + 1) * 4, hipMemcpyHostToDevice,
                                    ctx->stream));
        if (!units.empty())
            GBX_HIP(ctx, hipMemcpyAsync(bp->graph_uidx.p, uidx.data(), units.size() * 4, hipMemcpyHostToDevice,
                                        ctx->stream));
    }
    GBX_TRY(dev_alloc(ctx, bp->units, units.size() * sizeof(BinUnit)));
    GBX_HIP(ctx, hipMemcpyAsync(bp->units.p, units.data(), units.size() * sizeof(BinUnit), hipMemcpyHostToDevice,
                                ctx->stream));
    GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (getenv("GBX_VERBOSE"))
        fprintf(stderr,
                "[gbx] binned layout: N=%u E=%u hot=%u (%.1f%% of edges) pad=%u chunks=%u bins=%u (<=%u rows) "
                "units=%u split_bins=%u segments=%u V=%u (%.3f per cold edge) hot_pairs=%u scatter_blocks=%u\n",
                N, E, H, E ? 100.0 * (double)Ehot / E : 0.0, pad, NC, NB, SB, bp->NU, n_multi, bp->nseg, bp->Vlen,
                Ecold ? (double)bp->Vlen / Ecold : 0.0, bp->Hlen, runs ? bp->NP : NC);
    return GBX_OK;
}

static int build_bplan(gbx_ctx *ctx, gbx_plan *p, const std::vector<uint32_t> *goff, const uint32_t (&cut)[8],
                       gbx_bplan **out)
{
    gbx_bplan *bp = new (std::nothrow) gbx_bplan();
    if (!bp) return set_error(ctx, GBX_ERR_NOMEM, "binned layout: host memory");
    memcpy(bp->cut, cut, sizeof(bp->cut));
    const int rc = build_bplan_into(ctx, p, cut[4], cut[5], cut[6], cut[7] != 0, goff, bp);
    if (rc) {
        delete bp;
        return rc;
    }
    *out = bp;
    return GBX_OK;
}

template <class K>
static void grant_lds(K *kernel, size_t shm)
{
    if (shm > 64 * 1024)
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)shm);
}

bool binned_supported(const gbx_pr *pr)
{
    // a graph with no edges has nothing to bin (the pull iteration runs it)
    if (pr->g->E == 0) return false;
    if (pr->kind == K_SEG) return pr->seg && pr->plan->segmented && pr->g->N && pr->elem == 4;
    if (pr->batch || pr->seg || pr->plan->segmented || pr->g->N == 0) return false;
    if (is_weighted_kind(pr->kind)) return pr->elem == 4;  // FP32 contributions and weights
    return pr->kind == K_PR_REF || pr->kind == K_PR_STOCH || is_int_kind(pr->kind);
}

// The plan's per-edge weights a weighted kind multiplies its values by.
static const float *kind_weights(const gbx_pr *pr)
{
    switch (pr->kind) {
    case K_PR_STOCH_W: return pr->plan->wn.as<float>();  // row-normalised transition weights
    case K_ADJ: return pr->plan->wr.as<float>();         // P itself, forward-plan rows
    case K_SPMV_W: return pr->plan->wt.as<float>();      // raw weights
    default: return nullptr;
    }
}

__global__ void gather_edge_weights(const uint32_t *__restrict__ eidx, uint32_t n, const float *__restrict__ w,
                                    float *__restrict__ out)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t e = eidx[i];
        out[i] = e == kNoEdge ? 0.0f : w[e];
    }
}

// The weights of a layout's src_local entries and hot pairs taken from the
// plan's weight array w (built once per layout and array).
static int layout_weights(gbx_ctx *ctx, gbx_bplan *bp, const float *w, const gbx_bplan::EdgeWeights **out)
{
    for (auto &lw : bp->weights)
        if (lw->src == w) {
            *out = lw.get();
            return GBX_OK;
        }
    if (!bp->runs || (bp->H && !bp->hot16))
        return set_error(ctx, GBX_ERR_STATE, "binned layout: weighted kinds need the run layout");
    auto lw = std::make_unique<gbx_bplan::EdgeWeights>();
    lw->src = w;
    GBX_TRY(dev_alloc(ctx, lw->cold, (size_t)bp->Plen * 4 + 64));
    GBX_TRY(dev_alloc(ctx, lw->hot, (size_t)bp->Hlen * 4 + 64));
    if (bp->Plen) {
        gather_edge_weights<<<grid_for(bp->Plen, 256, 16384), 256, 0, ctx->stream>>>(
            bp->cold_eidx.as<uint32_t>(), bp->Plen, w, lw->cold.as<float>());
        GBX_HIP(ctx, hipGetLastError());
    }
    if (bp->Hlen) {
        gather_edge_weights<<<grid_for(bp->Hlen, 256, 16384), 256, 0, ctx->stream>>>(
            bp->hot_eidx.as<uint32_t>(), bp->Hlen, w, lw->hot.as<float>());
        GBX_HIP(ctx, hipGetLastError());
    }
    GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *out = lw.get();
    bp->weights.push_back(std::move(lw));
    return GBX_OK;
}

// The scatter moves values without looking at them: one instance per width.
template <size_t W>
using UintOf = std::conditional_t<W == 1, uint8_t,
               std::conditional_t<W == 2, uint16_t, std::conditional_t<W == 4, uint32_t, uint64_t>>>;

// The scatter's instance: values moved as stored, or quantized on loading.
template <int KIND>
constexpr int scatter_qk() { return is_quant_kind(KIND) ? KIND : 0; }

template <int KIND, class T>
static void grant_kind(const gbx_binned_state *st)
{
    if constexpr (KIND == K_PR_REF || KIND == K_PR_STOCH || KIND == K_SEG)
        grant_lds(&bin_scatter_runs_kernel<T>, st->shm_scatter);
    if constexpr (is_weighted_kind(KIND)) grant_lds(&bin_scatter_runs_kernel<T, true>, st->shm_scatter);
    else grant_lds(&bin_scatter_kernel<UintOf<sizeof(T)>, scatter_qk<KIND>()>, st->shm_scatter);
    grant_lds(&bin_accum_kernel<KIND, T>, st->shm_accum);
}

// Bytes of one value in V and in the LDS copies: the contribution's own
// width, or the quantized width.
static size_t value_bytes(const gbx_pr *pr)
{
    switch (pr->kind) {
    case K_Q32: return 4;
    case K_Q16: return 2;
    case K_Q8: return 1;
    default: return pr->elem;
    }
}

int binned_create(gbx_pr *pr)
{
    gbx_ctx *ctx = pr->ctx;
    gbx_plan *p = pr->plan;
    const gbx_tuning &tu = ctx->tune;
    // The float kinds sum the cold edges of one chunk into one row in the
    // scatter (run layout; segments padded to 8; the weighted kinds always).
    // The integer kinds keep one V value per edge: segments padded to 32
    // values (a multiple of 32 is required by the per-block V positions),
    // whole 128-byte lines of FP32 V, 64-byte halves of 16-bit values
    // (measured as fast as whole lines).
    const bool runs = is_weighted_kind(pr->kind) || pr->kind == K_SEG ||
                      ((pr->kind == K_PR_REF || pr->kind == K_PR_STOCH) && tu.binned_edge_values == 0);
    const uint32_t pad = runs ? 8 : 32;
    // hot sources: the tuning value for contributions of up to 4 bytes (one
    // layout for FP32, BF16 and the 32/16/8-bit fixed-point kinds), half as
    // many 8-byte ones in the same LDS; a multiple of 8, at most N and 65528
    // (u16 source ids, 0xFFFF marks pads)
    const size_t vel = value_bytes(pr);
    const uint32_t hot_elem = std::max<uint32_t>((uint32_t)vel, 4u);
    const uint32_t H = std::min(std::min(tu.binned_hot_sources * 4u / hot_elem, runs ? kHot16Max : 65528u), p->N) &
                       ~7u;
    const uint32_t want[8] = {cut_edges_of(tu.binned_chunk_edges, ctx, p->E, runs, true),
                              (std::max(cut_edges_of(tu.binned_bin_edges, ctx, p->E, runs, false), 8u) + 7u) & ~7u,
                              std::min(tu.binned_bin_rows, kMaxBinRows),
                              chunk_sources_of(tu, hot_elem, runs),
                              pad,
                              H,
                              hot_elem,
                              runs ? 1u : 0u};
    std::shared_ptr<gbx_bplan> bp;
    for (auto &c : p->bplans)
        if (memcmp(c->cut, want, sizeof(want)) == 0) bp = c;
    if (!bp) {
        gbx_bplan *raw = nullptr;
        GBX_TRY(build_bplan(ctx, p, p->segmented ? &pr->g->seg_off : nullptr, want, &raw));
        bp = std::shared_ptr<gbx_bplan>(raw);
        p->bplans.push_back(bp);
    }
    gbx_binned_state *st = new (std::nothrow) gbx_binned_state();
    if (!st) return set_error(ctx, GBX_ERR_NOMEM, "binned state: host memory");
    pr->bin = st;
    st->bp = bp;
    if (is_weighted_kind(pr->kind)) {
        const float *w = kind_weights(pr);
        if (!w) return set_error(ctx, GBX_ERR_STATE, "binned: the plan has no weights for this kind");
        const gbx_bplan::EdgeWeights *lw = nullptr;
        GBX_TRY(layout_weights(ctx, bp.get(), w, &lw));
        st->wcold = lw->cold.as<float>();
        st->whot = lw->hot.as<float>();
    }
    // run layout: the dummy areas of the scatter's idle store lanes follow V
    const size_t vbytes = ((size_t)bp->Vlen * vel + 64 + kRunDummyStride - 1) / kRunDummyStride * kRunDummyStride;
    if (bp->runs && vbytes + (size_t)kRunDummyAreas * kRunDummyStride > 0xFFFFFFFFull)
        return set_error(ctx, GBX_ERR_RANGE, "binned layout: V of %zu bytes beyond 32-bit offsets", vbytes);
    st->dummy_off = (uint32_t)vbytes;
    GBX_TRY(dev_alloc(ctx, st->V, vbytes + (bp->runs ? (size_t)kRunDummyAreas * kRunDummyStride : 0)));
    // the run layout's pads are never written: zero once
    if (bp->runs) GBX_HIP(ctx, hipMemsetAsync(st->V.p, 0, vbytes, ctx->stream));
    GBX_TRY(dev_alloc(ctx, st->acc, (size_t)bp->n_multi * 8 + 8));
    GBX_TRY(dev_alloc(ctx, st->cnt, (size_t)bp->n_multi * 4 + 4));
    GBX_HIP(ctx, hipMemsetAsync(st->acc.p, 0, (size_t)bp->n_multi * 8 + 8, ctx->stream));
    GBX_HIP(ctx, hipMemsetAsync(st->cnt.p, 0, (size_t)bp->n_multi * 4 + 4, ctx->stream));
    GBX_TRY(dev_alloc(ctx, st->work, 16));
    GBX_HIP(ctx, hipMemsetAsync(st->work.p, 0, 16, ctx->stream));
    // one accumulate block per CU (the LDS holds one), never more than the units
    st->accum_grid = std::max(1u, std::min(bp->NU, (uint32_t)std::max(1, ctx->num_cus)));
    if (bp->NU > 16384) GBX_TRY(dev_alloc(ctx, pr->partials, (size_t)bp->NU * sizeof(Partial)));
    GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    // run layout: the chunk's contributions in FP32 (BF16 widened), a zero at
    // index SC, then the per-wave staging of run sums and V positions
    st->stage_off = bp->runs ? (uint32_t)((((size_t)bp->SC + 1) * 4u + 15) & ~(size_t)15) : 0;
    st->shm_scatter = bp->runs ? st->stage_off + (size_t)(kRunsBlock / 64) * kRunStage * sizeof(float)
                               : (size_t)bp->cut[3] * vel;
    st->shm_accum = (size_t)bp->SB * 8 + (size_t)(bp->hot16 ? kHot16Slots : bp->H) * vel;
    switch (pr->kind) {
    case K_PR_REF: pr->elem == 4 ? grant_kind<K_PR_REF, float>(st) : grant_kind<K_PR_REF, uint16_t>(st); break;
    case K_PR_STOCH: pr->elem == 4 ? grant_kind<K_PR_STOCH, float>(st) : grant_kind<K_PR_STOCH, uint16_t>(st); break;
    case K_PR_STOCH_W: grant_kind<K_PR_STOCH_W, float>(st); break;
    case K_SEG: grant_kind<K_SEG, float>(st); break;
    case K_ADJ: grant_kind<K_ADJ, float>(st); break;
    case K_SPMV_W: grant_kind<K_SPMV_W, float>(st); break;
    case K_FIX64: grant_kind<K_FIX64, uint64_t>(st); break;
    case K_FIX32: grant_kind<K_FIX32, uint32_t>(st); break;
    case K_FIX16: grant_kind<K_FIX16, uint16_t>(st); break;
    case K_FIX8: grant_kind<K_FIX8, uint8_t>(st); break;
    case K_Q32: grant_kind<K_Q32, uint32_t>(st); break;
    case K_Q16: grant_kind<K_Q16, uint16_t>(st); break;
    default: grant_kind<K_Q8, uint8_t>(st); break;
    }
    return GBX_OK;
}

template <int KIND, class T>
static void launch_binned(gbx_pr *pr, const IterView &A, const GCoef *coef = nullptr, double *uslots = nullptr)
{
    using U = UintOf<sizeof(T)>;
    const gbx_binned_state *st = pr->bin;
    const gbx_bplan *bp = st->bp.get();
    hipStream_t stream = pr->ctx->stream;
    ScatterView S;
    S.chunk_lo = bp->chunk_lo.as<uint32_t>();
    S.chunk_grp = bp->chunk_grp.as<uint32_t>();
    S.src_local = bp->src_local.as<uint16_t>();
    S.bdst = bp->bdst.as<uint32_t>();
    S.NC = bp->NC;
    if constexpr (KIND == K_PR_REF || KIND == K_PR_STOCH || KIND == K_SEG || is_weighted_kind(KIND)) {
        if (bp->runs) {
            RunsView RV;
            RV.chunk_lo = S.chunk_lo;
            RV.chunk_grp = S.chunk_grp;
            RV.pieces = bp->pieces.as<uint4>();
            RV.src_local = S.src_local;
            RV.blk_rec = bp->blk_rec.as<uint4>();
            RV.seg_v = bp->seg_v.as<uint32_t>();
            RV.nseg = bp->nseg;
            RV.pad = bp->SC;
            RV.stage_off = st->stage_off;
            RV.dummy = st->dummy_off;
            RV.wcold = st->wcold;
            if (bp->NP)
                bin_scatter_runs_kernel<T, is_weighted_kind(KIND)><<<bp->NP, kRunsBlock, st->shm_scatter, stream>>>(
                    RV, A.slots, A.t, reinterpret_cast<const T *>(A.c_in), st->V.as<T>());
            S.NC = 0;
        }
    }
    if constexpr (!is_weighted_kind(KIND)) {
        if (S.NC)
            bin_scatter_kernel<U, scatter_qk<KIND>()><<<bp->NC, kBlock, st->shm_scatter, stream>>>(
                S, A.slots, A.t, A.c_in, st->V.as<U>());
    }
    AccumView B;
    B.units = bp->units.as<BinUnit>();
    B.dst_local = bp->dst_local.as<uint16_t>();
    B.hot_pairs = bp->hot_pairs.as<uint32_t>();
    B.hot_base = bp->hot16 ? bp->hot_base.as<uint16_t>() : nullptr;
    B.H = bp->H;
    B.hot_off = bp->SB * 8u;
    B.acc = st->acc.as<long long>();
    B.cnt = st->cnt.as<uint32_t>();
    B.work = st->work.as<uint32_t>();
    B.NU = bp->NU;
    B.whot = st->whot;
    B.coef = coef;
    B.uslots = uslots;
    bin_accum_kernel<KIND, T><<<st->accum_grid, kBlock, st->shm_accum, stream>>>(
        B, pr->plan->outdeg.as<uint32_t>(), pr->plan->N, st->V.as<T>(), A);
}

int binned_launch_iter(gbx_pr *pr)
{
    IterView A;
    A.r = pr->r.as<float>();
    A.c_in = pr->c[pr->t & 1].p;
    A.c_out = pr->c[(pr->t + 1) & 1].p;
    A.tele = pr->has_tele ? pr->tele.as<float>() : nullptr;
    A.slots = pr->slots.as<Slot>();
    A.partials = pr->partials.as<Partial>();
    A.counter = pr->counter.as<uint32_t>();
    A.t = pr->t;
    A.hot = 0;
    A.damp = pr->p.damp;
    A.eps = pr->p.epsilon;
    A.stop = pr->p.stop;
    switch (pr->kind) {
    case K_PR_REF:
        if (pr->elem == 4) launch_binned<K_PR_REF, float>(pr, A);
        else launch_binned<K_PR_REF, uint16_t>(pr, A);
        break;
    case K_PR_STOCH:
        if (pr->elem == 4) launch_binned<K_PR_STOCH, float>(pr, A);
        else launch_binned<K_PR_STOCH, uint16_t>(pr, A);
        break;
    case K_PR_STOCH_W: launch_binned<K_PR_STOCH_W, float>(pr, A); break;
    case K_ADJ: launch_binned<K_ADJ, float>(pr, A); break;
    case K_SPMV_W:
        // the input vector stays in c[0]: every pass reads it
        A.c_in = pr->c[0].p;
        A.c_out = pr->c[1].p;
        launch_binned<K_SPMV_W, float>(pr, A);
        break;
    case K_FIX64: launch_binned<K_FIX64, uint64_t>(pr, A); break;
    case K_FIX32: launch_binned<K_FIX32, uint32_t>(pr, A); break;
    case K_FIX16: launch_binned<K_FIX16, uint16_t>(pr, A); break;
    case K_FIX8: launch_binned<K_FIX8, uint8_t>(pr, A); break;
    case K_Q32: launch_binned<K_Q32, uint32_t>(pr, A); break;
    case K_Q16: launch_binned<K_Q16, uint16_t>(pr, A); break;
    default: launch_binned<K_Q8, uint8_t>(pr, A); break;
    }
    GBX_HIP(pr->ctx, hipGetLastError());
    pr->t++;
    return GBX_OK;
}

int binned_seg_iter(gbx_pr *pr, float *r, const float *c_in, float *c_out, const GCoef *coef, double *uslots)
{
    IterView A;
    A.r = r;
    A.c_in = c_in;
    A.c_out = c_out;
    A.tele = nullptr;
    A.slots = pr->slots.as<Slot>();
    A.partials = nullptr;
    A.counter = pr->counter.as<uint32_t>();
    A.t = pr->t;
    A.hot = 0;
    A.damp = pr->p.damp;
    A.eps = pr->p.epsilon;
    A.stop = pr->p.stop;
    launch_binned<K_SEG, float>(pr, A, coef, uslots);
    GBX_HIP(pr->ctx, hipGetLastError());
    return GBX_OK;
}

uint32_t binned_units(const gbx_pr *pr) { return pr->bin ? pr->bin->bp->NU : 0u; }

void binned_graph_units(const gbx_pr *pr, const uint32_t **uptr, const uint32_t **uidx)
{
    *uptr = pr->bin->bp->graph_uptr.as<uint32_t>();
    *uidx = pr->bin->bp->graph_uidx.as<uint32_t>();
}

void binned_state_free(gbx_binned_state *s) { delete s; }

void binned_drop_unused(gbx_plan *p)
{
    auto &v = p->bplans;
    v.erase(std::remove_if(v.begin(), v.end(), [](const std::shared_ptr<gbx_bplan> &b) { return b.use_count() == 1; }),
            v.end());
}

}  // namespace gbx

gbx_plan::~gbx_plan() = default;
