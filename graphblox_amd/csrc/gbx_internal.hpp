// gbx_internal.hpp -- shared internals of libgbx_hip.so (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/gbx.h"

namespace gbx {

constexpr uint32_t kIdMask = 0x3FFFFFFFu;  // EXTRACT_VALUE, reorder.h:23,30
constexpr uint32_t kMaxVertices = 1u << 30;

// Error carrier: every failing step records a message in the context.
struct Status {
    int code = GBX_OK;
    std::string msg;
};

int set_error(gbx_ctx *ctx, int code, const char *fmt, ...);

#define GBX_HIP(ctx, expr)                                                          \
    do {                                                                            \
        hipError_t _e = (expr);                                                     \
        if (_e != hipSuccess)                                                       \
            return ::gbx::set_error((ctx), GBX_ERR_HIP, "%s: %s (%s:%d)", #expr,    \
                                    hipGetErrorString(_e), __FILE__, __LINE__);     \
    } while (0)

#define GBX_TRY(expr)                \
    do {                             \
        int _rc = (expr);            \
        if (_rc != GBX_OK) return _rc; \
    } while (0)

// Owning device buffer (hipMalloc / hipFree), move-only.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

int dev_alloc(gbx_ctx *ctx, DevBuf &b, size_t bytes);
// Grow-only scratch (keeps the larger of the current and requested size).
int dev_reserve(gbx_ctx *ctx, DevBuf &b, size_t bytes);

}  // namespace gbx

struct gbx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cus = 0;
    std::string err;
    gbx_tuning tune;
    // scratch reused across calls (grow-only)
    gbx::DevBuf scan_tmp, sort_hist, sort_tmp[3], misc;
    // pinned staging for copies into the caller's host memory (grow-only)
    void *h_stage = nullptr;
    size_t h_stage_bytes = 0;
};

// One CSR in the reference's layout, device resident.
struct gbx_csr_dev {
    gbx::DevBuf out_degree, in_degree, edges_idx;  // [N]
    gbx::DevBuf src, dst, w;                       // [E]
};

struct gbx_plan;  // degree-binned, relabelled pull plan (gbx_pr.hip)

struct gbx_graph {
    gbx_ctx *ctx = nullptr;
    uint32_t N = 0, E = 0;
    bool weighted = false;
    gbx_csr_dev fwd, inv;
    // per direction; PageRank objects hold a reference, so a plan rebuilt
    // after a tuning change stays valid for the objects made before it
    std::shared_ptr<gbx_plan> plan[2];
    // a batch of independent graphs stored block-diagonally: graph k owns
    // vertices [seg_off[k], seg_off[k+1]); empty for a single graph
    std::vector<uint32_t> seg_off;
    std::vector<uint32_t> seg_edge_off;
    gbx::DevBuf seg_gid;  // u32[N]: graph of each vertex
};

namespace gbx {

// ---- primitives (gbx_core.hip)
// Exclusive scan of n uint32 values (in may equal out); *total (nullable,
// device) receives the sum.  Uses ctx scratch.
int exclusive_scan_u32(gbx_ctx *ctx, const uint32_t *in, uint32_t *out, size_t n,
                       uint32_t *d_total);
int fill_u32(gbx_ctx *ctx, uint32_t *p, uint32_t v, size_t n);
int max_u32(gbx_ctx *ctx, const uint32_t *a, size_t n, uint32_t *h_max);
// Device -> host copy on ctx->stream into the caller's memory, which is
// often pageable; returns when the data is there.
int copy_to_host(gbx_ctx *ctx, void *dst, const void *src, size_t bytes);

// ---- sort / CSR (gbx_sort.hip)
// Stable LSD radix sort of (key, pay[, w]) by the low `bits` bits of key.
int radix_sort_pairs(gbx_ctx *ctx, uint32_t *key, uint32_t *pay, float *w, size_t n,
                     uint32_t bits);
// Build one CSR (sorted arrays + vertex arrays) from device edge arrays that
// are consumed (reordered in place).
int build_csr_device(gbx_ctx *ctx, uint32_t *src, uint32_t *dst, float *w, uint32_t E,
                     uint32_t N, int sort_mode, bool move_weights, uint32_t *out_degree,
                     uint32_t *edges_idx);
int graph_from_device_edges(gbx_ctx *ctx, DevBuf &src, DevBuf &dst, DevBuf &w, uint32_t E,
                            uint32_t N, int sort_mode, unsigned flags, gbx_graph **out,
                            const gbx_relabel *relabel = nullptr, uint32_t *labels_out = nullptr);
// Ids below 2^30 (the column mask) and below N; N = 0 becomes max id + 1
// (edgeList.c:679).
int check_ids(gbx_ctx *ctx, const DevBuf &ds, const DevBuf &dd, uint32_t E, uint32_t *N);
// ---- relabel levels (gbx_relabel.hip)
bool relabel_requested(const gbx_relabel *rl);
int relabel_validate(gbx_ctx *ctx, const gbx_relabel *rl);
// Renames the (sorted) edges level by level, sorting again after each; total
// [N] receives the composed new id of every original vertex.
int relabel_edges(gbx_ctx *ctx, DevBuf &src, DevBuf &dst, DevBuf &w, uint32_t E, uint32_t N, uint32_t avg,
                  int sort_mode, bool weighted, uint32_t *out_degree, uint32_t *edges_idx,
                  const gbx_relabel *rl, DevBuf &total);
uint32_t bits_for(uint32_t max_value);

// ---- tuning (gbx_core.hip)
void tuning_defaults(gbx_tuning *t);
bool tuning_valid(const gbx_tuning *t);
// The pull plan's hub threshold (rows of at least this degree are cut into
// chunks of 4x as many edges) when the tuning leaves it 0, for a graph whose
// largest member has E edges.  Measured with the pull kernel on R-MAT (us per
// iteration, threshold 2048 / 4096 / 8192): s16 22 / 26 / 31, s18 34 / 41 / 58,
// s20 79 / 80 / 88, s22 305 / 280 / 305, s24 1694 / 1320 / 1302.
inline uint32_t auto_hub_threshold(uint64_t E)
{
    return E < (1ull << 25) ? 2048u : E < (1ull << 27) ? 4096u : 8192u;
}
// The graph's edge count that sets the automatic cuts: its largest member's
// for a batch of graphs.
uint64_t cut_edges(const gbx_graph *g);


inline unsigned grid_for(size_t n, unsigned block, unsigned cap = 1u << 20) {
    size_t g = (n + block - 1) / block;
    if (g == 0) g = 1;
    return (unsigned)(g < cap ? g : cap);
}

}  // namespace gbx
