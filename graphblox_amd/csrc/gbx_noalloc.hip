// gbx_noalloc.hip -- gbx_spmv_pull_noalloc: the pull SpMV of gbx_spmv_pull
// on caller-owned device vectors, for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops).
//
// The pull kernels combine the chunks of a hub row (a row split over
// several blocks) through two scratch buffers: one FP32 partial per hub
// chunk and one arrival counter per hub row.  gbx_spmv_pull uses the plan's
// own copies; this entry point uses a file-scope device array instead, so a
// call allocates nothing.  The array lives on every device the library
// runs on until the process ends; its address on each device is looked up
// once, by gbx_init, with that device current.  Calls on one device share
// it: calls that overlap on one GPU (two streams, two contexts) race.
#include <mutex>

#include "gbx_iter.hpp"

namespace gbx {

constexpr uint32_t kScratchChunks = 65536;  // hub chunks of a plan (FP32 partial each)
constexpr uint32_t kScratchHubs = 65536;    // hub rows of a plan (u32 arrival counter each)
constexpr size_t kScratchBytes = (size_t)kScratchChunks * 4 + (size_t)kScratchHubs * 4;  // 512 KiB
constexpr int kMaxScratchDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxScratchDevices];  // device address of `scratch` on each device

int noalloc_setup(gbx_ctx *ctx)
{
    if (ctx->device < 0 || ctx->device >= kMaxScratchDevices)
        return set_error(ctx, GBX_ERR_ARG, "gbx_init: device %d beyond %d", ctx->device, kMaxScratchDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[ctx->device]) return GBX_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return set_error(ctx, GBX_ERR_HIP, "gbx_init: scratch address on device %d: %s", ctx->device,
                         hipGetErrorString(e));
    g_scratch_ptr[ctx->device] = ptr;
    return GBX_OK;
}

// x and y in the graph's vertex order: a plan column c (row i) is vertex
// perm[c] (perm[i]).  The products w * x are those of gbx_spmv_pull.
struct GatherPermW {
    const float *__restrict__ x;
    const uint32_t *__restrict__ perm;
    const float *__restrict__ w;  // per-edge weights in plan order, or null (the constant 0.0001f, SPMV.c:177)
    __device__ __forceinline__ float operator()(uint32_t j, uint32_t c) const
    {
        const float xv = x[perm[c]];
        return w ? ld_stream(w + j) * xv : 0.0001f * xv;
    }
};

struct SpmvPermEpi {
    float *__restrict__ y;
    const uint32_t *__restrict__ perm;
    __device__ __forceinline__ float init(uint32_t i) const { return y[perm[i]]; }
    __device__ __forceinline__ void operator()(uint32_t i, float acc) { y[perm[i]] = acc; }
};

__global__ void __launch_bounds__(kBlock)
spmv_perm_kernel(PlanView P, const float *__restrict__ x, const uint32_t *__restrict__ perm, float *__restrict__ y)
{
    SpmvPermEpi epi{y, perm};
    GatherPermW g{x, perm, P.wt};
    tile_loop<AccF32>(P, g, epi);
}

__global__ void zero_f32_u32(float *__restrict__ y, uint32_t n, uint32_t *__restrict__ cnt, uint32_t m)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) y[i] = 0.0f;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) cnt[i] = 0u;
}

}  // namespace gbx

using namespace gbx;

extern "C" int gbx_spmv_pull_noalloc(gbx_ctx *ctx, gbx_graph *g, uint32_t iterations, int weighted, int direction,
                                     const float *x, float *y)
{
    if (!ctx || !g || !x || !y) return set_error(ctx, GBX_ERR_ARG, "gbx_spmv_pull_noalloc: NULL argument");
    if (direction != GBX_PULL_IN && direction != GBX_PULL_FWD)
        return set_error(ctx, GBX_ERR_ARG, "gbx_spmv_pull_noalloc: direction %d not in {0,1}", direction);
    GBX_HIP(ctx, hipSetDevice(ctx->device));
    void *base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        base = g_scratch_ptr[ctx->device];
    }
    if (!base) return set_error(ctx, GBX_ERR_STATE, "gbx_spmv_pull_noalloc: context set up without scratch");
    const gbx_plan *plan = g->plan[direction].get();
    if (!plan)
        return set_error(ctx, GBX_ERR_STATE, "gbx_spmv_pull_noalloc: no plan for direction %d yet (building one "
                                             "allocates: run gbx_spmv_pull or gbx_pr_create on this graph first)",
                         direction);
    if (plan->segmented) return set_error(ctx, GBX_ERR_STATE, "gbx_spmv_pull_noalloc: not for a batch of graphs");
    if (weighted && !plan->has_wt)
        return set_error(ctx, GBX_ERR_STATE, "gbx_spmv_pull_noalloc: weighted SpMV needs a weighted graph");
    if (plan->n_chunks > kScratchChunks || plan->n_hubs > kScratchHubs)
        return set_error(ctx, GBX_ERR_RANGE,
                         "gbx_spmv_pull_noalloc: %u hub chunks / %u hub rows, scratch holds at most %u / %u",
                         plan->n_chunks, plan->n_hubs, kScratchChunks, kScratchHubs);
    float *part = static_cast<float *>(base);
    uint32_t *cnt = reinterpret_cast<uint32_t *>(static_cast<unsigned char *>(base) + (size_t)kScratchChunks * 4);
    PlanView P = plan_view(plan, weighted ? plan->wt.as<float>() : nullptr);
    P.hub_part = part;
    P.hub_cnt = cnt;
    const uint32_t N = plan->N;
    // y starts at 0; every hub counter starts at 0 (the last chunk of a hub
    // resets its counter, but a call may follow a failed one)
    zero_f32_u32<<<grid_for(std::max(N, plan->n_hubs), 256, 16384), 256, 0, ctx->stream>>>(y, N, cnt, plan->n_hubs);
    GBX_HIP(ctx, hipGetLastError());
    const unsigned grid = iter_grid(ctx, plan);
    const uint32_t *perm = plan->perm.as<uint32_t>();
    for (uint32_t it = 0; it < iterations; it++) {
        spmv_perm_kernel<<<grid, kBlock, 0, ctx->stream>>>(P, x, perm, y);
        GBX_HIP(ctx, hipGetLastError());
    }
    return GBX_OK;
}
