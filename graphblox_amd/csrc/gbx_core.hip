// gbx_core.hip -- context, errors, device buffers and the scan/fill/max
// primitives shared by the CSR build and the PageRank plan.
#include <cstdarg>
#include <cstdlib>
#include <mutex>
#include <thread>

#include "gbx_plan.hpp"

namespace gbx {

static thread_local std::string g_noctx_err = "no error";

int set_error(gbx_ctx *ctx, int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx)
        ctx->err = buf;
    else
        g_noctx_err = buf;
    return code;
}

int dev_alloc(gbx_ctx *ctx, DevBuf &b, size_t bytes)
{
    b.release();
    if (bytes == 0) bytes = 4;
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        b.p = nullptr;
        (void)hipGetLastError();
        return set_error(ctx, GBX_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes,
                         hipGetErrorString(e));
    }
    b.bytes = bytes;
    return GBX_OK;
}

int dev_reserve(gbx_ctx *ctx, DevBuf &b, size_t bytes)
{
    if (b.p && b.bytes >= bytes) return GBX_OK;
    // the stream may still use the old buffer
    GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return dev_alloc(ctx, b, bytes);
}

uint32_t bits_for(uint32_t max_value)
{
    uint32_t b = 0;
    while (b < 32 && (max_value >> b) != 0) b++;
    return b;
}

// ------------------------------------------------------------------ kernels

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}

// Exclusive scan of the 256 per-thread totals of a block; returns this
// thread's exclusive prefix and writes the block total to *total.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *total)
{
    __shared__ uint32_t wsum[kScanBlock / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t inc = wave_inclusive_scan(v);
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    uint32_t wbase = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kScanBlock / 64; w++) {
        uint32_t s = wsum[w];
        wbase += (w < wid) ? s : 0u;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return wbase + inc - v;
}

__global__ void __launch_bounds__(kScanBlock)
scan_reduce_kernel(const uint32_t *__restrict__ in, size_t n, uint32_t *__restrict__ sums)
{
    const size_t base = (size_t)blockIdx.x * kScanTile;
    uint32_t acc = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        size_t i = base + (size_t)k * kScanBlock + threadIdx.x;
        if (i < n) acc += in[i];
    }
    // wave reduce then block reduce
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    __shared__ uint32_t ws[kScanBlock / 64];
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kScanBlock / 64; w++) t += ws[w];
        sums[blockIdx.x] = t;
    }
}

// Exclusive scan of one tile per block, plus the tile's offset (nullable).
__global__ void __launch_bounds__(kScanBlock)
scan_tile_kernel(const uint32_t *__restrict__ in, uint32_t *__restrict__ out, size_t n,
                 const uint32_t *__restrict__ offsets, uint32_t *__restrict__ total)
{
    __shared__ uint32_t tile[kScanTile];
    const size_t base = (size_t)blockIdx.x * kScanTile;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        int li = k * kScanBlock + threadIdx.x;
        size_t i = base + li;
        tile[li] = (i < n) ? in[i] : 0u;
    }
    __syncthreads();
    uint32_t vals[kScanItems];
    uint32_t run = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        vals[k] = tile[threadIdx.x * kScanItems + k];
        run += vals[k];
    }
    uint32_t btot;
    uint32_t pre = block_exclusive_scan(run, &btot);
    const uint32_t off = offsets ? offsets[blockIdx.x] : 0u;
    pre += off;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        tile[threadIdx.x * kScanItems + k] = pre;
        pre += vals[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        int li = k * kScanBlock + threadIdx.x;
        size_t i = base + li;
        if (i < n) out[i] = tile[li];
    }
    if (total && threadIdx.x == 0 && blockIdx.x == gridDim.x - 1) *total = off + btot;
}

__global__ void fill_u32_kernel(uint32_t *p, uint32_t v, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void max_u32_kernel(const uint32_t *a, size_t n, uint32_t *out)
{
    uint32_t m = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        m = max(m, a[i]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = max(m, (uint32_t)__shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

// ------------------------------------------------------------------ host

static int scan_level(gbx_ctx *ctx, const uint32_t *in, uint32_t *out, size_t n,
                      uint32_t *d_total, uint32_t *tmp, size_t tmp_elems)
{
    const size_t nb = (n + kScanTile - 1) / kScanTile;
    if (nb <= 1) {
        scan_tile_kernel<<<1, kScanBlock, 0, ctx->stream>>>(in, out, n, nullptr, d_total);
        GBX_HIP(ctx, hipGetLastError());
        return GBX_OK;
    }
    if (tmp_elems < nb) return set_error(ctx, GBX_ERR_STATE, "scan scratch too small");
    uint32_t *sums = tmp;
    scan_reduce_kernel<<<(unsigned)nb, kScanBlock, 0, ctx->stream>>>(in, n, sums);
    GBX_HIP(ctx, hipGetLastError());
    GBX_TRY(scan_level(ctx, sums, sums, nb, d_total, tmp + nb, tmp_elems - nb));
    scan_tile_kernel<<<(unsigned)nb, kScanBlock, 0, ctx->stream>>>(in, out, n, sums, nullptr);
    GBX_HIP(ctx, hipGetLastError());
    return GBX_OK;
}

int exclusive_scan_u32(gbx_ctx *ctx, const uint32_t *in, uint32_t *out, size_t n,
                       uint32_t *d_total)
{
    if (n == 0) {
        if (d_total) return fill_u32(ctx, d_total, 0, 1);
        return GBX_OK;
    }
    // scratch: the tile sums of every level above the first
    size_t need = 16;
    for (size_t m = n; m > kScanTile;) {
        m = (m + kScanTile - 1) / kScanTile;
        need += m;
    }
    GBX_TRY(dev_reserve(ctx, ctx->scan_tmp, need * sizeof(uint32_t)));
    return scan_level(ctx, in, out, n, d_total, ctx->scan_tmp.as<uint32_t>(),
                      ctx->scan_tmp.bytes / sizeof(uint32_t));
}

int fill_u32(gbx_ctx *ctx, uint32_t *p, uint32_t v, size_t n)
{
    if (n == 0) return GBX_OK;
    fill_u32_kernel<<<grid_for(n, 256, 8192), 256, 0, ctx->stream>>>(p, v, n);
    GBX_HIP(ctx, hipGetLastError());
    return GBX_OK;
}

int max_u32(gbx_ctx *ctx, const uint32_t *a, size_t n, uint32_t *h_max)
{
    GBX_TRY(dev_reserve(ctx, ctx->misc, 256));
    uint32_t *d = ctx->misc.as<uint32_t>();
    GBX_TRY(fill_u32(ctx, d, 0, 1));
    if (n) {
        max_u32_kernel<<<grid_for(n, 256, 4096), 256, 0, ctx->stream>>>(a, n, d);
        GBX_HIP(ctx, hipGetLastError());
    }
    GBX_HIP(ctx, hipMemcpyAsync(h_max, d, 4, hipMemcpyDeviceToHost, ctx->stream));
    GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GBX_OK;
}

// Large copies go through the context's pinned staging buffer, two chunks
// in flight: the DMA of chunk i + 1 runs while host threads copy chunk i
// out (a copy straight into pageable memory runs at about 6 GB/s: 11 ms for
// the 16.7M ranks of R-MAT s24).
int copy_to_host(gbx_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    constexpr size_t kDirect = 1u << 20, kChunk = 32u << 20;
    const size_t chunk = std::min(bytes, kChunk);
    if (bytes > kDirect && ctx->h_stage_bytes < 2 * chunk) {
        if (ctx->h_stage) (void)hipHostFree(ctx->h_stage);
        ctx->h_stage = nullptr;
        ctx->h_stage_bytes = 0;
        if (hipHostMalloc(&ctx->h_stage, 2 * chunk, 0) == hipSuccess) {
            ctx->h_stage_bytes = 2 * chunk;
        } else {
            ctx->h_stage = nullptr;  // no pinned memory to spare: the runtime's own copy below
            (void)hipGetLastError();
        }
    }
    if (bytes <= kDirect || !ctx->h_stage) {
        GBX_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        return GBX_OK;
    }
    hipEvent_t ev[2] = {nullptr, nullptr};
    int rc = GBX_OK;
    for (auto &e : ev)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
            rc = set_error(ctx, GBX_ERR_HIP, "copy_to_host: event");
    char *stage = static_cast<char *>(ctx->h_stage);
    const char *s = static_cast<const char *>(src);
    char *d = static_cast<char *>(dst);
    const size_t nch = (bytes + chunk - 1) / chunk;
    auto issue = [&](size_t i) {
        const size_t off = i * chunk, n = std::min(chunk, bytes - off);
        if (hipMemcpyAsync(stage + (i & 1) * chunk, s + off, n, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipEventRecord(ev[i & 1], ctx->stream) != hipSuccess)
            rc = set_error(ctx, GBX_ERR_HIP, "copy_to_host: %s", hipGetErrorString(hipGetLastError()));
    };
    const unsigned nthr = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
    if (!rc) issue(0);
    for (size_t i = 0; i < nch && !rc; i++) {
        if (i + 1 < nch) issue(i + 1);  // its slot was copied out in the previous round
        if (rc) break;
        if (hipEventSynchronize(ev[i & 1]) != hipSuccess) {
            rc = set_error(ctx, GBX_ERR_HIP, "copy_to_host: %s", hipGetErrorString(hipGetLastError()));
            break;
        }
        const size_t off = i * chunk, n = std::min(chunk, bytes - off), part = (n + nthr - 1) / nthr;
        const char *from = stage + (i & 1) * chunk;
        std::vector<std::thread> th;
        for (unsigned t = 1; t < nthr && t * part < n; t++)
            th.emplace_back([=] { std::memcpy(d + off + t * part, from + t * part, std::min(part, n - t * part)); });
        std::memcpy(d + off, from, std::min(part, n));
        for (auto &x : th) x.join();
    }
    if (rc) (void)hipStreamSynchronize(ctx->stream);  // no DMA left in flight into the staging
    for (auto &e : ev)
        if (e) (void)hipEventDestroy(e);
    return rc;
}

void tuning_defaults(gbx_tuning *t)
{
    // hub threshold and chunk 0: automatic, by the graph's size (auto_hub_threshold)
    static const uint32_t thr[7] = {0, 256, 128, 64, 32, 16, 8};
    memset(t, 0, sizeof(*t));
    memcpy(t->bin_thresholds, thr, sizeof(thr));
    t->hub_chunk = 0;
    t->lds_bytes = 159744;  // 156 KiB of the 160 KiB LDS of a CU (1024-thread block, one per CU)
    t->blocks_per_cu = 1;
    // batched form, measured at R-MAT s22 with 32 vectors (ms per
    // iteration, HIP graph replays): rows of 4,096+ edges chunked in 8,192;
    // four contributions a lane with one, two, four blocks per CU 1.64,
    // 1.40, 1.27 (one a lane: 1.39 with four, 1.38 with two); rows from
    // 1,024 edges chunked: 1.39 at one to four blocks per CU (the blocks
    // per CU also size a batch of graphs' tile grid)
    t->ppr_hub = 4096;
    t->ppr_chunk = 8192;
    t->ppr_blocks_per_cu = 4;
    t->ppr_vectors_per_lane = 4;
    t->binned_min_vertices = 786432;  // measured crossover: pull faster at R-MAT s19, binned from s20
    t->binned_chunk_edges = 0;  // automatic
    t->binned_bin_edges = 0;
    t->binned_bin_rows = 20000;
    t->binned_chunk_rows = 32768;
    t->binned_hot_sources = 16384;
}

bool tuning_valid(const gbx_tuning *t)
{
    // the hub threshold may be 0 (automatic: at least 2048, above the lane thresholds' 256 limit)
    if (t->bin_thresholds[0] != 0 && t->bin_thresholds[1] > t->bin_thresholds[0]) return false;
    if (t->bin_thresholds[1] == 0 || (t->bin_thresholds[0] == 0 && t->bin_thresholds[1] > 2048)) return false;
    for (int i = 2; i < 7; i++)
        if (t->bin_thresholds[i] == 0 || t->bin_thresholds[i] > t->bin_thresholds[i - 1]) return false;
    if (t->hub_chunk != 0 && (t->hub_chunk < 256 || t->hub_chunk > (1u << 24))) return false;
    if (t->lds_bytes > 160 * 1024 || t->blocks_per_cu < 1 || t->blocks_per_cu > 64) return false;
    if (t->ppr_hub < 64 || t->ppr_chunk < 256 || t->ppr_blocks_per_cu < 1 || t->ppr_blocks_per_cu > 64)
        return false;
    if (t->binned_bin_edges != 0 && t->binned_bin_edges < 8) return false;
    if (t->binned_bin_rows < 1 || t->binned_bin_rows > 20000) return false;
    if (t->binned_chunk_rows < 1 || t->binned_chunk_rows > 32768) return false;
    if (t->binned_hot_sources > 65528 || t->binned_edge_values > 1) return false;
    const uint32_t v = t->ppr_vectors_per_lane;
    return v == 1 || v == 2 || v == 4;
}

// Environment overrides, read once per context.
static void tuning_from_env(gbx_tuning *t)
{
    gbx_tuning c = *t;
    if (const char *s = getenv("GBX_BIN_THRESHOLDS"))
        sscanf(s, "%u,%u,%u,%u,%u,%u,%u", &c.bin_thresholds[0], &c.bin_thresholds[1], &c.bin_thresholds[2],
               &c.bin_thresholds[3], &c.bin_thresholds[4], &c.bin_thresholds[5], &c.bin_thresholds[6]);
    struct {
        const char *name;
        uint32_t *field;
    } vars[] = {{"GBX_HUB_CHUNK", &c.hub_chunk},
                {"GBX_LDS_BYTES", &c.lds_bytes},
                {"GBX_BLOCKS_PER_CU", &c.blocks_per_cu},
                {"GBX_PPR_HUB", &c.ppr_hub},
                {"GBX_PPR_CHUNK", &c.ppr_chunk},
                {"GBX_PPR_BLOCKS_PER_CU", &c.ppr_blocks_per_cu},
                {"GBX_PPR_VPL", &c.ppr_vectors_per_lane},
                {"GBX_BINNED_MIN_VERTICES", &c.binned_min_vertices},
                {"GBX_BINNED_CHUNK_EDGES", &c.binned_chunk_edges},
                {"GBX_BINNED_BIN_EDGES", &c.binned_bin_edges},
                {"GBX_BINNED_BIN_ROWS", &c.binned_bin_rows},
                {"GBX_BINNED_CHUNK_ROWS", &c.binned_chunk_rows},
                {"GBX_BINNED_HOT_SOURCES", &c.binned_hot_sources},
                {"GBX_BINNED_EDGE_VALUES", &c.binned_edge_values}};
    for (auto &v : vars)
        if (const char *s = getenv(v.name)) *v.field = (uint32_t)strtoul(s, nullptr, 10);
    if (tuning_valid(&c)) *t = c;
}

}  // namespace gbx

// ------------------------------------------------------------------ C ABI

extern "C" {

int gbx_init(int device, gbx_ctx **out)
{
    if (!out) return gbx::set_error(nullptr, GBX_ERR_ARG, "gbx_init: out is NULL");
    *out = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0)
        return gbx::set_error(nullptr, GBX_ERR_HIP, "gbx_init: no HIP device (%s)",
                              hipGetErrorString(e));
    if (device < 0 || device >= ndev)
        return gbx::set_error(nullptr, GBX_ERR_ARG, "gbx_init: device %d out of range [0,%d)",
                              device, ndev);
    gbx_ctx *ctx = new (std::nothrow) gbx_ctx();
    if (!ctx) return gbx::set_error(nullptr, GBX_ERR_NOMEM, "gbx_init: out of host memory");
    ctx->device = device;
    e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipDeviceGetAttribute(&ctx->num_cus, hipDeviceAttributeMultiprocessorCount, device);
    if (e != hipSuccess) {
        int rc = gbx::set_error(nullptr, GBX_ERR_HIP, "gbx_init: %s", hipGetErrorString(e));
        delete ctx;
        return rc;
    }
    ctx->err = "no error";
    gbx::tuning_defaults(&ctx->tune);
    gbx::tuning_from_env(&ctx->tune);
    // the device address of the allocation-free entry point's scratch array
    if (int src = gbx::noalloc_setup(ctx)) {
        int rc = gbx::set_error(nullptr, src, "%s", ctx->err.c_str());
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return rc;
    }
    *out = ctx;
    return GBX_OK;
}

int gbx_tuning_get(const gbx_ctx *ctx, gbx_tuning *t)
{
    if (!ctx || !t) return gbx::set_error(nullptr, GBX_ERR_ARG, "gbx_tuning_get: NULL argument");
    *t = ctx->tune;
    return GBX_OK;
}

int gbx_tuning_set(gbx_ctx *ctx, const gbx_tuning *t)
{
    if (!ctx || !t) return gbx::set_error(ctx, GBX_ERR_ARG, "gbx_tuning_set: NULL argument");
    if (!gbx::tuning_valid(t)) return gbx::set_error(ctx, GBX_ERR_ARG, "gbx_tuning_set: invalid tuning");
    ctx->tune = *t;
    return GBX_OK;
}

int gbx_free(gbx_ctx *ctx)
{
    if (!ctx) return GBX_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamDestroy(ctx->stream);
    }
    if (ctx->h_stage) (void)hipHostFree(ctx->h_stage);
    delete ctx;
    return GBX_OK;
}

const char *gbx_last_error(const gbx_ctx *ctx)
{
    return ctx ? ctx->err.c_str() : gbx::g_noctx_err.c_str();
}

const char *gbx_status_string(int status)
{
    switch (status) {
    case GBX_OK: return "ok";
    case GBX_ERR_ARG: return "invalid argument";
    case GBX_ERR_NOMEM: return "out of memory";
    case GBX_ERR_HIP: return "HIP runtime error";
    case GBX_ERR_RANGE: return "vertex id out of range";
    case GBX_ERR_STATE: return "invalid state";
    case GBX_ERR_IO: return "file error";
    default: return "unknown status";
    }
}

int gbx_sync(gbx_ctx *ctx)
{
    if (!ctx) return gbx::set_error(nullptr, GBX_ERR_ARG, "gbx_sync: ctx is NULL");
    GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GBX_OK;
}

void *gbx_stream(gbx_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

const char *gbx_version(void) { return "gbx_hip 0.1 gfx950"; }

}  // extern "C"
