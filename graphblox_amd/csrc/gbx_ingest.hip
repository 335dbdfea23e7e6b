// gbx_ingest.hip -- readEdgeListsbin (edgeList.c:450-706) on the GPU.
//
// The raw records of a .bin / .wbin file go to the device as they are
// (64 MiB pieces through two pinned host buffers: a piece is read or copied
// into one buffer while the other's DMA runs), and one kernel deinterleaves
// them into the src / dst / weight arrays of a device edge list, applying
// the reader's inverse swap and symmetric append; the vertex count
// N = max id + 1 and the largest weight are device reductions.
#include <algorithm>
#include <cerrno>
#include <cstring>
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include "gbx_internal.hpp"

struct gbx_edges {
    gbx_ctx *ctx = nullptr;
    uint32_t N = 0, E = 0;
    float max_weight = 0.0f;
    gbx::DevBuf src, dst, w;
};

namespace gbx {

constexpr size_t kIngestPiece = 64u << 20;  // bytes per staged piece (a multiple of 8 and 12)

// Records r in [0, E): (s, d[, w]) at words r*stride ..; inverse swaps s and
// d; the symmetric copy (d, s) goes to r + E (edgeList.c:551-578).
template <int STRIDE>
__global__ void deinterleave_records(const uint32_t *__restrict__ raw, uint32_t E, int inverse, int symmetric,
                                     uint32_t *__restrict__ src, uint32_t *__restrict__ dst, float *__restrict__ w)
{
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < E; r += gridDim.x * blockDim.x) {
        uint32_t s, d;
        float x = 1.0f;
        if constexpr (STRIDE == 2) {
            const uint2 v = reinterpret_cast<const uint2 *>(raw)[r];
            s = v.x;
            d = v.y;
        } else {
            const size_t o = (size_t)r * 3u;
            s = raw[o];
            d = raw[o + 1];
            x = __uint_as_float(raw[o + 2]);
        }
        if (inverse) {
            const uint32_t t = s;
            s = d;
            d = t;
        }
        src[r] = s;
        dst[r] = d;
        w[r] = x;
        if (symmetric) {
            src[r + E] = d;
            dst[r + E] = s;
            w[r + E] = x;
        }
    }
}

// The reader's running maximum (`if (w > max) max = w`, from 0.0f): per
// block, then over blocks, with the same comparison (NaN never larger; only
// values above +0.0f replace the start, so -0.0f never enters).
__global__ void max_weight_partial(const float *__restrict__ w, uint32_t n, float *__restrict__ part)
{
    float m = 0.0f;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float x = w[i];
        if (x > m) m = x;
    }
    __shared__ float sm[256];
    sm[threadIdx.x] = m;
    __syncthreads();
    for (unsigned o = blockDim.x / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o && sm[threadIdx.x + o] > sm[threadIdx.x]) sm[threadIdx.x] = sm[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

// The records' source: host memory or a file descriptor.
struct RecordSource {
    const unsigned char *mem = nullptr;
    int fd = -1;
    const char *path = nullptr;
};

static int fill_piece(gbx_ctx *ctx, const RecordSource &in, uint64_t off, size_t n, unsigned char *buf)
{
    if (in.mem) {
        memcpy(buf, in.mem + off, n);
        return GBX_OK;
    }
    size_t got = 0;
    while (got < n) {
        const ssize_t k = pread(in.fd, buf + got, n - got, (off_t)(off + got));
        if (k < 0 && errno == EINTR) continue;
        if (k <= 0)
            return set_error(ctx, GBX_ERR_IO, "gbx_edges_read_file: read %s at byte %llu: %s", in.path,
                             (unsigned long long)(off + got), k < 0 ? strerror(errno) : "unexpected end of file");
        got += (size_t)k;
    }
    return GBX_OK;
}

// The first `bytes` bytes of the source into device memory `raw`.
static int stage_records(gbx_ctx *ctx, const RecordSource &in, uint64_t bytes, unsigned char *raw)
{
    if (!bytes) return GBX_OK;
    unsigned char *pin[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    const size_t piece = (size_t)std::min<uint64_t>(bytes, kIngestPiece);
    int rc = GBX_OK;
    for (int k = 0; k < 2 && !rc; k++) {
        if (hipHostMalloc((void **)&pin[k], piece, hipHostMallocDefault) != hipSuccess) {
            pin[k] = nullptr;
            rc = set_error(ctx, GBX_ERR_NOMEM, "edge ingest: %zu bytes of pinned host memory", piece);
        } else if (hipEventCreateWithFlags(&ev[k], hipEventDisableTiming) != hipSuccess) {
            ev[k] = nullptr;
            rc = set_error(ctx, GBX_ERR_HIP, "edge ingest: event creation");
        }
    }
    bool used[2] = {false, false};
    for (uint64_t off = 0, k = 0; off < bytes && !rc; off += piece, k++) {
        const int b = (int)(k & 1u);
        const size_t n = (size_t)std::min<uint64_t>(piece, bytes - off);
        // the buffer's previous copy has left it
        if (used[b] && hipEventSynchronize(ev[b]) != hipSuccess) {
            rc = set_error(ctx, GBX_ERR_HIP, "edge ingest: copy wait");
            break;
        }
        if ((rc = fill_piece(ctx, in, off, n, pin[b]))) break;
        if (hipMemcpyAsync(raw + off, pin[b], n, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
            hipEventRecord(ev[b], ctx->stream) != hipSuccess)
            rc = set_error(ctx, GBX_ERR_HIP, "edge ingest: copy to the device");
        used[b] = true;
    }
    // every copy has finished before the pinned buffers go
    if (hipStreamSynchronize(ctx->stream) != hipSuccess && !rc) rc = set_error(ctx, GBX_ERR_HIP, "edge ingest: sync");
    for (int k = 0; k < 2; k++) {
        if (ev[k]) (void)hipEventDestroy(ev[k]);
        if (pin[k]) (void)hipHostFree(pin[k]);
    }
    return rc;
}

static int edges_ingest(gbx_ctx *ctx, const RecordSource &in, uint64_t num_bytes, int weighted_records, int inverse,
                        int symmetric, gbx_edges **out)
{
    const uint64_t rec = weighted_records ? 12u : 8u;
    const uint64_t n64 = num_bytes / rec;
    if (n64 > 0xFFFFFFFFull || (symmetric && n64 > 0x7FFFFFFFull))
        return set_error(ctx, GBX_ERR_RANGE, "edge ingest: %llu records are too many%s", (unsigned long long)n64,
                         symmetric ? " to symmetrise" : "");
    const uint32_t E = (uint32_t)n64, total = symmetric ? 2u * E : E;
    GBX_HIP(ctx, hipSetDevice(ctx->device));
    gbx_edges *e = new (std::nothrow) gbx_edges();
    if (!e) return set_error(ctx, GBX_ERR_NOMEM, "edge ingest: host memory");
    e->ctx = ctx;
    auto fail = [&](int rc) {
        delete e;
        return rc;
    };
    int rc;
    DevBuf raw;
    if ((rc = dev_alloc(ctx, raw, (size_t)(n64 * rec) + 16))) return fail(rc);
    if ((rc = stage_records(ctx, in, n64 * rec, raw.as<unsigned char>()))) return fail(rc);
    if ((rc = dev_alloc(ctx, e->src, (size_t)total * 4 + 4))) return fail(rc);
    if ((rc = dev_alloc(ctx, e->dst, (size_t)total * 4 + 4))) return fail(rc);
    if ((rc = dev_alloc(ctx, e->w, (size_t)total * 4 + 4))) return fail(rc);
    if (E) {
        const unsigned grid = grid_for(E, 256, 16384);
        if (weighted_records)
            deinterleave_records<3><<<grid, 256, 0, ctx->stream>>>(raw.as<uint32_t>(), E, inverse, symmetric,
                                                                  e->src.as<uint32_t>(), e->dst.as<uint32_t>(),
                                                                  e->w.as<float>());
        else
            deinterleave_records<2><<<grid, 256, 0, ctx->stream>>>(raw.as<uint32_t>(), E, inverse, symmetric,
                                                                  e->src.as<uint32_t>(), e->dst.as<uint32_t>(),
                                                                  e->w.as<float>());
        if (hipGetLastError() != hipSuccess) return fail(set_error(ctx, GBX_ERR_HIP, "edge ingest: deinterleave"));
    }
    raw.release();
    // N = max id + 1 (the symmetric copy holds the same ids)
    uint32_t N = 0;
    if ((rc = check_ids(ctx, e->src, e->dst, E, &N))) return fail(rc);
    e->N = N;
    e->E = total;
    if (weighted_records && E) {
        const unsigned blocks = grid_for(E, 256, 1024);
        DevBuf part;
        std::vector<float> h(blocks);
        if ((rc = dev_alloc(ctx, part, (size_t)blocks * 4))) return fail(rc);
        max_weight_partial<<<blocks, 256, 0, ctx->stream>>>(e->w.as<float>(), E, part.as<float>());
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(h.data(), part.p, (size_t)blocks * 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
            return fail(set_error(ctx, GBX_ERR_HIP, "edge ingest: weight maximum"));
        float m = 0.0f;
        for (float x : h)
            if (x > m) m = x;
        e->max_weight = m;
    } else {
        e->max_weight = E ? 1.0f : 0.0f;
    }
    *out = e;
    return GBX_OK;
}

}  // namespace gbx

using namespace gbx;

extern "C" {

int gbx_edges_from_records(gbx_ctx *ctx, const void *records, uint64_t num_bytes, int weighted_records, int inverse,
                           int symmetric, gbx_edges **out)
{
    if (!ctx || !out) return set_error(ctx, GBX_ERR_ARG, "gbx_edges_from_records: NULL argument");
    *out = nullptr;
    if (num_bytes && !records) return set_error(ctx, GBX_ERR_ARG, "gbx_edges_from_records: records is NULL");
    RecordSource in;
    in.mem = static_cast<const unsigned char *>(records);
    return edges_ingest(ctx, in, num_bytes, weighted_records, inverse, symmetric, out);
}

int gbx_edges_read_file(gbx_ctx *ctx, const char *path, int weighted_records, int inverse, int symmetric,
                        gbx_edges **out)
{
    if (!ctx || !out || !path) return set_error(ctx, GBX_ERR_ARG, "gbx_edges_read_file: NULL argument");
    *out = nullptr;
    RecordSource in;
    in.path = path;
    in.fd = open(path, O_RDONLY);
    if (in.fd < 0) return set_error(ctx, GBX_ERR_IO, "gbx_edges_read_file: open %s: %s", path, strerror(errno));
    struct stat st;
    int rc;
    if (fstat(in.fd, &st) != 0)
        rc = set_error(ctx, GBX_ERR_IO, "gbx_edges_read_file: stat %s: %s", path, strerror(errno));
    else
        rc = edges_ingest(ctx, in, (uint64_t)st.st_size, weighted_records, inverse, symmetric, out);
    close(in.fd);
    return rc;
}

int gbx_edges_info(const gbx_edges *e, uint32_t *num_vertices, uint32_t *num_edges, float *max_weight)
{
    if (!e) return set_error(nullptr, GBX_ERR_ARG, "gbx_edges_info: NULL");
    if (num_vertices) *num_vertices = e->N;
    if (num_edges) *num_edges = e->E;
    if (max_weight) *max_weight = e->max_weight;
    return GBX_OK;
}

int gbx_edges_export(const gbx_edges *e, uint32_t *src, uint32_t *dst, float *w)
{
    if (!e) return set_error(nullptr, GBX_ERR_ARG, "gbx_edges_export: NULL");
    gbx_ctx *ctx = e->ctx;
    GBX_HIP(ctx, hipSetDevice(ctx->device));
    const size_t eb = (size_t)e->E * 4;
    if (eb) {
        if (src) GBX_HIP(ctx, hipMemcpyAsync(src, e->src.p, eb, hipMemcpyDeviceToHost, ctx->stream));
        if (dst) GBX_HIP(ctx, hipMemcpyAsync(dst, e->dst.p, eb, hipMemcpyDeviceToHost, ctx->stream));
        if (w) GBX_HIP(ctx, hipMemcpyAsync(w, e->w.p, eb, hipMemcpyDeviceToHost, ctx->stream));
    }
    GBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GBX_OK;
}

int gbx_graph_create_from_edges(gbx_ctx *ctx, gbx_edges *e, int sort_mode, unsigned flags, const gbx_relabel *relabel,
                                uint32_t *labels_out, gbx_graph **out)
{
    if (!ctx || !e || !out) return set_error(ctx, GBX_ERR_ARG, "gbx_graph_create_from_edges: NULL argument");
    *out = nullptr;
    if (e->ctx != ctx) return set_error(ctx, GBX_ERR_ARG, "gbx_graph_create_from_edges: edge list of another context");
    if (sort_mode < 0 || sort_mode > 3)
        return set_error(ctx, GBX_ERR_ARG, "gbx_graph_create_from_edges: sort_mode %d not in 0..3", sort_mode);
    if (flags & GBX_GRAPH_SYMMETRIC)
        return set_error(ctx, GBX_ERR_ARG, "gbx_graph_create_from_edges: the symmetric copy is made at ingest");
    GBX_TRY(relabel_validate(ctx, relabel));
    if (!e->src.p || !e->dst.p || !e->w.p)
        return set_error(ctx, GBX_ERR_STATE, "gbx_graph_create_from_edges: the edge list was consumed");
    GBX_HIP(ctx, hipSetDevice(ctx->device));
    const uint32_t E = e->E, N = e->N;
    e->E = 0;  // consumed, whatever the outcome
    DevBuf s = std::move(e->src), d = std::move(e->dst), w = std::move(e->w);
    return graph_from_device_edges(ctx, s, d, w, E, N, sort_mode, flags, out, relabel, labels_out);
}

int gbx_edges_free(gbx_edges *e)
{
    if (!e) return GBX_OK;
    if (e->ctx) (void)hipSetDevice(e->ctx->device);
    delete e;
    return GBX_OK;
}

}  // extern "C"
