// gbx_tune.hip -- autotuner for the pull kernels' launch shapes and the
// binned iteration's layout cut.
//
// Coordinate descent from the context's current tuning: for each field in
// turn, try its candidate values on short timed runs of the real kernels on
// the caller's graph (HIP events around every iteration), keep a value only
// when it is faster by more than 1 %, and repeat the sweep once; the pick
// is then timed again against the starting setting and kept only if
// faster.  Which fields are searched follows the method the parameters
// select on this graph: for the pull iteration the degree-bin thresholds, hub chunk, LDS
// prefix and blocks per CU (bin thresholds and hub chunks change the plan,
// rebuilt by the next object creation); for the binned iteration the edges
// per source chunk and per destination bin, the rows per bin and the hot
// sources (each cut builds a layout; those not chosen are dropped at the
// end); for a batch of graphs the bin thresholds, hub chunk and blocks per
// CU of its tile kernel; for the batched form its lanes, hub cut and blocks
// per CU.
#include <algorithm>
#include <cstring>
#include <vector>

#include "gbx_plan.hpp"

namespace gbx {

static int trial(gbx_ctx *ctx, gbx_graph *g, const gbx_pr_params *p, uint32_t k, uint32_t iters,
                 const gbx_tuning &t, double *ms, int *method = nullptr)
{
    ctx->tune = t;
    gbx_pr_params q = *p;
    q.stop = GBX_STOP_NONE;
    q.max_iters = iters + 2;
    if (k) q.mode = GBX_PR_STOCHASTIC;
    gbx_pr *pr = nullptr;
    int rc = k ? gbx_ppr_create(ctx, g, &q, k, nullptr, nullptr, &pr) : gbx_pr_create(ctx, g, &q, &pr);
    if (rc) return rc;
    if (method) *method = pr->bin ? 1 : pr->seg ? 2 : 0;  // binned, batch of graphs, pull
    rc = gbx_pr_reset(pr, nullptr);
    if (!rc) rc = gbx_pr_launch(pr, 2);
    if (!rc) rc = gbx_pr_launch_timed(pr, iters, ms);
    gbx_pr_free(pr);
    return rc;
}

}  // namespace gbx

using namespace gbx;

extern "C" int gbx_autotune(gbx_ctx *ctx, gbx_graph *g, const gbx_pr_params *p, uint32_t num_vectors,
                            uint32_t iters_per_trial, gbx_tuning *best, double *best_ms)
{
    if (!ctx || !g || !p) return set_error(ctx, GBX_ERR_ARG, "gbx_autotune: NULL argument");
    if (iters_per_trial == 0) iters_per_trial = 5;
    const gbx_tuning start = ctx->tune;
    gbx_tuning cur = start;
    double cur_ms = 0.0;
    int method = 0;
    int rc = trial(ctx, g, p, num_vectors, iters_per_trial, cur, &cur_ms, &method);
    const bool binned = method == 1, seg = method == 2;
    if (rc) {
        ctx->tune = start;
        return rc;
    }
    struct Field {
        uint32_t *(*get)(gbx_tuning &);
        std::vector<uint32_t> values;
    };
    std::vector<Field> fields;
    std::vector<std::vector<uint32_t>> lane_sets = {
        {256, 128, 64, 32, 16, 8}, {512, 256, 128, 64, 32, 16}, {128, 64, 32, 16, 8, 4}};
    if (num_vectors == 0 && binned) {
        // 0: the automatic cut (7/8 of E / CUs in [64K, 1M] for the run layout)
        const std::vector<uint32_t> cuts = {0, 393216, 524288, 786432, 1048576, 1310720};
        fields.push_back({[](gbx_tuning &t) { return &t.binned_chunk_edges; }, cuts});
        fields.push_back({[](gbx_tuning &t) { return &t.binned_bin_edges; }, cuts});
        fields.push_back({[](gbx_tuning &t) { return &t.binned_hot_sources; }, {8192, 12288, 16384}});
        fields.push_back({[](gbx_tuning &t) { return &t.binned_bin_rows; }, {4096, 8192, 20000}});
    } else if (num_vectors == 0 && seg) {
        // a batch of graphs: its graph-pure tiles follow the bin thresholds
        // and hub chunk; its grid, the batched form's blocks per CU
        fields.push_back({[](gbx_tuning &t) { return &t.bin_thresholds[0]; }, {2048, 4096, 8192, 16384, 32768}});
        fields.push_back({[](gbx_tuning &t) { return &t.hub_chunk; }, {8192, 16384, 32768, 65536}});
        fields.push_back({[](gbx_tuning &t) { return &t.ppr_blocks_per_cu; }, {1, 2, 4, 8}});
    } else if (num_vectors == 0) {
        fields.push_back({[](gbx_tuning &t) { return &t.bin_thresholds[0]; }, {2048, 4096, 8192, 16384, 32768}});
        fields.push_back({[](gbx_tuning &t) { return &t.hub_chunk; }, {8192, 16384, 32768, 65536}});
        fields.push_back({[](gbx_tuning &t) { return &t.lds_bytes; }, {0, 65536, 159744}});
        fields.push_back({[](gbx_tuning &t) { return &t.blocks_per_cu; }, {1, 2}});
    } else {
        fields.push_back({[](gbx_tuning &t) { return &t.ppr_vectors_per_lane; }, {1, 2, 4}});
        fields.push_back({[](gbx_tuning &t) { return &t.ppr_hub; }, {256, 1024, 4096}});
        fields.push_back({[](gbx_tuning &t) { return &t.ppr_chunk; }, {1024, 2048, 8192}});
        fields.push_back({[](gbx_tuning &t) { return &t.ppr_blocks_per_cu; }, {1, 2, 4}});
    }
    for (int pass = 0; pass < 2; pass++) {
        if (num_vectors == 0 && !binned) {
            // the lane thresholds move together
            for (const auto &ls : lane_sets) {
                gbx_tuning t = cur;
                const uint32_t hub = t.bin_thresholds[0] ? t.bin_thresholds[0] : auto_hub_threshold(cut_edges(g));
                for (int i = 0; i < 6; i++) t.bin_thresholds[i + 1] = std::min(ls[i], hub);
                if (!tuning_valid(&t) || memcmp(&t, &cur, sizeof(t)) == 0) continue;
                double ms = 0.0;
                if ((rc = trial(ctx, g, p, num_vectors, iters_per_trial, t, &ms))) break;
                if (ms < cur_ms * 0.99) {
                    cur = t;
                    cur_ms = ms;
                }
            }
            if (rc) break;
        }
        for (auto &f : fields) {
            for (uint32_t v : f.values) {
                gbx_tuning t = cur;
                if (*f.get(t) == v) continue;
                *f.get(t) = v;
                if (!tuning_valid(&t)) continue;
                double ms = 0.0;
                if ((rc = trial(ctx, g, p, num_vectors, iters_per_trial, t, &ms))) break;
                if (ms < cur_ms * 0.99) {
                    cur = t;
                    cur_ms = ms;
                }
            }
            if (rc) break;
        }
        if (rc) break;
    }
    // confirmation: every value above was kept on one short run, so a noisy
    // trial can leave a pick slower than where the search started; the
    // search's pick and its start are timed again, alternately, on runs twice
    // as long, and the pick stays only if its best run beats the start's
    if (!rc && memcmp(&cur, &start, sizeof(cur)) != 0) {
        double best_cur = 1e300, best_start = 1e300;
        for (int rep = 0; rep < 2 && !rc; rep++) {
            double ms = 0.0;
            if ((rc = trial(ctx, g, p, num_vectors, 2 * iters_per_trial, start, &ms))) break;
            best_start = std::min(best_start, ms);
            if ((rc = trial(ctx, g, p, num_vectors, 2 * iters_per_trial, cur, &ms))) break;
            best_cur = std::min(best_cur, ms);
        }
        if (!rc) {
            if (best_start <= best_cur) cur = start;
            cur_ms = std::min(best_start, best_cur);
        }
    }
    // layouts of the cuts not chosen (no object holds them) are dropped;
    // the next object on the chosen cut reuses or rebuilds its own
    if (binned && (unsigned)p->direction < 2 && g->plan[p->direction]) binned_drop_unused(g->plan[p->direction].get());
    if (rc) {
        ctx->tune = start;
        return rc;
    }
    ctx->tune = cur;
    if (best) *best = cur;
    if (best_ms) *best_ms = cur_ms;
    return GBX_OK;
}
