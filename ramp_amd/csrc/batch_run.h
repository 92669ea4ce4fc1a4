/* batch_run.h - many instrument settings in one persistent launch.
 *
 * The reference runs one Instrument per setting (a triple-axis scan point,
 * a chopper phase, an Ei) and calls execute() on each in turn
 * (mcramp.py:403-455), so a scan of small settings is dominated by per-run
 * launches and synchronisations.  Here every setting keeps its own plan
 * (components, tables, detectors: the physics is unchanged) and a batch runs
 * all of them in one launch of k_batch:
 *
 *  - the work is cut into tiles (setting, first history, length) of
 *    64 x npt x 8 histories; a wave takes the next tile from a device
 *    counter (one atomic per tile), runs its histories through the
 *    setting's plan, and repeats until the tiles are exhausted, so uneven
 *    counts balance by themselves;
 *  - a tile runs through the setting's plan in one fused pass; the rare,
 *    long reflection branch of a monochromator is deferred through a
 *    per-wave LDS queue and runs for 64 neutrons at once;
 *  - a history of setting s is generated exactly as in a plain run of
 *    plan s with count N_s as its only chunk and the setting's seed (key =
 *    history index, weight 1/N_s), so each setting's histograms equal a
 *    plain run's (bit for bit in deterministic mode);
 *  - the launch sequence (counter reset, tally reset, k_batch) is captured
 *    once as a HIP graph and replayed; per-run inputs (counts, seeds) live
 *    in device tables written before each replay.
 */
#ifndef RAMP_BATCH_RUN_H
#define RAMP_BATCH_RUN_H

#include "batch_kernel.h"

static batch_fn batch_kernel(int v) { return kern_table(v).batch; }

struct ramp_batch {
    ramp_ctx *ctx = nullptr;
    std::vector<ramp_plan *> plans;
    int variant = kNumVariants - 1;
    DevBuf sets, tiles, next, cnt;
    std::vector<SegArgs> hsets;
    std::vector<BatchTile> htiles;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    /* configuration the graph was captured with */
    unsigned g_grid = 0, g_wg = 0;
    uint32_t g_ntiles = 0;
    void *g_sets = nullptr, *g_tiles = nullptr;
    ramp_run_stats stats;
    uint32_t last_wg = 0, last_npt = 0, last_grid = 0;
};

static void batch_drop_graph(ramp_batch *b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

extern "C" int ramp_batch_create(ramp_ctx *ctx, ramp_plan *const *plans, uint32_t n_settings,
                                 ramp_batch **out)
{
    if (!out) return fail(RAMP_E_INVALID, "null output pointer");
    *out = nullptr;
    if (!ctx) return fail(RAMP_E_INVALID, "null context");
    if (!plans || n_settings == 0) return fail(RAMP_E_INVALID, "no settings");
    uint32_t feat = 0;
    for (uint32_t s = 0; s < n_settings; s++) {
        if (!plans[s]) return fail(RAMP_E_INVALID, "setting %u: null plan", s);
        if (plans[s]->ctx != ctx) return fail(RAMP_E_INVALID, "setting %u: plan of another context", s);
        if (plans[s]->need_full)
            return fail(RAMP_E_INVALID, "setting %u: SMCPLIn/SMCPLOut/SRescal/SAnalyser plans run with ramp_run only", s);
        for (uint32_t r = 0; r < s; r++)
            if (plans[r] == plans[s]) return fail(RAMP_E_INVALID, "settings %u and %u share a plan", r, s);
        feat |= plans[s]->feat;
    }
    ramp_batch *b = new ramp_batch();
    b->ctx = ctx;
    b->plans.assign(plans, plans + n_settings);
    for (int v = 0; v < kNumVariants; v++)
        if ((feat & ~kVariants[v]) == 0) { b->variant = v; break; }
    std::memset(&b->stats, 0, sizeof b->stats);
    int rc = dev_alloc(b->sets, n_settings * sizeof(SegArgs));
    if (!rc) rc = dev_alloc(b->next, 256);
    if (!rc) rc = dev_alloc(b->cnt, CNT_N * sizeof(unsigned long long));
    if (rc) {
        dev_free(b->sets); dev_free(b->next); dev_free(b->cnt);
        delete b;
        return rc;
    }
    *out = b;
    return 0;
}

extern "C" int ramp_batch_destroy(ramp_batch *b)
{
    if (!b) return 0;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    batch_drop_graph(b);
    dev_free(b->sets);
    dev_free(b->tiles);
    dev_free(b->next);
    dev_free(b->cnt);
    delete b;
    return 0;
}

extern "C" int ramp_batch_run(ramp_batch *b, const uint64_t *counts, const uint32_t *seeds,
                              const double *weight_bounds, const ramp_run_opts *opts)
{
    if (!b) return fail(RAMP_E_INVALID, "null batch");
    if (!counts || !seeds) return fail(RAMP_E_INVALID, "null counts or seeds");
    ramp_run_opts o;
    std::memset(&o, 0, sizeof o);
    if (opts) o = *opts;
    if (o.hist_mode != RAMP_HIST_ATOMIC && o.hist_mode != RAMP_HIST_DETERMINISTIC)
        return fail(RAMP_E_INVALID, "unknown histogram mode %u", o.hist_mode);
    if (o.keep_state) return fail(RAMP_E_INVALID, "a batch keeps no per-neutron state");
    if (o.flags) return fail(RAMP_E_INVALID, "run flags 0x%x are not available in a batch", o.flags);
    if (o.key_base || o.shard_total) return fail(RAMP_E_INVALID, "a batch does not run shards");
    unsigned wg = o.wg ? o.wg : 256;
    if (wg != 64 && wg != 128 && wg != 256 && wg != 512) return fail(RAMP_E_INVALID, "workgroup size %u", wg);
    uint32_t npt = o.npt ? o.npt : 4;
    if (npt != 1 && npt != 2 && npt != 4 && npt != 8 && npt != 16)
        return fail(RAMP_E_INVALID, "histories per thread %u", npt);
    const uint32_t ns = (uint32_t)b->plans.size();
    uint64_t total = 0;
    for (uint32_t s = 0; s < ns; s++) {
        if (counts[s] > 0x7fffffffull)
            return fail(RAMP_E_RANGE, "setting %u: %llu histories exceed 2^31-1", s, (unsigned long long)counts[s]);
        if (seeds[s] > 0x7fffffffu) return fail(RAMP_E_INVALID, "setting %u: seed %u out of range", s, seeds[s]);
        total += counts[s];
    }
    ramp_ctx *ctx = b->ctx;
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    /* per-setting arguments: the whole plan as one fused pass */
    b->hsets.assign(ns, SegArgs());
    for (uint32_t s = 0; s < ns; s++) {
        ramp_plan *p = b->plans[s];
        uint64_t c = counts[s] ? counts[s] : 1;
        double wb = weight_bounds ? weight_bounds[s] : 0.0;
        int rc = prepare_dets(p, (int)o.hist_mode, wb, c, c, counts[s]);
        if (rc) return rc;
        SegArgs &a = b->hsets[s];
        std::memset(&a, 0, sizeof a);
        a.revive_k = -1;
        a.comps = (const DComp *)p->comps.p;
        a.steps = (const DStep *)p->steps.p;
        a.src = (const DPart *)p->src.p;
        a.dets = (const DDet *)p->dets.p;
        a.cnt = (unsigned long long *)p->cnt.p;
        a.cnt = (unsigned long long *)b->cnt.p;
        a.hist_mode = (int)o.hist_mode;
        a.s0 = 0;
        a.s1 = (int)p->hsteps.size();
        a.defer_k = p->defer_k;
        a.last = 1;
        a.pilot = 0;
        a.hist_mode = (int)o.hist_mode;
        a.num_sim = (int)counts[s];
        a.seed = (int)seeds[s];
        a.n = counts[s];
    }
    /* tiles, setting by setting */
    const uint64_t tile = 64ull * npt * 8ull;
    b->htiles.clear();
    for (uint32_t s = 0; s < ns; s++)
        for (uint64_t beg = 0; beg < counts[s]; beg += tile) {
            BatchTile t;
            t.set = s;
            t.beg = (uint32_t)beg;
            t.len = (uint32_t)std::min<uint64_t>(tile, counts[s] - beg);
            t.pad = 0;
            b->htiles.push_back(t);
        }
    const uint32_t ntiles = (uint32_t)b->htiles.size();
    if (ntiles == 0) return 0;
    if (b->tiles.bytes < ntiles * sizeof(BatchTile)) {
        batch_drop_graph(b);
        HIPCHK(hipStreamSynchronize(st));
        dev_free(b->tiles);
        int rc = dev_alloc(b->tiles, ntiles * sizeof(BatchTile) * 5 / 4 + 64);
        if (rc) return rc;
    }
    HIPCHK(hipMemcpyAsync(b->sets.p, b->hsets.data(), ns * sizeof(SegArgs), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(b->tiles.p, b->htiles.data(), ntiles * sizeof(BatchTile), hipMemcpyHostToDevice, st));
    batch_fn kfn = batch_kernel(b->variant);
    const size_t lds = (kVariants[b->variant] & FEAT_MONO) ? (size_t)(wg / 64) * DQ_BYTES_PER_WAVE : 0;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kfn, (int)wg, lds) != hipSuccess || nb < 1) nb = 1;
    unsigned grid = (unsigned)ctx->n_cu * (unsigned)nb;
    uint64_t waves_needed = ntiles;
    grid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(grid, (waves_needed * 64 + wg - 1) / wg));
    if (!b->exec || b->g_grid != grid || b->g_wg != wg || b->g_ntiles != ntiles || b->g_sets != b->sets.p ||
        b->g_tiles != b->tiles.p) {
        batch_drop_graph(b);
        BatchArgs ba;
        ba.sets = (const SegArgs *)b->sets.p;
        ba.tiles = (const BatchTile *)b->tiles.p;
        ba.ntiles = ntiles;
        ba.pad = 0;
        ba.next = (uint32_t *)b->next.p;
        ba.cnt = (unsigned long long *)b->cnt.p;
        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        /* counters are cleared by kernels: no memset nodes in the graph */
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, (uint32_t *)b->next.p, (uint64_t)1);
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(256), 0, st, (unsigned long long *)b->cnt.p, (uint64_t)CNT_N);
        hipLaunchKernelGGL(kfn, dim3(grid), dim3(wg), lds, st, ba);
        hipError_t e = hipStreamEndCapture(st, &b->graph);
        if (e != hipSuccess) {
            b->graph = nullptr;
            return fail(RAMP_E_DEVICE, "batch graph capture: %s", hipGetErrorString(e));
        }
        e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            batch_drop_graph(b);
            return fail(RAMP_E_DEVICE, "batch graph instantiation: %s", hipGetErrorString(e));
        }
        b->g_grid = grid;
        b->g_wg = wg;
        b->g_ntiles = ntiles;
        b->g_sets = b->sets.p;
        b->g_tiles = b->tiles.p;
    }
    HIPCHK(hipEventRecord(ctx->ev[0], st));
    HIPCHK(hipGraphLaunch(b->exec, st));
    HIPCHK(hipEventRecord(ctx->ev[1], st));
    HIPCHK(hipEventSynchronize(ctx->ev[1]));
    std::vector<unsigned long long> cnt(CNT_N);
    HIPCHK(hipMemcpy(cnt.data(), b->cnt.p, CNT_N * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
    std::memset(&b->stats, 0, sizeof b->stats);
    b->stats.n_hist = total;
    b->stats.n_chunks = ns;
    b->stats.n_events = cnt[CNT_EVENTS];
    b->stats.n_oob = cnt[CNT_OOB];
    b->stats.n_det_overflow = cnt[CNT_OVERFLOW];
    b->stats.n_segments = 1;
    b->stats.ms_total = ms;
    b->stats.ms_source_segment = ms;
    b->last_wg = wg;
    b->last_npt = npt;
    b->last_grid = grid;
    return 0;
}

extern "C" int ramp_batch_stats_read(const ramp_batch *b, ramp_run_stats *out)
{
    if (!b || !out) return fail(RAMP_E_INVALID, "null argument");
    *out = b->stats;
    return 0;
}

#endif
