/* adjoint_run.h - gradient of a histogram loss (no reference counterpart).
 *
 * ramp_adjoint re-runs the histories of a forward ramp_run (same plan,
 * counts, chunking and seeds, so every trajectory and weight is the same)
 * in one fused pass per chunk, and instead of filling the loss detector's
 * histogram it accumulates dL/dh[bin] * w * d log w / d theta (see
 * dev_adjoint.h for the terms): the scalar parameters per lane in FP32, the
 * per-line sums per wave in FP64 LDS.  At the end each wave reduces its
 * lanes in FP64 and each workgroup sums its waves in FP64 into one partial
 * per workgroup (the only FP64 arithmetic on the GPU); the host adds the
 * partials in FP64 in workgroup order.
 */
#ifndef RAMP_ADJOINT_RUN_H
#define RAMP_ADJOINT_RUN_H

#include "adjoint_kernel.h"

static adj_fn adjoint_kernel(int v)
{
    /* the adjoint needs SPowderDiff (a sample feature); other features ride along */
    const uint32_t f = kVariants[v] | FEAT_SAMPLE;
    for (int u = 0; u < kNumVariants; u++)
        if ((f & ~kVariants[u]) == 0 && kern_table(u).adjoint) return kern_table(u).adjoint;
    return kern_table(kNumVariants - 1).adjoint;
}

extern "C" int ramp_adjoint(ramp_plan *p, uint64_t n_hist, uint64_t chunk, const uint32_t *seeds,
                            uint32_t n_seeds, uint32_t det, const float *dL_dh, uint64_t n_bins,
                            double *grad, uint32_t n_grad, const ramp_run_opts *opts)
{
    if (!p) return fail(RAMP_E_INVALID, "null plan");
    if (!dL_dh || !grad) return fail(RAMP_E_INVALID, "null dL/dh or gradient buffer");
    if (p->need_full) return fail(RAMP_E_INVALID, "SMCPLIn/SMCPLOut/SRescal/SAnalyser plans run with ramp_run only");
    ramp_run_opts o;
    std::memset(&o, 0, sizeof o);
    if (opts) o = *opts;
    if (o.flags || o.keep_state || o.key_base || o.shard_total)
        return fail(RAMP_E_INVALID, "run flags 0x%x, keep_state and shards are not available in the adjoint", o.flags);
    if (det >= p->hdets.size()) return fail(RAMP_E_INVALID, "detector %u of %zu", det, p->hdets.size());
    if (n_bins != p->hdets[det].n)
        return fail(RAMP_E_INVALID, "detector %u has %llu bins, dL/dh has %llu", det,
                    (unsigned long long)p->hdets[det].n, (unsigned long long)n_bins);
    /* the differentiated sample: at most one SPowderDiff */
    int npd = 0;
    uint32_t nl = 0;
    for (const DComp &c : p->hcomps)
        if (c.s.type == RAMP_S_POWDERDIFF) { npd++; nl = c.s.ni[PDI_NLINES]; }
    if (npd > 1) return fail(RAMP_E_INVALID, "the adjoint handles one SPowderDiff per plan, found %d", npd);
    if (nl > ADJ_MAXL) return fail(RAMP_E_RANGE, "SPowderDiff with %u lines exceeds %d", nl, ADJ_MAXL);
    if (n_grad != 4u + nl)
        return fail(RAMP_E_INVALID, "gradient has %u entries: T, sigma_inc, sigma_abs, U and %u d-spacings",
                    4u + nl, nl);
    if (n_hist == 0) {
        std::fill(grad, grad + n_grad, 0.0);
        return 0;
    }
    if (chunk == 0 || chunk > n_hist) chunk = n_hist;
    if (chunk > 0x7fffffffull) return fail(RAMP_E_RANGE, "chunk %llu exceeds 2^31-1", (unsigned long long)chunk);
    uint64_t n_chunks = (n_hist + chunk - 1) / chunk;
    if (!seeds || n_seeds < n_chunks)
        return fail(RAMP_E_INVALID, "need %llu seeds, got %u", (unsigned long long)n_chunks, n_seeds);
    for (uint64_t c = 0; c < n_chunks; c++)
        if (seeds[c] > 0x7fffffffu) return fail(RAMP_E_INVALID, "seed %u out of range", seeds[c]);
    unsigned wg = o.wg ? o.wg : 256;
    if (wg != 64 && wg != 128 && wg != 256) return fail(RAMP_E_INVALID, "workgroup size %u", wg);
    ramp_ctx *ctx = p->ctx;
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    adj_fn kfn = adjoint_kernel(p->variant);
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kfn, (int)wg, 0) != hipSuccess || nb < 1) nb = 1;
    unsigned grid = (unsigned)ctx->n_cu * (unsigned)nb;
    grid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(grid, (chunk + wg - 1) / wg));
    DevBuf dl, part;
    int rc = dev_alloc(dl, n_bins * sizeof(float));
    if (!rc) rc = dev_alloc(part, (size_t)grid * ADJ_NPART * sizeof(double));
    if (rc) { dev_free(dl); dev_free(part); return rc; }
    bool carry = n_chunks > 1;
    if (carry) {
        rc = dev_alloc(p->slot_ctr, chunk * sizeof(float));
        if (!rc) rc = dev_alloc(p->slot_iidx, chunk * sizeof(uint32_t));
        if (rc) { dev_free(dl); dev_free(part); return rc; }
        launch_fill_f32(s, p->slot_ctr.p, chunk);
        launch_fill_u32(s, p->slot_iidx.p, chunk);
    }
    HIPCHK(hipMemcpyAsync(dl.p, dL_dh, n_bins * sizeof(float), hipMemcpyHostToDevice, s));
    SegArgs a;
    std::memset(&a, 0, sizeof a);
    a.revive_k = -1;
    a.comps = (const DComp *)p->comps.p;
    a.steps = (const DStep *)p->steps.p;
    a.src = (const DPart *)p->src.p;
    a.dets = (const DDet *)p->dets.p;
    a.cnt = (unsigned long long *)p->cnt.p;
    a.s0 = 0;
    a.s1 = (int)p->hsteps.size();
    a.last = 1;
    a.carry = carry ? 1 : 0;
    a.defer_k = -1;
    a.slot_ctr = (float *)p->slot_ctr.p;
    a.slot_iidx = (uint32_t *)p->slot_iidx.p;
    AdjArgs x;
    x.det = (int)det;
    x.pad = 0;
    x.nbins = n_bins;
    x.dLdh = (const float *)dl.p;
    x.partial = (double *)part.p;
    std::vector<double> acc(ADJ_NPART, 0.0), hp((size_t)grid * ADJ_NPART);
    launch_fill_u64(s, p->cnt.p, CNT_N);
    HIPCHK(hipEventRecord(ctx->ev[0], s));
    for (uint64_t ck = 0; ck < n_chunks; ck++) {
        uint64_t torun = std::min(chunk, n_hist - ck * chunk);
        a.num_sim = (int)torun;
        a.seed = (int)seeds[ck];
        a.n = torun;
        unsigned gr = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(grid, (torun + wg - 1) / wg));
        hipLaunchKernelGGL(kfn, dim3(gr), dim3(wg), 0, s, a, x);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(hp.data(), part.p, (size_t)gr * ADJ_NPART * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        for (unsigned b = 0; b < gr; b++)
            for (int i = 0; i < ADJ_NPART; i++) acc[i] += hp[(size_t)b * ADJ_NPART + i];
    }
    HIPCHK(hipEventRecord(ctx->ev[1], s));
    HIPCHK(hipEventSynchronize(ctx->ev[1]));
    std::vector<unsigned long long> cnt(CNT_N);
    HIPCHK(hipMemcpy(cnt.data(), p->cnt.p, CNT_N * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
    dev_free(dl);
    dev_free(part);
    for (int i = 0; i < ADJ_NG; i++) grad[i] = acc[i];
    /* line j: -ssv2_j dlogsig_j * sum_{m >= c_j} A[m] + B[j] */
    const double *A = acc.data() + ADJ_NG, *B = acc.data() + ADJ_NG + ADJ_MAXL + 1;
    for (uint32_t j = 0; j < nl; j++) {
        const float *Lj = p->pd_lines.data() + (size_t)j * PD_LW;
        uint32_t cj = 0;
        for (uint32_t i = 0; i < nl; i++) cj += p->pd_lines[(size_t)i * PD_LW + PDL_Q0] <= Lj[PDL_Q0] ? 1u : 0u;
        double suf = 0.0;
        for (uint32_t m = cj; m <= nl; m++) suf += A[m];
        grad[ADJ_NG + j] = -(double)Lj[PDL_SSV2] * (double)Lj[PDL_DLS] * suf + B[j];
    }
    std::memset(&p->stats, 0, sizeof p->stats);
    p->stats.n_hist = n_hist;
    p->stats.n_chunks = n_chunks;
    p->stats.n_events = cnt[CNT_EVENTS];
    p->stats.n_oob = cnt[CNT_OOB];
    p->stats.n_segments = n_chunks;
    p->stats.ms_total = ms;
    p->stats.ms_source_segment = ms;
    return 0;
}

#endif
