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

struct AdjArgs {
    int det;
    uint32_t pad;
    uint64_t nbins;
    const float *dLdh;
    double *partial;   /* [grid][ADJ_NPART] */
};

DEV double wave_sum_f64(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

/* d log w0 / dT1 of an MCont history (cont_mod.cl:37-47): w0 is
 * I1 M(E;T)/M(kT;T)/N, and M(E;T)/M(kT;T) = sqrt(E/kT) e^{1 - E/kT}; on
 * the 1/E tail w0 is proportional to T.  Only a single-temperature source
 * (the second and third terms not evaluated) has this form; otherwise 0. */
DEV float source_dlogw_dT(const NS &n, const DPart &s)
{
    if (s.type != RAMP_SRC_CONT || s.d[16] == 0.0f || s.d[17] == 0.0f) return 0.0f;
    float T = s.p[SP_T1];
    float v = v3_len(nvel(n));
    float E_J = RC_MEV2J * (RC_VS2E * (v * v));
    if (s.p[SP_TAIL_JOIN] > 0.0f && E_J > s.d[9]) return 1.0f / T;
    return (E_J * s.d[0] - 0.5f) / T;
}

template <uint32_t F>
__global__ __launch_bounds__(256) void k_adjoint(SegArgs a, AdjArgs x)
{
    __shared__ double red[4][ADJ_NG];
    __shared__ double lsum[4][ADJ_NLDS];
    AdjCtx hc;
    hc.dets = a.dets;
    hc.mode = 0;
    hc.enabled = 1;
    hc.events = hc.oob = hc.overflow = 0;
    hc.det = x.det;
    hc.nbins = x.nbins;
    hc.dLdh = x.dLdh;
#pragma unroll
    for (int i = 0; i < ADJ_NG; i++) hc.G[i] = 0.0f;
    adj_history_reset(hc, 0.0f);
    const uint32_t lane = __lane_id();
    const uint32_t wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (uint32_t i = lane; i < ADJ_NLDS; i += 64) lsum[wid][i] = 0.0;
    hc.la = lsum[wid];
    wave_lds_sync();
    const uint64_t G = (uint64_t)gridDim.x * (blockDim.x >> 6);
    const uint64_t g = (uint64_t)blockIdx.x * (blockDim.x >> 6) + wid;
    const uint64_t per = (a.n + G - 1) / G;
    const uint64_t beg = g * per;
    const uint64_t end = beg + per < a.n ? beg + per : a.n;
    const int ns = a.s1;
    for (uint64_t base = beg; base < end; base += 64) {
        const uint64_t i = base + lane;
        const bool valid = i < end;
        NS n;
        if (valid) {
            float sb_old = a.carry ? a.slot_ctr[i] : 0.0f;
            n.iidx = a.carry ? a.slot_iidx[i] : 0u;
            n.key = (uint32_t)i;
            src_generate<false, F>(n, plan_at(a.src, 0), a.num_sim, a.seed, sb_old);
        } else {
            n.dead = true;
            n.pend = false;
            n.key = 0;
            n.iidx = 0;
        }
        adj_history_reset(hc, valid ? source_dlogw_dT(n, plan_at(a.src, 0)) : 0.0f);
        for (int k = 0; k < ns; k++) {
            if (!__any(!n.dead)) break;
            const DStep st = plan_at(a.steps, k);
            if (st.kind == 0) {
                const DComp &c = plan_at(a.comps, st.first);
                tr_in<false>(n, c);
                geom_dispatch<F>(n, c);
                terminate(n, c.restore);
                scat_dispatch<false, F>(n, c, hc);
                tr_out<false>(n, c);
            } else {
                for (uint32_t ev = 0; ev < st.multi; ev++) {
                    for (uint32_t j = 0; j < st.count; j++) {
                        const DComp &c = plan_at(a.comps, st.first + j);
                        tr_in<false>(n, c);
                        geom_dispatch<F>(n, c);
                        tr_out<false>(n, c);
                    }
                    for (uint32_t j = 0; j < st.count; j++) {
                        const DComp &c = plan_at(a.comps, st.first + j);
                        tr_in<false>(n, c);
                        scat_dispatch<false, F>(n, c, hc);
                        tr_out<false>(n, c);
                    }
                }
            }
        }
        if (valid && a.carry) {
            a.slot_ctr[n.key] = n.ctr;
            a.slot_iidx[n.key] = n.iidx;
        }
    }
#pragma unroll
    for (int i = 0; i < ADJ_NG; i++) {
        double v = wave_sum_f64((double)hc.G[i]);
        if (lane == 0) red[wid][i] = v;
    }
    __syncthreads();
    const uint32_t nw = blockDim.x >> 6;
    for (uint32_t i = threadIdx.x; i < ADJ_NPART; i += blockDim.x) {
        double t = 0.0;
        if (i < ADJ_NG)
            for (uint32_t w = 0; w < nw; w++) t += red[w][i];
        else
            for (uint32_t w = 0; w < nw; w++) t += lsum[w][i - ADJ_NG];
        x.partial[(uint64_t)blockIdx.x * ADJ_NPART + i] = t;
    }
    uint32_t ev = wave_sum(hc.events), oob = wave_sum(hc.oob);
    if (lane == 0) {
        if (ev) atomicAdd(&a.cnt[CNT_EVENTS], (unsigned long long)ev);
        if (oob) atomicAdd(&a.cnt[CNT_OOB], (unsigned long long)oob);
    }
}

typedef void (*adj_fn)(SegArgs, AdjArgs);

static adj_fn adjoint_kernel(int v)
{
    /* the adjoint needs SPowderDiff (a sample feature); other features ride along */
    uint32_t f = kVariants[v] | FEAT_SAMPLE;
    if (f == FEAT_SAMPLE) return k_adjoint<FEAT_SAMPLE>;
    if (f == (FEAT_MESH | FEAT_SAMPLE)) return k_adjoint<FEAT_MESH | FEAT_SAMPLE>;
    return k_adjoint<FEAT_ALL>;
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
