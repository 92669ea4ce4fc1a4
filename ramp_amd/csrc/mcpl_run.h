/* mcpl_run.h - ramp_mcpl_read: the records of one SMCPLOut component,
 * compacted on the device in history order; and the per-chunk append of
 * SRescal's detected events to their device-resident lists.
 *
 * The reference copies the whole neutron buffer of the chunk to the host
 * and filters it particle by particle in Python (mcpl_out.py:99-107).
 * Here lane i of the last chunk wrote record i and tagged it with the
 * chunk's tag; a three-kernel stream compaction keeps the tagged records in
 * order: per-wave counts over contiguous ranges, one exclusive scan, and a
 * second pass in which every wave writes its records at its offset (skipped
 * when the caller only asks for the count).  Only the compacted records
 * cross PCIe.  Included by ramp_gpu.hip ahead of ramp_run.
 */

/* waves of a compaction pass: the waves of its writing kernel that fit on
 * the GPU at once (a persistent grid; more would leave a partly filled last
 * round of blocks), but no more than one per SEL_GROUP candidates.  A wave
 * takes SEL_GROUP = 256 candidates per step, four per lane. */
#define SEL_GROUP 256u

static uint32_t sel_waves(const ramp_ctx *ctx, const void *kernel, uint64_t n)
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, 256, 0) != hipSuccess || nb < 1) nb = 1;
    const uint64_t resident = (uint64_t)ctx->n_cu * (uint64_t)nb * 4;
    const uint64_t need = (n + SEL_GROUP - 1) / SEL_GROUP;
    return (uint32_t)std::max<uint64_t>(1, std::min(resident, need));
}

/* waves over the n tags of one slot; each takes a contiguous range */
DEV void sel_range(uint64_t n, uint32_t G, uint32_t g, uint64_t &b, uint64_t &e)
{
    const uint64_t per = ((n + G - 1) / G + SEL_GROUP - 1) & ~(uint64_t)(SEL_GROUP - 1);
    b = (uint64_t)g * per;
    e = b + per < n ? b + per : n;
    if (b > n) b = n;
}

/* one group [i0, i0 + 64 * PER) of a wave's range: the offsets (from i0)
 * of its tagged entries, in order, go to src[0..cnt) in LDS; returns cnt */
template <int PER>
DEV uint32_t sel_group(const uint32_t *tags, uint64_t i0, uint64_t e, uint32_t tag, uint32_t *src)
{
    const uint32_t lane = __lane_id();
    bool h[PER];
    uint32_t c = 0;
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const uint64_t i = i0 + PER * lane + k;
        h[k] = i < e && tags[i] == tag;
        c += h[k] ? 1u : 0u;
    }
    uint32_t x = c; /* inclusive prefix over the lanes */
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= (uint32_t)o) x += y;
    }
    uint32_t r = x - c;
#pragma unroll
    for (int k = 0; k < PER; k++)
        if (h[k]) src[r++] = PER * lane + k;
    wave_lds_sync();
    return __shfl(x, 63);
}

__global__ __launch_bounds__(256) void k_sel_count(const uint32_t *tags, uint64_t n, uint32_t tag,
                                                   uint32_t G, uint32_t *counts)
{
    const uint32_t g = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (g >= G) return;
    uint64_t b, e;
    sel_range(n, G, g, b, e);
    uint32_t c = 0;
    for (uint64_t i = b + 4 * __lane_id(); i < e; i += SEL_GROUP)
#pragma unroll
        for (int k = 0; k < 4; k++) c += (i + k < e && tags[i + k] == tag) ? 1u : 0u;
    c = wave_sum(c);
    if (__lane_id() == 0) counts[g] = c;
}

/* exclusive prefix of G counts in one block; prefix[G] is the total */
__global__ __launch_bounds__(1024) void k_sel_scan(const uint32_t *counts, uint32_t G, uint64_t *prefix)
{
    __shared__ uint64_t part[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (G + 1023) / 1024;
    const uint32_t b = t * per, e = b + per < G ? b + per : G;
    uint64_t sum = 0;
    for (uint32_t i = b; i < e; i++) sum += counts[i];
    part[t] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        uint64_t v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint64_t run = part[t] - sum;
    for (uint32_t i = b; i < e; i++) {
        prefix[i] = run;
        run += counts[i];
    }
    if (t == 1023) prefix[G] = part[1023];
}

/* The records sit field-major in the dump (field k of record i at
 * k * cap + i); the output is the file's record-major layout.  Per group
 * of 128 candidates the wave ranks its hits into LDS (sel_group), gathers
 * their fields with coalesced loads into an LDS tile in record-major
 * order, and then streams the tile out, so that every load and store
 * instruction touches contiguous memory. */
#define SEL_WGROUP 128u
__global__ __launch_bounds__(256) void k_sel_write(const uint32_t *tags, const float *recs, uint64_t n,
                                                   uint64_t cap, uint32_t tag, uint32_t G,
                                                   const uint64_t *prefix, float *out)
{
    __shared__ uint32_t src[4][SEL_WGROUP];
    __shared__ float tile[4][SEL_WGROUP * MCR_N];
    const uint32_t wv = threadIdx.x >> 6;
    const uint32_t g = blockIdx.x * (blockDim.x >> 6) + wv;
    if (g >= G) return;
    uint64_t b, e;
    sel_range(n, G, g, b, e);
    uint64_t o = prefix[g];
    const uint32_t lane = __lane_id();
    for (uint64_t base = b; base < e; base += SEL_WGROUP) {
        const uint32_t cnt = sel_group<2>(tags, base, e, tag, src[wv]);
        /* a lane gathers its (at most two) records' fields with all loads in
         * flight at once, then writes them to the tile */
        float v[2][MCR_N];
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const uint32_t r = lane + 64u * q;
            if (r < cnt) {
                const float *rp = recs + base + src[wv][r];
#pragma unroll
                for (int k = 0; k < MCR_N; k++) v[q][k] = rp[(size_t)k * cap];
            }
        }
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const uint32_t r = lane + 64u * q;
            if (r < cnt) {
#pragma unroll
                for (int k = 0; k < MCR_N; k++) tile[wv][r * MCR_N + k] = v[q][k];
            }
        }
        wave_lds_sync();
        float *w = out + (size_t)MCR_N * o;
        for (uint32_t j = lane; j < cnt * MCR_N; j += 64) w[j] = tile[wv][j];
        wave_lds_sync();
        o += cnt;
    }
}

/* SRescal: after each chunk, the events of its histories tagged as
 * detected are appended, in history order, to the slot's field-major list
 * [5][list_cap] at the device-held count *list_n (rescal.py:51-62). */
__global__ __launch_bounds__(256) void k_ev_append(const uint32_t *tags, const float *ev, uint64_t n,
                                                   uint64_t cap, uint32_t tag, uint32_t G,
                                                   const uint64_t *prefix, float *list, uint64_t list_cap,
                                                   const unsigned long long *list_n)
{
    __shared__ uint32_t src[4][SEL_GROUP];
    const uint32_t wv = threadIdx.x >> 6;
    const uint32_t g = blockIdx.x * (blockDim.x >> 6) + wv;
    if (g >= G) return;
    uint64_t b, e;
    sel_range(n, G, g, b, e);
    uint64_t o = *list_n + prefix[g];
    const uint32_t lane = __lane_id();
    for (uint64_t base = b; base < e; base += SEL_GROUP) {
        const uint32_t cnt = sel_group<4>(tags, base, e, tag, src[wv]);
        /* a lane gathers its (at most four) events with all loads in flight
         * at once, then stores them */
        float v[4][RS_NEV];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint32_t r = lane + 64u * q;
            if (r < cnt) {
                const float *ep = ev + base + src[wv][r];
#pragma unroll
                for (int f = 0; f < RS_NEV; f++) v[q][f] = ep[(size_t)f * cap];
            }
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint32_t r = lane + 64u * q;
            if (r < cnt) {
#pragma unroll
                for (int f = 0; f < RS_NEV; f++) list[(size_t)f * list_cap + o + r] = v[q][f];
            }
        }
        wave_lds_sync();
        o += cnt;
    }
}

__global__ void k_ev_advance(unsigned long long *list_n, const uint64_t *total)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *list_n += *total;
}

/* the append of one chunk's events, for every SRescal slot (stream order) */
static int rescal_append(ramp_plan *p, uint64_t n, uint32_t tag, hipStream_t s)
{
    const uint32_t G = sel_waves(p->ctx, (const void *)k_ev_append, n);
    int rc = dev_alloc(p->sel_counts, G * sizeof(uint32_t));
    if (!rc) rc = dev_alloc(p->sel_prefix, (G + 1) * sizeof(uint64_t));
    if (rc) return rc;
    const unsigned blocks = (G + 3) / 4;
    for (uint32_t sl = 0; sl < p->n_rescal; sl++) {
        const uint32_t *tags = (const uint32_t *)p->ev_tag.p + (size_t)sl * p->ev_cap;
        const float *ev = (const float *)p->ev.p + (size_t)sl * RS_NEV * p->ev_cap;
        float *list = (float *)p->ev_list.p + (size_t)sl * RS_NEV * p->ev_list_cap;
        unsigned long long *ln = (unsigned long long *)p->ev_list_n.p + sl;
        hipLaunchKernelGGL(k_sel_count, dim3(blocks), dim3(256), 0, s, tags, n, tag, G,
                           (uint32_t *)p->sel_counts.p);
        hipLaunchKernelGGL(k_sel_scan, dim3(1), dim3(1024), 0, s, (const uint32_t *)p->sel_counts.p, G,
                           (uint64_t *)p->sel_prefix.p);
        hipLaunchKernelGGL(k_ev_append, dim3(blocks), dim3(256), 0, s, tags, ev, n, p->ev_cap, tag, G,
                           (const uint64_t *)p->sel_prefix.p, list, p->ev_list_cap, ln);
        hipLaunchKernelGGL(k_ev_advance, dim3(1), dim3(64), 0, s, ln, (const uint64_t *)p->sel_prefix.p + G);
        HIPCHK(hipGetLastError());
    }
    return 0;
}

/* event buffers of a run: per-chunk events and tags (stride `chunk`), and
 * lists with room for every history of this run on top of what they hold */
static int rescal_prepare(ramp_plan *p, uint64_t chunk, uint64_t n_hist, hipStream_t s)
{
    const uint32_t ns = p->n_rescal;
    if (p->ev_cap < chunk) {
        int rc = dev_alloc(p->ev, (size_t)ns * RS_NEV * chunk * sizeof(float));
        if (!rc) rc = dev_alloc(p->ev_tag, (size_t)ns * chunk * sizeof(uint32_t));
        if (rc) return rc;
        launch_fill_u32(s, p->ev_tag.p, (uint64_t)ns * chunk);
        p->ev_cap = chunk;
        for (uint32_t sl = 0; sl < ns; sl++) {
            DPart &d = p->hcomps[p->rescal_comp[sl]].s;
            d.aux = (const float *)p->ev.p + (size_t)sl * RS_NEV * chunk;
            d.ni[RSI_CAP_LO] = (uint32_t)chunk;
            d.ni[RSI_CAP_HI] = (uint32_t)(chunk >> 32);
        }
        HIPCHK(hipMemcpyAsync(p->comps.p, p->hcomps.data(), p->hcomps.size() * sizeof(DComp),
                              hipMemcpyHostToDevice, s));
        dev_free(p->comps_brute); /* rebuilt from hcomps when asked for */
    }
    if (!p->ev_list_n.p) {
        int rc = dev_alloc(p->ev_list_n, ns * sizeof(unsigned long long));
        if (rc) return rc;
        launch_fill_u64(s, p->ev_list_n.p, ns);
    }
    uint64_t need = n_hist;
    for (uint64_t c : p->ev_count) need = std::max(need, c + n_hist);
    if (p->ev_list_cap < need) {
        const uint64_t cap = std::max(need, 2 * p->ev_list_cap);
        DevBuf nb;
        int rc = dev_alloc(nb, (size_t)ns * RS_NEV * cap * sizeof(float));
        if (rc) return rc;
        for (uint32_t sl = 0; sl < ns; sl++)
            for (int f = 0; f < RS_NEV && p->ev_count[sl]; f++)
                HIPCHK(hipMemcpyAsync((float *)nb.p + ((size_t)sl * RS_NEV + f) * cap,
                                      (const float *)p->ev_list.p + ((size_t)sl * RS_NEV + f) * p->ev_list_cap,
                                      p->ev_count[sl] * sizeof(float), hipMemcpyDeviceToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
        dev_free(p->ev_list);
        p->ev_list = nb;
        nb.p = nullptr;
        p->ev_list_cap = cap;
    }
    return 0;
}

extern "C" int ramp_rescal_read(ramp_plan *p, uint32_t comp, float *events, uint64_t cap, uint64_t *count)
{
    if (!p || !count) return fail(RAMP_E_INVALID, "null argument");
    if (comp >= p->n_rescal) return fail(RAMP_E_INVALID, "SRescal %u of %u", comp, p->n_rescal);
    if (cap && !events) return fail(RAMP_E_INVALID, "null event buffer");
    HIPCHK(hipSetDevice(p->ctx->device));
    HIPCHK(hipStreamSynchronize(p->ctx->stream));
    *count = p->ev_count[comp];
    const uint64_t nc = std::min(cap, *count);
    for (int f = 0; f < RS_NEV && nc; f++)
        HIPCHK(hipMemcpy(events + (size_t)f * cap,
                         (const float *)p->ev_list.p + ((size_t)comp * RS_NEV + f) * p->ev_list_cap,
                         nc * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int ramp_mcpl_read(ramp_plan *p, uint32_t comp, float *records, uint64_t cap, uint64_t *count)
{
    if (!p || !count) return fail(RAMP_E_INVALID, "null argument");
    if (comp >= p->n_mcpl_out) return fail(RAMP_E_INVALID, "SMCPLOut %u of %u", comp, p->n_mcpl_out);
    if (cap && !records) return fail(RAMP_E_INVALID, "null record buffer");
    *count = 0;
    if (!p->dump.p || p->dump_n == 0) return 0;
    HIPCHK(hipSetDevice(p->ctx->device));
    hipStream_t s = p->ctx->stream;
    const uint64_t n = p->dump_n;
    const uint32_t G = sel_waves(p->ctx, (const void *)k_sel_write, n);
    int rc = dev_alloc(p->sel_counts, G * sizeof(uint32_t));
    if (!rc) rc = dev_alloc(p->sel_prefix, (G + 1) * sizeof(uint64_t));
    if (rc) return rc;
    const uint32_t *tags = (const uint32_t *)p->dump_tag.p + (size_t)comp * p->dump_cap;
    const float *recs = (const float *)p->dump.p + (size_t)comp * MCR_N * p->dump_cap;
    const uint32_t tag = p->dump_chunk_tag;
    const unsigned blocks = (G + 3) / 4;
    ramp_ctx *ctx = p->ctx;
    HIPCHK(hipEventRecord(ctx->ev[0], s));
    hipLaunchKernelGGL(k_sel_count, dim3(blocks), dim3(256), 0, s, tags, n, tag, G,
                       (uint32_t *)p->sel_counts.p);
    hipLaunchKernelGGL(k_sel_scan, dim3(1), dim3(1024), 0, s, (const uint32_t *)p->sel_counts.p, G,
                       (uint64_t *)p->sel_prefix.p);
    HIPCHK(hipEventRecord(ctx->ev[1], s));
    HIPCHK(hipGetLastError());
    uint64_t total = 0;
    HIPCHK(hipMemcpyAsync(&total, (const uint64_t *)p->sel_prefix.p + G, sizeof total, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    *count = total;
    if (total && cap) {
        rc = dev_alloc(p->sel_out, total * MCR_N * sizeof(float));
        if (rc) return rc;
        HIPCHK(hipEventRecord(ctx->ev[2], s));
        hipLaunchKernelGGL(k_sel_write, dim3(blocks), dim3(256), 0, s, tags, recs, n, p->dump_cap, tag,
                           G, (const uint64_t *)p->sel_prefix.p, (float *)p->sel_out.p);
        HIPCHK(hipEventRecord(ctx->ev[3], s));
        HIPCHK(hipGetLastError());
    }
    const uint64_t ncopy = std::min(total, cap);
    if (ncopy)
        HIPCHK(hipMemcpyAsync(records, p->sel_out.p, ncopy * MCR_N * sizeof(float), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    float ms = 0, ms_w = 0;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
    if (total && cap) HIPCHK(hipEventElapsedTime(&ms_w, ctx->ev[2], ctx->ev[3]));
    p->stats.ms_mcpl_compact = ms + ms_w;
    p->stats.ms_mcpl_write = ms_w;
    return 0;
}
