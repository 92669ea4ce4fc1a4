/* batch_kernel.h - k_batch<F>: the persistent work-queue kernel of a
 * batched run (host side and the scheduling description in batch_run.h).
 * Instances are compiled per feature set in kern_v.hip. */
#ifndef RAMP_BATCH_KERNEL_H
#define RAMP_BATCH_KERNEL_H

#include "seg_kernel.h"

struct BatchTile {
    uint32_t set;
    uint32_t beg;
    uint32_t len;
    uint32_t pad;
};

struct BatchArgs {
    const SegArgs *sets;
    const BatchTile *tiles;
    uint32_t ntiles;
    uint32_t pad;
    uint32_t *next;
    unsigned long long *cnt;
};

/* A wave takes tiles from the work queue and runs each 64-history group
 * of the tile through the setting's whole plan in one fused pass.  Its
 * deferral queue (monochromator reflections, see run_range) is drained
 * whenever it holds 64 and at the end of each tile, so it only ever holds
 * neutrons of one setting. */
template <uint32_t F>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu((F) == FEAT_MONO ? RAMP_W_BATCH_MONO : 1)))
void k_batch(BatchArgs b)
{
    extern __shared__ float batch_lds[];
    HistCtx hc;
    hc.enabled = true;
    hc.events = hc.oob = hc.overflow = 0;
    hc.dets = nullptr;
    hc.mode = 0;
    hc.acc = nullptr;  /* settings change per tile: global atomics */
    const uint32_t lane = __lane_id();
    DeferQ dq;
    dq.b = batch_lds;
    dq.n = 0;
    if constexpr ((F & FEAT_MONO) != 0) defer_init(dq, batch_lds, DQ_CAP * DQ_WORDS);
    /* every wave leaves once the counter has passed ntiles */
    for (;;) {
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(b.next, 1u);
        t = __shfl(t, 0);
        if (t >= b.ntiles) break;
        const BatchTile tl = b.tiles[t];
        const SegArgs &a = b.sets[tl.set];
        bool defer = false;
        if constexpr ((F & FEAT_MONO) != 0) defer = a.defer_k >= 0;
        hc.dets = a.dets;
        hc.mode = a.hist_mode;
        for (uint32_t o = 0; o < tl.len; o += 64) {
            const uint32_t j = o + lane;
            NS n;
            if (j < tl.len) {
                n.iidx = 0u;
                n.key = tl.beg + j;
                src_generate<false, F>(n, plan_at(a.src, 0), a.num_sim, a.seed, 0.0f);
            } else {
                n.dead = true;
                n.pend = false;
                n.key = 0;
                n.iidx = 0;
            }
            run_range<false, F>(n, a, hc, a.s0, a.s1, dq, defer);
            if constexpr ((F & FEAT_MONO) != 0) {
                if (defer && dq.n >= 64u) {
                    NS r;
                    defer_pop<DQ_CAP, DQ_WORDS>(dq, r);
                    defer_finish<F>(r, a, hc, a.s1, dq);
                }
            }
        }
        if constexpr ((F & FEAT_MONO) != 0) {
            while (defer && dq.n > 0u) {
                NS r;
                defer_pop<DQ_CAP, DQ_WORDS>(dq, r);
                defer_finish<F>(r, a, hc, a.s1, dq);
            }
        }
    }
    uint32_t ev = wave_sum(hc.events), oob = wave_sum(hc.oob), ovf = wave_sum(hc.overflow);
    if (lane == 0) {
        if (ev) atomicAdd(&b.cnt[CNT_EVENTS], (unsigned long long)ev);
        if (oob) atomicAdd(&b.cnt[CNT_OOB], (unsigned long long)oob);
        if (ovf) atomicAdd(&b.cnt[CNT_OVERFLOW], (unsigned long long)ovf);
    }
}

typedef void (*batch_fn)(BatchArgs);

#endif
