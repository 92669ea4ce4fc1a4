/* dev_queue.h - the structure-of-arrays neutron queue between segment
 * launches: one neutron's load and store, and the map from a dense index to
 * its slot among the producing segment's per-wave regions.  Shared by the
 * segment kernels and by run-time compiled plugin kernels. */
#ifndef RAMP_DEV_QUEUE_H
#define RAMP_DEV_QUEUE_H

/* status word of the queue: iidx in bits 0..23, bit 24: intersection is all
 * zeros (GBox miss with restore) instead of the sentinel.  The full-state
 * mode stores the whole intersection instead (components that leave it as
 * their geometry set it, such as the MCPL steps, need that).  Bit 25: the
 * neutron died in a plugin step, which rewrites the queue in place; the
 * next segment finalizes it. */
#define ST_ZERO_ISECT (1u << 24)
#define ST_DEAD (1u << 25)

template <bool FULL>
DEV void queue_load(NS &n, const DQueue &q, uint64_t i)
{
    n.x = q.x[i]; n.y = q.y[i]; n.z = q.z[i];
    n.vx = q.vx[i]; n.vy = q.vy[i]; n.vz = q.vz[i];
    n.w = q.w[i]; n.t = q.t[i]; n.ctr = q.ctr[i];
    n.key = q.key[i];
    uint32_t st = q.status[i];
    n.iidx = st & 0x00ffffffu;
    n.dead = FULL && (st & ST_DEAD) != 0u;
    n.pend = false;
    n.aux = 0u;
    if (FULL) {
        for (int k = 0; k < 8; k++) n.is[k] = q.is[k][i];
    } else if (st & ST_ZERO_ISECT) {
        for (int k = 0; k < 8; k++) n.is[k] = 0.0f;
    } else {
        isect_reset(n);
    }
    if (FULL) {
        n.px = q.px[i]; n.py = q.py[i]; n.pz = q.pz[i];
        n.sc = q.sc[i]; n.sd = q.sd[i]; n.se = q.se[i];
    } else {
        n.px = n.py = n.pz = n.sc = n.sd = n.se = 0.0f;
    }
}

template <bool FULL>
DEV void queue_store(const NS &n, const DQueue &q, uint64_t i)
{
    q.x[i] = n.x; q.y[i] = n.y; q.z[i] = n.z;
    q.vx[i] = n.vx; q.vy[i] = n.vy; q.vz[i] = n.vz;
    q.w[i] = n.w; q.t[i] = n.t; q.ctr[i] = n.ctr;
    q.key[i] = n.key;
    bool zero = (n.is[3] == 0.0f && n.is[7] == 0.0f);
    q.status[i] = (n.iidx & 0x00ffffffu) | (zero ? ST_ZERO_ISECT : 0u) | ((FULL && n.dead) ? ST_DEAD : 0u);
    if (FULL) {
        q.px[i] = n.px; q.py[i] = n.py; q.pz[i] = n.pz;
        q.sc[i] = n.sc; q.sd[i] = n.sd; q.se[i] = n.se;
        for (int k = 0; k < 8; k++) q.is[k][i] = n.is[k];
    }
}

/* Map dense input indices to slots in the previous segment's wave
 * regions (region r holds inputs [prefix[r], prefix[r+1]) at slots
 * r * stride + ...).  A region may be empty; the region of v is the last
 * one whose prefix is <= v. */

/* the region of v, given prefix[lo] <= v < prefix[hi] */
DEV uint64_t region_find(const uint64_t *prefix, uint64_t lo, uint64_t hi, uint64_t v)
{
    while (hi - lo > 1) {
        uint64_t mid = (lo + hi) >> 1;
        if (prefix[mid] <= v) lo = mid;
        else hi = mid;
    }
    return lo;
}

/* the region of v from a region r with prefix[r] <= v: steps of 1, 2,
 * 4 ... until past v, then bisection (prefix[nreg] is the total, > v) */
DEV uint64_t region_advance(const uint64_t *prefix, uint64_t nreg, uint64_t r, uint64_t v)
{
    uint64_t step = 1;
    while (r + step < nreg && prefix[r + step] <= v) step <<= 1;
    return region_find(prefix, r + (step >> 1), r + step < nreg ? r + step : nreg, v);
}

/* A wave walks its inputs in order, 64 at a time: `cur` (wave-uniform)
 * is the region of the group's first input `first`, advanced from the
 * previous group's; each lane then looks up its own input v among the few
 * regions up to the group's last input `last`. */
DEV uint64_t region_slot_group(const uint64_t *prefix, uint64_t nreg, uint64_t stride, uint64_t &cur,
                               uint64_t first, uint64_t last, uint64_t v)
{
    cur = region_advance(prefix, nreg, cur, first);
    const uint64_t r1 = region_advance(prefix, nreg, cur, last);
    const uint64_t r = region_find(prefix, cur, r1 + 1, v);
    return r * stride + (v - prefix[r]);
}

/* Trace points (full-state runs with RAMP_RUN_TRACE): for traced component
 * `slot`, field-major [slot][4][cap] at the history's run-wide index h,
 * the intersection's entry point (x1, y1, z1) in the component frame after
 * terminate, and 1.0 if the neutron is alive after the component.  The
 * buffer is zero-filled before the run, so histories that never reach the
 * component read "not alive". */
DEV void trace_entry(const NS &n, float *tr, uint64_t cap, uint64_t h, int slot)
{
    float *o = tr + (size_t)slot * 4 * cap + h;
    o[0] = n.is[0];
    o[cap] = n.is[1];
    o[2 * cap] = n.is[2];
}

DEV void trace_alive(const NS &n, float *tr, uint64_t cap, uint64_t h, int slot)
{
    tr[(size_t)slot * 4 * cap + 3 * cap + h] = n.dead ? 0.0f : 1.0f;
}

#endif
