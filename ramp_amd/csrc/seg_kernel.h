/* seg_kernel.h - the fused segment kernel k_segment<FULL, FIRST, F> and the
 * per-wave machinery it shares with the batched and adjoint kernels
 * (monochromator deferral queue, step loop, queue tail).  Instances are
 * compiled per feature set in kern_v.hip. */
#ifndef RAMP_SEG_KERNEL_H
#define RAMP_SEG_KERNEL_H

#include "dev_components.h"

/* ======================================================================== */
/* segment kernel                                                           */
/* ======================================================================== */
#define WALK_EMIT 1u
#define WALK_USE 2u

/* a walk's outcome (mesh_walk_code, or 3: the ray missed the mesh's box)
 * with its entering and exiting hit distances, one 16-byte record per
 * queue slot */
struct WalkHit {
    float t1, t2;
    uint32_t code;
    uint32_t pad;
};

/* a ray listed for k_mesh_walk: position and velocity in the mesh's frame,
 * and its queue slot (32 bytes, two 16-byte loads) */
struct WalkEntry {
    float pos[3];
    uint32_t slot;
    float vel[3];
    uint32_t pad;
};

struct SegArgs {
    const DComp *__restrict__ comps;
    const DStep *__restrict__ steps;
    const DPart *__restrict__ src;
    const DDet *__restrict__ dets;
    unsigned long long *cnt;
    int s0, s1;
    int last;
    int pilot;
    int hist_mode;
    int carry;
    int keep;
    int defer_k;                /* step whose monochromator reflections are deferred (-1: none) */
    uint32_t defer_lds;         /* LDS bytes per wave of the deferral queue (0: none) */
    /* mesh walks in their own launch (k_mesh_walk, fast mode).  WALK_EMIT:
     * the segment prepares each output neutron for the walk against
     * component walk_ci (its walk-list entries per wave in walk_counts;
     * a ray that misses the mesh's box gets outcome 3 at once).  WALK_USE:
     * the segment starts at that GMesh step and takes the hits instead of
     * walking.  walk_res: a WalkHit per queue slot; walk_list: a
     * WalkEntry per listed ray, walk_keys its ordering key.
     * Zero-initialised arguments switch both off. */
    uint32_t walk_role;
    uint32_t walk_ci;
    struct WalkHit *walk_res;
    struct WalkEntry *walk_list;
    uint32_t *walk_keys;
    uint32_t *walk_counts;
    int num_sim;
    int seed;
    uint64_t n;                 /* items of the first segment */
    /* input of later segments: wave regions of the previous segment */
    const uint64_t *in_prefix;  /* [n_in + 1] exclusive prefix of region counts */
    const uint64_t *in_meta;    /* [0] region stride, [1] number of regions */
    /* output: one region per wave, stride written to out_meta[0] */
    uint32_t *out_counts;
    uint64_t *out_meta;
    DQueue in, out;
    float *slot_ctr;
    uint32_t *slot_iidx;
    float *st16, *st8;
    uint32_t *stii;
    /* MCPL steps: SMCPLOut records [slot][dump_cap][MCR_N] with their chunk
     * tags; revive_k is the last SMCPLIn step (-1: none), which every lane
     * must reach even when all of its neutrons are dead */
    float *dump;
    uint32_t *dump_tag;
    uint64_t dump_cap;
    uint32_t chunk_tag;
    int revive_k;
    /* SRescal: a finished neutron with se > 0 tags its events [slot][ev_cap] */
    uint32_t *ev_tag;
    uint64_t ev_cap;
    uint32_t n_rescal;
    /* small detectors accumulated in LDS (DDet.acc >= 0): their bins and
     * the bytes after the deferral queues; the detector count */
    uint32_t acc_bins, acc_bytes;
    int n_dets;
    /* trace points [slot][4][trace_cap] (full-state, RAMP_RUN_TRACE), or
     * null; trace_base is the run-wide index of the chunk's first history */
    float *trace;
    uint64_t trace_cap, trace_base;
    /* a shard of a chunk: the first history's key (its chunk-local index) */
    uint64_t key_base;
};

#include "dev_queue.h"

/* Deferral of the monochromator's reflection branch.  s_mono is cheap
 * for the many neutrons it does not reflect and long (mosaic search,
 * 15-point quadrature, ~100 transcendentals) for the few it reflects, so
 * in a wave of 64 that branch runs at one or two live lanes in most groups.
 * At the deferred step (SegArgs.defer_k) the reflecting neutrons instead
 * park in a per-wave LDS queue (state in the slab frame, slab in aux);
 * when 64 are waiting the wave runs s_mono_b and the remaining steps for
 * all of them at once.  Each neutron's arithmetic and draws are unchanged,
 * only the order in which a wave processes them differs. */
#define DQ_CAP 128
#define DQ_WORDS 12
#define DQ_BYTES_PER_WAVE (DQ_CAP * DQ_WORDS * 4)

struct DeferQ {
    float *b;     /* this wave's words x capacity of LDS (wave-uniform) */
    uint32_t n;   /* parked neutrons, wave-uniform, below the capacity */
};

/* order this wave's LDS accesses: the compiler may not move loads of
 * other lanes' entries across it (a wave executes LDS operations in order) */
DEV void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

DEV void defer_init(DeferQ &d, float *lds, uint32_t words_per_wave)
{
    const uint32_t wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    d.b = lds + wid * words_per_wave;
    d.n = 0;
}

/* the record of a parked neutron: the segment-queue record (x y z vx vy vz
 * w t ctr key status) and aux (the slab of a monochromator; for a mesh,
 * bit 0: the intersection was all zeros, bits 1-2 the walk's outcome) */
template <uint32_t CAP>
DEV void defer_store(const NS &n, float *b, uint32_t i)
{
    b[0 * CAP + i] = n.x; b[1 * CAP + i] = n.y; b[2 * CAP + i] = n.z;
    b[3 * CAP + i] = n.vx; b[4 * CAP + i] = n.vy; b[5 * CAP + i] = n.vz;
    b[6 * CAP + i] = n.w; b[7 * CAP + i] = n.t; b[8 * CAP + i] = n.ctr;
    uint32_t *u = (uint32_t *)b;
    u[9 * CAP + i] = n.key;
    u[10 * CAP + i] = n.iidx & 0x00ffffffu;
    u[11 * CAP + i] = n.aux;
}

template <uint32_t CAP>
DEV void defer_load(NS &n, const float *b, uint32_t i)
{
    n.x = b[0 * CAP + i]; n.y = b[1 * CAP + i]; n.z = b[2 * CAP + i];
    n.vx = b[3 * CAP + i]; n.vy = b[4 * CAP + i]; n.vz = b[5 * CAP + i];
    n.w = b[6 * CAP + i]; n.t = b[7 * CAP + i]; n.ctr = b[8 * CAP + i];
    const uint32_t *u = (const uint32_t *)b;
    n.key = u[9 * CAP + i];
    n.iidx = u[10 * CAP + i];
    n.aux = u[11 * CAP + i];
    n.dead = false;
    n.pend = true;
    isect_reset(n);
    n.px = n.py = n.pz = n.sc = n.sd = n.se = 0.0f;
}

/* Steps [k0, k1); with defer, the step defer_k parks reflecting neutrons
 * in dq (they leave the range dead with pend set). */
template <bool FULL, uint32_t F>
DEV void run_range(NS &n, const SegArgs &a, HistCtx &hc, int k0, int k1, DeferQ &dq, bool defer)
{
    for (int k = k0; k < k1; k++) {
        /* SMCPLIn revives dead slots; its plans run in full-state mode only */
        if (!a.pilot && (!FULL || k > a.revive_k) && !__any(!n.dead)) break;
        const DStep st = plan_at(a.steps, k);
        if (st.kind == 0) {
            const DComp &c = plan_at(a.comps, st.first);
            tr_in<FULL>(n, c);
            /* a mesh step at the start of a segment after a walk launch
             * takes the walk's hits (at the neutron's input queue slot,
             * held in aux) instead of walking (k_mesh_walk) */
            bool walked = false;
            if constexpr ((F & FEAT_MESH) != 0) {
                if ((a.walk_role & WALK_USE) != 0u && k == a.s0 && !n.dead) {
                    const WalkHit h = a.walk_res[n.aux];
                    if (h.code == 1u || h.code == 2u) mesh_commit(n, npos(n), nvel(n), h.t1, h.t2, h.code, c.comp_idx);
                    walked = true;
                }
            }
            if (!walked) geom_dispatch<F>(n, c);
            terminate(n, c.restore);
            bool traced = false, mcpl = false;
            if constexpr (FULL) {
                traced = a.trace && !a.pilot && c.trace >= 0 && n.key != 0xffffffffu;
                if (traced) trace_entry(n, a.trace, a.trace_cap, a.trace_base + n.key, c.trace);
                /* plans holding MCPL components run in full-state mode only */
                mcpl = c.s.type == RAMP_S_MCPL_IN || c.s.type == RAMP_S_MCPL_OUT;
                if (mcpl) {
                    if (c.s.type == RAMP_S_MCPL_IN) mcpl_in_step(n, c, (uint32_t)a.seed);
                    else if (!a.pilot) mcpl_out_step(n, c, a.dump, a.dump_tag, a.dump_cap, a.chunk_tag);
                    tr_out<FULL>(n, c);
                }
            }
            /* (the pilot never defers: the deferred step's live count is
             * taken at its end like every other step's) */
            if constexpr (!FULL && (F & FEAT_MONO) != 0) {
                if (defer && k == a.defer_k) {
                    bool park = false;
                    if (n.iidx == c.comp_idx && !n.dead) park = s_mono_a(n, c, n.aux);
                    if (!park) tr_out<FULL>(n, c);
                    const unsigned long long m = __ballot(park);
                    if (m) {
                        if (park) {
                            const unsigned long long below = (1ull << __lane_id()) - 1ull;
                            defer_store<DQ_CAP>(n, dq.b, dq.n + (uint32_t)__popcll(m & below));
                            n.pend = true;
                            n.dead = true;
                        }
                        dq.n += (uint32_t)__popcll(m);
                    }
                    continue;
                }
            }
            if (!mcpl) {
                scat_dispatch<FULL, F>(n, c, hc);
                tr_out<FULL>(n, c);
            }
            if constexpr (FULL) {
                if (traced) trace_alive(n, a.trace, a.trace_cap, a.trace_base + n.key, c.trace);
            }
        } else if constexpr ((F & FEAT_NLIN) != 0u) {
            /* (a plan with a non-linear block always runs an instance that
             * carries FEAT_NLIN: the host picks the instance from the plan's
             * feature bits, so other instances never meet such a step) */
            for (uint32_t ev = 0; ev < st.multi; ev++) {
                for (uint32_t j = 0; j < st.count; j++) {
                    const DComp &c = plan_at(a.comps, st.first + j);
                    tr_in<FULL>(n, c);
                    geom_dispatch<F>(n, c);
                    tr_out<FULL>(n, c);
                }
                for (uint32_t j = 0; j < st.count; j++) {
                    const DComp &c = plan_at(a.comps, st.first + j);
                    tr_in<FULL>(n, c);
                    scat_dispatch<FULL, F>(n, c, hc);
                    tr_out<FULL>(n, c);
                }
            }
        }
        if (a.pilot) {
            unsigned long long m = __ballot(!n.dead);
            if (__lane_id() == 0 && m) atomicAdd(&a.cnt[CNT_NSTEP0 + k], (unsigned long long)__popcll(m));
        }
    }
}

template <bool FULL, uint32_t F>
DEV void run_steps(NS &n, const SegArgs &a, HistCtx &hc, DeferQ &dq, bool defer)
{
    run_range<FULL, F>(n, a, hc, a.s0, a.s1, dq, defer);
}

/* Take up to 64 parked neutrons (lane i gets entry i); returns how many.
 * The entries beyond them move down, 64 per round. */
template <uint32_t CAP, uint32_t WORDS>
DEV uint32_t defer_pop(DeferQ &dq, NS &r)
{
    const uint32_t lane = __lane_id();
    const uint32_t take = dq.n < 64u ? dq.n : 64u;
    wave_lds_sync();
    if (lane < take) {
        defer_load<CAP>(r, dq.b, lane);
    } else {
        r.dead = true;
        r.pend = false;
        r.key = 0;
        r.iidx = 0;
    }
    wave_lds_sync();
    uint32_t *u = (uint32_t *)dq.b;
    for (uint32_t j0 = 0; j0 < dq.n - take; j0 += 64u) {
        const uint32_t j = j0 + lane;
        if (j < dq.n - take)
            for (uint32_t f = 0; f < WORDS; f++) u[f * CAP + j] = u[f * CAP + take + j];
        wave_lds_sync();
    }
    dq.n -= take;
    return take;
}

/* reflection half of the deferred step, then the rest of the range */
template <uint32_t F>
DEV void defer_finish(NS &r, const SegArgs &a, HistCtx &hc, int k1, DeferQ &dq)
{
    if constexpr ((F & FEAT_MONO) != 0) {
        const DComp &c = plan_at(a.comps, plan_at(a.steps, a.defer_k).first);
        if (r.pend && !r.dead) {
            s_mono_b(r, c, r.aux);
            tr_out<false>(r, c);
        }
        r.pend = false;
        run_range<false, F>(r, a, hc, a.defer_k + 1, k1, dq, false);
    }
}

template <bool FULL>
DEV void finalize(const NS &n, const SegArgs &a)
{
    if (FULL && a.n_rescal && n.se > 0.0f)
        for (uint32_t s = 0; s < a.n_rescal; s++) a.ev_tag[s * a.ev_cap + n.key] = a.chunk_tag;
    if (a.carry) {
        a.slot_ctr[n.key] = n.ctr;
        a.slot_iidx[n.key] = n.iidx;
    }
    if (FULL && a.keep) {
        float *o = a.st16 + 16 * (size_t)n.key;
        o[0] = n.x; o[1] = n.y; o[2] = n.z; o[3] = n.vx; o[4] = n.vy; o[5] = n.vz;
        o[6] = n.px; o[7] = n.py; o[8] = n.pz; o[9] = n.w; o[10] = n.t; o[11] = n.ctr;
        o[12] = n.sc; o[13] = n.sd; o[14] = n.se; o[15] = n.dead ? 1.0f : 0.0f;
        float *q = a.st8 + 8 * (size_t)n.key;
        for (int i = 0; i < 8; i++) q[i] = n.is[i];
        a.stii[n.key] = n.iidx;
    }
}

DEV uint32_t wave_sum(uint32_t v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

/* The ordering key of a ray for k_mesh_walk: its direction class (the
 * BVH layout it walks) above an 8+8-bit Morton code of its position across
 * the dominant axis, quantised over the mesh's box.  Rays walked in key
 * order run through nearly the same nodes side by side in a wave; the key
 * orders work only and does not touch any result. */
DEV uint32_t walk_key(const DPart &g, rf3 pos, rf3 vel)
{
    const float ax = rm_fabs(vel.x), ay = rm_fabs(vel.y), az = rm_fabs(vel.z);
    const int a = (ax >= ay && ax >= az) ? 0 : (ay >= az) ? 1 : 2;
    const float neg = a == 0 ? vel.x : a == 1 ? vel.y : vel.z;
    const int lay = 2 * a + (neg < 0.0f ? 1 : 0);
    const int u = a == 0 ? 1 : 0, v = a == 2 ? 1 : 2;
    const float pu = u == 0 ? pos.x : pos.y, pv = v == 1 ? pos.y : pos.z;
    const float lu = g.p[GP_XMIN + 2 * u], hu = g.p[GP_XMAX + 2 * u];
    const float lv = g.p[GP_XMIN + 2 * v], hv = g.p[GP_XMAX + 2 * v];
    const float su = hu > lu ? 256.0f / (hu - lu) : 0.0f, sv = hv > lv ? 256.0f / (hv - lv) : 0.0f;
    uint32_t qu = (uint32_t)rm_fmin(rm_fmax((pu - lu) * su, 0.0f), 255.0f);
    uint32_t qv = (uint32_t)rm_fmin(rm_fmax((pv - lv) * sv, 0.0f), 255.0f);
    uint32_t m = 0;
    for (int b = 0; b < 8; b++) m |= (((qu >> b) & 1u) << (2 * b)) | (((qv >> b) & 1u) << (2 * b + 1));
    return ((uint32_t)lay << 16) | m;
}

/* Prepare a stored neutron (queue slot `slot`) for k_mesh_walk: bring it
 * into the mesh's frame as the next segment's transform does (the general
 * rotation, which equals its identity shortcut wherever that applies),
 * run the reference's box precheck (mesh.cl:23-42), and list the rays that
 * pass in the wave's walk-list region, with their ordering keys; the
 * others get outcome 3. */
DEV void walk_emit(const NS &n, bool live, const SegArgs &a, uint64_t beg, uint64_t slot, uint32_t &wl_count)
{
    const DComp &c = plan_at(a.comps, (size_t)a.walk_ci);
    bool pass = false;
    rf3 pos = rf3_make(0.0f, 0.0f, 0.0f), vel = pos;
    if (live) {
        pos = mat_derot(c.R, rf3_make(n.x - c.pos[0], n.y - c.pos[1], n.z - c.pos[2]));
        vel = mat_derot(c.R, nvel(n));
        pass = mesh_box_precheck(c.g, pos, vel);
        if (!pass) {
            WalkHit h;
            h.t1 = h.t2 = 1e8f;
            h.code = 3u;
            h.pad = 0u;
            a.walk_res[slot] = h;
        }
    }
    const unsigned long long m = __ballot(pass);
    if (pass) {
        const uint64_t e = beg + wl_count + (uint64_t)__popcll(m & ((1ull << __lane_id()) - 1ull));
        WalkEntry w;
        w.pos[0] = pos.x; w.pos[1] = pos.y; w.pos[2] = pos.z;
        w.slot = (uint32_t)slot;
        w.vel[0] = vel.x; w.vel[1] = vel.y; w.vel[2] = vel.z;
        w.pad = 0u;
        a.walk_list[e] = w;
        a.walk_keys[e] = walk_key(c.g, pos, vel);
    }
    wl_count += (uint32_t)__popcll(m);
}

/* end of a group of 64 in a segment: survivors to the wave's output
 * region, finished neutrons to finalize (parked ones are neither) */
template <bool FULL, uint32_t F>
DEV void seg_tail(const NS &n, bool valid, const SegArgs &a, uint64_t beg, uint32_t &wcount, uint32_t &wl_count)
{
    const bool live = valid && !n.dead && !a.last;
    if (!a.last) {
        const unsigned long long m = __ballot(live);
        const uint64_t slot = beg + wcount + (uint64_t)__popcll(m & ((1ull << __lane_id()) - 1ull));
        if (live) queue_store<FULL>(n, a.out, slot);
        wcount += (uint32_t)__popcll(m);
        if constexpr (!FULL && (F & FEAT_MESH) != 0) {
            if (a.walk_role & WALK_EMIT) walk_emit(n, live, a, beg, slot, wl_count);
        }
    }
    if (valid && !live && !n.pend) finalize<FULL>(n, a);
}

/* One segment.  Work is split statically: wave g of G handles the dense
 * item range [g*per, (g+1)*per) and appends its survivors to its own region
 * [g*per, ...) of the output queue, so compaction needs no atomics; the
 * region counts are prefix-summed by k_scan before the next segment.  A
 * segment holding the deferred monochromator step drains the wave's
 * deferral queue whenever it holds 64, and at the end. */
/* Waves per SIMD each instance is compiled for (amdgpu_waves_per_eu; 1
 * leaves the register count to the compiler).  The fast-mode kernels are
 * latency-bound at the occupancy their register use gives by itself (111 to
 * 161 VGPRs: three or four waves), so they are held to more waves, trading
 * a few spills outside the inner loops for latency hiding; each value is
 * the best of the measured sweep on its bench workload.  The instance with
 * every feature (ISIS sources, non-linear blocks) keeps the compiler's two
 * waves: an ISIS S(q,w) run of 1e8 histories took 136.7 ms so and 146.2 ms
 * held to three (168 VGPRs, 120 B of spills).  Of the full-state
 * instances the plain one (155 VGPRs by itself) runs the MCPL, SRescal and
 * plugin workloads 1-6 % faster at four waves; the others keep the count
 * the compiler picks, as do the batch and adjoint kernels of the instances
 * no bench workload measures. */
#ifndef RAMP_W_SEG_F0
#define RAMP_W_SEG_F0 7      /* LET-like, minimal: later segments */
#endif
#ifndef RAMP_W_SEG_F0_FIRST
#define RAMP_W_SEG_F0_FIRST 7  /* LET-like, minimal: source segment */
#endif
#ifndef RAMP_W_SEG_SAMPLE
#define RAMP_W_SEG_SAMPLE 6  /* powder forward */
#endif
#ifndef RAMP_W_SEG_MONO
#define RAMP_W_SEG_MONO 4    /* TAS per-setting runs */
#endif
#ifndef RAMP_W_SEG_MESH
#define RAMP_W_SEG_MESH 1    /* imaging */
#endif
#ifndef RAMP_W_BATCH_MONO
#define RAMP_W_BATCH_MONO 4  /* TAS batch */
#endif
#ifndef RAMP_W_ADJ
#define RAMP_W_ADJ 6         /* powder adjoint */
#endif
#ifndef RAMP_W_FULL_F0
#define RAMP_W_FULL_F0 4     /* plain full-state (MCPL, SRescal, plugins, trace) */
#endif
constexpr int seg_min_waves(bool full, bool first, uint32_t f)
{
    return full ? (f == 0u ? RAMP_W_FULL_F0 : 1)
                : f == FEAT_MONO ? RAMP_W_SEG_MONO
                : f == 0u ? (first ? RAMP_W_SEG_F0_FIRST : RAMP_W_SEG_F0)
                : f == FEAT_SAMPLE ? RAMP_W_SEG_SAMPLE
                : f == FEAT_MESH ? RAMP_W_SEG_MESH : 1;
}
#define SEG_MIN_WAVES(FULL, FIRST, F) seg_min_waves((FULL), (FIRST), (F))

template <bool FULL, bool FIRST, uint32_t F>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SEG_MIN_WAVES(FULL, FIRST, F))))
void k_segment(SegArgs a)
{
    extern __shared__ float seg_lds[];
    HistCtx hc;
    hc.dets = a.dets;
    hc.mode = a.hist_mode;
    hc.enabled = !a.pilot;
    hc.events = hc.oob = hc.overflow = 0;
    hc.acc = nullptr;
    if (a.acc_bins) {
        /* the detectors' LDS copy follows the deferral queues */
        const size_t defer_bytes = FULL ? 0 : (size_t)(blockDim.x >> 6) * a.defer_lds;
        hist_acc_begin(hc, (char *)seg_lds + defer_bytes, a.acc_bytes);
    }
    const uint32_t lane = __lane_id();
    const uint64_t G = (uint64_t)gridDim.x * (blockDim.x >> 6);
    const uint64_t g = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    uint64_t total, nreg = 0, stride = 0;
    if (FIRST) {
        total = a.n;
    } else {
        nreg = a.in_meta[1];
        stride = a.in_meta[0];
        total = a.in_prefix[nreg];
    }
    const uint64_t per = (total + G - 1) / G;
    if (g == 0 && lane == 0 && !a.pilot) a.out_meta[0] = per;
    const uint64_t beg = g * per;
    const uint64_t end = beg + per < total ? beg + per : total;
    uint32_t wcount = 0, wl_count = 0;
    DeferQ dq;
    dq.b = seg_lds;
    dq.n = 0;
    bool defer = false;
    if constexpr (!FULL && (F & FEAT_MONO) != 0) {
        defer = !a.pilot && a.defer_k >= a.s0 && a.defer_k < a.s1;
        if (defer) defer_init(dq, seg_lds, DQ_CAP * DQ_WORDS);
    }
    uint64_t rcur = 0; /* region of the group's first input (later segments) */
    for (uint64_t base = beg; base < end; base += 64) {
        uint64_t i = base + lane;
        bool valid = i < end;
        const uint64_t last = base + 63 < end ? base + 63 : end - 1;
        NS n;
        if (valid) {
            if (FIRST) {
                float sb_old = a.carry ? a.slot_ctr[i] : 0.0f;
                n.iidx = a.carry ? a.slot_iidx[i] : 0u;
                n.key = (uint32_t)(a.key_base + i);
                src_generate<FULL, F>(n, plan_at(a.src, 0), a.num_sim, a.seed, sb_old);
            } else {
                const uint64_t slot = region_slot_group(a.in_prefix, nreg, stride, rcur, base, last, i);
                queue_load<FULL>(n, a.in, slot);
                if (a.walk_role & WALK_USE) n.aux = (uint32_t)slot;   /* where its walk result is */
            }
        } else {
            n.dead = true;
            n.pend = false;
            n.key = 0xffffffffu; /* no history on this lane */
            n.iidx = 0;
        }
        run_steps<FULL, F>(n, a, hc, dq, defer);
        if (a.pilot) continue;
        seg_tail<FULL, F>(n, valid, a, beg, wcount, wl_count);
        if constexpr (!FULL && (F & FEAT_MONO) != 0) {
            if (defer && dq.n >= 64u) {
                NS r;
                const uint32_t take = defer_pop<DQ_CAP, DQ_WORDS>(dq, r);
                defer_finish<F>(r, a, hc, a.s1, dq);
                seg_tail<FULL, F>(r, lane < take, a, beg, wcount, wl_count);
            }
        }
    }
    if (a.pilot) return;
    if constexpr (!FULL && (F & FEAT_MONO) != 0) {
        while (defer && dq.n > 0u) {
            NS r;
            const uint32_t take = defer_pop<DQ_CAP, DQ_WORDS>(dq, r);
            defer_finish<F>(r, a, hc, a.s1, dq);
            seg_tail<FULL, F>(r, lane < take, a, beg, wcount, wl_count);
        }
    }
    if (!a.last && lane == 0) a.out_counts[g] = wcount;
    if constexpr (!FULL && (F & FEAT_MESH) != 0) {
        if ((a.walk_role & WALK_EMIT) && lane == 0) a.walk_counts[g] = wl_count;
    }
    if (a.acc_bins) hist_acc_end(hc, a.n_dets, a.acc_bins);
    uint32_t ev = wave_sum(hc.events), oob = wave_sum(hc.oob), ovf = wave_sum(hc.overflow);
    if (lane == 0) {
        if (ev) atomicAdd(&a.cnt[CNT_EVENTS], (unsigned long long)ev);
        if (oob) atomicAdd(&a.cnt[CNT_OOB], (unsigned long long)oob);
        if (ovf) atomicAdd(&a.cnt[CNT_OVERFLOW], (unsigned long long)ovf);
    }
}

typedef void (*seg_fn)(SegArgs);

#endif
