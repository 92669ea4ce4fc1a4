/* plugin_kernel.h - the kernel around a hipRTC-compiled scatter plugin.
 * A plugin step is a segment of its own: the previous segment leaves the
 * live neutrons in the queue, ramp_plugin_kernel runs the component's
 * transform, geometry, terminate, the plugin and the untransform on each
 * of them in place (full-state queue), and the next segment reads the
 * queue as it would have read the previous segment's output. */
#ifndef RAMP_PLUGIN_KERNEL_H
#define RAMP_PLUGIN_KERNEL_H

#include "dev_components.h"
#include "dev_queue.h"
#include "ramp_plugin.h"

struct PluginArgs {
    DQueue q;
    const uint64_t *prefix;  /* [nreg + 1] region prefix of the queue */
    const uint64_t *meta;    /* [0] region stride, [1] number of regions */
    const DComp *comp;
    float *trace;            /* trace points (RAMP_RUN_TRACE), or null */
    uint64_t trace_cap, trace_base;
};

typedef void (*ramp_scatter_fn)(RampNeutron *, RampIntersection *, uint32_t *, uint32_t, const float *,
                                uint32_t);

template <bool TRACE, uint32_t GF>
DEV void plugin_one(NS &n, const DComp &c, ramp_scatter_fn f, const PluginArgs &a)
{
    tr_in<true>(n, c);
    geom_dispatch<GF>(n, c);
    terminate(n, c.restore);
    const bool traced = TRACE && c.trace >= 0;
    if (traced) trace_entry(n, a.trace, a.trace_cap, a.trace_base + n.key, c.trace);
    RampNeutron r;
    r.x = n.x; r.y = n.y; r.z = n.z; r.vx = n.vx; r.vy = n.vy; r.vz = n.vz;
    r.px = n.px; r.py = n.py; r.pz = n.pz; r.p = n.w; r.tof = n.t; r.sb = n.ctr;
    r.sc = n.sc; r.sd = n.sd; r.se = n.se; r.die = n.dead ? 1.0f : 0.0f;
    RampIntersection is;
    is.x1 = n.is[0]; is.y1 = n.is[1]; is.z1 = n.is[2]; is.t1 = n.is[3];
    is.x2 = n.is[4]; is.y2 = n.is[5]; is.z2 = n.is[6]; is.t2 = n.is[7];
    uint32_t ii = n.iidx;
    f(&r, &is, &ii, c.comp_idx, c.s.p, n.key);
    n.x = r.x; n.y = r.y; n.z = r.z; n.vx = r.vx; n.vy = r.vy; n.vz = r.vz;
    n.px = r.px; n.py = r.py; n.pz = r.pz; n.w = r.p; n.t = r.tof; n.ctr = r.sb;
    n.sc = r.sc; n.sd = r.sd; n.se = r.se; n.dead = r.die > 0.0f;
    n.is[0] = is.x1; n.is[1] = is.y1; n.is[2] = is.z1; n.is[3] = is.t1;
    n.is[4] = is.x2; n.is[5] = is.y2; n.is[6] = is.z2; n.is[7] = is.t2;
    n.iidx = ii;
    tr_out<true>(n, c);
    if (traced) trace_alive(n, a.trace, a.trace_cap, a.trace_base + n.key, c.trace);
}

/* the plugin's kernels: ramp_plugin_kernel, ramp_plugin_kernel_trace for
 * runs that record trace points, and the _mesh forms of both for a plugin
 * component whose geometry is a GMesh (the mesh traversal costs registers
 * that every other geometry would pay for) */
template <bool TRACE, uint32_t GF>
DEV void plugin_queue(const PluginArgs &a, ramp_scatter_fn f)
{
    const uint64_t nreg = a.meta[1], stride = a.meta[0], total = a.prefix[nreg];
    const DComp &c = plan_at(a.comp, 0);
    /* wave g of G takes the inputs [g * per, (g + 1) * per) in order */
    const uint64_t G = (uint64_t)gridDim.x * (blockDim.x >> 6);
    const uint64_t g = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const uint64_t per = (total + G - 1) / G;
    const uint64_t beg = g * per, end = beg + per < total ? beg + per : total;
    uint64_t rcur = 0;
    for (uint64_t base = beg; base < end; base += 64) {
        const uint64_t i = base + __lane_id();
        const uint64_t last = base + 63 < end ? base + 63 : end - 1;
        if (i < end) {
            const uint64_t slot = region_slot_group(a.prefix, nreg, stride, rcur, base, last, i);
            NS n;
            queue_load<true>(n, a.q, slot);
            if (!n.dead) plugin_one<TRACE, GF>(n, c, f, a);
            queue_store<true>(n, a.q, slot);
        }
    }
}

#define RAMP_PLUGIN_KERNEL(entry)                                                               \
    extern "C" __global__ __launch_bounds__(256) void ramp_plugin_kernel(PluginArgs a)          \
    {                                                                                           \
        plugin_queue<false, 0u>(a, entry);                                                      \
    }                                                                                           \
    extern "C" __global__ __launch_bounds__(256) void ramp_plugin_kernel_trace(PluginArgs a)    \
    {                                                                                           \
        plugin_queue<true, 0u>(a, entry);                                                       \
    }                                                                                           \
    extern "C" __global__ __launch_bounds__(256) void ramp_plugin_kernel_mesh(PluginArgs a)     \
    {                                                                                           \
        plugin_queue<false, FEAT_MESH>(a, entry);                                               \
    }                                                                                           \
    extern "C" __global__ __launch_bounds__(256) void ramp_plugin_kernel_mesh_trace(PluginArgs a) \
    {                                                                                           \
        plugin_queue<true, FEAT_MESH>(a, entry);                                                \
    }

#endif
