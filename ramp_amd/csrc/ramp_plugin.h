/* ramp_plugin.h - the interface of a user-written scatter plugin.
 *
 * The reference lets users add components as OpenCL kernels resolved by
 * name (mcramp.py:13-17, 80-92; doc/writing/scatter.rst:9-69).  Here a
 * plugin is one HIP device function, compiled for the GPU at run time with
 * hipRTC (ramp_plugin_compile) and run by the library on every neutron
 * that reaches its component, after the component's own geometry and
 * terminate step.  It sees one neutron the way the reference's kernels
 * do - the float16 record of consts.h:91-109, the float8 intersection,
 * iidx and the component index - and keeps the reference's contract: it
 * checks iidx == comp_idx and the dead flag itself, and when it scatters,
 * sets iidx = 0, sc = comp_idx and resets the intersection
 * (ramp_isect_reset).  params are the component's 48 float parameters.
 *
 *   RAMP_SCATTER void my_scatter(RampNeutron *n, RampIntersection *is,
 *                                uint32_t *iidx, uint32_t comp_idx,
 *                                const float *params, uint32_t tid);
 *
 * ramp_rand(n, tid) draws from the neutron's own stream (rand.h:4-20).
 * The same source also compiles on the host (C or C++), so that a plugin
 * can be checked against itself on the CPU.
 */
#ifndef RAMP_PLUGIN_H
#define RAMP_PLUGIN_H

typedef struct {
    float x, y, z, vx, vy, vz, px, py, pz;
    float p;     /* weight */
    float tof;
    float sb;    /* RNG counter */
    float sc;    /* previous component */
    float sd, se;
    float die;   /* > 0: dead */
} RampNeutron;

typedef struct {
    float x1, y1, z1, t1, x2, y2, z2, t2;
} RampIntersection;

#if defined(__HIPCC_RTC__) || defined(__HIPCC__)
#define RAMP_FN __device__ static inline
#define RAMP_SCATTER extern "C" __device__
#elif defined(__cplusplus)
#define RAMP_FN static inline
#define RAMP_SCATTER extern "C"
#else
#define RAMP_FN static inline
#define RAMP_SCATTER
#endif

RAMP_FN float ramp_rand(RampNeutron *n, uint32_t tid) { return rr_draw(&n->sb, tid); }

RAMP_FN void ramp_isect_reset(RampIntersection *is)
{
    is->x1 = is->y1 = is->z1 = 0.0f;
    is->t1 = 100000.0f;
    is->x2 = is->y2 = is->z2 = 0.0f;
    is->t2 = 100000.0f;
}

#endif
