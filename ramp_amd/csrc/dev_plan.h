/* dev_plan.h - device-side form of an instrument plan.
 *
 * The host library converts the C-ABI descriptors (include/ramp.h) into
 * these structs once per plan: rotation matrices are formed once with the
 * shared math library (bit-identical to re-forming them per neutron),
 * parameter-only expressions are evaluated once on the host (in double
 * where the reference evaluates them in double), and tables become device
 * pointers.  Kernels read the plan through uniform (scalar) loads.
 */
#ifndef RAMP_DEV_PLAN_H
#define RAMP_DEV_PLAN_H

#include <stdint.h>

#include "rdefs.h"

#define DP_ND 32

struct DPart {
    uint32_t type;
    uint32_t ni[8];
    float p[48];
    const float *tab[4];
    uint64_t tab_len[4];
    float d[DP_ND];          /* derived constants, per type (see host code) */
    const void *aux;         /* mesh: BVH nodes */
    const void *aux2;        /* mesh: triangle order */
};

/* a double kept in two consecutive slots of a derived-constant table */
RHD double dp_f64(const float *d, int k)
{
    double v;
    __builtin_memcpy(&v, d + k, sizeof v);
    return v;
}

struct DComp {
    DPart g;
    DPart s;
    uint32_t comp_idx;
    uint32_t restore;
    int32_t det;
    uint32_t pad;            /* 1: zero Euler angles (identity up to signed zeros) */
    float pos[3];
    float R[9];
    int32_t trace;           /* trace slot (RAMP_RUN_TRACE), or -1 */
    uint32_t pad2;
};

/* one step of a segment: a linear-block component, or a whole non-linear
 * block (kind 1) swept `multi` times */
struct DStep {
    uint32_t kind;
    uint32_t first;
    uint32_t count;
    uint32_t multi;
};

struct DDet {
    float *w;                    /* atomic mode */
    float *w2;
    unsigned long long *qw;      /* deterministic mode: fixed point */
    unsigned long long *qw2;
    uint64_t n;
    int32_t sh_w;                /* fixed-point exponents: q = rint(w * 2^sh) */
    int32_t sh_w2;
    int32_t acc;                 /* small detector: its first bin in the workgroup's LDS copy, else -1 */
    int32_t pad;
};

/* detectors small enough (all together within this many bytes of LDS per
 * workgroup: 8 B per bin for float atomics, 16 B for fixed point)
 * accumulate per workgroup in LDS and are added to HBM once per workgroup,
 * instead of one same-address atomic per event */
#define DET_LDS_BYTES 32768

/* SoA neutron queue */
struct DQueue {
    float *x, *y, *z, *vx, *vy, *vz, *w, *t, *ctr;
    uint32_t *key, *status;
    float *px, *py, *pz, *sc, *sd, *se; /* full-state mode only */
    float *is[8];                       /* full-state mode only: the intersection */
};

/* counters written by the kernels */
#define CNT_EVENTS 0
#define CNT_OOB 1
#define CNT_OVERFLOW 2
#define CNT_QOUT 3
#define CNT_QOUT0 4
#define CNT_QIN 5
#define CNT_NSTEP0 8                     /* pilot: alive after step k at [8+k] */
#define CNT_N 264

#endif
