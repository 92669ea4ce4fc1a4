/* kern_table.h - the feature-set instances of the large kernels.
 *
 * k_segment, k_batch and k_adjoint are compiled once per feature set
 * (FEAT_* bits: the mesh, sample and monochromator code a plan needs), each
 * set in three translation units (kern_v.hip built with RAMP_KV = 0..5 and
 * RAMP_KPART = 0 for the fast-mode segment kernels, 1 for the full-state
 * ones, 2 for the batch and adjoint kernels), so the builds run in
 * parallel.  The host picks an instance through the
 * table the unit fills in.
 */
#ifndef RAMP_KERN_TABLE_H
#define RAMP_KERN_TABLE_H

#include "batch_kernel.h"
#include "adjoint_kernel.h"

/* feature-set instances, cheapest first (index = RAMP_KV of the unit) */
#define RAMP_KV_FEATS {0u, FEAT_SAMPLE, FEAT_MONO, FEAT_MESH, FEAT_MESH | FEAT_SAMPLE, FEAT_ALL}
#define RAMP_NUM_KV 6

struct KernTable {
    seg_fn seg[2][2];   /* [FULL][FIRST] */
    batch_fn batch;
    adj_fn adjoint;     /* null where the set lacks FEAT_SAMPLE (the adjoint needs SPowderDiff) */
};

/* the fast-mode unit fills seg[0][*], the full-state unit seg[1][*], the
 * third unit batch and adjoint */
void ramp_kern_v0(KernTable *t);
void ramp_kern_v1(KernTable *t);
void ramp_kern_v2(KernTable *t);
void ramp_kern_v3(KernTable *t);
void ramp_kern_v4(KernTable *t);
void ramp_kern_v5(KernTable *t);
void ramp_kern_full_v0(KernTable *t);
void ramp_kern_full_v1(KernTable *t);
void ramp_kern_full_v2(KernTable *t);
void ramp_kern_full_v3(KernTable *t);
void ramp_kern_full_v4(KernTable *t);
void ramp_kern_full_v5(KernTable *t);
void ramp_kern_ba_v0(KernTable *t);
void ramp_kern_ba_v1(KernTable *t);
void ramp_kern_ba_v2(KernTable *t);
void ramp_kern_ba_v3(KernTable *t);
void ramp_kern_ba_v4(KernTable *t);
void ramp_kern_ba_v5(KernTable *t);

#endif
