/* oracle.h - CPU oracle of the RAMP neutron ray-tracing path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the
 * reference's OpenCL kernels and host loop (the .cl files under mcramp/ and mcramp/mcramp.py),
 * one neutron at a time in the reference's AoS layout, used by the tests,
 * by __graft_entry__.smoke() and by bench.py's CPU-baseline leg as the
 * checker of the HIP path.  The product (ramp_amd, libramp.so) never links,
 * imports or calls it.
 *
 * Arithmetic follows the reference: float where the kernels use float,
 * double where the kernels promote to double (M_PI and unsuffixed literals
 * are double on fp64-capable OpenCL devices).  OpenCL builtins are the
 * project math library (ramp_amd/csrc/rmath.h, rvec.h), the RNG is the
 * Threefry stream of ramp_amd/csrc/rrng.h (pinned against the reference's
 * own Random123 header by tests/test_rng.py).
 */
#ifndef RAMP_ORACLE_H
#define RAMP_ORACLE_H

#include <stdint.h>
#include "../include/ramp.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Run n_hist histories of `desc` in chunks (chunk = 0: one chunk).
 * hist_w / hist_w2: arrays of n_dets pointers to caller buffers of
 * dets[d].n_bins doubles, accumulated into (not cleared).
 * neutrons16 / isect8 / iidx (may be NULL): final state of the last chunk.
 * n_threads <= 0: OpenMP default.  Returns 0 or a negative error. */
/* Operation counts of the ORACLE_COUNT build (liboracle_count.so; zeros in
 * the others), accumulated over single-threaded runs since the last reset:
 * out[0] Threefry evaluations, [1] FP32 divisions, [2] square roots,
 * [3..12] calls of exp, log, pow, tanh, atan2, sin, cos, asin, acos, fmod,
 * [13] guide bounce iterations, [14] guide exit iterations.  reset != 0
 * clears them after reading. */
void oracle_opcount(uint64_t *out, int reset);

int oracle_run(const ramp_plan_desc *desc, uint64_t n_hist, uint64_t chunk,
               const uint32_t *seeds, uint32_t n_seeds, double **hist_w, double **hist_w2,
               float *neutrons16, float *isect8, uint32_t *iidx, int n_threads,
               uint64_t *n_oob_out);

/* FP64 reference gradient of a loss of detector `det`'s histogram (dL_dh:
 * dL/dh per bin) over the same histories as oracle_run: grad =
 * [T1, sigma_inc, sigma_abs, U, d_0 .. d_{nl-1}] of the plan's SPowderDiff
 * and MCont source, n_grad = 4 + nl. */
int oracle_adjoint(const ramp_plan_desc *desc, uint64_t n_hist, uint64_t chunk,
                   const uint32_t *seeds, uint32_t n_seeds, uint32_t det, const float *dL_dh,
                   double *grad, uint32_t n_grad, int n_threads);

/* Where SMCPLOut components write their records: records[slot][cap][11]
 * and tags[slot][cap] (chunk number + 1 of the chunk that wrote them); slot
 * is the component's index among the plan's SMCPLOut components.  NULL
 * records: SMCPLOut writes nothing. */
void oracle_set_mcpl_dump(float *records, uint32_t *tags, uint64_t cap);

/* Where SRescal event lists go: events[slot][5][cap] (qx, qy, qz, dE, p),
 * appended at counts[slot] (not cleared), in chunk and history order. */
void oracle_set_rescal_out(float *events, uint64_t cap, uint64_t *counts);

/* The host build of the plugin bound to component `comp` (its index in
 * desc->comps): a function (float *neutron16, float *isect8, uint32_t
 * *iidx, uint32_t comp_idx, const float *params, uint32_t tid). */
void oracle_set_plugin(uint32_t comp, void *fn);

/* Where trace points go (Instrument.execute(trace=True), mcramp.py:118-165,
 * 222-251): points[slot][4][cap] = x1, y1, z1 of the intersection after
 * terminate (component frame) and 1.0 if alive after the component, at the
 * history's run-wide index, for the drawn (vis) components of linear blocks
 * numbered in plan order.  Every history that reaches the component writes
 * its entry, dead or alive.  NULL: no trace. */
void oracle_set_trace(float *points, uint64_t cap);

/* Test hook: the shared rotate-about-axis restatement (geom.h:1-19). */
void oracle_rotate_axis(float angle, const float *axis, const float *v, float *out);

const char *oracle_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
