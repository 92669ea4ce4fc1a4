/* ramp.h - C-ABI of libramp.so, the MI355X implementation of RAMP's
 * per-neutron Monte Carlo ray-tracing path.
 *
 * The reference drives its OpenCL kernels from Python (PyOpenCL):
 *   Instrument.execute(N)            mcramp/mcramp.py:403-455
 *   source.gen_prg(queue, N, ...)    mcramp/moderator/flat_mod.py:24-32
 *   geom.intersect_prg(...)          mcramp/geom/plane.py:48-58
 *   scat.scatter_prg(...)            mcramp/scat/chopper.py:58-70
 *   scat.data(queue)                 mcramp/scat/detector1D.py:114-116
 * Those per-kernel launches over AoS float16/float8 buffers are replaced
 * here by one plan (the instrument flattened into POD descriptors) and a
 * run call that executes fused, stream-compacted HIP segment kernels over
 * structure-of-arrays FP32 state.  All functions return 0 on success and a
 * negative RAMP_E_* code on failure; ramp_last_error() gives the message of
 * the last failure on the calling thread.  No function throws or exits.
 *
 * Memory: device memory is owned by the context/plan, host buffers by the
 * caller.  A context is bound to one HIP device and one stream and is not
 * thread-safe; separate contexts are independent.
 */
#ifndef RAMP_H
#define RAMP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes -------------------------------------------------------- */
#define RAMP_OK 0
#define RAMP_E_INVALID (-1)   /* bad argument / descriptor */
#define RAMP_E_DEVICE (-2)    /* HIP runtime failure */
#define RAMP_E_NOMEM (-3)     /* allocation failure */
#define RAMP_E_RANGE (-4)     /* size beyond a documented limit */
#define RAMP_E_STATE (-5)     /* call not valid in the current state */

/* ---- component type codes ---------------------------------------------- */
/* sources (mcramp/moderator) */
#define RAMP_SRC_FLAT 1      /* MFlat     flat_mod.cl:4-51 */
#define RAMP_SRC_CONT 2      /* MCont     cont_mod.cl:11-73 (+ optional 1/E tail) */
#define RAMP_SRC_ISIS 3      /* MISIS     isis_mod.cl:15-105 */
#define RAMP_SRC_LAMBDA 4    /* MLambda   wavelength-uniform, weight ~ 1/lambda (no reference counterpart) */
#define RAMP_SRC_PARALLEL 5  /* MParallel parallel beam (no reference counterpart) */

/* geometry (mcramp/geom) */
#define RAMP_G_ARM 1         /* GArm         geom/arm.cl:3-23 */
#define RAMP_G_PLANE 2       /* GPlane       geom/plane.cl:3-53 */
#define RAMP_G_BOX 3         /* GBox         geom/box.cl:3-54 */
#define RAMP_G_SPHERE 4      /* GSphere      geom/sphere.cl:3-44 */
#define RAMP_G_SPHERE_INT 5  /* GSphereInt   geom/sphere_int.cl:3-43 */
#define RAMP_G_CYLINDER 6    /* GCylinderExt geom/cylinder_ext.cl:3-53 */
#define RAMP_G_BANANA 7      /* GBanana      geom/banana.cl:3-55 */
#define RAMP_G_MESH 8        /* GMesh        geom/mesh.cl:3-130 (BVH on the GPU) */

/* scattering (mcramp/scat) */
#define RAMP_S_ARM 1         /* SArm           scat/arm.cl:3-30 */
#define RAMP_S_BEAMSTOP 2    /* SBeamstop      scat/beamstop.cl:3-32 */
#define RAMP_S_CHOPPER 3     /* SChopper       scat/chopper.cl:4-59 */
#define RAMP_S_GUIDE 4       /* SGuide         scat/guide.cl:9-150 */
#define RAMP_S_ISOTROPIC 5   /* SIsotropic     scat/isotropic.cl:75-323 */
#define RAMP_S_POWDER1 6     /* SPowder1       scat/powder1.cl:5-101 */
#define RAMP_S_POWDERN 7     /* SPowderN       scat/powderN.cl:5-157 */
#define RAMP_S_MONO 8        /* SMonochromator scat/monochromator.cl:22-197 */
#define RAMP_S_DET1D 9       /* SDetector1D    scat/detector1D.cl:20-77 */
#define RAMP_S_DET2D 10      /* SDetector2D    scat/detector2D.cl:31-186 */
#define RAMP_S_COUNTER 11    /* SCounter       scat/counter.cl:18-49 */
#define RAMP_S_INCOHERENT 12 /* SIncoherent    constant-Sigma incoherent sample (no reference counterpart) */
#define RAMP_S_ABSORBER 13   /* SAbsorber      Beer-Lambert absorber (no reference counterpart) */
#define RAMP_S_DETPIX 14     /* SDetectorTubes tube x pixel x TOF bank (no reference counterpart) */
#define RAMP_S_POWDERDIFF 15 /* SPowderDiff    differentiable powder + incoherent sample (no reference counterpart) */
#define RAMP_S_MCPL_IN 16    /* SMCPLIn        scat/mcpl_in.py:9-76 (particles from an MCPL file) */
#define RAMP_S_MCPL_OUT 17   /* SMCPLOut       scat/mcpl_out.py:7-117 (live particles to an MCPL file) */
#define RAMP_S_RESCAL 18     /* SRescal        scat/rescal.cl:5-77, rescal.py:9-66 (TAS resolution events) */
#define RAMP_S_POLARISER 19  /* SPolariser     scat/polariser.cl:3-32 */
#define RAMP_S_ANALYSER 20   /* SAnalyser      scat/analyser.cl:1-31 (with the consts.h include it lacks) */
#define RAMP_S_COLLIMATOR 21 /* SLinearCollimator scat/collimator_lin.cl:3-55 (with the ';' it lacks at :30) */
#define RAMP_S_PLUGIN 22     /* user scatter compiled at run time (ramp_plugin_compile; mcramp.py:13-17, 80-92) */

/* ---- plan descriptors --------------------------------------------------- */
#define RAMP_NP 48  /* float parameters per component part */
#define RAMP_NI 8   /* integer parameters per component part */
#define RAMP_NT 4   /* table references per component part */

/* Parameters are stored as float32 exactly as the reference wrappers convert
 * them (np.float32 / np.uint32); the per-type slot meaning is listed in
 * ramp_amd/components.py next to each class. */
typedef struct {
    uint32_t type;
    uint32_t ni[RAMP_NI];
    float p[RAMP_NP];
    uint64_t tab_off[RAMP_NT]; /* offsets (in floats) into ramp_plan_desc.tables */
    uint64_t tab_len[RAMP_NT];
} ramp_part_desc;

typedef struct {
    ramp_part_desc geom;
    ramp_part_desc scat;
    uint32_t comp_idx;   /* reference component index (mcramp.py:275,112) */
    uint32_t restore;    /* component-level restore_neutron (terminator.cl:17) */
    int32_t det;         /* histogram slot written by the scatter part, or -1 */
    uint32_t vis;        /* 1: drawn and traced (the component's 'visualise', default
                            true; mcramp.py:67, 141, 159) */
    float pos[3];        /* component position (after 'relative' resolution) */
    float rot[3];        /* ZYX Euler angles in radians */
} ramp_component_desc;

typedef struct {
    uint32_t linear;     /* 1: linear block, 0: non-linear ('multi') block */
    uint32_t multi;      /* sweeps of a non-linear block */
    uint32_t first;      /* first component of the block */
    uint32_t count;      /* number of components */
} ramp_block_desc;

typedef struct {
    uint64_t n_bins;     /* bins per moment (w and w^2 are stored separately) */
} ramp_det_desc;

typedef struct {
    ramp_part_desc source;
    uint32_t source_block;   /* block holding the source (execution starts there) */
    uint32_t n_blocks;
    uint32_t n_comps;
    uint32_t n_dets;
    const ramp_block_desc *blocks;
    const ramp_component_desc *comps;
    const ramp_det_desc *dets;
    const float *tables;
    uint64_t n_tables;
} ramp_plan_desc;

/* ---- run options -------------------------------------------------------- */
#define RAMP_HIST_ATOMIC 0        /* float atomics, order-dependent sums */
#define RAMP_HIST_DETERMINISTIC 1 /* int64 fixed-point sums, order-independent */
/* run flags */
#define RAMP_RUN_MESH_BRUTE_FORCE 1  /* test every triangle of every GMesh (the
                                        reference's loop, mesh.cl:52-112) instead
                                        of traversing the BVH; for checking the BVH */
#define RAMP_RUN_TRACE 2             /* record the trace points of every traced
                                        component (ramp_trace_read); the run is
                                        in full-state mode */

typedef struct {
    uint32_t hist_mode;      /* RAMP_HIST_* */
    uint32_t keep_state;     /* 1: keep the final per-neutron state of the last chunk */
    uint32_t wg;             /* workgroup size (0: default) */
    uint32_t npt;            /* neutrons per thread in the source segment (0: default) */
    uint32_t max_segments;   /* compaction boundaries allowed (0: automatic) */
    uint32_t flags;          /* RAMP_RUN_* bits */
    uint32_t pad;
    double weight_bound;     /* deterministic mode: bound on one event's weight (0: derive) */
    /* One shard of a chunk (shard_total > 0): the run's n_hist histories are
     * histories [key_base, key_base + n_hist) of a chunk of shard_total, with
     * those indices as Threefry keys and weights normalised to shard_total,
     * so shards run in turn (or on several GPUs) add up to the histograms of
     * the whole chunk.  One chunk only (chunk = 0 or n_hist), fast-mode plans
     * only, no keep_state or trace.  0, 0: an ordinary run. */
    uint64_t key_base;
    uint64_t shard_total;
} ramp_run_opts;

typedef struct {
    uint64_t n_hist;         /* histories run */
    uint64_t n_chunks;
    uint64_t n_events;       /* detector events accumulated */
    uint64_t n_oob;          /* detector events dropped as out of range (reference writes them out of bounds) */
    uint64_t n_segments;     /* segment launches */
    uint64_t state_in;       /* neutrons read from compacted queues */
    uint64_t state_out;      /* neutrons written to compacted queues */
    uint64_t n_det_overflow; /* deterministic mode: events beyond the fixed-point range */
    uint64_t state_out_first;/* neutrons the first (source) segment wrote to its queue */
    double ms_total;         /* device time of the last run (HIP events) */
    double ms_source_segment;/* device time of the first (source) segment of the last run */
    double ms_mcpl_compact;  /* last ramp_mcpl_read: device time of its three compaction kernels */
    double ms_mcpl_write;    /* last ramp_mcpl_read: device time of the record-writing pass alone */
    double ms_rescal_append; /* device time of the first chunk's SRescal event append (count, scan, append) */
    double ms_plugin;        /* device time of the first chunk's first plugin step */
    uint64_t plugin_items;   /* neutrons that step processed (the queue it rewrote) */
    double ms_mesh_walk;     /* device time of the first chunk's separate mesh-walk launch (k_mesh_walk) */
    double ms_walk_sort;     /* device time of the first chunk's walk-list ordering (keys, sort, gather) */
    uint64_t walk_rays;      /* rays that walk launch walked */
} ramp_run_stats;

typedef struct ramp_ctx ramp_ctx;
typedef struct ramp_plan ramp_plan;

/* Context bound to HIP device `device`.  Replaces the PyOpenCL context and
 * queue passed to Instrument(fn, ctx, queue) (mcramp.py:271). */
int ramp_ctx_create(int device, ramp_ctx **out);
int ramp_ctx_destroy(ramp_ctx *ctx);

/* Upload an instrument plan.  Replaces the per-component program builds and
 * parameter/histogram buffer uploads done by the wrapper constructors
 * (mcramp.py:46-116, e.g. detector1D.py:58-70). */
int ramp_plan_create(ramp_ctx *ctx, const ramp_plan_desc *desc, ramp_plan **out);
int ramp_plan_destroy(ramp_plan *plan);

/* Run n_hist histories in chunks of `chunk` (0: one chunk).  seeds[c] is the
 * source seed of chunk c (the reference draws randint(0,4096) per chunk,
 * flat_mod.py:32); n_seeds must cover ceil(n_hist/chunk) chunks.
 * Replaces Instrument.execute (mcramp.py:403-455).  Histograms accumulate
 * across runs, as the reference's do (detector1D.py:59-70).  With n_hist =
 * 0 nothing runs; the plan's histograms of opts->hist_mode are allocated
 * (zeroed) if they did not exist. */
int ramp_run(ramp_plan *plan, uint64_t n_hist, uint64_t chunk, const uint32_t *seeds,
             uint32_t n_seeds, const ramp_run_opts *opts);

/* ramp_run without any device or pinned allocation: no counterpart in the
 * reference.  For callers that cannot allocate (HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The segment queues,
 * region counts and prefixes and the per-slot carry buffers of several
 * chunks come from one static device array per GPU (RAMP_NOALLOC_SCRATCH_BYTES,
 * held until the process ends) instead of plan allocations.  Limits:
 * chunk (n_hist if chunk = 0) at most RAMP_NOALLOC_MAX_CHUNK (RAMP_E_RANGE
 * otherwise); fast-mode plans only (no SMCPLIn/SMCPLOut, SRescal, SAnalyser
 * or plugin components) and no keep_state, trace or brute-force flags
 * (RAMP_E_INVALID); the plan's histograms of the mode must exist (ramp_run
 * with n_hist = 0 allocates them; RAMP_E_STATE otherwise); a deterministic
 * run that would need its sums folded to the host returns RAMP_E_RANGE.
 * Results equal ramp_run's.  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (other threads, other contexts' streams)
 * race and must be serialised by the caller. */
#define RAMP_NOALLOC_MAX_CHUNK 262144
#define RAMP_NOALLOC_SCRATCH_BYTES 25985280
int ramp_run_noalloc(ramp_plan *plan, uint64_t n_hist, uint64_t chunk, const uint32_t *seeds,
                     uint32_t n_seeds, const ramp_run_opts *opts);

/* Read detector slot `det` (w and w^2 sums, n values each).  Replaces the
 * wrappers' _cached_copy readback (detector1D.py:130-138). */
int ramp_hist_read(ramp_plan *plan, uint32_t det, double *w, double *w2, uint64_t n);

/* Zero every histogram of the plan (the reference has no counterpart). */
int ramp_hist_reset(ramp_plan *plan);

/* Final per-neutron state of the last chunk of the last run with
 * keep_state=1, in the reference's AoS layouts: neutrons as 16 floats
 * (consts.h:91-121), intersections as 8 floats, iidx as uint32.  Replaces
 * the trace/debug copies of mcramp.py:141-165. */
int ramp_state_read(ramp_plan *plan, float *neutrons16, float *isect8, uint32_t *iidx, uint64_t n);

int ramp_stats_read(const ramp_plan *plan, ramp_run_stats *out);

/* ---- trace -----------------------------------------------------------------
 * Trace points of the last run made with RAMP_RUN_TRACE.  The traced
 * components are those with vis = 1 in linear blocks, numbered in plan
 * order from 0 (`slot`).  For every history h of the run (h = chunk *
 * chunk_size + chunk-local index, n = the run's n_hist) `out` receives, as
 * four arrays of n floats: the entry point (x1, y1, z1) of the intersection
 * the component's geometry produced, in the component's frame, as it stood
 * after terminate; and 1.0 where the neutron was alive after the
 * component's scatter and untransform, else 0.0 (then the point is
 * meaningless; histories that never reached the component read 0.0 too).
 * Replaces the per-component copies of the intersection and neutron
 * buffers that Instrument.execute(trace=True) makes (mcramp.py:141-161)
 * and from which _get_trace_lines (:222-251) draws each surviving
 * neutron's line; the host forms the lines.  Components in non-linear
 * blocks are not traced (the reference keeps no trace positions there). */
int ramp_trace_read(ramp_plan *plan, uint32_t slot, float *out, uint64_t n);

/* ---- MCPL particle output -------------------------------------------------
 * Records of SMCPLOut component `comp` (its index among the plan's
 * SMCPLOut components, in plan order) from the last chunk of the last run:
 * the neutrons alive at that component, in history order, each packed on
 * the GPU into the 11-float record the reference writes (polarisation,
 * position in cm, packed direction and kinetic energy, time in ms, weight;
 * mcpl_out.py:55-73).  The records are compacted on the device; *count
 * receives their number and min(count, cap) of them are copied to
 * `records` (11 floats each; may be NULL with cap = 0 to query the count).
 * Replaces the host loop of SMCPLOut.scatter_prg (mcpl_out.py:99-117),
 * which copies the whole neutron buffer back and filters it in Python;
 * the file header is written by the caller. */
int ramp_mcpl_read(ramp_plan *plan, uint32_t comp, float *records, uint64_t cap, uint64_t *count);

/* ---- user-written scatter plugins -----------------------------------------
 * The reference resolves unknown component names to user modules whose
 * OpenCL kernels it builds at run time (build_kernel, mcramp.py:13-17;
 * importlib fallback :80-92; template doc/writing/scatter.rst:51-69).
 * Here a plugin is a HIP device function with the signature of
 * ramp_amd/csrc/ramp_plugin.h, compiled for the context's GPU with hipRTC.
 * `source` holds the function (and anything it needs), `entry` its name.
 * A component whose scatter part has type RAMP_S_PLUGIN runs the plugin
 * bound to it after its geometry and terminate steps; every such
 * component needs a bound plugin before ramp_run.  Plans with plugins run
 * in full-state mode; the plugin step is a segment of its own.  The
 * compiler's log of a failed build is the error message. */
typedef struct ramp_plugin ramp_plugin;
int ramp_plugin_compile(ramp_ctx *ctx, const char *source, const char *entry, ramp_plugin **out);
int ramp_plugin_destroy(ramp_plugin *plugin);
/* Bind `plugin` (of the plan's context) to component `comp` (its index in
 * ramp_plan_desc.comps). */
int ramp_plan_bind_plugin(ramp_plan *plan, uint32_t comp, ramp_plugin *plugin);

/* ---- SRescal event lists --------------------------------------------------
 * Events of SRescal component `comp` (its index among the plan's SRescal
 * components, in plan order), accumulated over every chunk of every run
 * since the plan was created or ramp_hist_reset was called: for each
 * history that scattered there and ended its block detected (se > 0),
 * (qx, qy, qz) = ki - kf, the energy transfer and the weight, in history
 * order within a chunk and chunk order.  The device appends each chunk's
 * events after the chunk (a compaction kernel), so nothing crosses PCIe
 * during a run.  *count receives the number of events; min(count, cap) of
 * them are copied field-major to `events`: qx[cap], qy[cap], qz[cap],
 * dE[cap], p[cap] (events may be NULL with cap = 0).  Replaces
 * SRescal.data_reduce (rescal.py:51-62), which copies the event and neutron
 * buffers of every chunk back and filters them on the host. */
int ramp_rescal_read(ramp_plan *plan, uint32_t comp, float *events, uint64_t cap, uint64_t *count);

/* ---- gradients ----------------------------------------------------------
 * Gradient of a loss of one detector's histogram with respect to the model
 * parameters of the plan's SPowderDiff sample and its MCont source
 * (no reference counterpart; the forward it differentiates is ramp_run with
 * the same n_hist, chunk and seeds).  dL_dh holds dL/dh per bin of detector
 * `det` (n_bins floats, the layout of ramp_hist_read); grad receives, in
 * FP64, [dL/dT1, dL/dsigma_inc, dL/dsigma_abs, dL/dU, dL/dd_0 ..
 * dL/dd_{n-1}] (n_grad = 4 + number of lines; T1 only for a
 * single-temperature MCont, else 0).  opts: wg only (flags and keep_state
 * must be 0).  The run statistics
 * afterwards describe the adjoint pass. */
int ramp_adjoint(ramp_plan *plan, uint64_t n_hist, uint64_t chunk, const uint32_t *seeds,
                 uint32_t n_seeds, uint32_t det, const float *dL_dh, uint64_t n_bins, double *grad,
                 uint32_t n_grad, const ramp_run_opts *opts);

/* ---- batches of instrument settings --------------------------------------
 * A batch runs many plans - one per instrument setting (scan point, chopper
 * phase, Ei), all created on one context - in one persistent launch with a
 * device work queue, replayed as a HIP graph.  Replaces the reference's
 * per-setting loop of Instrument(...).execute(N) calls (one Instrument per
 * setting, mcramp.py:271, 403-455).  Setting s runs counts[s] histories as
 * one chunk with seeds[s]; its histograms accumulate in plan s (read them
 * with ramp_hist_read) and equal those of ramp_run(plan s, counts[s],
 * counts[s], &seeds[s], 1) - bit for bit in deterministic mode.
 * weight_bounds (may be NULL: 1/counts[s]) bounds one event's weight per
 * setting for the deterministic fixed-point scale.  opts: hist_mode, wg
 * (64/128/256/512), npt (histories per lane per work tile: 1/2/4/8/16);
 * keep_state and flags must be 0. */
typedef struct ramp_batch ramp_batch;
int ramp_batch_create(ramp_ctx *ctx, ramp_plan *const *plans, uint32_t n_settings, ramp_batch **out);
int ramp_batch_destroy(ramp_batch *batch);
int ramp_batch_run(ramp_batch *batch, const uint64_t *counts, const uint32_t *seeds,
                   const double *weight_bounds, const ramp_run_opts *opts);
int ramp_batch_stats_read(const ramp_batch *batch, ramp_run_stats *out);

/* Evaluate the shared FP32 math library (fn = RAMP_FN_*) on n inputs, on the
 * host (on_device = 0) or on the GPU of ctx (on_device = 1).  Test hook for
 * the bit-identity of host and device arithmetic; no reference counterpart. */
#define RAMP_FN_SIN 1
#define RAMP_FN_COS 2
#define RAMP_FN_EXP 3
#define RAMP_FN_LOG 4
#define RAMP_FN_POW 5
#define RAMP_FN_TANH 6
#define RAMP_FN_ATAN2 7
#define RAMP_FN_ASIN 8
#define RAMP_FN_ACOS 9
#define RAMP_FN_FMOD 10
#define RAMP_FN_ROUND 11
#define RAMP_FN_SQRT 12
#define RAMP_FN_DIV 13
int ramp_math_eval(ramp_ctx *ctx, int fn, const float *x, const float *y, float *out, uint64_t n,
                   int on_device);

/* Draws of the per-neutron stream (mcramp/include/rand.h:4-20): out[i] is
 * the u01 value of thread id tid[i] at counter ctr[i].  Host or device. */
int ramp_rng_eval(ramp_ctx *ctx, const uint32_t *tid, const float *ctr, float *out, uint64_t n,
                  int on_device);

const char *ramp_last_error(void);
const char *ramp_version(void);

#ifdef __cplusplus
}
#endif

#endif
