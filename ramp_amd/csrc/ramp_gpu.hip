/* ramp_gpu.hip - libramp.so: fused, stream-compacted HIP segment kernels
 * for the RAMP neutron path on gfx950, and the C-ABI of include/ramp.h.
 *
 * Execution model
 *   The instrument is a list of steps (one per linear-block component, one
 *   per non-linear block).  A run splits every chunk's histories over a
 *   few segment launches.  Segment 0 generates neutrons from the source
 *   (the Threefry key is the chunk-local index) and carries each through
 *   its steps in registers; at the end of a segment the surviving
 *   neutrons are appended to a structure-of-arrays queue with one
 *   wave-wide ballot and one atomic per wave, so the next segment runs
 *   dense waves.  Segment boundaries are placed where a pilot run shows
 *   the live fraction dropping (after choppers, slits).  Detector events
 *   are float atomics (reference-like, order dependent) or int64 fixed
 *   point atomics (order independent, bit-reproducible).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cstdarg>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <mutex>
#include <vector>

#include <hip/hiprtc.h>

#include "../../include/ramp.h"
#include "dev_components.h"
#include "bvh_build.h"
#include "bvh_sah.h"
#include "det_scale.h"

/* ======================================================================== */
/* error handling                                                           */
/* ======================================================================== */
static thread_local std::string g_err;

static int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
static int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(x)                                                                     \
    do {                                                                              \
        hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess)                                                         \
            return fail(RAMP_E_DEVICE, "%s failed: %s", #x, hipGetErrorString(e_));   \
    } while (0)

extern "C" const char *ramp_last_error(void) { return g_err.c_str(); }
extern "C" const char *ramp_version(void) { return "ramp_amd 0.1.0 gfx950"; }

#include "seg_kernel.h"
#include "kern_table.h"

/* exclusive prefix of the wave-region counts of one segment (one block);
 * also records the totals for the run statistics */
__global__ __launch_bounds__(1024) void k_scan(const uint32_t *counts, uint64_t nreg,
                                                const uint64_t *prod_meta, uint64_t *prefix,
                                                uint64_t *cons_meta, unsigned long long *cnt,
                                                int first)
{
    __shared__ uint64_t part[1024];
    const uint32_t t = threadIdx.x;
    const uint64_t per = (nreg + 1023) / 1024;
    const uint64_t b = t * per, e = b + per < nreg ? b + per : nreg;
    uint64_t sum = 0;
    for (uint64_t i = b; i < e; i++) sum += counts[i];
    part[t] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        uint64_t v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint64_t run = part[t] - sum;
    for (uint64_t i = b; i < e; i++) {
        prefix[i] = run;
        run += counts[i];
    }
    if (t == 1023) {
        prefix[nreg] = part[1023];
        cons_meta[0] = prod_meta[0];
        cons_meta[1] = nreg;
        if (!cnt) return;   /* a walk list: no queue counters */
        cnt[CNT_QOUT] += part[1023];
        cnt[CNT_QIN] += part[1023];
        if (first) cnt[CNT_QOUT0] += part[1023];
    }
}

/* ======================================================================== */
/* small kernels                                                            */
/* ======================================================================== */
__global__ void k_fill_u64(unsigned long long *p, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

__global__ void k_fill_f32(float *p, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0.0f;
}

__global__ void k_fill_u32(uint32_t *p, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

__device__ static float math_eval(int fn, float x, float y)
{
    switch (fn) {
    case RAMP_FN_SIN: return rm_sin(x);
    case RAMP_FN_COS: return rm_cos(x);
    case RAMP_FN_EXP: return rm_exp(x);
    case RAMP_FN_LOG: return rm_log(x);
    case RAMP_FN_POW: return rm_pow(x, y);
    case RAMP_FN_TANH: return rm_tanh(x);
    case RAMP_FN_ATAN2: return rm_atan2(x, y);
    case RAMP_FN_ASIN: return rm_asin(x);
    case RAMP_FN_ACOS: return rm_acos(x);
    case RAMP_FN_FMOD: return rm_fmod(x, y);
    case RAMP_FN_ROUND: return rm_round(x);
    case RAMP_FN_SQRT: return rm_sqrt(x);
    case RAMP_FN_DIV: return x / y;
    default: return 0.0f;
    }
}

static float math_eval_host(int fn, float x, float y)
{
    switch (fn) {
    case RAMP_FN_SIN: return rm_sin(x);
    case RAMP_FN_COS: return rm_cos(x);
    case RAMP_FN_EXP: return rm_exp(x);
    case RAMP_FN_LOG: return rm_log(x);
    case RAMP_FN_POW: return rm_pow(x, y);
    case RAMP_FN_TANH: return rm_tanh(x);
    case RAMP_FN_ATAN2: return rm_atan2(x, y);
    case RAMP_FN_ASIN: return rm_asin(x);
    case RAMP_FN_ACOS: return rm_acos(x);
    case RAMP_FN_FMOD: return rm_fmod(x, y);
    case RAMP_FN_ROUND: return rm_round(x);
    case RAMP_FN_SQRT: return rm_sqrt(x);
    case RAMP_FN_DIV: return x / y;
    default: return 0.0f;
    }
}

__global__ void k_math(int fn, const float *x, const float *y, float *out, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = math_eval(fn, x[i], y ? y[i] : 0.0f);
}

__global__ void k_rng(const uint32_t *tid, const float *ctr, float *out, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        float c = ctr[i];
        out[i] = rr_draw(&c, tid[i]);
    }
}

/* ---- mesh walks in their own launch ----------------------------------------
 * The walk of a ray over a mesh's BVH varies several-fold in length from
 * ray to ray, so walks run in lockstep inside a segment leave most lanes of
 * a wave idle.  For the plan's first linear GMesh step, the segment before
 * it prepares every output neutron (walk_emit: the mesh frame, the box
 * precheck) and lists the rays that need a walk; k_mesh_walk walks
 * them with lane refill, and the segment starting at the step takes the
 * hits from the neutron's queue slot.  Every ray runs the same walk and
 * arithmetic as in g_mesh.  The listed rays are walked in the order of
 * their keys (walk_key: direction class, then position across the dominant
 * axis), radix-sorted and gathered between the two launches, so the rays
 * side by side in a wave visit nearly the same nodes and share their loads.
 *
 * A lane whose walk is over takes the next ray from its wave's pool of
 * claimed rays, which the wave refills from a device counter, one atomic
 * per WALK_POOL rays (device-scope atomics on one address are serialised
 * across the XCDs).  Lanes alternate between box steps and triangle tests:
 * a lane whose box step meets a leaf parks the leaf and waits; the wave
 * tests the parked leaves' triangles together once WALK_TRI_BATCH lanes
 * hold one (or no lane can step), so the long triangle test runs with most
 * lanes active instead of once per step for the few lanes at a leaf.  Refills
 * likewise run once WALK_REFILL lanes are idle. */
struct WalkArgs {
    const WalkEntry *list;     /* the listed rays in key order, [n] */
    uint64_t n;                /* rays listed */
    const DComp *comp;         /* the mesh component */
    WalkHit *res;              /* the outcome per queue slot */
    unsigned long long *next;  /* work counter, zero at launch */
};

/* the listed rays in key order (order: their list indices after the sort) */
__global__ void k_walk_gather(const WalkEntry *list, const uint32_t *order, uint64_t n, WalkEntry *out)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = list[order[i]];
}

/* The keys in dense order: one block per producing wave, item i = prefix[r]
 * + k of region r is list entry e = r * stride + k. */
__global__ void k_walk_order(const uint32_t *keys, const uint64_t *prefix, const uint64_t *meta, uint32_t *okeys,
                             uint32_t *oidx)
{
    const uint64_t r = blockIdx.x, stride = meta[0];
    if (r >= meta[1]) return;
    const uint64_t b = prefix[r], cnt = prefix[r + 1] - b;
    for (uint64_t k = threadIdx.x; k < cnt; k += blockDim.x) {
        const uint64_t e = r * stride + k;
        okeys[b + k] = keys[e];
        oidx[b + k] = (uint32_t)e;
    }
}

#define WALK_KEY_BITS 19     /* walk_key: 3 bits of direction class, 16 of position */

#ifndef WALK_POOL
#define WALK_POOL 1024u     /* rays a wave claims per atomic on the work counter */
#endif
#ifndef WALK_REFILL
#define WALK_REFILL 24u     /* idle lanes that trigger a refill pass */
#endif
#ifndef WALK_TRI_BATCH
#define WALK_TRI_BATCH 24u  /* lanes with a parked leaf that trigger the tests */
#endif
#ifndef WALK_TRI_Q4
#define WALK_TRI_Q4 3u      /* ... or this many quarters of the walking lanes */
#endif
#define IDLE 0xffffffffu    /* a lane without a ray (queue slots stay below 2^32 - 1) */

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_mesh_walk(WalkArgs a)
{
    const DComp &c = plan_at(a.comp, 0);
    const float *pts = c.g.tab[GT_POINTS];
    const BvhTri *lt = (const BvhTri *)c.g.aux2;
    const BvhFlat *nodes = (const BvhFlat *)c.g.aux;
    const bool brute = nodes == nullptr;
    const uint32_t ntri = c.g.ni[GI_NTRI], interior = c.g.ni[GI_INTERIOR];
    const uint32_t lane = __lane_id();
    const unsigned long long below = (1ull << lane) - 1ull;
    const uint64_t total = a.n;

    /* wave-uniform: the wave's pool of claimed rays [pool, pool_end) and
     * whether the counter has passed the total */
    uint64_t pool = 0, pool_end = 0;
    bool drained = false;
    uint32_t my = IDLE;        /* this lane's queue slot */
    int32_t tri = -1;          /* its parked leaf (brute force: triangle); -1 none */
    /* A refilled lane loads its entry (fresh 2), then starts the walk with
     * its first node load (0): each load is used an iteration after it is
     * issued, so the wave never waits on one. */
    uint32_t fresh = 0;
    MeshWalk w;
    BvhFlat nd;                /* the lane's next node, loaded one step ahead */
    for (;;) {
        const unsigned long long im = __ballot(my == IDLE);
        if (!drained && im && ((uint32_t)__popcll(im) >= WALK_REFILL || im == __ballot(true))) {
            if (pool >= pool_end) {
                unsigned long long base = 0;
                if (lane == 0) base = atomicAdd(a.next, (unsigned long long)WALK_POOL);
                /* every lane runs here: lane 0 is the first active one, and
                 * the claimed base lives in scalar registers */
                base = ((unsigned long long)(uint32_t)__builtin_amdgcn_readfirstlane((int)(base >> 32)) << 32) |
                       (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)base);
                if (base >= total) {
                    drained = true;
                } else {
                    pool = base;
                    pool_end = base + WALK_POOL < total ? base + WALK_POOL : total;
                }
            }
            if (!drained) {
                const uint64_t avail = pool_end - pool;
                const uint32_t cnt = (uint32_t)__popcll(im);
                const uint32_t take = (uint64_t)cnt < avail ? cnt : (uint32_t)avail;
                const uint32_t rank = (uint32_t)__popcll(im & below);
                if (((im >> lane) & 1ull) != 0ull && rank < take) {
                    const WalkEntry &we = a.list[pool + rank];
                    w.pos = rf3_make(we.pos[0], we.pos[1], we.pos[2]);
                    w.vel = rf3_make(we.vel[0], we.vel[1], we.vel[2]);
                    my = we.slot;
                    tri = -1;
                    fresh = 2u;
                }
                pool += take;
            }
        }
        const unsigned long long walking = __ballot(my != IDLE);
        if (!walking) {
            if (drained) break;
            continue;
        }
        /* the parked triangles */
        const unsigned long long parked = __ballot(tri >= 0);
        const uint32_t npark = (uint32_t)__popcll(parked);
        if (npark >= WALK_TRI_BATCH || (parked && 4u * npark >= WALK_TRI_Q4 * (uint32_t)__popcll(walking))) {
            if (tri >= 0) {
                if (brute) mesh_walk_brute(w, pts, (uint32_t)tri);
                else mesh_walk_leaf(w, lt, tri);
                tri = -1;
            }
        }
        /* one box step of every walking lane without a parked triangle */
        if (fresh == 2u) {
            mesh_walk_init(w, c.g, w.pos, w.vel);
            if (!brute && w.cur < w.end) nd = nodes[w.cur];
            fresh = 0u;
        } else if (my != IDLE && tri < 0) {
            if (w.cur < w.end) {
                if (!brute) {
                    tri = mesh_walk_node(w, nd, ntri, interior != 0);
                    if (w.cur < w.end) nd = nodes[w.cur];   /* consumed at the lane's next step */
                } else {
                    tri = w.cur++;
                }
            }
            if (tri < 0 && w.cur >= w.end) {
                WalkHit h;
                h.t1 = w.t[0];
                h.t2 = w.t[1];
                h.code = mesh_walk_code(w, interior);
                h.pad = 0u;
                a.res[my] = h;
                my = IDLE;
            }
        }
    }
}

/* ======================================================================== */
/* host: context and plan                                                   */
/* ======================================================================== */
struct ramp_ctx {
    int device;
    int n_cu;
    hipStream_t stream;
    hipEvent_t ev[12];
    std::string arch;   /* gcnArchName, for run-time compiled plugins */
    void *scratch = nullptr; /* this device's scratch array (ramp_run_noalloc) */
};

struct ramp_plugin {
    ramp_ctx *ctx = nullptr;
    hipModule_t mod = nullptr;
    /* ramp_plugin_kernel, ramp_plugin_kernel_trace (runs with RAMP_RUN_TRACE)
     * and their _mesh forms (for a GMesh geometry), indexed [mesh][trace] */
    hipFunction_t fn[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    unsigned grid[2][2] = {{0, 0}, {0, 0}}; /* persistent grids: the co-resident blocks of 256 */
};

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

static int dev_alloc(DevBuf &b, size_t bytes)
{
    if (b.bytes >= bytes && b.p) return 0;
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
    if (bytes == 0) return 0;
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) return fail(RAMP_E_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    b.bytes = bytes;
    return 0;
}

static void dev_free(DevBuf &b)
{
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
}

struct QueueBufs {
    DevBuf f[23];   /* x y z vx vy vz w t ctr px py pz sc sd se, intersection[8] */
    DevBuf key, status;
};

/* feature-set instances of the segment kernel, cheapest first */
static const uint32_t kVariants[] = RAMP_KV_FEATS;
static const int kNumVariants = RAMP_NUM_KV;

/* the instances of feature set v (filled on first use) */
static const KernTable &kern_table(int v)
{
    static KernTable tab[RAMP_NUM_KV];
    static bool init = false;
    if (!init) {
        ramp_kern_v0(&tab[0]);
        ramp_kern_v1(&tab[1]);
        ramp_kern_v2(&tab[2]);
        ramp_kern_v3(&tab[3]);
        ramp_kern_v4(&tab[4]);
        ramp_kern_v5(&tab[5]);
        ramp_kern_full_v0(&tab[0]);
        ramp_kern_full_v1(&tab[1]);
        ramp_kern_full_v2(&tab[2]);
        ramp_kern_full_v3(&tab[3]);
        ramp_kern_full_v4(&tab[4]);
        ramp_kern_full_v5(&tab[5]);
        ramp_kern_ba_v0(&tab[0]);
        ramp_kern_ba_v1(&tab[1]);
        ramp_kern_ba_v2(&tab[2]);
        ramp_kern_ba_v3(&tab[3]);
        ramp_kern_ba_v4(&tab[4]);
        ramp_kern_ba_v5(&tab[5]);
        init = true;
    }
    return tab[v < 0 || v >= RAMP_NUM_KV ? RAMP_NUM_KV - 1 : v];
}

/* the largest trace buffer a run may allocate (16 B per traced component
 * and history) */
static const double TRACE_MAX_BYTES = 32.0 * 1073741824.0;

/* the largest pinned staging buffer a plan keeps for histogram reads */
static const size_t PIN_MAX = (size_t)64 << 20;

struct ramp_plan {
    ramp_ctx *ctx = nullptr;
    std::vector<DComp> hcomps;
    std::vector<DStep> hsteps;
    DPart hsrc;
    std::vector<DDet> hdets;
    std::vector<ramp_det_desc> det_desc;
    DevBuf comps, steps, src, dets, tables, cnt, count;
    std::vector<DevBuf> hw, hw2, hq, hq2;
    std::vector<DevBuf> mesh_nodes;
    DevBuf comps_brute;  /* comps with every mesh's BVH removed (RAMP_RUN_MESH_BRUTE_FORCE) */
    QueueBufs q[2];
    uint64_t qcap = 0;
    bool qfull = false;
    DevBuf slot_ctr, slot_iidx, st16, st8, stii, rcounts, prefix;
    uint64_t st_n = 0;
    DetScale ds;              /* deterministic mode's fixed-point scale and int64 budget */
    std::vector<std::vector<double>> fold_w, fold_w2; /* deterministic sums folded to the host (FP64) */
    bool has_atomic = false, has_det = false;
    std::vector<int> bounds;  /* segment boundaries (step indices) */
    int defer_k = -1;         /* first linear monochromator step (reflections deferred) */
    int walk_k = -1;          /* first linear GMesh step: its walks run in k_mesh_walk (fast mode) */
    /* the walks: hits [3][queue slots], the walk list and its keys [queue
     * slots], its region counts, [prefix (waves + 1), meta (2), work counter
     * (1)], the keys and list indices in dense order and sorted, and the
     * sort's scratch */
    DevBuf walk_res, walk_list, walk_keys, walk_counts, walk_aux, walk_sort[4], walk_temp, walk_sorted;
    int revive_k = -1;        /* last SMCPLIn step: segment 0 runs at least through it */
    bool has_mcpl = false;    /* SMCPLIn/SMCPLOut present */
    bool need_full = false;   /* MCPL, SRescal or SAnalyser present: full-state execution */
    uint32_t n_rescal = 0;    /* SRescal components (event slots) */
    uint32_t acc_bins = 0, acc_bytes = 0; /* small detectors' LDS copy (set per run by the mode) */
    uint32_t n_trace = 0;     /* traced components (vis = 1 in linear blocks) */
    DevBuf trace;             /* their trace points [slot][4][trace_n] */
    uint64_t trace_n = 0;     /* histories of the last traced run (0: the last run was not traced) */
    void *pin = nullptr;      /* pinned host staging for histogram reads (hipHostMalloc) */
    size_t pin_bytes = 0;
    unsigned long long *cnt_host = nullptr;  /* pinned copy of the run counters (hipHostMalloc) */
    std::vector<int> plugin_step;           /* per step: the component number of a plugin step, else -1 */
    std::vector<ramp_plugin *> plugin_of;   /* per component: the bound plugin */
    std::vector<uint32_t> rescal_comp; /* their component numbers */
    DevBuf ev, ev_tag, ev_list, ev_list_n; /* per-chunk events, tags; accumulated lists [slot][5][list_cap], counts */
    uint64_t ev_cap = 0, ev_list_cap = 0;
    std::vector<uint64_t> ev_count; /* host copy of the list counts after the last run */
    uint32_t n_mcpl_out = 0;  /* SMCPLOut components (dump slots) */
    DevBuf dump, dump_tag, sel_counts, sel_prefix, sel_out; /* SMCPLOut records and their compaction */
    uint64_t dump_cap = 0;
    uint32_t dump_chunk_tag = 0; /* tag of the last chunk of the last run */
    uint64_t dump_n = 0;         /* histories of that chunk */
    std::vector<float> pd_lines; /* host copy of the SPowderDiff line table (adjoint) */
    bool bounds_set = false;
    uint64_t bounds_chunk = 0;   /* the chunk the boundaries were chosen for */
    int bounds_max_seg = 0;      /* and the segment limit (0: none) */
    int variant = kNumVariants - 1;
    uint32_t feat = FEAT_ALL;  /* FEAT_* bits the plan's components need */
    ramp_run_stats stats;
};

/* ---- scratch of the no-allocation run (ramp_run_noalloc) ----------------
 * A fast-mode run needs, besides the plan: two segment queues (9 float
 * fields, key and status: 44 B per record), the per-wave region counts and
 * their prefix, and with several chunks each slot's RNG counter and
 * component index.  ramp_run allocates them per plan; ramp_run_noalloc
 * takes them from `scratch`, one static device array per GPU, sized for
 * chunks of at most RAMP_NOALLOC_MAX_CHUNK histories and NA_MAX_WAVES
 * co-resident waves (256 CUs x 32 waves).  Its device address is taken once
 * per device with hipGetSymbolAddress when the first context on that device
 * is created; host code uses only that address. */
static const uint64_t NA_MAX_WAVES = 8192;
static const uint64_t NA_CAP = RAMP_NOALLOC_MAX_CHUNK + NA_MAX_WAVES;   /* queue records */
static const size_t NA_ALIGN = 256;
static constexpr size_t na_round(size_t b) { return (b + NA_ALIGN - 1) / NA_ALIGN * NA_ALIGN; }
static const size_t NA_QUEUE_FIELD = na_round(NA_CAP * 4);             /* one field of one queue */
static const size_t NA_OFF_COUNTS = 2 * 11 * NA_QUEUE_FIELD;
static const size_t NA_OFF_PREFIX = NA_OFF_COUNTS + na_round(NA_MAX_WAVES * 4);
static const size_t NA_OFF_CTR = NA_OFF_PREFIX + na_round((NA_MAX_WAVES + 1) * 8);
static const size_t NA_OFF_IIDX = NA_OFF_CTR + na_round(RAMP_NOALLOC_MAX_CHUNK * 4);
static const size_t NA_BYTES = NA_OFF_IIDX + na_round(RAMP_NOALLOC_MAX_CHUNK * 4);
static_assert(NA_BYTES == RAMP_NOALLOC_SCRATCH_BYTES, "RAMP_NOALLOC_SCRATCH_BYTES in ramp.h");

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

static std::mutex g_scratch_mu;
static std::vector<void *> g_scratch_ptr;   /* per device: the address of `scratch` */

/* the device address of this device's `scratch` (device current) */
static int scratch_setup(int device, void **out)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((size_t)device >= g_scratch_ptr.size()) g_scratch_ptr.resize(device + 1, nullptr);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(RAMP_E_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", device,
                        hipGetErrorString(e));
        g_scratch_ptr[device] = ptr;
    }
    *out = g_scratch_ptr[device];
    return 0;
}

extern "C" int ramp_ctx_destroy(ramp_ctx *ctx);

extern "C" int ramp_ctx_create(int device, ramp_ctx **out)
{
    if (!out) return fail(RAMP_E_INVALID, "null output pointer");
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(RAMP_E_INVALID, "device %d of %d", device, ndev);
    HIPCHK(hipSetDevice(device));
    ramp_ctx *c = new ramp_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    c->n_cu = prop.multiProcessorCount;
    c->arch = prop.gcnArchName;
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (int i = 0; i < 12; i++) HIPCHK(hipEventCreate(&c->ev[i]));
    int rc = scratch_setup(device, &c->scratch);
    if (rc) {
        (void)ramp_ctx_destroy(c);
        return rc;
    }
    *out = c;
    return 0;
}

extern "C" int ramp_ctx_destroy(ramp_ctx *ctx)
{
    if (!ctx) return 0;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (int i = 0; i < 12; i++) (void)hipEventDestroy(ctx->ev[i]);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return 0;
}

/* ---- parameter-only expressions of the reference, evaluated once -------- */
static float maxwell_ref_host(float E, float T)
{
    /* cont_mod.cl:4-9 with its double promotion (2.0f*rootE/sqrt(M_PI)) */
    float rootE = rm_sqrt(E);
    float overKT = 1.0f / (RC_K_B * T);
    double r = (double)(2.0f * rootE) / sqrt(RC_M_PI_D);
    r = r * (double)rm_pow(overKT, (float)(3.0f / 2.0f));
    r = r * (double)rm_exp(-E * overKT);
    return (float)r;
}

/* a double in slots i, i+1 (read by dp_f64) */
static void set_d64(float *d, int i, double v)
{
    std::memcpy(d + i, &v, sizeof v);
}

/* smallest float threshold T with: refl < T  <=>  (double)refl < 1e-12 */
static float thresh_lt_double(double c)
{
    float f = (float)c;
    if ((double)f < c) f = nextafterf(f, INFINITY);
    return f;
}

static void derive_source(DPart &s)
{
    std::memset(s.d, 0, sizeof s.d);
    if (s.type == RAMP_SRC_CONT) {
        float T[3] = {s.p[SP_T1], s.p[SP_T2], s.p[SP_T3]};
        for (int i = 0; i < 3; i++) {
            float overKT = 1.0f / (RC_K_B * T[i]);
            s.d[i] = overKT;
            s.d[3 + i] = rm_pow(overKT, (float)(3.0f / 2.0f));
            s.d[6 + i] = maxwell_ref_host(RC_K_B * T[i], T[i]);
        }
        float EjJ = s.p[SP_TAIL_JOIN] * (RC_K_B * T[0]);
        s.d[9] = EjJ;
        s.d[10] = s.p[SP_TAIL_JOIN] > 0.0f ? maxwell_ref_host(EjJ, T[0]) : 0.0f;
        /* sqrt(M_PI) and its reciprocal, doubles (maxwell_dev) */
        set_d64(s.d, 18, sqrt(RC_M_PI_D));
        set_d64(s.d, 20, 1.0 / sqrt(RC_M_PI_D));
        /* terms 2 and 3 with I == 0 add exactly +-0 when the Maxwellian is
         * finite, which holds for 1 K <= T <= 1e5 K and E <= 1e4 meV */
        for (int i = 1; i < 3; i++) {
            float I = i == 1 ? s.p[SP_I2] : s.p[SP_I3];
            bool finite = T[i] >= 1.0f && T[i] <= 1e5f && s.p[SP_EMAX] <= 1e4f && s.p[SP_EMIN] >= 0.0f;
            s.d[15 + i] = (I == 0.0f && finite && s.d[6 + i] > 0.0f && std::isfinite(s.d[6 + i])) ? 1.0f : 0.0f;
        }
    }
}

static void derive_geom(DPart &g)
{
    std::memset(g.d, 0, sizeof g.d);
    switch (g.type) {
    case RAMP_G_PLANE:
        g.d[0] = g.p[GP_W] / 2.0f;
        g.d[1] = g.p[GP_H] / 2.0f;
        break;
    case RAMP_G_BOX:
        g.d[0] = -g.p[GP_W] / 2.0f; g.d[1] = -g.p[GP_H] / 2.0f; g.d[2] = -g.p[GP_D] / 2.0f;
        g.d[3] = g.p[GP_W] / 2.0f; g.d[4] = g.p[GP_H] / 2.0f; g.d[5] = g.p[GP_D] / 2.0f;
        break;
    case RAMP_G_BANANA: {
        float mx = g.p[GP_MAXTH];
        /* (mintheta < theta2) is 0 or 1, then compared with maxtheta */
        g.d[0] = (mx > 1.0f) ? 0.0f : (mx <= 0.0f ? 1.0f : 2.0f);
        break;
    }
    default: break;
    }
}

static void derive_scat(DPart &s)
{
    std::memset(s.d, 0, sizeof s.d);
    switch (s.type) {
    case RAMP_S_CHOPPER:
        s.d[0] = (float)((double)2.0f * RC_M_PI_D / (double)(float)s.ni[CI_NSLITS]);
        break;
    case RAMP_S_GUIDE: {
        float w1 = s.p[GU_W1], h1 = s.p[GU_H1], w2 = s.p[GU_W2], h2 = s.p[GU_H2], l = s.p[GU_L];
        float ww = (float)(0.5 * (double)(w2 - w1)), hh = (float)(0.5 * (double)(h2 - h1));
        s.d[0] = ww;
        s.d[1] = hh;
        s.d[2] = (float)(0.5 * (double)w1);
        s.d[3] = (float)(0.5 * (double)h1);
        float m = s.p[GU_M], W = s.p[GU_W], alpha = s.p[GU_ALPHA], beta = 0.0f;
        if (W == 0.0 && alpha == 0.0) {
            m = (float)((double)m * 0.9853 + 0.1978);
            W = (float)(-0.0002 * (double)m + 0.0022);
            alpha = (float)(0.2304 * (double)m + 5.0944);
            beta = (float)(-7.6251 * (double)m + 68.1137);
            if (m <= 3) { alpha = m; beta = 0; }
        }
        s.d[4] = m; s.d[5] = W; s.d[6] = alpha; s.d[7] = beta;
        s.d[8] = thresh_lt_double(1e-12);
        s.d[9] = RC_V2K * -2.0f;
        float nv = l * l + ww * ww, nh = l * l + hh * hh;
        s.d[10] = rm_sqrt(nv); s.d[11] = rm_sqrt(nh);
        s.d[12] = nv; s.d[13] = nh;
        /* reciprocals of the roots, the double-precision divisors of q
         * (rd_div_rc) */
        set_d64(s.d, 18, 1.0 / (double)s.d[10]);
        set_d64(s.d, 20, 1.0 / (double)s.d[11]);
        float Qc = s.p[GU_QC];
        float m2 = m;
        bool zero = (m <= 0 || Qc <= 0 || s.p[GU_R0] <= 0);
        if (m2 < 1) { Qc *= m2; m2 = 1; }
        s.d[14] = Qc; s.d[15] = m2; s.d[16] = zero ? 1.0f : 0.0f;
        break;
    }
    case RAMP_S_ISOTROPIC:
        set_d64(s.d, 0, 2 * RC_M_PI_D);
        break;
    case RAMP_S_POWDER1:
    case RAMP_S_POWDERN: {
        float d_phi = s.type == RAMP_S_POWDER1 ? s.p[P1_DPHI] : s.p[PN_DPHI];
        set_d64(s.d, 0, (double)d_phi * RC_M_PI_D / 180.0f);
        break;
    }
    case RAMP_S_MONO: {
        float l8 = rm_sqrt(8.0f * rm_log(2.0f));
        float mry = (float)(RC_MIN2RAD_D * (double)s.p[MO_MOS_H] / (double)l8);
        float mrz = (float)(RC_MIN2RAD_D * (double)s.p[MO_MOS_V] / (double)l8);
        s.d[0] = mry; s.d[1] = mrz; s.d[2] = mry > mrz ? mry : mrz;
        /* GAUSS constants per mosaic width (mono_gauss_c): 1/(2*rms*rms)
         * of the float product, sqrt(2.0f*M_PI)*rms and its reciprocal */
        for (int j = 0; j < 2; j++) {
            const float rms = j == 0 ? mry : mrz;
            const float c2 = 2.0f * rms * rms;
            const double dn = RC_SQRT_2PI_D * (double)rms;
            set_d64(s.d, 6 + 6 * j, 1.0 / (double)c2);
            set_d64(s.d, 8 + 6 * j, dn);
            set_d64(s.d, 10 + 6 * j, 1.0 / dn);
        }
        s.d[3] = (float)((double)2.0f * RC_M_PI_D / (double)s.p[MO_D]);
        float sw = s.p[MO_SLAB_W], sh = s.p[MO_SLAB_H], gap = s.p[MO_GAP];
        s.d[4] = (((float)s.ni[MOI_NH] * (sw + gap)) - gap) / 2.0f;
        s.d[5] = (((float)s.ni[MOI_NV] * (sh + gap)) - gap) / 2.0f;
        break;
    }
    case RAMP_S_DET1D:
    case RAMP_S_DET2D:
        set_d64(s.d, 0, 2 * RC_M_PI_D);
        break;
    case RAMP_S_RESCAL: {
        /* rescal.cl:36-50, the parameter-only part: cos of the cone
         * half-angle and target x target_perp */
        rf3 t = rf3_make(s.p[RS_T], s.p[RS_T + 1], s.p[RS_T + 2]);
        float fr = s.p[RS_FOCUS_R];
        float l2 = v3_len(t) * v3_len(t);
        s.d[0] = rm_sqrt(l2 / (fr * fr + l2));
        rf3 tp = (t.x == 0.0f && t.z == 0.0f) ? rf3_make(1.0f, 0.0f, 0.0f) : rf3_make(-t.z, 0.0f, t.x);
        rf3 tpp = v3_cross(t, tp);
        s.d[1] = tpp.x; s.d[2] = tpp.y; s.d[3] = tpp.z;
        break;
    }
    default: break;
    }
}

/* ---- BVH over the triangles of a mesh, built on the GPU (bvh_build.h) ---- */
__global__ void k_fill_i32(int *p, uint64_t n, int v)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

/* The BVH of a GMesh: a binned-SAH tree built on the host (bvh_sah.h) up
 * to BVH_SAH_MAX_TRIS triangles, the GPU LBVH (bvh_build.h) above that or
 * when the host triangles are not at hand; RAMP_BVH_BUILDER=lbvh or =sah
 * forces one.  Either tree is written as the six preorder layouts. */
#define BVH_SAH_MAX_TRIS (1u << 21)

/* the six preorder layouts of a build-layout tree of nn nodes (on the
 * device), into a new buffer held by the plan */
static int bvh_layouts(ramp_plan *p, DPart &g, const BvhNode *nodes, const int *parent, int nn)
{
    hipStream_t s = p->ctx->stream;
    DevBuf flat;
    int rc = dev_alloc(flat, BVH_LAYOUTS * (uint64_t)nn * sizeof(BvhFlat));
    if (rc) return rc;
    for (int lay = 0; lay < BVH_LAYOUTS; lay++)
        hipLaunchKernelGGL(k_bvh_flatten, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, nodes, parent, nn,
                           lay, (BvhFlat *)flat.p + (size_t)lay * nn);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        dev_free(flat);
        return fail(RAMP_E_DEVICE, "BVH layouts: %s", hipGetErrorString(e));
    }
    g.aux = flat.p;
    g.ni[GI_NNODES] = (uint32_t)nn;
    p->mesh_nodes.push_back(flat);
    return 0;
}

static int build_mesh_sah(ramp_plan *p, DPart &g, const float *hpts, uint32_t ntri, float margin)
{
    hipStream_t s = p->ctx->stream;
    const int n = (int)ntri;
    SahBuilder b;
    b.P = hpts;
    b.n = n;
    b.margin = margin;
    b.init();
    int nn = 0;
    b.build(0, n, -1, nn, 0);
    /* the leaf-order triangle records */
    std::vector<BvhTri> lt(n);
    for (int k = 0; k < n; k++) lt[k] = bvh_tri_record(hpts + 9 * (size_t)b.order[k], b.order[k]);
    DevBuf nodes, parent, tris;
    int rc = dev_alloc(nodes, (uint64_t)nn * sizeof(BvhNode));
    if (!rc) rc = dev_alloc(parent, (uint64_t)nn * sizeof(int));
    if (!rc) rc = dev_alloc(tris, (uint64_t)n * sizeof(BvhTri));
    hipError_t e = hipSuccess;
    if (!rc) {
        e = hipMemcpyAsync(nodes.p, b.nodes.data(), (size_t)nn * sizeof(BvhNode), hipMemcpyHostToDevice, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(parent.p, b.parent.data(), (size_t)nn * sizeof(int), hipMemcpyHostToDevice, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(tris.p, lt.data(), (size_t)n * sizeof(BvhTri), hipMemcpyHostToDevice, s);
        if (e != hipSuccess) rc = fail(RAMP_E_DEVICE, "BVH upload: %s", hipGetErrorString(e));
    }
    if (!rc) rc = bvh_layouts(p, g, (const BvhNode *)nodes.p, (const int *)parent.p, nn);
    dev_free(nodes);
    dev_free(parent);
    if (rc) {
        dev_free(tris);
        return rc;
    }
    g.aux2 = tris.p;
    p->mesh_nodes.push_back(tris);
    return 0;
}

static int build_mesh(ramp_plan *p, DPart &g, const float *dpts, const float *hpts, uint32_t ntri)
{
    g.aux = nullptr;
    g.aux2 = nullptr;
    if (ntri == 0) return 0;
    if (ntri > BVH_LEAF_POS_MASK) return fail(RAMP_E_RANGE, "mesh of %u triangles", ntri);
    hipStream_t s = p->ctx->stream;
    int n = (int)ntri;
    bool sah = hpts != nullptr && ntri <= BVH_SAH_MAX_TRIS;
    if (const char *f = std::getenv("RAMP_BVH_BUILDER")) {
        if (std::strcmp(f, "lbvh") == 0) sah = false;
        else if (std::strcmp(f, "sah") == 0 && hpts != nullptr) sah = true;
    }
    if (sah) {
        float ext = 0.0f;
        for (int a = 0; a < 3; a++) ext = std::max(ext, g.p[GP_XMAX + 2 * a] - g.p[GP_XMIN + 2 * a]);
        return build_mesh_sah(p, g, hpts, ntri, std::max(1e-3f * ext, 1e-6f));
    }
    DevBuf keys, vals, skeys, order, nodes, parent, flags, temp, tris;
    int rc = dev_alloc(keys, n * 4ull);
    if (!rc) rc = dev_alloc(vals, n * 4ull);
    if (!rc) rc = dev_alloc(skeys, n * 4ull);
    if (!rc) rc = dev_alloc(order, n * 4ull);
    if (!rc) rc = dev_alloc(nodes, (2ull * n - 1) * sizeof(BvhNode));
    if (!rc) rc = dev_alloc(parent, (2ull * n - 1) * 4ull);
    if (!rc) rc = dev_alloc(flags, (uint64_t)std::max(1, n - 1) * 4ull);
    if (!rc) rc = dev_alloc(tris, (uint64_t)n * sizeof(BvhTri));
    auto cleanup = [&]() {
        dev_free(keys); dev_free(vals); dev_free(skeys); dev_free(parent); dev_free(flags); dev_free(temp);
        dev_free(order); dev_free(nodes);
    };
    if (rc) { cleanup(); dev_free(tris); return rc; }
    float lo[3] = {g.p[GP_XMIN], g.p[GP_YMIN], g.p[GP_ZMIN]};
    float hi[3] = {g.p[GP_XMAX], g.p[GP_YMAX], g.p[GP_ZMAX]};
    float ext = 0.0f, ie[3];
    for (int a = 0; a < 3; a++) {
        float e = hi[a] - lo[a];
        ext = std::max(ext, e);
        ie[a] = e > 0.0f ? 1.0f / e : 0.0f;
    }
    float margin = std::max(1e-3f * ext, 1e-6f);
    unsigned b = 256, gblk = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_bvh_morton, dim3(gblk), dim3(b), 0, s, dpts, ntri, make_float3(lo[0], lo[1], lo[2]),
                       make_float3(ie[0], ie[1], ie[2]), (uint32_t *)keys.p, (uint32_t *)vals.p);
    size_t tbytes = 0;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tbytes, (uint32_t *)keys.p, (uint32_t *)skeys.p,
                                              (uint32_t *)vals.p, (uint32_t *)order.p, n, 0, 30, s));
    rc = dev_alloc(temp, tbytes + 16);
    if (rc) { cleanup(); dev_free(tris); return rc; }
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(temp.p, tbytes, (uint32_t *)keys.p, (uint32_t *)skeys.p,
                                              (uint32_t *)vals.p, (uint32_t *)order.p, n, 0, 30, s));
    hipLaunchKernelGGL(k_fill_i32, dim3(64), dim3(256), 0, s, (int *)parent.p, 2ull * n - 1, -1);
    hipLaunchKernelGGL(k_fill_i32, dim3(64), dim3(256), 0, s, (int *)flags.p, (uint64_t)std::max(1, n - 1), 0);
    if (n > 1)
        hipLaunchKernelGGL(k_bvh_tree, dim3((n - 1 + 255) / 256), dim3(256), 0, s, (const uint32_t *)skeys.p,
                           n, (BvhNode *)nodes.p, (int *)parent.p);
    hipLaunchKernelGGL(k_bvh_boxes, dim3(gblk), dim3(b), 0, s, dpts, (const uint32_t *)order.p, n, margin,
                       (BvhNode *)nodes.p, (const int *)parent.p, (int *)flags.p);
    hipLaunchKernelGGL(k_bvh_leaf_tris, dim3(gblk), dim3(b), 0, s, dpts, (const uint32_t *)order.p, n,
                       (BvhTri *)tris.p);
    /* six preorder layouts, one per ray direction class (bvh_build.h) */
    rc = bvh_layouts(p, g, (const BvhNode *)nodes.p, (const int *)parent.p, 2 * n - 1);
    cleanup();
    if (rc) {
        dev_free(tris);
        return rc;
    }
    g.aux2 = tris.p;
    p->mesh_nodes.push_back(tris);
    return 0;
}

static int copy_part(DPart &d, const ramp_part_desc &s, const float *dtab, uint64_t n_tables)
{
    std::memset(&d, 0, sizeof d);
    d.type = s.type;
    for (int i = 0; i < 8; i++) d.ni[i] = s.ni[i];
    for (int i = 0; i < 48; i++) d.p[i] = s.p[i];
    for (int k = 0; k < RAMP_NT; k++) {
        if (s.tab_len[k] == 0) continue;
        if (s.tab_off[k] + s.tab_len[k] > n_tables)
            return fail(RAMP_E_INVALID, "table %d out of range", k);
        d.tab[k] = dtab + s.tab_off[k];
        d.tab_len[k] = s.tab_len[k];
    }
    return 0;
}

static void rotation_matrix(const float rot[3], float R[9])
{
    rmat3 M = rmat3_euler(rot[0], rot[1], rot[2]);
    for (int i = 0; i < 9; i++) R[i] = M.m[i];
}

static int validate_plan(const ramp_plan_desc *d)
{
    if (!d) return fail(RAMP_E_INVALID, "null plan descriptor");
    if (d->n_blocks == 0 || !d->blocks) return fail(RAMP_E_INVALID, "plan has no blocks");
    if (d->n_comps && !d->comps) return fail(RAMP_E_INVALID, "null component array");
    if (d->n_dets && !d->dets) return fail(RAMP_E_INVALID, "null detector array");
    if (d->source.type < RAMP_SRC_FLAT || d->source.type > RAMP_SRC_PARALLEL)
        return fail(RAMP_E_INVALID, "unknown source type %u", d->source.type);
    if (d->source.type == RAMP_SRC_ISIS) {
        const ramp_part_desc &s = d->source;
        uint64_t nt = s.ni[SI_ISIS_NT], ne = s.ni[SI_ISIS_NE];
        if (nt < 2 || ne < 2 || s.tab_len[ST_ISIS_FLUX] < (ne - 1) * nt || s.tab_len[ST_ISIS_TBINS] < nt ||
            s.tab_len[ST_ISIS_EBINS] < ne || s.tab_len[ST_ISIS_EINT] < ne)
            return fail(RAMP_E_INVALID, "ISIS source tables inconsistent with their sizes");
    }
    for (uint32_t b = 0; b < d->n_blocks; b++) {
        const ramp_block_desc &bl = d->blocks[b];
        if ((uint64_t)bl.first + bl.count > d->n_comps)
            return fail(RAMP_E_INVALID, "block %u out of range", b);
        if (!bl.linear && bl.multi > 100000) return fail(RAMP_E_RANGE, "block %u: multi too large", b);
        if (!bl.linear)
            for (uint32_t k = bl.first; k < bl.first + bl.count; k++)
                if (d->comps[k].scat.type == RAMP_S_MCPL_IN || d->comps[k].scat.type == RAMP_S_MCPL_OUT ||
                    d->comps[k].scat.type == RAMP_S_PLUGIN)
                    return fail(RAMP_E_INVALID, "component %u: SMCPLIn/SMCPLOut/plugins must be in a linear block", k);
        /* the event filter reads the detected flag at the end of the block
         * (rescal.py:51-55 is called when its block ends) */
        if (b + 1 < d->n_blocks)
            for (uint32_t k = bl.first; k < bl.first + bl.count; k++)
                if (d->comps[k].scat.type == RAMP_S_RESCAL)
                    return fail(RAMP_E_INVALID, "component %u: SRescal must be in the last block", k);
    }
    for (uint32_t k = 0; k < d->n_comps; k++) {
        const ramp_component_desc &c = d->comps[k];
        if (c.geom.type < RAMP_G_ARM || c.geom.type > RAMP_G_MESH)
            return fail(RAMP_E_INVALID, "component %u: unknown geometry type %u", k, c.geom.type);
        if (c.scat.type < RAMP_S_ARM || c.scat.type > RAMP_S_PLUGIN)
            return fail(RAMP_E_INVALID, "component %u: unknown scatter type %u", k, c.scat.type);
        if (c.comp_idx == 0 || c.comp_idx > 0x00ffffffu)
            return fail(RAMP_E_INVALID, "component %u: index %u out of range", k, c.comp_idx);
        bool hist = c.scat.type == RAMP_S_DET1D || c.scat.type == RAMP_S_DET2D ||
                    c.scat.type == RAMP_S_COUNTER || c.scat.type == RAMP_S_DETPIX;
        if (hist && (c.det < 0 || (uint32_t)c.det >= d->n_dets))
            return fail(RAMP_E_INVALID, "component %u: detector slot %d invalid", k, c.det);
        for (int i = 0; i < 3; i++)
            if (!std::isfinite(c.pos[i]) || !std::isfinite(c.rot[i]))
                return fail(RAMP_E_INVALID, "component %u: non-finite position/rotation", k);
        if (c.scat.type == RAMP_S_CHOPPER && c.scat.ni[CI_NSLITS] == 0)
            return fail(RAMP_E_INVALID, "component %u: chopper with 0 slits", k);
        if (c.scat.type == RAMP_S_MONO && c.scat.p[MO_D] == 0.0f)
            return fail(RAMP_E_INVALID, "component %u: monochromator d-spacing 0", k);
        if (c.scat.type == RAMP_S_MONO && (c.scat.ni[MOI_NH] > 65535u || c.scat.ni[MOI_NV] > 65535u))
            return fail(RAMP_E_INVALID, "component %u: more than 65535 monochromator slabs in a row", k);
        if (c.geom.type == RAMP_G_MESH && c.geom.ni[GI_NTRI] * 9ull > c.geom.tab_len[GT_POINTS])
            return fail(RAMP_E_INVALID, "component %u: mesh table shorter than 9*n_tri", k);
        const ramp_part_desc &s = c.scat;
        if (s.type == RAMP_S_POWDERN &&
            (s.tab_len[PNT_SSCAT_V2] < s.ni[PNI_NLINES] || s.tab_len[PNT_Q] < s.ni[PNI_NLINES]))
            return fail(RAMP_E_INVALID, "component %u: powder tables shorter than the line count", k);
        if (s.type == RAMP_S_MONO &&
            (s.tab_len[MOT_GX] < s.ni[MOI_GLEN] || s.tab_len[MOT_GW] < s.ni[MOI_GLEN]))
            return fail(RAMP_E_INVALID, "component %u: quadrature tables too short", k);
        if (s.type == RAMP_S_POWDERDIFF &&
            (s.ni[PDI_NLINES] == 0 || s.tab_len[PDT_LINES] < (uint64_t)s.ni[PDI_NLINES] * PD_LW ||
             !(s.p[PD_STRAIN] > 0.0f) || !(s.p[PD_LREF] > 0.0f)))
            return fail(RAMP_E_INVALID, "component %u: SPowderDiff needs lines, a line table, strain > 0 "
                        "and lambda_ref > 0", k);
        if (s.type == RAMP_S_ISOTROPIC)
            for (int ch = 0; ch < 3; ch++) {
                uint64_t nq = s.ni[ISI_NQ(ch)], nw = s.ni[ISI_NW(ch)];
                if (nq * nw && s.tab_len[ch] < nq + 2 * nw + nw * nq)
                    return fail(RAMP_E_INVALID, "component %u: S(Q,w) table %d too short", k, ch);
            }
        if (s.type == RAMP_S_MCPL_IN) {
            const uint64_t words = (uint64_t)s.ni[MCI_N] * MCR_N * (s.ni[MCI_DOUBLE] ? 2u : 1u);
            if (s.tab_len[MCT_REC] < words || (s.ni[MCI_DOUBLE] && (s.tab_off[MCT_REC] & 1u)))
                return fail(RAMP_E_INVALID, "component %u: MCPL record table shorter than %u particles "
                            "or misaligned", k, s.ni[MCI_N]);
        }
        if (s.type == RAMP_S_DETPIX && (uint64_t)s.ni[DTI_NTUBE] * s.ni[DTI_NPIX] * s.ni[DTI_NTOF] >
                                           (c.det >= 0 ? d->dets[c.det].n_bins : 0))
            return fail(RAMP_E_INVALID, "component %u: tube bank larger than its histogram", k);
    }
    for (uint32_t k = 0; k < d->n_dets; k++)
        if (d->dets[k].n_bins == 0 || d->dets[k].n_bins > (1ull << 33))
            return fail(RAMP_E_RANGE, "detector %u: %llu bins", k, (unsigned long long)d->dets[k].n_bins);
    return 0;
}

extern "C" int ramp_plan_destroy(ramp_plan *p);

extern "C" int ramp_plan_create(ramp_ctx *ctx, const ramp_plan_desc *d, ramp_plan **out)
{
    int rc = validate_plan(d);
    if (rc) return rc;
    if (!ctx || !out) return fail(RAMP_E_INVALID, "null context or output");
    HIPCHK(hipSetDevice(ctx->device));
    ramp_plan *p = new ramp_plan();
    p->ctx = ctx;
    std::memset(&p->stats, 0, sizeof p->stats);
#define PCHK(x) do { int r_ = (x); if (r_) { ramp_plan_destroy(p); return r_; } } while (0)
    uint64_t ntab = d->n_tables ? d->n_tables : 1;
    PCHK(dev_alloc(p->tables, ntab * sizeof(float)));
    if (d->n_tables && d->tables)
        HIPCHK(hipMemcpy(p->tables.p, d->tables, d->n_tables * sizeof(float), hipMemcpyHostToDevice));
    const float *dtab = (const float *)p->tables.p;
    PCHK(copy_part(p->hsrc, d->source, dtab, d->n_tables));
    derive_source(p->hsrc);
    p->hcomps.resize(d->n_comps);
    for (uint32_t k = 0; k < d->n_comps; k++) {
        const ramp_component_desc &c = d->comps[k];
        DComp &dc = p->hcomps[k];
        std::memset(&dc, 0, sizeof dc);
        PCHK(copy_part(dc.g, c.geom, dtab, d->n_tables));
        PCHK(copy_part(dc.s, c.scat, dtab, d->n_tables));
        derive_geom(dc.g);
        derive_scat(dc.s);
        if (c.scat.type == RAMP_S_POWDERDIFF && d->tables)
            p->pd_lines.assign(d->tables + c.scat.tab_off[PDT_LINES],
                               d->tables + c.scat.tab_off[PDT_LINES] + (size_t)c.scat.ni[PDI_NLINES] * PD_LW);
        if (c.geom.type == RAMP_G_MESH)
            PCHK(build_mesh(p, dc.g, dtab + c.geom.tab_off[GT_POINTS],
                            d->tables ? d->tables + c.geom.tab_off[GT_POINTS] : nullptr, c.geom.ni[GI_NTRI]));
        dc.comp_idx = c.comp_idx;
        dc.restore = c.restore;
        dc.det = c.det;
        for (int i = 0; i < 3; i++) dc.pos[i] = c.pos[i];
        rotation_matrix(c.rot, dc.R);
        dc.pad = (c.rot[0] == 0.0f && c.rot[1] == 0.0f && c.rot[2] == 0.0f) ? 1u : 0u;
        if (c.scat.type == RAMP_S_MCPL_OUT) dc.s.ni[MCO_SLOT] = p->n_mcpl_out++;
        if (c.scat.type == RAMP_S_MCPL_IN || c.scat.type == RAMP_S_MCPL_OUT) p->has_mcpl = true;
        if (c.scat.type == RAMP_S_RESCAL) {
            dc.s.ni[RSI_SLOT] = p->n_rescal++;
            p->rescal_comp.push_back(k);
        }
    }
    uint32_t feat = 0;
    if (d->source.type == RAMP_SRC_ISIS) feat |= FEAT_ISIS;
    for (uint32_t k = 0; k < d->n_comps; k++) {
        uint32_t g = d->comps[k].geom.type, sc = d->comps[k].scat.type;
        if (g == RAMP_G_MESH) feat |= FEAT_MESH;
        if (sc == RAMP_S_ISOTROPIC || sc == RAMP_S_POWDER1 || sc == RAMP_S_POWDERN || sc == RAMP_S_POWDERDIFF)
            feat |= FEAT_SAMPLE;
        if (sc == RAMP_S_MONO) feat |= FEAT_MONO;
    }
    for (uint32_t b = 0; b < d->n_blocks; b++)
        if (d->blocks[b].count != 0 && !d->blocks[b].linear) feat |= FEAT_NLIN;
    p->feat = feat;
    for (int v = 0; v < kNumVariants; v++)
        if ((feat & ~kVariants[v]) == 0) { p->variant = v; break; }
    /* steps: blocks in order, starting from block 0 (mcramp.py:452-453) */
    for (uint32_t b = 0; b < d->n_blocks; b++) {
        const ramp_block_desc &bl = d->blocks[b];
        if (bl.count == 0) continue;
        if (bl.linear) {
            for (uint32_t k = 0; k < bl.count; k++) {
                DStep s = {0u, bl.first + k, 1u, 1u};
                p->hsteps.push_back(s);
            }
        } else {
            DStep s = {1u, bl.first, bl.count, bl.multi};
            p->hsteps.push_back(s);
        }
    }
    /* trace slots: components drawn (vis) in linear blocks, in plan order */
    for (DComp &c : p->hcomps) c.trace = -1;
    for (uint32_t b = 0; b < d->n_blocks; b++)
        if (d->blocks[b].linear)
            for (uint32_t k = d->blocks[b].first; k < d->blocks[b].first + d->blocks[b].count; k++)
                if (d->comps[k].vis) p->hcomps[k].trace = (int32_t)p->n_trace++;
    if (p->hsteps.size() > CNT_N - CNT_NSTEP0)
        PCHK(fail(RAMP_E_RANGE, "too many steps (%zu)", p->hsteps.size()));
    for (size_t k = 0; k < p->hsteps.size() && p->defer_k < 0; k++)
        if (p->hsteps[k].kind == 0 && p->hcomps[p->hsteps[k].first].s.type == RAMP_S_MONO)
            p->defer_k = (int)k;
    for (size_t k = 0; k < p->hsteps.size() && p->walk_k < 0; k++)
        if (p->hsteps[k].kind == 0 && p->hcomps[p->hsteps[k].first].g.type == RAMP_G_MESH)
            p->walk_k = (int)k;
    for (size_t k = 0; k < p->hsteps.size(); k++)
        if (p->hsteps[k].kind == 0 && p->hcomps[p->hsteps[k].first].s.type == RAMP_S_MCPL_IN)
            p->revive_k = (int)k;
    bool analyser = false, plugin = false;
    for (const DComp &c : p->hcomps) {
        analyser = analyser || c.s.type == RAMP_S_ANALYSER;
        plugin = plugin || c.s.type == RAMP_S_PLUGIN;
    }
    p->need_full = p->has_mcpl || p->n_rescal > 0 || analyser || plugin;
    p->plugin_of.assign(p->hcomps.size(), nullptr);
    p->plugin_step.assign(p->hsteps.size(), -1);
    for (size_t k = 0; k < p->hsteps.size(); k++)
        if (p->hsteps[k].kind == 0 && p->hcomps[p->hsteps[k].first].s.type == RAMP_S_PLUGIN)
            p->plugin_step[k] = (int)p->hsteps[k].first;
    p->ev_count.assign(p->n_rescal, 0);
    p->det_desc.assign(d->dets, d->dets + d->n_dets);
    p->hdets.resize(d->n_dets);
    p->hw.resize(d->n_dets);
    p->hw2.resize(d->n_dets);
    p->hq.resize(d->n_dets);
    p->hq2.resize(d->n_dets);
    for (uint32_t k = 0; k < d->n_dets; k++) {
        std::memset(&p->hdets[k], 0, sizeof(DDet));
        p->hdets[k].n = d->dets[k].n_bins;
        p->hdets[k].acc = -1;
    }
    PCHK(dev_alloc(p->comps, std::max<size_t>(1, p->hcomps.size()) * sizeof(DComp)));
    PCHK(dev_alloc(p->steps, std::max<size_t>(1, p->hsteps.size()) * sizeof(DStep)));
    PCHK(dev_alloc(p->src, sizeof(DPart)));
    PCHK(dev_alloc(p->dets, std::max<size_t>(1, p->hdets.size()) * sizeof(DDet)));
    PCHK(dev_alloc(p->cnt, CNT_N * sizeof(unsigned long long)));
    if (hipHostMalloc((void **)&p->cnt_host, CNT_N * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess)
        PCHK(fail(RAMP_E_DEVICE, "hipHostMalloc of the pinned counter buffer failed"));
    PCHK(dev_alloc(p->count, 64 * sizeof(uint32_t)));
    if (!p->hcomps.empty())
        HIPCHK(hipMemcpy(p->comps.p, p->hcomps.data(), p->hcomps.size() * sizeof(DComp), hipMemcpyHostToDevice));
    if (!p->hsteps.empty())
        HIPCHK(hipMemcpy(p->steps.p, p->hsteps.data(), p->hsteps.size() * sizeof(DStep), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(p->src.p, &p->hsrc, sizeof(DPart), hipMemcpyHostToDevice));
#undef PCHK
    *out = p;
    return 0;
}

extern "C" int ramp_plan_destroy(ramp_plan *p)
{
    if (!p) return 0;
    if (p->ctx) {
        (void)hipSetDevice(p->ctx->device);
        (void)hipStreamSynchronize(p->ctx->stream);
    }
    DevBuf *single[] = {&p->comps, &p->steps, &p->src, &p->dets, &p->tables, &p->cnt, &p->count,
                        &p->slot_ctr, &p->slot_iidx, &p->st16, &p->st8, &p->stii,
                        &p->rcounts, &p->prefix, &p->comps_brute, &p->dump, &p->dump_tag, &p->walk_res, &p->walk_list, &p->walk_keys, &p->walk_counts, &p->walk_aux,
                        &p->walk_sort[0], &p->walk_sort[1], &p->walk_sort[2], &p->walk_sort[3], &p->walk_temp,
                        &p->walk_sorted,
                        &p->sel_counts, &p->sel_prefix, &p->sel_out, &p->ev, &p->ev_tag,
                        &p->ev_list, &p->ev_list_n, &p->trace};
    for (DevBuf *b : single) dev_free(*b);
    if (p->pin) (void)hipHostFree(p->pin);
    if (p->cnt_host) (void)hipHostFree(p->cnt_host);
    for (auto *v : {&p->hw, &p->hw2, &p->hq, &p->hq2, &p->mesh_nodes})
        for (auto &b : *v) dev_free(b);
    for (int i = 0; i < 2; i++) {
        for (auto &b : p->q[i].f) dev_free(b);
        dev_free(p->q[i].key);
        dev_free(p->q[i].status);
    }
    delete p;
    return 0;
}

/* ---- run ---------------------------------------------------------------- */
static void launch_fill_u64(hipStream_t s, void *p, uint64_t n)
{
    if (!n) return;
    unsigned g = (unsigned)std::min<uint64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_u64, dim3(g), dim3(256), 0, s, (unsigned long long *)p, n);
}
static void launch_fill_f32(hipStream_t s, void *p, uint64_t n)
{
    if (!n) return;
    unsigned g = (unsigned)std::min<uint64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_f32, dim3(g), dim3(256), 0, s, (float *)p, n);
}
static void launch_fill_u32(hipStream_t s, void *p, uint64_t n)
{
    if (!n) return;
    unsigned g = (unsigned)std::min<uint64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_u32, dim3(g), dim3(256), 0, s, (uint32_t *)p, n);
}

static int ensure_hist(ramp_plan *p, int mode)
{
    hipStream_t s = p->ctx->stream;
    for (size_t k = 0; k < p->hdets.size(); k++) {
        uint64_t n = p->hdets[k].n;
        if (mode == RAMP_HIST_ATOMIC && !p->hw[k].p) {
            int rc = dev_alloc(p->hw[k], n * sizeof(float));
            if (!rc) rc = dev_alloc(p->hw2[k], n * sizeof(float));
            if (rc) return rc;
            launch_fill_f32(s, p->hw[k].p, n);
            launch_fill_f32(s, p->hw2[k].p, n);
            p->hdets[k].w = (float *)p->hw[k].p;
            p->hdets[k].w2 = (float *)p->hw2[k].p;
        }
        if (mode == RAMP_HIST_DETERMINISTIC && !p->hq[k].p) {
            int rc = dev_alloc(p->hq[k], n * sizeof(unsigned long long));
            if (!rc) rc = dev_alloc(p->hq2[k], n * sizeof(unsigned long long));
            if (rc) return rc;
            launch_fill_u64(s, p->hq[k].p, n);
            launch_fill_u64(s, p->hq2[k].p, n);
            p->hdets[k].qw = (unsigned long long *)p->hq[k].p;
            p->hdets[k].qw2 = (unsigned long long *)p->hq2[k].p;
        }
    }
    if (mode == RAMP_HIST_ATOMIC) p->has_atomic = true;
    else p->has_det = true;
    return 0;
}

static int pin_reserve(ramp_plan *p, uint64_t n);
static int read_fixed(ramp_plan *p, const DDet &d, double *w, double *w2, uint64_t n);

/* Fold the deterministic sums into the host's FP64 accumulators and clear
 * them on the device (before a run that would pass the int64 budget). */
static int det_fold(ramp_plan *p)
{
    hipStream_t s = p->ctx->stream;
    if (p->fold_w.size() != p->hdets.size()) {
        p->fold_w.assign(p->hdets.size(), std::vector<double>());
        p->fold_w2.assign(p->hdets.size(), std::vector<double>());
    }
    for (size_t k = 0; k < p->hdets.size(); k++) {
        const DDet &d = p->hdets[k];
        if (!d.qw) continue;
        if (p->fold_w[k].empty()) {
            p->fold_w[k].assign(d.n, 0.0);
            p->fold_w2[k].assign(d.n, 0.0);
        }
        int rc = read_fixed(p, d, p->fold_w[k].data(), p->fold_w2[k].data(), d.n);
        if (rc) return rc;
        launch_fill_u64(s, d.qw, d.n);
        launch_fill_u64(s, d.qw2, d.n);
    }
    return 0;
}

/* histogram buffers of the mode, the fixed-point scale and budget of
 * deterministic mode (det_scale.h: n_scale histories fix the scale, n_run
 * are added by this run), and the detector table upload */
static int prepare_dets(ramp_plan *p, int mode, double weight_bound, uint64_t n_scale, uint64_t chunk,
                        uint64_t n_run)
{
    int rc = ensure_hist(p, mode);
    if (rc) return rc;
    /* small detectors, in plan order while they fit, get an LDS copy */
    const uint32_t per = mode == RAMP_HIST_DETERMINISTIC ? 16u : 8u;
    p->acc_bins = 0;
    for (auto &d : p->hdets) {
        d.acc = -1;
        if ((uint64_t)(p->acc_bins + d.n) * per <= DET_LDS_BYTES) {
            d.acc = (int32_t)p->acc_bins;
            p->acc_bins += (uint32_t)d.n;
        }
    }
    p->acc_bytes = p->acc_bins * per;
    if (mode == RAMP_HIST_DETERMINISTIC) {
        const double wb = weight_bound > 0 ? weight_bound : 1.0 / (double)chunk;
        if (det_scale_run(p->ds, wb, (double)n_scale, (double)n_run)) {
            rc = det_fold(p);
            if (rc) return rc;
        }
        for (auto &d : p->hdets) { d.sh_w = p->ds.sh; d.sh_w2 = p->ds.sh2; }
    }
    if (!p->hdets.empty())
        HIPCHK(hipMemcpyAsync(p->dets.p, p->hdets.data(), p->hdets.size() * sizeof(DDet), hipMemcpyHostToDevice,
                              p->ctx->stream));
    return 0;
}

/* the two segment queues, `cap` records each: the 9 fast-mode fields, or
 * all 23 when this run or an earlier one is full-state (every field in use
 * grows with the capacity, so qcap holds for all of them) */
static int ensure_queues(ramp_plan *p, uint64_t cap, bool full)
{
    const bool want_full = full || p->qfull;
    if (cap <= p->qcap && want_full == p->qfull) return 0;
    uint64_t c = std::max(cap, p->qcap);
    int nf = want_full ? 23 : 9;
    for (int i = 0; i < 2; i++) {
        for (int f = 0; f < nf; f++) {
            int rc = dev_alloc(p->q[i].f[f], c * sizeof(float));
            if (rc) return rc;
        }
        int rc = dev_alloc(p->q[i].key, c * sizeof(uint32_t));
        if (!rc) rc = dev_alloc(p->q[i].status, c * sizeof(uint32_t));
        if (rc) return rc;
    }
    p->qcap = c;
    p->qfull = want_full;
    return 0;
}

/* every queue field a run of this mode touches holds `cap` records */
static int queues_hold(const ramp_plan *p, uint64_t cap, bool full)
{
    const int nf = full ? 23 : 9;
    for (int i = 0; i < 2; i++) {
        for (int f = 0; f < nf; f++)
            if (p->q[i].f[f].bytes < cap * sizeof(float))
                return fail(RAMP_E_STATE, "queue %d field %d holds %zu B, the run needs %llu records", i, f,
                            p->q[i].f[f].bytes, (unsigned long long)cap);
        if (p->q[i].key.bytes < cap * sizeof(uint32_t) || p->q[i].status.bytes < cap * sizeof(uint32_t))
            return fail(RAMP_E_STATE, "queue %d keys hold fewer than %llu records", i, (unsigned long long)cap);
    }
    return 0;
}

static DQueue queue_view(ramp_plan *p, int i)
{
    DQueue q;
    float **f[23] = {&q.x, &q.y, &q.z, &q.vx, &q.vy, &q.vz, &q.w, &q.t, &q.ctr,
                     &q.px, &q.py, &q.pz, &q.sc, &q.sd, &q.se,
                     &q.is[0], &q.is[1], &q.is[2], &q.is[3], &q.is[4], &q.is[5], &q.is[6], &q.is[7]};
    for (int k = 0; k < 23; k++) *f[k] = (float *)p->q[i].f[k].p;
    q.key = (uint32_t *)p->q[i].key.p;
    q.status = (uint32_t *)p->q[i].status.p;
    return q;
}


static seg_fn pick_kernel(bool full, bool first, int v)
{
    return kern_table(v).seg[full ? 1 : 0][first ? 1 : 0];
}

/* LDS per wave of a plan's monochromator deferral queue in fast mode */
static uint32_t defer_lds_per_wave(const ramp_plan *p, bool full)
{
    return (full || p->defer_k < 0) ? 0u : (uint32_t)DQ_BYTES_PER_WAVE;
}

static void launch_segment(const SegArgs &a, bool full, bool first, int v, unsigned grid,
                           unsigned wg, hipStream_t s)
{
    hipLaunchKernelGGL(pick_kernel(full, first, v), dim3(grid), dim3(wg),
                       (size_t)(wg / 64) * a.defer_lds + a.acc_bytes, s, a);
}

/* resident blocks per CU of a kernel instance (for a persistent grid).  In
 * a plan with a GMesh the source segment, which writes the queue and the
 * listed walk records (about 90 B per history), runs faster at three waves
 * per SIMD than at the four its registers allow; its grid is held there. */
#ifndef RAMP_MESH_FIRST_WAVES
#define RAMP_MESH_FIRST_WAVES 3
#endif
static unsigned blocks_per_cu(bool full, bool first, int v, unsigned wg, uint32_t defer_lds, uint32_t acc_bytes,
                              bool mesh)
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, pick_kernel(full, first, v), (int)wg,
                                                     (size_t)(wg / 64) * defer_lds + acc_bytes) != hipSuccess ||
        nb < 1)
        nb = 1;
    if (!full && first && mesh) {
        const int cap = (int)(RAMP_MESH_FIRST_WAVES * 4 * 64 / wg);  /* 4 SIMDs per CU */
        if (cap >= 1 && nb > cap) nb = cap;
    }
    return (unsigned)nb;
}

/* choose segment boundaries from a pilot pass: a boundary follows every step
 * after which the live fraction fell below half of the segment's start */
/* fewest wave-steps (64 histories through one step) behind a compaction
 * boundary for it to pay: on MI355X a 1e6-history run of a three-step
 * instrument is 26 % faster in one segment, a 1e6-history run of the
 * 28-step LET-like one 10 % faster with its boundaries */
static const double MIN_COMPACT_WAVE_STEPS = 2.0e5;

static int choose_bounds(ramp_plan *p, const SegArgs &base, uint64_t chunk, unsigned wg, int max_seg)
{
    int ns = (int)p->hsteps.size();
    p->bounds.clear();
    p->bounds.push_back(0);
    if (ns > 1 && max_seg != 1) {
        hipStream_t s = p->ctx->stream;
        uint64_t pn = std::min<uint64_t>(chunk, 65536);
        SegArgs a = base;
        a.s0 = 0;
        a.s1 = ns;
        a.pilot = 1;
        a.walk_role = 0;
        a.last = 1;
        a.carry = 0;
        a.keep = 0;
        a.n = pn;
        a.num_sim = (int)pn;
        launch_fill_u64(s, p->cnt.p, CNT_N);
        unsigned grid = (unsigned)std::max<uint64_t>(1, (pn + wg - 1) / wg);
        launch_segment(a, p->need_full, true, p->variant, grid, wg, s);
        HIPCHK(hipGetLastError());
        std::vector<unsigned long long> cnt(CNT_N);
        HIPCHK(hipMemcpyAsync(cnt.data(), p->cnt.p, CNT_N * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        double start = (double)pn;
        int lim = max_seg > 0 ? max_seg : 16;
        /* SMCPLIn revives dead slots, which only the first segment holds */
        for (int k = std::max(0, p->revive_k); k < ns - 1 && (int)p->bounds.size() < lim; k++) {
            double alive = (double)cnt[CNT_NSTEP0 + k];
            /* a boundary costs a launch, a scan and the queue traffic; it
             * pays when enough wave-steps remain behind it (waves whose
             * lanes are all dead leave the range early anyway) */
            const bool worth = (double)chunk / 64.0 * (double)(ns - k - 1) >= MIN_COMPACT_WAVE_STEPS;
            if (alive < 0.5 * start && worth) {
                p->bounds.push_back(k + 1);
                start = std::max(alive, 1.0);
            }
        }
    }
    p->bounds.push_back(ns);
    /* a segment starts at the first mesh step: its walks run before it in
     * k_mesh_walk over the queue of the segment before */
    bool first_walk = max_seg != 1 && p->walk_k == 0;
    if (max_seg != 1 && p->walk_k > 0) p->bounds.push_back(p->walk_k);
    /* a plugin step is a segment of its own (run by its plugin kernel) */
    bool first_plugin = false;
    for (int k = 0; k < ns; k++)
        if (p->plugin_step[k] >= 0) {
            p->bounds.push_back(k);
            p->bounds.push_back(k + 1);
            first_plugin = first_plugin || k == 0;
        }
    std::sort(p->bounds.begin(), p->bounds.end());
    p->bounds.erase(std::unique(p->bounds.begin(), p->bounds.end()), p->bounds.end());
    if (first_plugin || first_walk) p->bounds.insert(p->bounds.begin(), 0);  /* the source alone first */
    if (p->bounds.size() == 1) p->bounds.push_back(ns);  /* no steps: the source segment alone */
    const size_t nb = p->bounds.size();
    if (nb >= 2 && p->bounds[nb - 1] == p->bounds[nb - 2] + 1 && p->plugin_step[p->bounds[nb - 2]] >= 0)
        p->bounds.push_back(ns);  /* a last, empty segment finalizes the neutrons */
    p->bounds_set = true;
    p->bounds_chunk = chunk;
    p->bounds_max_seg = max_seg;
    return 0;
}

#include "mcpl_run.h"

/* ---- run-time compiled scatter plugins (hipRTC) ----------------------- */
#include "plugin_headers.inc"

struct PluginArgs {  /* the kernel argument of ramp_plugin_kernel (plugin_kernel.h) */
    DQueue q;
    const uint64_t *prefix;
    const uint64_t *meta;
    const DComp *comp;
    float *trace;
    uint64_t trace_cap, trace_base;
};

extern "C" int ramp_plugin_compile(ramp_ctx *ctx, const char *source, const char *entry, ramp_plugin **out)
{
    if (!ctx || !source || !entry || !out) return fail(RAMP_E_INVALID, "null argument");
    if (!*entry) return fail(RAMP_E_INVALID, "empty plugin entry name");
    for (const char *c = entry; *c; c++)
        if (!(std::isalnum((unsigned char)*c) || *c == '_'))
            return fail(RAMP_E_INVALID, "plugin entry '%s' is not an identifier", entry);
    std::string prog = std::string("#include \"plugin_kernel.h\"\n#line 1 \"plugin\"\n") + source +
                       "\nRAMP_PLUGIN_KERNEL(" + entry + ")\n";
    hiprtcProgram pr;
    if (hiprtcCreateProgram(&pr, prog.c_str(), "ramp_plugin.hip", kPluginHeaderCount, kPluginHeaderTexts,
                            kPluginHeaderNames) != HIPRTC_SUCCESS)
        return fail(RAMP_E_DEVICE, "hiprtcCreateProgram failed");
    std::string arch = "--offload-arch=" + ctx->arch;
    const char *opts[] = {arch.c_str(), "-O3", "-std=c++17", "-ffp-contract=off",
                          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    hiprtcResult r = hiprtcCompileProgram(pr, 8, opts);
    if (r != HIPRTC_SUCCESS) {
        size_t n = 0;
        hiprtcGetProgramLogSize(pr, &n);
        std::string log(n, '\0');
        if (n) hiprtcGetProgramLog(pr, &log[0]);
        hiprtcDestroyProgram(&pr);
        g_err = "plugin compilation failed:\n" + log;
        return RAMP_E_INVALID;
    }
    size_t n = 0;
    hiprtcGetCodeSize(pr, &n);
    std::vector<char> code(n);
    hiprtcGetCode(pr, code.data());
    hiprtcDestroyProgram(&pr);
    HIPCHK(hipSetDevice(ctx->device));
    ramp_plugin *pl = new ramp_plugin();
    pl->ctx = ctx;
    hipError_t e = hipModuleLoadData(&pl->mod, code.data());
    static const char *const names[2][2] = {{"ramp_plugin_kernel", "ramp_plugin_kernel_trace"},
                                            {"ramp_plugin_kernel_mesh", "ramp_plugin_kernel_mesh_trace"}};
    for (int m = 0; m < 2; m++)
        for (int t = 0; t < 2; t++) {
            int nb = 0;
            if (e == hipSuccess) e = hipModuleGetFunction(&pl->fn[m][t], pl->mod, names[m][t]);
            if (e == hipSuccess) e = hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&nb, pl->fn[m][t], 256, 0);
            pl->grid[m][t] = (unsigned)ctx->n_cu * (unsigned)std::max(nb, 1);
        }
    if (e != hipSuccess) {
        if (pl->mod) (void)hipModuleUnload(pl->mod);
        delete pl;
        return fail(RAMP_E_DEVICE, "loading the plugin: %s", hipGetErrorString(e));
    }
    *out = pl;
    return 0;
}

extern "C" int ramp_plugin_destroy(ramp_plugin *pl)
{
    if (!pl) return 0;
    if (pl->mod) {
        (void)hipSetDevice(pl->ctx->device);
        (void)hipStreamSynchronize(pl->ctx->stream);
        (void)hipModuleUnload(pl->mod);
    }
    delete pl;
    return 0;
}

extern "C" int ramp_plan_bind_plugin(ramp_plan *p, uint32_t comp, ramp_plugin *pl)
{
    if (!p || !pl) return fail(RAMP_E_INVALID, "null argument");
    if (comp >= p->hcomps.size()) return fail(RAMP_E_INVALID, "component %u of %zu", comp, p->hcomps.size());
    if (p->hcomps[comp].s.type != RAMP_S_PLUGIN)
        return fail(RAMP_E_INVALID, "component %u is not a plugin component", comp);
    if (pl->ctx != p->ctx) return fail(RAMP_E_INVALID, "plugin of another context");
    p->plugin_of[comp] = pl;
    return 0;
}

/* the device buffers a run works in besides the plan's tables and
 * histograms: the plan's own (ramp_run) or the device's scratch array
 * (ramp_run_noalloc) */
struct RunBufs {
    DQueue q[2];
    uint32_t *rcounts = nullptr;
    uint64_t *prefix = nullptr;
    float *slot_ctr = nullptr;
    uint32_t *slot_iidx = nullptr;
};

static int run_plan(ramp_plan *p, uint64_t n_hist, uint64_t chunk, const uint32_t *seeds, uint32_t n_seeds,
                    const ramp_run_opts *opts, bool noalloc)
{
    if (!p) return fail(RAMP_E_INVALID, "null plan");
    ramp_run_opts o;
    std::memset(&o, 0, sizeof o);
    if (opts) o = *opts;
    if (o.hist_mode != RAMP_HIST_ATOMIC && o.hist_mode != RAMP_HIST_DETERMINISTIC)
        return fail(RAMP_E_INVALID, "unknown histogram mode %u", o.hist_mode);
    if (o.flags & ~(uint32_t)(RAMP_RUN_MESH_BRUTE_FORCE | RAMP_RUN_TRACE))
        return fail(RAMP_E_INVALID, "unknown run flags 0x%x", o.flags);
    if (n_hist == 0) {
        /* no histories: the histograms of the mode exist afterwards (what
         * ramp_run_noalloc needs), nothing else happens */
        if (noalloc) return 0;
        HIPCHK(hipSetDevice(p->ctx->device));
        int rc = ensure_hist(p, (int)o.hist_mode);
        if (!rc) HIPCHK(hipStreamSynchronize(p->ctx->stream));
        return rc;
    }
    const bool shard = o.shard_total != 0;
    if (!shard && o.key_base) return fail(RAMP_E_INVALID, "key_base without shard_total");
    if (shard) {
        if (o.shard_total > 0x7fffffffull || o.key_base + n_hist > o.shard_total)
            return fail(RAMP_E_RANGE, "shard [%llu, %llu) of a chunk of %llu",
                        (unsigned long long)o.key_base, (unsigned long long)(o.key_base + n_hist),
                        (unsigned long long)o.shard_total);
        if (chunk != 0 && chunk < n_hist) return fail(RAMP_E_INVALID, "a shard runs as one chunk");
        if (p->need_full || o.keep_state || (o.flags & RAMP_RUN_TRACE))
            return fail(RAMP_E_INVALID, "shards run fast-mode plans without keep_state or trace");
    }
    const bool trace = (o.flags & RAMP_RUN_TRACE) != 0 && p->n_trace > 0;
    if (trace && (double)p->n_trace * 16.0 * (double)n_hist > TRACE_MAX_BYTES)
        return fail(RAMP_E_RANGE, "a trace of %u components x %llu histories needs %.3g GB (limit %.3g GB)",
                    p->n_trace, (unsigned long long)n_hist, (double)p->n_trace * 16.0 * (double)n_hist * 1e-9,
                    TRACE_MAX_BYTES * 1e-9);
    if (chunk == 0 || chunk > n_hist) chunk = n_hist;
    if (chunk > 0x7fffffffull) return fail(RAMP_E_RANGE, "chunk %llu exceeds 2^31-1", (unsigned long long)chunk);
    uint64_t n_chunks = (n_hist + chunk - 1) / chunk;
    if (!seeds || n_seeds < n_chunks)
        return fail(RAMP_E_INVALID, "need %llu seeds, got %u", (unsigned long long)n_chunks, n_seeds);
    for (uint64_t c = 0; c < n_chunks; c++)
        if (seeds[c] > 0x7fffffffu) return fail(RAMP_E_INVALID, "seed %u out of range", seeds[c]);
    for (size_t k = 0; k < p->hcomps.size(); k++)
        if (p->hcomps[k].s.type == RAMP_S_PLUGIN && !p->plugin_of[k])
            return fail(RAMP_E_STATE, "component %zu: no plugin bound (ramp_plan_bind_plugin)", k);
    ramp_ctx *ctx = p->ctx;
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    unsigned wg = o.wg ? o.wg : 256;
    if (wg != 64 && wg != 128 && wg != 256) return fail(RAMP_E_INVALID, "workgroup size %u", wg);
    const bool keep = o.keep_state != 0;
    /* the MCPL steps need the polarisation and the whole intersection,
     * SRescal's filter the detected flag */
    const bool full = keep || p->need_full || trace;
    if (noalloc) {
        if (full || (o.flags & RAMP_RUN_MESH_BRUTE_FORCE))
            return fail(RAMP_E_INVALID, "ramp_run_noalloc runs fast-mode plans (no MCPL, SRescal, SAnalyser or "
                                        "plugin components) without keep_state, trace or brute force");
        if (chunk > RAMP_NOALLOC_MAX_CHUNK)
            return fail(RAMP_E_RANGE, "ramp_run_noalloc: chunk %llu exceeds RAMP_NOALLOC_MAX_CHUNK (%llu)",
                        (unsigned long long)chunk, (unsigned long long)RAMP_NOALLOC_MAX_CHUNK);
        if (!ctx->scratch) return fail(RAMP_E_STATE, "no scratch array on device %d", ctx->device);
        for (size_t k = 0; k < p->hdets.size(); k++)
            if ((o.hist_mode == RAMP_HIST_ATOMIC && !p->hw[k].p) ||
                (o.hist_mode == RAMP_HIST_DETERMINISTIC && !p->hq[k].p))
                return fail(RAMP_E_STATE, "ramp_run_noalloc: the plan's histograms of this mode do not exist; "
                                          "call ramp_run with n_hist = 0 first");
        if (o.hist_mode == RAMP_HIST_DETERMINISTIC) {
            DetScale t = p->ds;
            const double wb = o.weight_bound > 0 ? o.weight_bound : 1.0 / (double)chunk;
            if (det_scale_run(t, wb, (double)(shard ? o.shard_total : n_hist), (double)n_hist))
                return fail(RAMP_E_RANGE, "ramp_run_noalloc: the deterministic histograms' int64 budget is spent; "
                                          "read them and reset, or run with ramp_run");
        }
    }
    /* a shard takes the fixed-point scale of its whole chunk, so that the
     * shards' sums add up exactly to the unsharded run's */
    int rc = prepare_dets(p, (int)o.hist_mode, o.weight_bound, shard ? o.shard_total : n_hist,
                          shard ? o.shard_total : chunk, n_hist);
    if (rc) return rc;
    RunBufs bufs;
    std::memset(bufs.q, 0, sizeof bufs.q);
    bool carry = n_chunks > 1;
    if (carry) {
        if (noalloc) {
            bufs.slot_ctr = (float *)((char *)ctx->scratch + NA_OFF_CTR);
            bufs.slot_iidx = (uint32_t *)((char *)ctx->scratch + NA_OFF_IIDX);
        } else {
            rc = dev_alloc(p->slot_ctr, chunk * sizeof(float));
            if (!rc) rc = dev_alloc(p->slot_iidx, chunk * sizeof(uint32_t));
            if (rc) return rc;
            bufs.slot_ctr = (float *)p->slot_ctr.p;
            bufs.slot_iidx = (uint32_t *)p->slot_iidx.p;
        }
        launch_fill_f32(s, bufs.slot_ctr, chunk);
        launch_fill_u32(s, bufs.slot_iidx, chunk);
    }
    if (keep) {
        rc = dev_alloc(p->st16, chunk * 16 * sizeof(float));
        if (!rc) rc = dev_alloc(p->st8, chunk * 8 * sizeof(float));
        if (!rc) rc = dev_alloc(p->stii, chunk * sizeof(uint32_t));
        if (rc) return rc;
    }
    if (p->n_mcpl_out && p->dump_cap < chunk) {
        const uint64_t nslot = (uint64_t)p->n_mcpl_out * chunk;
        rc = dev_alloc(p->dump, nslot * MCR_N * sizeof(float));
        if (!rc) rc = dev_alloc(p->dump_tag, nslot * sizeof(uint32_t));
        if (rc) return rc;
        launch_fill_u32(s, (uint32_t *)p->dump_tag.p, nslot);
        p->dump_cap = chunk;
        p->dump_chunk_tag = 0;
    }
    if (p->n_rescal) {
        rc = rescal_prepare(p, chunk, n_hist, s);
        if (rc) return rc;
    }
    p->trace_n = 0;
    if (trace) {
        rc = dev_alloc(p->trace, (uint64_t)p->n_trace * 4 * n_hist * sizeof(float));
        if (rc) return rc;
        launch_fill_f32(s, p->trace.p, (uint64_t)p->n_trace * 4 * n_hist);
    }
    SegArgs a;
    std::memset(&a, 0, sizeof a);
    a.comps = (const DComp *)p->comps.p;
    if (o.flags & RAMP_RUN_MESH_BRUTE_FORCE) {
        if (!p->comps_brute.p) {
            std::vector<DComp> bc = p->hcomps;
            for (auto &c : bc)
                if (c.g.type == RAMP_G_MESH) { c.g.aux = nullptr; c.g.aux2 = nullptr; }
            rc = dev_alloc(p->comps_brute, bc.size() * sizeof(DComp));
            if (rc) return rc;
            HIPCHK(hipMemcpy(p->comps_brute.p, bc.data(), bc.size() * sizeof(DComp), hipMemcpyHostToDevice));
        }
        a.comps = (const DComp *)p->comps_brute.p;
    }
    a.steps = (const DStep *)p->steps.p;
    a.src = (const DPart *)p->src.p;
    a.dets = (const DDet *)p->dets.p;
    a.cnt = (unsigned long long *)p->cnt.p;
    a.hist_mode = (int)o.hist_mode;
    a.carry = carry ? 1 : 0;
    a.keep = keep ? 1 : 0;
    a.defer_k = full ? -1 : p->defer_k;
    a.defer_lds = defer_lds_per_wave(p, full);
    a.revive_k = p->revive_k;
    a.dump = (float *)p->dump.p;
    a.dump_tag = (uint32_t *)p->dump_tag.p;
    a.dump_cap = p->dump_cap;
    a.ev_tag = (uint32_t *)p->ev_tag.p;
    a.ev_cap = p->ev_cap;
    a.n_rescal = p->n_rescal;
    a.acc_bins = p->acc_bins;
    a.acc_bytes = p->acc_bytes;
    a.n_dets = (int)p->hdets.size();
    a.slot_ctr = bufs.slot_ctr;
    a.slot_iidx = bufs.slot_iidx;
    a.st16 = (float *)p->st16.p;
    a.st8 = (float *)p->st8.p;
    a.stii = (uint32_t *)p->stii.p;
    a.trace = trace ? (float *)p->trace.p : nullptr;
    a.trace_cap = n_hist;
    /* the boundaries depend on the chunk (MIN_COMPACT_WAVE_STEPS) and the
     * segment limit: choose them again when the chunk moves by more than 4x,
     * when the limit is lowered below them, or raised or cleared */
    const bool rechoose = !p->bounds_set ||
                          (o.max_segments && (int)p->bounds.size() - 1 > (int)o.max_segments) ||
                          chunk > 4 * p->bounds_chunk || 4 * chunk < p->bounds_chunk ||
                          (p->bounds_max_seg && (!o.max_segments || (int)o.max_segments > p->bounds_max_seg));
    if (rechoose) {
        rc = choose_bounds(p, a, chunk, wg, (int)o.max_segments);
        if (rc) return rc;
    }
    int nseg = (int)p->bounds.size() - 1;
    /* persistent grids: as many blocks as are co-resident */
    const bool mesh = (p->feat & FEAT_MESH) != 0u;
    unsigned grid_first = (unsigned)ctx->n_cu * blocks_per_cu(full, true, p->variant, wg, a.defer_lds, p->acc_bytes, mesh);
    unsigned grid_rest = (unsigned)ctx->n_cu * blocks_per_cu(full, false, p->variant, wg, a.defer_lds, p->acc_bytes, mesh);
    uint64_t max_waves = (uint64_t)std::max(grid_first, grid_rest) * (wg / 64);
    if (nseg > 1 && noalloc) {
        if (max_waves > NA_MAX_WAVES || chunk + max_waves > NA_CAP)
            return fail(RAMP_E_RANGE, "ramp_run_noalloc: %llu waves and %llu queue records exceed the scratch "
                                      "array (%llu, %llu)", (unsigned long long)max_waves,
                        (unsigned long long)(chunk + max_waves), (unsigned long long)NA_MAX_WAVES,
                        (unsigned long long)NA_CAP);
        char *base = (char *)ctx->scratch;
        for (int i = 0; i < 2; i++) {
            DQueue &q = bufs.q[i];
            std::memset(&q, 0, sizeof q);
            float **f[9] = {&q.x, &q.y, &q.z, &q.vx, &q.vy, &q.vz, &q.w, &q.t, &q.ctr};
            for (int k = 0; k < 9; k++) *f[k] = (float *)(base + (size_t)(i * 11 + k) * NA_QUEUE_FIELD);
            q.key = (uint32_t *)(base + (size_t)(i * 11 + 9) * NA_QUEUE_FIELD);
            q.status = (uint32_t *)(base + (size_t)(i * 11 + 10) * NA_QUEUE_FIELD);
        }
        bufs.rcounts = (uint32_t *)(base + NA_OFF_COUNTS);
        bufs.prefix = (uint64_t *)(base + NA_OFF_PREFIX);
    } else if (nseg > 1) {
        rc = ensure_queues(p, chunk + max_waves, full);
        if (!rc) rc = queues_hold(p, chunk + max_waves, full);
        if (!rc) rc = dev_alloc(p->rcounts, max_waves * sizeof(uint32_t));
        if (!rc) rc = dev_alloc(p->prefix, (max_waves + 1) * sizeof(uint64_t));
        if (rc) return rc;
        bufs.q[0] = queue_view(p, 0);
        bufs.q[1] = queue_view(p, 1);
        bufs.rcounts = (uint32_t *)p->rcounts.p;
        bufs.prefix = (uint64_t *)p->prefix.p;
    }
    /* mesh walks in their own launch (fast mode): the segment before the
     * first mesh step lists the rays, k_mesh_walk walks them, the segment
     * starting at the step takes their hits.  Queue slots index the hits
     * (32-bit, in the neutron's aux), and the listing segment must be an
     * ordinary one (not a plugin's). */
    int walk_seg = -1;
    unsigned walk_grid = 0;
    const uint64_t walk_cap = chunk + max_waves;   /* the queues' slots */
    const bool walk_ok = !full && !noalloc && p->walk_k >= 0 && p->plugin_step[p->walk_k] < 0 &&
                         walk_cap < (1ull << 31);
    if (walk_ok && p->walk_k > 0 && nseg > 1) {
        for (int sg = 1; sg < nseg; sg++)
            if (p->bounds[sg] == p->walk_k) walk_seg = sg;
    } else if (walk_ok && p->walk_k == 0 && nseg > 1 && p->bounds[1] == 0) {
        walk_seg = 1;
    }
    if (walk_seg > 1) {
        const int ps = walk_seg - 1;
        if (p->bounds[ps + 1] == p->bounds[ps] + 1 && p->plugin_step[p->bounds[ps]] >= 0) walk_seg = -1;
    }
    size_t walk_tbytes = 0;
    if (walk_seg > 0) {
        rc = dev_alloc(p->walk_res, walk_cap * sizeof(WalkHit));
        if (!rc) rc = dev_alloc(p->walk_list, walk_cap * sizeof(WalkEntry));
        if (!rc) rc = dev_alloc(p->walk_keys, walk_cap * sizeof(uint32_t));
        if (!rc) rc = dev_alloc(p->walk_counts, max_waves * sizeof(uint32_t));
        if (!rc) rc = dev_alloc(p->walk_aux, (max_waves + 4) * sizeof(uint64_t));
        for (int i = 0; i < 4 && !rc; i++) rc = dev_alloc(p->walk_sort[i], walk_cap * sizeof(uint32_t));
        if (rc) return rc;
        HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, walk_tbytes, (const uint32_t *)p->walk_sort[0].p,
                                                  (uint32_t *)p->walk_sort[1].p, (const uint32_t *)p->walk_sort[2].p,
                                                  (uint32_t *)p->walk_sort[3].p, (int)walk_cap, 0, WALK_KEY_BITS, s));
        rc = dev_alloc(p->walk_temp, walk_tbytes + 16);
        if (!rc) rc = dev_alloc(p->walk_sorted, walk_cap * sizeof(WalkEntry));
        if (rc) return rc;
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_mesh_walk, 256, 0) != hipSuccess || nb < 1) nb = 1;
        walk_grid = (unsigned)ctx->n_cu * (unsigned)nb;
    }
    a.walk_role = 0;
    a.walk_ci = walk_seg > 0 ? (uint32_t)p->hsteps[p->walk_k].first : 0u;
    a.walk_res = walk_seg > 0 ? (WalkHit *)p->walk_res.p : nullptr;
    a.walk_list = walk_seg > 0 ? (WalkEntry *)p->walk_list.p : nullptr;
    a.walk_keys = walk_seg > 0 ? (uint32_t *)p->walk_keys.p : nullptr;
    a.walk_counts = walk_seg > 0 ? (uint32_t *)p->walk_counts.p : nullptr;
    uint64_t *meta = (uint64_t *)p->count.p; /* [0..1] producer, [2..3] consumer */
    a.out_counts = bufs.rcounts;
    a.out_meta = meta;
    a.in_meta = meta + 2;
    a.in_prefix = bufs.prefix;
    launch_fill_u64(s, p->cnt.p, CNT_N);
    HIPCHK(hipEventRecord(ctx->ev[0], s));
    uint64_t segs = 0;
    bool plugin_timed = false, walk_timed = false;
    uint64_t walk_rays = 0;
    uint64_t plugin_items = 0, plugin_nreg = 0;
    for (uint64_t ck = 0; ck < n_chunks; ck++) {
        uint64_t torun = std::min(chunk, n_hist - ck * chunk);
        a.num_sim = shard ? (int)o.shard_total : (int)torun;
        a.key_base = shard ? o.key_base : 0;
        a.seed = (int)seeds[ck];
        a.n = torun;
        a.trace_base = ck * chunk;
        a.chunk_tag = ++p->dump_chunk_tag ? p->dump_chunk_tag : ++p->dump_chunk_tag;
        if (p->n_rescal) launch_fill_f32(s, p->ev.p, (uint64_t)p->n_rescal * RS_NEV * p->ev_cap);
        int cur = 0; /* the queue holding the live neutrons */
        for (int sg = 0; sg < nseg; sg++) {
            a.s0 = p->bounds[sg];
            a.s1 = p->bounds[sg + 1];
            if (sg > 0 && a.s1 == a.s0 + 1 && p->plugin_step[a.s0] >= 0) {
                /* a plugin step rewrites the queue in place */
                ramp_plugin *pl = p->plugin_of[p->plugin_step[a.s0]];
                PluginArgs pa;
                pa.q = bufs.q[cur];
                pa.prefix = bufs.prefix;
                pa.meta = meta + 2;
                pa.comp = (const DComp *)p->comps.p + p->plugin_step[a.s0];
                pa.trace = a.trace;
                pa.trace_cap = a.trace_cap;
                pa.trace_base = a.trace_base;
                void *args[] = {&pa};
                const bool timed = ck == 0 && !plugin_timed;
                if (timed) HIPCHK(hipEventRecord(ctx->ev[6], s));
                const int mesh = p->hcomps[p->plugin_step[a.s0]].g.type == RAMP_G_MESH ? 1 : 0;
                const int tr = a.trace ? 1 : 0;
                HIPCHK(hipModuleLaunchKernel(pl->fn[mesh][tr], pl->grid[mesh][tr], 1, 1, 256, 1, 1, 0, s, args,
                                             nullptr));
                if (timed) {
                    HIPCHK(hipEventRecord(ctx->ev[7], s));
                    HIPCHK(hipMemcpyAsync(&plugin_items, bufs.prefix + plugin_nreg,
                                          sizeof plugin_items, hipMemcpyDeviceToHost, s));
                    plugin_timed = true;
                }
                segs++;
                continue;
            }
            a.last = (sg == nseg - 1) ? 1 : 0;
            a.pilot = 0;
            a.walk_role = walk_seg < 0 ? 0u : sg == walk_seg ? WALK_USE : sg + 1 == walk_seg ? WALK_EMIT : 0u;
            if (sg >= 1) a.in = bufs.q[cur];
            a.out = bufs.q[sg == 0 ? 0 : cur ^ 1];
            unsigned grid;
            if (sg == 0) {
                /* the fewest waves that still take the histories in as few
                 * 64-history rounds each as the full persistent grid would:
                 * every wave then runs whole rounds (a small run does not
                 * spread a last partial round over every resident wave) */
                const uint64_t wpb = wg / 64, wmax = (uint64_t)grid_first * wpb;
                const uint64_t rounds = std::max<uint64_t>(1, (torun + 64 * wmax - 1) / (64 * wmax));
                const uint64_t waves = (torun + 64 * rounds - 1) / (64 * rounds);
                grid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((waves + wpb - 1) / wpb, grid_first));
            } else {
                grid = grid_rest;
            }
            if (sg == 0 && ck == 0) HIPCHK(hipEventRecord(ctx->ev[2], s));
            launch_segment(a, full, sg == 0, p->variant, grid, wg, s);
            if (sg == 0 && ck == 0) HIPCHK(hipEventRecord(ctx->ev[3], s));
            HIPCHK(hipGetLastError());
            segs++;
            cur = sg == 0 ? 0 : cur ^ 1;
            if (sg + 1 < nseg) {
                uint64_t nreg = (uint64_t)grid * (wg / 64);
                plugin_nreg = nreg;
                hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, (const uint32_t *)bufs.rcounts,
                                   nreg, (const uint64_t *)meta, bufs.prefix, meta + 2,
                                   (unsigned long long *)p->cnt.p, sg == 0 ? 1 : 0);
                HIPCHK(hipGetLastError());
                if (sg + 1 == walk_seg) {
                    uint64_t *wp = (uint64_t *)p->walk_aux.p;   /* prefix, meta, counter */
                    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, (const uint32_t *)p->walk_counts.p,
                                       nreg, (const uint64_t *)meta, wp, wp + max_waves + 1,
                                       (unsigned long long *)nullptr, 0);
                    HIPCHK(hipGetLastError());
                    if (ck == 0) HIPCHK(hipEventRecord(ctx->ev[10], s));
                    hipLaunchKernelGGL(k_walk_order, dim3((unsigned)nreg), dim3(256), 0, s,
                                       (const uint32_t *)p->walk_keys.p, (const uint64_t *)wp,
                                       (const uint64_t *)(wp + max_waves + 1), (uint32_t *)p->walk_sort[0].p,
                                       (uint32_t *)p->walk_sort[2].p);
                    HIPCHK(hipGetLastError());
                    launch_fill_u64(s, wp + max_waves + 3, 1);
                    /* the sort needs the number of listed rays on the host */
                    uint64_t nwalk = 0;
                    HIPCHK(hipMemcpyAsync(&nwalk, wp + nreg, sizeof nwalk, hipMemcpyDeviceToHost, s));
                    HIPCHK(hipStreamSynchronize(s));
                    if (nwalk > walk_cap) return fail(RAMP_E_DEVICE, "walk list of %llu rays", (unsigned long long)nwalk);
                    if (nwalk > 0) {
                        size_t tb = walk_tbytes;
                        HIPCHK(hipcub::DeviceRadixSort::SortPairs(p->walk_temp.p, tb, (const uint32_t *)p->walk_sort[0].p,
                                                                  (uint32_t *)p->walk_sort[1].p,
                                                                  (const uint32_t *)p->walk_sort[2].p,
                                                                  (uint32_t *)p->walk_sort[3].p, (int)nwalk, 0,
                                                                  WALK_KEY_BITS, s));
                        hipLaunchKernelGGL(k_walk_gather, dim3(4096), dim3(256), 0, s,
                                           (const WalkEntry *)p->walk_list.p, (const uint32_t *)p->walk_sort[3].p,
                                           nwalk, (WalkEntry *)p->walk_sorted.p);
                        HIPCHK(hipGetLastError());
                        if (ck == 0) HIPCHK(hipEventRecord(ctx->ev[11], s));
                        WalkArgs wa;
                        wa.list = (const WalkEntry *)p->walk_sorted.p;
                        wa.n = nwalk;
                        wa.comp = a.comps + a.walk_ci;   /* the brute-force copy when flagged */
                        wa.res = (WalkHit *)p->walk_res.p;
                        wa.next = (unsigned long long *)(wp + max_waves + 3);
                        if (ck == 0) HIPCHK(hipEventRecord(ctx->ev[8], s));
                        hipLaunchKernelGGL(k_mesh_walk, dim3(walk_grid), dim3(256), 0, s, wa);
                        if (ck == 0) HIPCHK(hipEventRecord(ctx->ev[9], s));
                        HIPCHK(hipGetLastError());
                    }
                    if (ck == 0) {
                        walk_timed = nwalk > 0;
                        walk_rays = nwalk;
                    }
                }
            }
        }
        if (p->n_rescal) {
            if (ck == 0) HIPCHK(hipEventRecord(ctx->ev[4], s));
            rc = rescal_append(p, torun, a.chunk_tag, s);
            if (rc) return rc;
            if (ck == 0) HIPCHK(hipEventRecord(ctx->ev[5], s));
        }
    }
    HIPCHK(hipEventRecord(ctx->ev[1], s));
    /* the counters come back on the stream into pinned memory: one wait */
    HIPCHK(hipMemcpyAsync(p->cnt_host, p->cnt.p, CNT_N * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    const unsigned long long *cnt = p->cnt_host;
    float ms = 0, ms0 = 0;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
    HIPCHK(hipEventElapsedTime(&ms0, ctx->ev[2], ctx->ev[3]));
    p->stats.n_hist = n_hist;
    p->stats.n_chunks = n_chunks;
    p->stats.n_events = cnt[CNT_EVENTS];
    p->stats.n_oob = cnt[CNT_OOB];
    p->stats.n_det_overflow = cnt[CNT_OVERFLOW];
    p->stats.n_segments = segs;
    p->stats.state_out = cnt[CNT_QOUT];
    p->stats.state_out_first = cnt[CNT_QOUT0];
    p->stats.state_in = cnt[CNT_QIN];
    p->stats.ms_total = ms;
    p->stats.ms_source_segment = ms0;
    float ms_ev = 0, ms_pl = 0;
    if (p->n_rescal) HIPCHK(hipEventElapsedTime(&ms_ev, ctx->ev[4], ctx->ev[5]));
    if (plugin_timed) HIPCHK(hipEventElapsedTime(&ms_pl, ctx->ev[6], ctx->ev[7]));
    p->stats.ms_rescal_append = ms_ev;
    p->stats.ms_plugin = ms_pl;
    p->stats.plugin_items = plugin_items;
    float ms_wk = 0, ms_ws = 0;
    if (walk_timed) {
        HIPCHK(hipEventElapsedTime(&ms_wk, ctx->ev[8], ctx->ev[9]));
        HIPCHK(hipEventElapsedTime(&ms_ws, ctx->ev[10], ctx->ev[11]));
    }
    p->stats.ms_mesh_walk = ms_wk;
    p->stats.ms_walk_sort = ms_ws;
    p->stats.walk_rays = walk_rays;
    if (p->n_rescal)
        HIPCHK(hipMemcpy(p->ev_count.data(), p->ev_list_n.p, p->n_rescal * sizeof(uint64_t),
                         hipMemcpyDeviceToHost));
    p->st_n = keep ? std::min(chunk, n_hist - (n_chunks - 1) * chunk) : 0;
    p->trace_n = trace ? n_hist : 0;
    p->dump_n = std::min(chunk, n_hist - (n_chunks - 1) * chunk);
    return 0;
}

extern "C" int ramp_run(ramp_plan *p, uint64_t n_hist, uint64_t chunk, const uint32_t *seeds,
                        uint32_t n_seeds, const ramp_run_opts *opts)
{
    return run_plan(p, n_hist, chunk, seeds, n_seeds, opts, false);
}

extern "C" int ramp_run_noalloc(ramp_plan *p, uint64_t n_hist, uint64_t chunk, const uint32_t *seeds,
                                uint32_t n_seeds, const ramp_run_opts *opts)
{
    return run_plan(p, n_hist, chunk, seeds, n_seeds, opts, true);
}

extern "C" int ramp_trace_read(ramp_plan *p, uint32_t slot, float *out, uint64_t n)
{
    if (!p || !out) return fail(RAMP_E_INVALID, "null argument");
    if (p->trace_n == 0) return fail(RAMP_E_STATE, "no trace: run with RAMP_RUN_TRACE first");
    if (slot >= p->n_trace) return fail(RAMP_E_INVALID, "trace slot %u of %u", slot, p->n_trace);
    if (n != p->trace_n)
        return fail(RAMP_E_INVALID, "the traced run had %llu histories, asked %llu",
                    (unsigned long long)p->trace_n, (unsigned long long)n);
    HIPCHK(hipSetDevice(p->ctx->device));
    HIPCHK(hipStreamSynchronize(p->ctx->stream));
    HIPCHK(hipMemcpy(out, (const float *)p->trace.p + (size_t)slot * 4 * n, 4 * n * sizeof(float),
                     hipMemcpyDeviceToHost));
    return 0;
}

/* the pinned staging buffer for histogram reads: at most PIN_MAX bytes,
 * enough for both moments of n bins where that fits (DMA instead of the
 * runtime's pageable bounce copies) */
static int pin_reserve(ramp_plan *p, uint64_t n)
{
    const size_t want = std::min<size_t>(PIN_MAX, std::max<size_t>(2 * n * sizeof(long long), 4096));
    if (p->pin_bytes < want) {
        if (p->pin) (void)hipHostFree(p->pin);
        p->pin = nullptr;
        p->pin_bytes = 0;
        HIPCHK(hipHostMalloc(&p->pin, want, hipHostMallocDefault));
        p->pin_bytes = want;
    }
    return 0;
}

/* add a deterministic histogram's sums, q * 2^-sh, into w and w2 */
static int read_fixed(ramp_plan *p, const DDet &d, double *w, double *w2, uint64_t n)
{
    int rc = pin_reserve(p, n);
    if (rc) return rc;
    hipStream_t s = p->ctx->stream;
    const uint64_t step = p->pin_bytes / (2 * sizeof(long long));
    /* a product with a power of two in the normal range: exact, as ldexp
     * would be */
    const double sw = std::ldexp(1.0, -d.sh_w), sw2 = std::ldexp(1.0, -d.sh_w2);
    for (uint64_t o = 0; o < n; o += step) {
        const uint64_t m = std::min(step, n - o);
        long long *a = (long long *)p->pin, *b = a + m;
        HIPCHK(hipMemcpyAsync(a, d.qw + o, m * sizeof(long long), hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(b, d.qw2 + o, m * sizeof(long long), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        for (uint64_t i = 0; i < m; i++) {
            w[o + i] += (double)a[i] * sw;
            w2[o + i] += (double)b[i] * sw2;
        }
    }
    return 0;
}

extern "C" int ramp_hist_read(ramp_plan *p, uint32_t det, double *w, double *w2, uint64_t n)
{
    if (!p || !w || !w2) return fail(RAMP_E_INVALID, "null argument");
    if (det >= p->hdets.size()) return fail(RAMP_E_INVALID, "detector %u of %zu", det, p->hdets.size());
    const DDet &d = p->hdets[det];
    if (n != d.n) return fail(RAMP_E_INVALID, "detector %u has %llu bins, asked %llu", det,
                              (unsigned long long)d.n, (unsigned long long)n);
    HIPCHK(hipSetDevice(p->ctx->device));
    hipStream_t s = p->ctx->stream;
    int rc = pin_reserve(p, n);
    if (rc) return rc;
    std::fill(w, w + n, 0.0);
    std::fill(w2, w2 + n, 0.0);
    if (p->has_atomic && d.w) {
        const uint64_t step = p->pin_bytes / (2 * sizeof(float));
        for (uint64_t o = 0; o < n; o += step) {
            const uint64_t m = std::min(step, n - o);
            float *a = (float *)p->pin, *b = a + m;
            HIPCHK(hipMemcpyAsync(a, d.w + o, m * sizeof(float), hipMemcpyDeviceToHost, s));
            HIPCHK(hipMemcpyAsync(b, d.w2 + o, m * sizeof(float), hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            for (uint64_t i = 0; i < m; i++) { w[o + i] += a[i]; w2[o + i] += b[i]; }
        }
    }
    if (det < p->fold_w.size() && !p->fold_w[det].empty())
        for (uint64_t i = 0; i < n; i++) { w[i] += p->fold_w[det][i]; w2[i] += p->fold_w2[det][i]; }
    if (p->has_det && d.qw) {
        rc = read_fixed(p, d, w, w2, n);
        if (rc) return rc;
    }
    HIPCHK(hipStreamSynchronize(s));
    return 0;
}

extern "C" int ramp_hist_reset(ramp_plan *p)
{
    if (!p) return fail(RAMP_E_INVALID, "null plan");
    HIPCHK(hipSetDevice(p->ctx->device));
    hipStream_t s = p->ctx->stream;
    for (size_t k = 0; k < p->hdets.size(); k++) {
        if (p->hw[k].p) { launch_fill_f32(s, p->hw[k].p, p->hdets[k].n); launch_fill_f32(s, p->hw2[k].p, p->hdets[k].n); }
        if (p->hq[k].p) { launch_fill_u64(s, p->hq[k].p, p->hdets[k].n); launch_fill_u64(s, p->hq2[k].p, p->hdets[k].n); }
    }
    if (p->ev_list_n.p) launch_fill_u64(s, p->ev_list_n.p, p->n_rescal);
    /* the fold accumulators and the int64 budget restart (the scale stays) */
    p->fold_w.clear();
    p->fold_w2.clear();
    p->ds.used = p->ds.used2 = 0;
    std::fill(p->ev_count.begin(), p->ev_count.end(), 0);
    HIPCHK(hipStreamSynchronize(s));
    return 0;
}

extern "C" int ramp_state_read(ramp_plan *p, float *n16, float *i8, uint32_t *ii, uint64_t n)
{
    if (!p) return fail(RAMP_E_INVALID, "null plan");
    if (p->st_n == 0) return fail(RAMP_E_STATE, "no state kept: run with keep_state=1 first");
    if (n > p->st_n) return fail(RAMP_E_RANGE, "asked %llu states, %llu kept", (unsigned long long)n,
                                 (unsigned long long)p->st_n);
    HIPCHK(hipSetDevice(p->ctx->device));
    HIPCHK(hipStreamSynchronize(p->ctx->stream));
    if (n16) HIPCHK(hipMemcpy(n16, p->st16.p, n * 16 * sizeof(float), hipMemcpyDeviceToHost));
    if (i8) HIPCHK(hipMemcpy(i8, p->st8.p, n * 8 * sizeof(float), hipMemcpyDeviceToHost));
    if (ii) HIPCHK(hipMemcpy(ii, p->stii.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int ramp_stats_read(const ramp_plan *p, ramp_run_stats *out)
{
    if (!p || !out) return fail(RAMP_E_INVALID, "null argument");
    *out = p->stats;
    return 0;
}

extern "C" int ramp_math_eval(ramp_ctx *ctx, int fn, const float *x, const float *y, float *out,
                              uint64_t n, int on_device)
{
    if (!x || !out) return fail(RAMP_E_INVALID, "null argument");
    if (fn < RAMP_FN_SIN || fn > RAMP_FN_DIV) return fail(RAMP_E_INVALID, "unknown function %d", fn);
    if (!on_device) {
        for (uint64_t i = 0; i < n; i++) out[i] = math_eval_host(fn, x[i], y ? y[i] : 0.0f);
        return 0;
    }
    if (!ctx) return fail(RAMP_E_INVALID, "device evaluation needs a context");
    HIPCHK(hipSetDevice(ctx->device));
    DevBuf bx, by, bo;
    int rc = dev_alloc(bx, n * 4 + 4);
    if (!rc) rc = dev_alloc(by, n * 4 + 4);
    if (!rc) rc = dev_alloc(bo, n * 4 + 4);
    if (rc) { dev_free(bx); dev_free(by); dev_free(bo); return rc; }
    (void)hipMemcpy(bx.p, x, n * 4, hipMemcpyHostToDevice);
    if (y) (void)hipMemcpy(by.p, y, n * 4, hipMemcpyHostToDevice);
    unsigned g = (unsigned)std::min<uint64_t>((n + 255) / 256 + 1, 4096);
    hipLaunchKernelGGL(k_math, dim3(g), dim3(256), 0, ctx->stream, fn, (const float *)bx.p,
                       y ? (const float *)by.p : nullptr, (float *)bo.p, n);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(out, bo.p, n * 4, hipMemcpyDeviceToHost);
    dev_free(bx); dev_free(by); dev_free(bo);
    if (e != hipSuccess) return fail(RAMP_E_DEVICE, "math eval: %s", hipGetErrorString(e));
    return 0;
}

extern "C" int ramp_rng_eval(ramp_ctx *ctx, const uint32_t *tid, const float *ctr, float *out,
                             uint64_t n, int on_device)
{
    if (!tid || !ctr || !out) return fail(RAMP_E_INVALID, "null argument");
    if (!on_device) {
        for (uint64_t i = 0; i < n; i++) {
            float c = ctr[i];
            out[i] = rr_draw(&c, tid[i]);
        }
        return 0;
    }
    if (!ctx) return fail(RAMP_E_INVALID, "device evaluation needs a context");
    HIPCHK(hipSetDevice(ctx->device));
    DevBuf bt, bc, bo;
    int rc = dev_alloc(bt, n * 4 + 4);
    if (!rc) rc = dev_alloc(bc, n * 4 + 4);
    if (!rc) rc = dev_alloc(bo, n * 4 + 4);
    if (rc) { dev_free(bt); dev_free(bc); dev_free(bo); return rc; }
    (void)hipMemcpy(bt.p, tid, n * 4, hipMemcpyHostToDevice);
    (void)hipMemcpy(bc.p, ctr, n * 4, hipMemcpyHostToDevice);
    unsigned g = (unsigned)std::min<uint64_t>((n + 255) / 256 + 1, 4096);
    hipLaunchKernelGGL(k_rng, dim3(g), dim3(256), 0, ctx->stream, (const uint32_t *)bt.p,
                       (const float *)bc.p, (float *)bo.p, n);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(out, bo.p, n * 4, hipMemcpyDeviceToHost);
    dev_free(bt); dev_free(bc); dev_free(bo);
    if (e != hipSuccess) return fail(RAMP_E_DEVICE, "rng eval: %s", hipGetErrorString(e));
    return 0;
}

#include "batch_run.h"
#include "adjoint_run.h"
