/* kern_v.hip - one feature set's instances of the large kernels (see
 * kern_table.h); built eighteen times, with RAMP_KV = 0..5 and RAMP_KPART
 * = 0 (fast-mode segment kernels), 1 (full-state segment kernels) or 2
 * (batch and adjoint kernels). */
#include <hip/hip_runtime.h>

#include "../../include/ramp.h"
#include "kern_table.h"

#if !defined(RAMP_KV) || !defined(RAMP_KPART)
#error "RAMP_KV (the feature-set index, 0..5) and RAMP_KPART (0, 1 or 2) must be defined"
#endif

#define KV_CAT2(a, b) a##b
#define KV_CAT(a, b) KV_CAT2(a, b)

static constexpr uint32_t kFeats[RAMP_NUM_KV] = RAMP_KV_FEATS;
static constexpr uint32_t KV_F = kFeats[RAMP_KV];

#if RAMP_KPART == 0
void KV_CAT(ramp_kern_v, RAMP_KV)(KernTable *t)
{
    t->seg[0][0] = k_segment<false, false, KV_F>;
    t->seg[0][1] = k_segment<false, true, KV_F>;
}
#elif RAMP_KPART == 1
void KV_CAT(ramp_kern_full_v, RAMP_KV)(KernTable *t)
{
    t->seg[1][0] = k_segment<true, false, KV_F>;
    t->seg[1][1] = k_segment<true, true, KV_F>;
}
#else
void KV_CAT(ramp_kern_ba_v, RAMP_KV)(KernTable *t)
{
    t->batch = k_batch<KV_F>;
    if constexpr ((KV_F & FEAT_SAMPLE) != 0) t->adjoint = k_adjoint<KV_F>;
    else t->adjoint = nullptr;
}
#endif
