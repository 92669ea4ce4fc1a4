/* adjoint_kernel.h - k_adjoint<F>: the fused re-run of a forward that
 * accumulates the gradient terms (host side in adjoint_run.h).  Instances
 * are compiled per feature set in kern_v.hip. */
#ifndef RAMP_ADJOINT_KERNEL_H
#define RAMP_ADJOINT_KERNEL_H

#include "seg_kernel.h"

struct AdjArgs {
    int det;
    uint32_t pad;
    uint64_t nbins;
    const float *dLdh;
    double *partial;   /* [grid][ADJ_NPART] */
};

DEV double wave_sum_f64(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

/* d log w0 / dT1 of an MCont history (cont_mod.cl:37-47): w0 is
 * I1 M(E;T)/M(kT;T)/N, and M(E;T)/M(kT;T) = sqrt(E/kT) e^{1 - E/kT}; on
 * the 1/E tail w0 is proportional to T.  Only a single-temperature source
 * (the second and third terms not evaluated) has this form; otherwise 0. */
DEV float source_dlogw_dT(const NS &n, const DPart &s)
{
    if (s.type != RAMP_SRC_CONT || s.d[16] == 0.0f || s.d[17] == 0.0f) return 0.0f;
    float T = s.p[SP_T1];
    float v = v3_len(nvel(n));
    float E_J = RC_MEV2J * (RC_VS2E * (v * v));
    if (s.p[SP_TAIL_JOIN] > 0.0f && E_J > s.d[9]) return 1.0f / T;
    return (E_J * s.d[0] - 0.5f) / T;
}

template <uint32_t F>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu((F) == FEAT_SAMPLE ? RAMP_W_ADJ : 1))) void k_adjoint(SegArgs a, AdjArgs x)
{
    __shared__ double red[4][ADJ_NG];
    __shared__ double lsum[4][ADJ_NLDS];
    AdjCtx hc;
    hc.dets = a.dets;
    hc.mode = 0;
    hc.enabled = 1;
    hc.events = hc.oob = hc.overflow = 0;
    hc.det = x.det;
    hc.nbins = x.nbins;
    hc.dLdh = x.dLdh;
#pragma unroll
    for (int i = 0; i < ADJ_NG; i++) hc.G[i] = 0.0f;
    adj_history_reset(hc, 0.0f);
    const uint32_t lane = __lane_id();
    const uint32_t wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (uint32_t i = lane; i < ADJ_NLDS; i += 64) lsum[wid][i] = 0.0;
    hc.la = lsum[wid];
    wave_lds_sync();
    const uint64_t G = (uint64_t)gridDim.x * (blockDim.x >> 6);
    const uint64_t g = (uint64_t)blockIdx.x * (blockDim.x >> 6) + wid;
    const uint64_t per = (a.n + G - 1) / G;
    const uint64_t beg = g * per;
    const uint64_t end = beg + per < a.n ? beg + per : a.n;
    const int ns = a.s1;
    for (uint64_t base = beg; base < end; base += 64) {
        const uint64_t i = base + lane;
        const bool valid = i < end;
        NS n;
        if (valid) {
            float sb_old = a.carry ? a.slot_ctr[i] : 0.0f;
            n.iidx = a.carry ? a.slot_iidx[i] : 0u;
            n.key = (uint32_t)i;
            src_generate<false, F>(n, plan_at(a.src, 0), a.num_sim, a.seed, sb_old);
        } else {
            n.dead = true;
            n.pend = false;
            n.key = 0;
            n.iidx = 0;
        }
        adj_history_reset(hc, valid ? source_dlogw_dT(n, plan_at(a.src, 0)) : 0.0f);
        for (int k = 0; k < ns; k++) {
            if (!__any(!n.dead)) break;
            const DStep st = plan_at(a.steps, k);
            if (st.kind == 0) {
                const DComp &c = plan_at(a.comps, st.first);
                tr_in<false>(n, c);
                geom_dispatch<F>(n, c);
                terminate(n, c.restore);
                scat_dispatch<false, F>(n, c, hc);
                tr_out<false>(n, c);
            } else if constexpr ((F & FEAT_NLIN) != 0u) {
                for (uint32_t ev = 0; ev < st.multi; ev++) {
                    for (uint32_t j = 0; j < st.count; j++) {
                        const DComp &c = plan_at(a.comps, st.first + j);
                        tr_in<false>(n, c);
                        geom_dispatch<F>(n, c);
                        tr_out<false>(n, c);
                    }
                    for (uint32_t j = 0; j < st.count; j++) {
                        const DComp &c = plan_at(a.comps, st.first + j);
                        tr_in<false>(n, c);
                        scat_dispatch<false, F>(n, c, hc);
                        tr_out<false>(n, c);
                    }
                }
            }
        }
        if (valid && a.carry) {
            a.slot_ctr[n.key] = n.ctr;
            a.slot_iidx[n.key] = n.iidx;
        }
    }
#pragma unroll
    for (int i = 0; i < ADJ_NG; i++) {
        double v = wave_sum_f64((double)hc.G[i]);
        if (lane == 0) red[wid][i] = v;
    }
    __syncthreads();
    const uint32_t nw = blockDim.x >> 6;
    for (uint32_t i = threadIdx.x; i < ADJ_NPART; i += blockDim.x) {
        double t = 0.0;
        if (i < ADJ_NG)
            for (uint32_t w = 0; w < nw; w++) t += red[w][i];
        else
            for (uint32_t w = 0; w < nw; w++) t += lsum[w][i - ADJ_NG];
        x.partial[(uint64_t)blockIdx.x * ADJ_NPART + i] = t;
    }
    uint32_t ev = wave_sum(hc.events), oob = wave_sum(hc.oob);
    if (lane == 0) {
        if (ev) atomicAdd(&a.cnt[CNT_EVENTS], (unsigned long long)ev);
        if (oob) atomicAdd(&a.cnt[CNT_OOB], (unsigned long long)oob);
    }
}

typedef void (*adj_fn)(SegArgs, AdjArgs);

#endif
