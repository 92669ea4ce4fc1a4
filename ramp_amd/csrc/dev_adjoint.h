/* dev_adjoint.h - histogram context of the adjoint pass (no reference
 * counterpart).
 *
 * The adjoint re-runs every history with the same seeds.  Model parameters
 * enter only through weights (SPowderDiff draws from a fixed proposal, the
 * MCont energy is uniform), so a history's final weight w is a smooth
 * function of theta with d w/d theta = w * d log w/d theta, and for a loss
 * L(h) of a histogram h = sum of event weights per bin,
 *     dL/d theta = sum over events e of dL/dh[bin(e)] * w_e * dlog w_e/d theta.
 * AdjCtx carries the per-history log-derivative terms (set by the source and
 * by s_powderdiff) and the per-lane FP32 gradient sums; hist_add on the loss
 * detector adds one event's contribution instead of filling a histogram.
 *
 * Gradient order: [T (MCont T1), sigma_inc, sigma_abs, U, d_0 .. d_{n-1}].
 */
#ifndef RAMP_DEV_ADJOINT_H
#define RAMP_DEV_ADJOINT_H

#define ADJ_MAXL 48
#define ADJ_NG 4                          /* per-lane scalar gradients: T, sigma_inc, sigma_abs, U */
#define ADJ_NLDS (2 * ADJ_MAXL + 1)       /* per-wave FP64 LDS sums: A[0..MAXL], B[0..MAXL-1] */
#define ADJ_NPART (ADJ_NG + ADJ_NLDS)     /* doubles per workgroup partial */
#define ADJ_NO_POWDER (-2)

/* The d-spacing gradient of line j has a term for every event whose
 * history scattered with line j active (q0_j < 2k):
 *   -x * (ssv2_j / k^2) * dlog(sigma_j)/dd_j,
 * and the chosen line's own term cd.  The active lines at a given k are
 * the m lines of smallest q0, so the first term is -ssv2_j dlogsig_j times
 * the sum of A[m] = dL/dh w x / k^2 over events with m >= c_j (c_j = number
 * of lines with q0 <= q0_j); the kernel sums A[m] and B[ch] = dL/dh w cd
 * per wave in LDS (FP64, one wave's lanes in a fixed order) and the host
 * forms the per-line gradients. */
struct AdjCtx {
    static constexpr bool adjoint = true;
    const DDet *dets;
    int mode;
    int enabled;
    uint32_t events, oob, overflow;
    int det;              /* detector slot of the loss */
    uint64_t nbins;
    const float *dLdh;    /* dL/dh of that detector, one float per bin */
    double *la;           /* this wave's LDS sums A[0..MAXL], B[0..MAXL-1] */
    /* per history */
    float gT;             /* d log w / dT from the source */
    float cs, ca, cU;     /* d log w / d sigma_inc, sigma_abs, U */
    float xivv;           /* depth (m) / k^2 of the powder scattering */
    float cd;             /* d log w / d d_ch, terms of the chosen line only */
    int ch;               /* chosen line, -1 incoherent, ADJ_NO_POWDER: none */
    int m;                /* active lines at the scattering k */
    /* per-lane gradient sums */
    float G[ADJ_NG];
};

DEV void adj_history_reset(AdjCtx &hc, float gT)
{
    hc.gT = gT;
    hc.cs = hc.ca = hc.cU = 0.0f;
    hc.xivv = hc.cd = 0.0f;
    hc.ch = ADJ_NO_POWDER;
    hc.m = 0;
}

DEV void hist_add(AdjCtx &hc, int det, uint64_t idx, float w)
{
    if (det != hc.det) return;
    if (idx >= hc.nbins) {
        hc.oob++;
        return;
    }
    const float sw = hc.dLdh[idx] * w;
    hc.events++;
    hc.G[0] += sw * hc.gT;
    hc.G[1] += sw * hc.cs;
    hc.G[2] += sw * hc.ca;
    hc.G[3] += sw * hc.cU;
    if (hc.ch != ADJ_NO_POWDER) {
        __hip_atomic_fetch_add(&hc.la[hc.m], (double)(sw * hc.xivv), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WAVEFRONT);
        if (hc.ch >= 0)
            __hip_atomic_fetch_add(&hc.la[ADJ_MAXL + 1 + hc.ch], (double)(sw * hc.cd), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WAVEFRONT);
    }
}

#endif
