/* dev_samples.h - sample and crystal components (isotropic S(Q,w), powder,
 * mosaic monochromator) of the fused segment kernels.  Included by
 * dev_components.h. */
#ifndef RAMP_DEV_SAMPLES_H
#define RAMP_DEV_SAMPLES_H

/* sample_distn (scat/isotropic.cl:5-73) over one packed channel table
 * [q(nq) | w(nw) | pw_cdf(nw) | pq_cdf(nw*nq)] */
DEV void iso_sample(NS &n, const float *base, uint32_t nq, uint32_t nw, float &omega, float &Q)
{
    const float *q = base, *w = base + nq, *pw_cdf = base + nq + nw, *pq_cdf = base + nq + 2 * nw;
    float u = drw(n), v = drw(n);
    float mindiff = 1.0f;
    uint32_t w_index = 0;
    omega = 0.0f;
    Q = 0.0f;
    for (uint32_t i = 0; i < nw; i++) {
        if (pw_cdf[i] == 1.0f) break;
        float dd = rm_fabs(pw_cdf[i] - v);
        if (dd < mindiff) {
            mindiff = dd;
            float wprev = (i > 0 ? w[i - 1] : 0.0f);
            omega = wprev + drw(n) * (w[i] - wprev);
            w_index = i;
        }
    }
    mindiff = 1.0f;
    for (uint32_t i = 0; i < nq; i++) {
        float cdf = pq_cdf[w_index * nq + i];
        if (cdf == 1.0f) break;
        float dd = rm_fabs(cdf - u);
        if (dd < mindiff) {
            mindiff = dd;
            float qprev = (i > 0 ? q[i - 1] : 0.0f);
            Q = qprev + drw(n) * (q[i] - qprev);
        }
    }
}

/* scat/isotropic.cl:75-323.  d[0..1]: 2*pi as a double.  The reference's
 * double expressions (the cross-sections over 2*ki*ki and the velocity, the
 * path length, the channel thresholds, 2*M_PI*rand() and the detailed
 * balance) are evaluated in double as it evaluates them. */
template <bool FULL>
DEV void s_isotropic(NS &n, const DComp &c)
{
    const DPart &s = c.s;
    rf3 path = sample_path(n);
    float vlen = v3_len(nvel(n));
    float ki = RC_V2K * vlen;
    rf3 normvel = v3_norm(nvel(n));
    float cs = s.p[IS_SSCAT(0)], is_ = s.p[IS_SSCAT(1)], ms = s.p[IS_SSCAT(2)];
    const double twok2 = 2. * (double)ki * (double)ki;
    float sigma_s = (float)((double)((cs + is_) + ms) / twok2);
    float sigma_a = (float)((double)((s.p[IS_SABS(0)] + s.p[IS_SABS(1)]) + s.p[IS_SABS(2)]) * 2200. /
                            (double)vlen);
    float sigma_tot = sigma_a + sigma_s;
    float mu = (((s.p[IS_RHO(0)] + s.p[IS_RHO(1)]) + s.p[IS_RHO(2)]) * sigma_tot) * 100.0f;
    float plen = v3_len(path);
    float p_trans = rm_exp(-mu * plen);
    if (drw(n) < p_trans) {
        n.w *= p_trans;
        setpos(n, v3_add(ipos2(n), v3_scale(0.0001f, v3_norm(path))));
        n.t += n.is[7];
        if (s.ni[ISI_TRANSMIT] == 0) n.dead = true;
        n.sc = (float)c.comp_idx;
        n.iidx = 0;
        isect_reset(n);
        return;
    }
    n.w *= (1 - p_trans) * sigma_s / sigma_tot;
    float eta = drw(n);
    float lg = rm_log(1 - eta * (1 - rm_exp(-mu * plen)));
    float path_length = (float)((-1. / (double)mu) * (double)lg);
    float u = drw(n);
    u *= sigma_s;
    uint32_t flag = 0;
    const double th1 = (double)cs / twok2, th2 = (double)(cs + is_) / twok2;
    if ((double)u < th1) flag = 1;
    else if ((double)u > th1 && (double)u < th2) flag = 2;
    else if ((double)u > th2) flag = 3;
    float omega = 0.0f, Q = 0.0f;
    if (flag >= 1) {
        int ch = (int)flag - 1;
        iso_sample(n, in_const(s.tab[ch]), s.ni[ISI_NQ(ch)], s.ni[ISI_NW(ch)], omega, Q);
    }
    if (ki * ki < (RC_E2KS * omega)) {
        n.dead = true;
        n.iidx = 0;
        isect_reset(n);
        return;
    }
    float kf;
    /* isotropic.cl:247 in double: kB is a double expression (consts.h:36);
     * the exponential is the shared rd_exp */
    double db = 0.5 * rd_exp(-(double)omega / (RC_KB_D * (double)s.p[IS_TEMP]));
    if ((double)drw(n) > db) {
        kf = rm_sqrt(ki * ki - RC_E2KS * omega);
        n.se = omega;
    } else {
        kf = rm_sqrt(ki * ki + RC_E2KS * omega);
        n.se = -omega;
    }
    float arg = (ki * ki + kf * kf - Q * Q) / (2 * ki * kf);
    if (rm_fabs(arg) > 1) {
        n.dead = true;
        n.iidx = 0;
        isect_reset(n);
        return;
    }
    float theta = rm_acos(arg);
    rf3 oldvel = nvel(n);
    float z = -(n.vx + n.vy) / (n.vz);
    rf3 perp = v3_norm(rf3_make(1.0f, 1.0f, z));
    float alpha = (float)(dp_f64(s.d, 0) * (double)drw(n));
    perp = rot_about_axis(alpha, normvel, perp);
    normvel = rot_about_axis(theta, perp, normvel);
    float TOF = path_length / vlen;
    n.sd = Q;
    if (v3_eq(ipos1(n), ipos2(n))) {
        setpos(n, v3_scale(path_length, path));
        n.t += TOF;
    } else {
        setpos(n, v3_add(ipos1(n), v3_scale(path_length, path)));
        n.t += n.is[3] + TOF;
    }
    setvel(n, v3_mulv(v3_mulv(normvel, kf), RC_K2V));
    if (FULL) {
        rf3 q_norm = v3_norm(v3_mulv(v3_sub(oldvel, nvel(n)), RC_V2K));
        if (flag == 2) {
            n.px = n.px * (-1.0f / 3.0f);
            n.py = n.py * (-1.0f / 3.0f);
            n.pz = n.pz * (-1.0f / 3.0f);
        } else if (flag == 3) {
            float dp = v3_dot(rf3_make(n.px, n.py, n.pz), q_norm);
            n.px = -q_norm.x * dp;
            n.py = -q_norm.y * dp;
            n.pz = -q_norm.z * dp;
        }
    }
    n.sc = (float)c.comp_idx;
    n.iidx = 0;
    isect_reset(n);
}

/* azimuth of the Debye-Scherrer cone (powder1.cl:68-79, powderN.cl:117-130),
 * its double expressions in double; d[0..1] span = d_phi*M_PI/180.0f as a
 * double */
DEV float powder_phi(NS &n, const DPart &s, float d_phi)
{
    if (d_phi != 0.0f) {
        float ph = (float)((double)(2.0f * drw(n) - 1.0f) * dp_f64(s.d, 0));
        if (drw(n) > 0.5f) ph = (float)((double)ph + RC_M_PI_D);
        return (float)((double)ph + RC_M_PI_D / 2.0f);
    }
    return (float)(2.0f * RC_M_PI_D * (double)drw(n));
}

DEV void powder_dirs(rf3 vel, rf3 &para, rf3 &perp)
{
    para = v3_norm(vel);
    perp = v3_cross(para, rf3_make(0.0f, 1.0f, 0.0f));
    if (v3_len(perp) < 1e-3f) perp = v3_norm(v3_cross(para, rf3_make(1.0f, 0.0f, 0.0f)));
    else perp = v3_norm(perp);
}

/* sigma_abs_v * 2200.0 / vel, a double expression (powder1.cl:37) */
DEV float powder_sabs(float sav, float vel) { return (float)((double)sav * 2200.0 / (double)vel); }

/* scat/powder1.cl:5-101 */
DEV void s_powder1(NS &n, const DComp &c)
{
    const DPart &s = c.s;
    n.t += n.is[3];
    rf3 path = v3_sub(ipos2(n), ipos1(n));
    float full = v3_len(path);
    float vel = v3_len(nvel(n));
    float sigma_scat = s.p[P1_SSCAT_V2] / (vel * vel * RC_V2K * RC_V2K);
    float sigma_abs = powder_sabs(s.p[P1_SABS_V], vel);
    float p_scat = sigma_scat / (sigma_abs + sigma_scat);
    float p_inter = 1 - rm_exp(-(sigma_scat + sigma_abs) * full);
    float arg = s.p[P1_Q] * RC_K2V / (2.0f * vel);
    if (arg > 1.0f) {
        s_done(n, c.comp_idx);
        return;
    }
    float twotheta = 2.0f * rm_asin(arg);
    n.w *= p_scat;
    float pen = -rm_log(1.0f - drw(n) * (1.0f - p_inter)) / (sigma_abs + sigma_scat);
    rf3 para, perp;
    powder_dirs(nvel(n), para, perp);
    float phi = powder_phi(n, s, s.p[P1_DPHI]);
    perp = rot_about_axis(phi, para, perp);
    para = rot_about_axis(twotheta, perp, para);
    setpos(n, v3_add(ipos1(n), v3_scale(pen, path)));
    setvel(n, v3_scale(vel, v3_norm(para)));
    n.t += n.is[7] - n.is[3];
    n.w *= full * sigma_scat * rm_exp(-(sigma_scat + sigma_abs) * full);
    s_done(n, c.comp_idx);
}

/* scat/powderN.cl:5-157 (line scan bounded by the line count) */
DEV void s_powderN(NS &n, const DComp &c)
{
    const DPart &s = c.s;
    const float *ssv2 = in_const(s.tab[PNT_SSCAT_V2]), *qs = in_const(s.tab[PNT_Q]);
    uint32_t nl = s.ni[PNI_NLINES];
    rf3 path = sample_path(n);
    float full = v3_len(path);
    float vel = v3_len(nvel(n));
    float sigma_abs = powder_sabs(s.p[PN_SABS_V], vel);
    float vv = vel * vel * RC_V2K * RC_V2K;
    float sum = 0.0f;
    for (uint32_t i = 0; i < nl; i++) sum += ssv2[i] / vv;
    float p_inter = 1 - rm_exp(-(sum + sigma_abs) * full);
    if (drw(n) > p_inter) {
        n.w *= 1 - p_inter;
        setpos(n, v3_add(ipos2(n), v3_scale(0.0001f, v3_norm(path))));
        n.t += n.is[7];
        if (s.ni[PNI_TRANSMIT] == 0) n.dead = true;
        n.sc = (float)c.comp_idx;
        n.iidx = 0;
        isect_reset(n);
        return;
    }
    uint32_t attempts = 0;
    float q_chosen = 0.0f, arg;
    while (true) {
        float u = drw(n), v = 0.0f;
        uint32_t i = 0;
        while (v < u) {
            if (i >= nl) break;
            float sg = ssv2[i] / vv;
            q_chosen = qs[i];
            v += sg / sum;
            i++;
        }
        arg = q_chosen * RC_K2V / (2.0f * vel);
        if (arg < 1.0f) break;
        else if (arg > 1.0f && attempts > 2 * nl) {
            n.dead = true;
            n.iidx = 0;
            isect_reset(n);
            return;
        }
        attempts++;
    }
    float twotheta = 2.0f * rm_asin(arg);
    float pen = -rm_log(1.0f - drw(n) * (1.0f - p_inter)) / (sigma_abs + sum);
    rf3 para, perp;
    powder_dirs(nvel(n), para, perp);
    float phi = powder_phi(n, s, s.p[PN_DPHI]);
    perp = rot_about_axis(phi, para, perp);
    para = rot_about_axis(twotheta, perp, para);
    if (v3_eq(ipos1(n), ipos2(n))) {
        rf3 dp = v3_scale(pen, path);
        setpos(n, v3_add(npos(n), dp));
        n.t += v3_len(dp) / v3_len(nvel(n));
    } else {
        setpos(n, v3_add(ipos1(n), v3_scale(pen, path)));
        n.t += n.is[7] - n.is[3];
    }
    setvel(n, v3_scale(vel, v3_norm(para)));
    n.w *= (1 - rm_exp(-(sum + sigma_abs) * pen)) * sum / (sum + sigma_abs);
    s_done(n, c.comp_idx);
}

/* GAUSS of monochromator.cl:5-8 with mean 0: the exponential in float,
 * divided in double by sqrt(2.0f*M_PI) (double) times rms, as the
 * reference's double promotion evaluates it */
DEV double mono_gauss(float x, float rms)
{
    float e = rm_exp(-(x) * (x) / (2.0f * (rms) * (rms)));
    return (double)e / (RC_SQRT_2PI_D * (double)rms);
}

/* the same for one of the component's two mosaic widths, from the host's
 * double constants (dp_f64 slots k..k+5): 1/(2.0f*rms*rms), the divisor
 * sqrt(2.0f*M_PI)*rms and its reciprocal.  The exponent's float quotient
 * is the float nearest the product with the reciprocal when it is normal
 * (a float over a float: the double product is within 2^-52 of it, closer
 * than any rounding boundary); a subnormal exponent gives exp = 1 either
 * way.  The double quotient is exact (rd_div_rc). */
DEV double mono_gauss_c(float x, const float *d, int k)
{
    float e = rm_exp((float)((double)(-(x) * (x)) * dp_f64(d, k)));
    return rd_div_rc((double)e, dp_f64(d, k + 2), dp_f64(d, k + 4));
}

/* scat/monochromator.cl:22-197.  d[0] mos_rms_y, d[1] mos_rms_z,
 * d[2] mos_rms_max, d[3] Q, d[4] zmax, d[5] ymax; doubles for mono_gauss_c
 * at d[6..11] (mos_rms_y) and d[12..17] (mos_rms_z).  The reference's double
 * sites (GAUSS's normalisation and the products and quotient of GAUSS
 * values, Box-Muller's cos(2.0f*M_PI*u2)) are evaluated in FP64, the cos
 * with the shared rd_cos: only reflected neutrons reach them. */
/* Differentiable powder + incoherent sample (no reference counterpart).
 * Depth, channel and grain d-spacing are drawn from the proposal; the
 * weight is multiplied by model / proposal probability:
 *   depth x and channel ch:  sig_ch e^{-mu x} (1 - e^{-mu0 L}) mus0 / (mu0 e^{-mu0 x} sig0_ch)
 *   grain d-spacing D of line j:  N(D; d_j, s d_j) / N(D; d0_j, s d0_j)
 * with mu = mus + mu_abs, mus = incoherent + active lines; at the proposal
 * the first factor is the forced-scattering weight (1 - e^{-mu L}) mus/mu
 * and the second is 1.  The incoherent direction is uniform on the sphere;
 * a line scatters by 2 asin(pi / (k D)) on a uniform Debye-Scherrer cone.
 * In the adjoint pass the log-derivative terms go to the context. */
template <class HC>
DEV void s_powderdiff(NS &n, const DComp &c, HC &hc)
{
    const DPart &s = c.s;
    const float *Lt = in_const(s.tab[PDT_LINES]);
    const uint32_t nl = s.ni[PDI_NLINES];
    rf3 path = sample_path(n);
    float Lp = v3_len(path);
    float v = v3_len(nvel(n));
    float k = RC_V2K * v;
    float vv = v * v * RC_V2K * RC_V2K;
    float twok = 2.0f * k;
    float lam = 6.28318530717958647692f / k;
    float inc = 100.0f * s.p[PD_SINC], inc0 = 100.0f * s.p[PD_SINC0];
    float ab = 100.0f * s.p[PD_SABS] * (lam / s.p[PD_LREF]);
    float ab0 = 100.0f * s.p[PD_SABS0] * (lam / s.p[PD_LREF]);
    /* the active lines' sigma*v^2 summed, then divided by v^2 once (one
     * reciprocal per neutron, not a division per line) */
    const float rvv = 1.0f / vv;
    float scoh = 0.0f, scoh0 = 0.0f;
    for (uint32_t j = 0; j < nl; j++) {
        const float *L = Lt + j * PD_LW;
        if (L[PDL_Q0] < twok) {
            scoh += L[PDL_SSV2];
            scoh0 += L[PDL_SSV2_0];
        }
    }
    float coh = scoh * rvv, coh0 = scoh0 * rvv;
    float mus = inc + coh, mu = mus + ab;
    float mus0 = inc0 + coh0, mu0 = mus0 + ab0;
    if (!(mus0 > 0.0f) || !(Lp > 0.0f)) {
        n.dead = true;
        s_done(n, c.comp_idx);
        return;
    }
    float pint0 = 1.0f - rm_exp(-mu0 * Lp);
    float x = -rm_log(1.0f - drw(n) * pint0) / mu0;
    /* channel from the proposal: incoherent, then the active lines in order */
    float uc = drw(n) * mus0;
    int ch = -1;
    float sig = inc, sig0 = inc0;
    if (uc >= inc0) {
        /* the first active line whose running proposal sum of sigma*v^2
         * exceeds (uc - inc0) v^2 */
        const float tc = (uc - inc0) * vv;
        float acc = 0.0f;
        for (uint32_t j = 0; j < nl; j++) {
            const float *L = Lt + j * PD_LW;
            if (L[PDL_Q0] < twok) {
                ch = (int)j;
                acc += L[PDL_SSV2_0];
                if (tc < acc) break;
            }
        }
        if (ch >= 0) {
            sig = Lt[ch * PD_LW + PDL_SSV2] * rvv;
            sig0 = Lt[ch * PD_LW + PDL_SSV2_0] * rvv;
        } else {
            sig = inc;
            sig0 = inc0;
        }
    }
    n.w *= (sig * rm_exp(-mu * x)) * pint0 * mus0 / ((mu0 * rm_exp(-mu0 * x)) * sig0);
    float frac = x / Lp;
    bool inside = v3_eq(ipos1(n), ipos2(n));
    rf3 start = inside ? npos(n) : ipos1(n);
    float t_start = inside ? 0.0f : n.is[3];
    setpos(n, v3_add(start, v3_scale(frac, path)));
    n.t += t_start + x / v;
    float cd = 0.0f;
    if (ch < 0) {
        float ct = 1.0f - 2.0f * drw(n);
        float ph = 6.28318530717958647692f * drw(n);
        float st = rm_sqrt(rm_fmax(0.0f, 1.0f - ct * ct));
        n.vx = v * (st * rm_cos(ph));
        n.vy = v * (st * rm_sin(ph));
        n.vz = v * ct;
    } else {
        const float *L = Lt + ch * PD_LW;
        float d = L[PDL_D], d0 = L[PDL_D0], sr = s.p[PD_STRAIN];
        float u1 = rm_fmax(drw(n), 1e-30f), u2 = drw(n);
        float g = rm_sqrt(-2.0f * rm_log(u1)) * rm_cos(6.28318530717958647692f * u2);
        float D = d0 * (1.0f + sr * g);
        float z = (D - d) / (sr * d), z0 = (D - d0) / (sr * d0);
        n.w *= (d0 / d) * rm_exp(0.5f * (z0 * z0 - z * z));
        cd = L[PDL_DLS] - 1.0f / d + z * (D / (sr * d * d));
        float sth = 3.14159265358979323846f / (k * D);
        if (!(sth < 1.0f) || !(D > 0.0f)) {
            n.dead = true;
            s_done(n, c.comp_idx);
            return;
        }
        float tt = 2.0f * rm_asin(sth);
        rf3 para, perp;
        powder_dirs(nvel(n), para, perp);
        float phi = 6.28318530717958647692f * drw(n);
        perp = rot_about_axis(phi, para, perp);
        para = rot_about_axis(tt, perp, para);
        setvel(n, v3_scale(v, v3_norm(para)));
    }
    if constexpr (HC::adjoint) {
        /* d log w: sig_ch and -mu x depend on the model */
        float sq2 = 0.0f;
        int m = 0;
        for (uint32_t j = 0; j < nl; j++) {
            const float *L = Lt + j * PD_LW;
            if (L[PDL_Q0] < twok) {
                sq2 += L[PDL_SSV2] * L[PDL_Q2];
                m++;
            }
        }
        const float sumq2 = sq2 * rvv;
        hc.cs = (ch < 0 && s.p[PD_SINC] > 0.0f ? 1.0f / s.p[PD_SINC] : 0.0f) - 100.0f * x;
        hc.ca = -100.0f * x * (lam / s.p[PD_LREF]);
        hc.cU = (ch >= 0 ? -Lt[ch * PD_LW + PDL_Q2] : 0.0f) + x * sumq2;
        hc.xivv = x / vv;
        hc.cd = cd;
        hc.ch = ch;
        hc.m = m;
    }
    s_done(n, c.comp_idx);
}

/* slab of the monochromator grid (monochromator.cl:62-82): tilts, centre
 * and the rotation into the slab frame, from the slab's column and row */
struct MonoSlab {
    rmat3 SR;
    float cy, cz;
};

DEV MonoSlab mono_slab(const DPart &s, float col, float row)
{
    uint32_t nh = s.ni[MOI_NH], nv = s.ni[MOI_NV];
    float sw = s.p[MO_SLAB_W], sh = s.p[MO_SLAB_H], gap = s.p[MO_GAP];
    float zmin = -s.d[4], ymin = -s.d[5];
    float tilt_h = s.p[MO_RAD_H] > 0.0f
                       ? rm_asin(((float)col - (float)(nh + 1) / 2.0f) * (sw + gap) / s.p[MO_RAD_H])
                       : 0.0f;
    float tilt_v = s.p[MO_RAD_V] > 0.0f
                       ? -rm_asin(((float)row - (float)(nv + 1) / 2.0f) * (sh + gap) / s.p[MO_RAD_V])
                       : 0.0f;
    MonoSlab m;
    m.cz = zmin + (col - 0.5f) * (sw + gap) - gap / 2.0f;
    m.cy = ymin + (row - 0.5f) * (sh + gap) - gap / 2.0f;
    m.SR = rmat3_euler(0.0f, tilt_h, tilt_v);
    return m;
}

/* Bragg geometry of the neutron in the slab frame (monochromator.cl:90-110) */
struct MonoBragg {
    rf3 ki, ku;
    float kl, q0x, theta, delta, l12;
};

DEV MonoBragg mono_bragg(const NS &n, float Q)
{
    MonoBragg g;
    g.ki = v3_scale(RC_V2K, nvel(n));
    float ratio = -2.0f * g.ki.x / Q;
    float Q_order = rm_floor(ratio + 0.5f);
    if (Q_order == 0.0f) Q_order = ratio < 0.0f ? -1.0f : 1.0f;
    if (Q_order < 0.0f) Q_order = -Q_order;
    g.ku = v3_norm(g.ki);
    g.kl = v3_len(g.ki);
    if (Q_order > 2.0f * g.kl / Q) Q_order -= 1.0f;
    float q0 = Q_order * Q;
    g.q0x = ratio < 0.0f ? -q0 : q0;
    g.theta = rm_asin(q0 / (2.0f * g.kl));
    g.delta = rm_asin(rm_fabs(g.ku.x)) - g.theta;
    g.l12 = v2_len(rf2_make(g.ki.y, g.ki.z));
    return g;
}

DEV void mono_leave(NS &n, const MonoSlab &m)
{
    setpos(n, rmat3_rotate(&m.SR, npos(n)));
    setpos(n, rf3_make(n.x + 0.0f, n.y + m.cy, n.z + m.cz));
    setvel(n, rmat3_rotate(&m.SR, nvel(n)));
}

/* monochromator.cl:10-197, first half: slab selection and the reflection
 * test (one draw).  Returns true when the neutron reflects; it is then
 * left in the slab frame, aux = column | row << 16, and s_mono_b finishes
 * it.  Otherwise the component is complete. */
DEV bool s_mono_a(NS &n, const DComp &c, uint32_t &aux)
{
    const DPart &s = c.s;
    setpos(n, ipos2(n));
    n.t += n.is[7];
    float Q = s.d[3];
    float sw = s.p[MO_SLAB_W], sh = s.p[MO_SLAB_H], gap = s.p[MO_GAP];
    float zmax = s.d[4], zmin = -zmax, ymax = s.d[5], ymin = -ymax;
    if (n.z > zmin && n.z < zmax && n.y > ymin && n.y < ymax) {
        float col = rm_ceil((n.z - zmin) / (sw + gap));
        float row = rm_ceil((n.y - ymin) / (sh + gap));
        MonoSlab m = mono_slab(s, col, row);
        setvel(n, rmat3_derotate(&m.SR, nvel(n)));
        setpos(n, rf3_make(n.x - 0.0f, n.y - m.cy, n.z - m.cz));
        setpos(n, rmat3_derotate(&m.SR, npos(n)));
        if (rm_fabs(n.z) <= sw / 2.0f && rm_fabs(n.y) <= sh / 2.0f) {
            MonoBragg g = mono_bragg(n, Q);
            /* the exponents' last division, by 2*mry*mry and 2*mrz*mrz, as
             * a product with the reciprocal (see mono_gauss_c) */
            float p_reflect = rm_fabs(s.p[MO_R0]) *
                              rm_exp((float)((double)(-g.ki.z * g.ki.z / (g.l12 * g.l12) * g.delta * g.delta) *
                                             dp_f64(s.d, 6))) *
                              rm_exp((float)((double)(-g.ki.y * g.ki.y / (g.l12 * g.l12) * g.delta * g.delta) *
                                             dp_f64(s.d, 12)));
            if (drw(n) <= p_reflect) {
                aux = (uint32_t)col | ((uint32_t)row << 16);
                return true;
            }
        }
        mono_leave(n, m);
    }
    s_done(n, c.comp_idx);
    return false;
}

/* second half: the mosaic reflection of a neutron s_mono_a selected (two
 * draws), back to the component frame */
DEV void s_mono_b(NS &n, const DComp &c, uint32_t aux)
{
    const DPart &s = c.s;
    const float *gx = in_const(s.tab[MOT_GX]), *gw = in_const(s.tab[MOT_GW]);
    uint32_t glen = s.ni[MOI_GLEN];
    float mry = s.d[0], mrz = s.d[1], mrmax = s.d[2], Q = s.d[3];
    MonoBragg g = mono_bragg(n, Q);
    rf3 ki = g.ki, ku = g.ku;
    float kl = g.kl, theta = g.theta;
    float cos_2theta = rm_cos(2.0f * theta);
    float k_sin_2theta = kl * rm_sin(2.0f * theta);
    rf3 b = v3_mulv(v3_norm(v3_cross(ki, rf3_make(g.q0x, 0.0f, 0.0f))), k_sin_2theta);
    rf3 a = v3_cross(b, ku);
    float total = 0.0f, w = 0.0f;
    float mos_sample = mrmax / rm_cos(theta);
    rf3 c1 = v3_mulv(ki, (cos_2theta - 1.0f));
    rf3 qv;
    for (uint32_t i = 0; i < 100; i++) {
        w = 5.0f * mos_sample;
        float cphi = rm_cos(w), sphi = rm_sin(w);
        qv = v3_add(v3_add(c1, v3_scale(cphi, a)), v3_scale(sphi, b));
        qv.y = qv.y / mrz;
        qv.z = qv.z / mry;
        if (qv.z * qv.z + qv.y * qv.y < (25.0f / (2.0f / 3.0f)) * (qv.x * qv.x)) break;
        mos_sample *= (2.0f / 3.0f);
    }
    for (uint32_t i = 0; i < glen; i++) {
        float phi = w * gx[i];
        float cphi = rm_cos(phi), sphi = rm_sin(phi);
        qv = v3_add(v3_add(c1, v3_scale(cphi, a)), v3_scale(sphi, b));
        float pr = (float)(mono_gauss_c(qv.z / qv.x, s.d, 6) * mono_gauss_c(qv.y / qv.x, s.d, 12));
        total += gw[i] * pr;
    }
    total *= w;
    /* randnorm (monochromator.cl:10-20) */
    float u1 = drw(n), u2 = drw(n);
    float rn = (float)((double)rm_sqrt(-2.0f * rm_log(u1)) * rd_cos(2.0f * RC_M_PI_D * (double)u2));
    float phi = mos_sample * rn;
    float cphi = rm_cos(phi), sphi = rm_sin(phi);
    qv = v3_add(v3_add(c1, v3_scale(cphi, a)), v3_scale(sphi, b));
    float p_reflect = (float)(mono_gauss_c(qv.z / qv.x, s.d, 6) * mono_gauss_c(qv.y / qv.x, s.d, 12));
    setvel(n, v3_scale(RC_K2V, v3_add(ki, qv)));
    p_reflect = (float)((double)p_reflect / ((double)total * mono_gauss(phi, mos_sample)));
    if (p_reflect > 1.0f) p_reflect = 1.0f;
    n.w *= p_reflect;
    /* the slab frame is rebuilt here rather than kept live through the
     * quadrature (register pressure) */
    MonoSlab m = mono_slab(s, (float)(aux & 0xffffu), (float)(aux >> 16));
    mono_leave(n, m);
    s_done(n, c.comp_idx);
}

DEV void s_mono(NS &n, const DComp &c)
{
    uint32_t aux = 0;
    if (s_mono_a(n, c, aux)) s_mono_b(n, c, aux);
}

#endif
