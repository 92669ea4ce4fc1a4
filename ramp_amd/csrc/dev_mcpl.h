/* dev_mcpl.h - the MCPL particle components as device steps, and SRescal,
 * whose event list uses the same tag-and-compact scheme.
 *
 * SMCPLIn (scat/mcpl_in.py:9-76).  The reference overwrites the whole
 * neutron buffer of a chunk on the host: slot i gets particle i of the file
 * (MCPL_to_RAMP, :9-33) and the slots beyond the file's particles are set
 * dead with all fields zero (:66-71).  It does this for every slot, dead or
 * alive and whatever the slot's iidx, and leaves iidx and the intersection
 * as the component's geometry set them.  Here each lane converts its own
 * particle record in FP64, as the host conversion does, and rounds once to
 * FP32.  The particle's RNG counter is float32(r_i + i) with r_i a random
 * integer in [0, 2^30) (:10); r_i is the low 30 bits of Threefry word 0 at
 * key i and counter 2^31 + chunk seed, so a run is reproducible.
 *
 * SMCPLOut (scat/mcpl_out.py:55-117).  Every live neutron of the chunk is
 * written as an 11-float record (write_particle, :55-73), whatever its
 * iidx; the state, iidx and intersection are left unchanged.  Lane i
 * writes record i of the component's dump (field-major: field k of record
 * i at k * cap + i) and tags it with the chunk number, so the records of
 * the last chunk are found in history order.
 */
#ifndef RAMP_DEV_MCPL_H
#define RAMP_DEV_MCPL_H

/* MCPL's adaptive-projection unit vector (format version 3): (u1, u2) and
 * the sign of the dropped component; |u1| > 1 means u1 = 1/uz with x
 * dropped, |u2| > 1 means u2 = 1/uz with y dropped, else z is dropped. */
RHD void mcpl_unpack_dir(double p0, double p1, double s, double *ux, double *uy, double *uz)
{
    /* written with selects: stores through branch-dependent pointers would
     * keep the outputs in scratch memory */
    const bool xd = fabs(p0) > 1.0;
    const bool yd = !xd && fabs(p1) > 1.0;
    const double a = xd ? p1 : p0;                           /* uy if x is dropped, else ux */
    const double b = xd ? 1.0 / p0 : (yd ? 1.0 / p1 : p1);   /* uz if x or y is dropped, else uy */
    const double d = s * sqrt(fmax(0.0, 1.0 - (a * a + b * b)));
    *ux = xd ? d : a;
    *uy = xd ? a : (yd ? d : b);
    *uz = (xd || yd) ? b : d;
}

/* the inverse, for a unit vector (z kept as the dropped component on ties) */
RHD void mcpl_pack_dir(double ux, double uy, double uz, double *p0, double *p1, double *s)
{
    double ax = fabs(ux), ay = fabs(uy), az = fabs(uz);
    if (az >= ax && az >= ay) {
        *p0 = ux; *p1 = uy; *s = uz;
    } else if (ax >= ay) {
        *p0 = 1.0 / uz; *p1 = uy; *s = ux;
    } else {
        *p0 = ux; *p1 = 1.0 / uz; *s = uy;
    }
    *s = signbit(*s) ? -1.0 : 1.0;
}

/* counter base of particle i (mcpl_in.py:10) */
RHD uint32_t mcpl_ctr_base(uint32_t i, uint32_t seed)
{
    return rr_word0(0x80000000u + seed, i) & 0x3fffffffu;
}

/* MCPL_to_RAMP (mcpl_in.py:9-33): record rec (canonical MCR_* layout, read
 * as doubles) -> the 16-float neutron (pos, vel, pol, w, t, counter) */
RHD void mcpl_to_state(const double rec[MCR_N], uint32_t i, uint32_t seed, float out[12])
{
    const double SE2V = 437.393377;
    double ekin = fabs(rec[MCR_U + 2]);
    double s = signbit(rec[MCR_U + 2]) ? -1.0 : 1.0;
    double ux, uy, uz;
    mcpl_unpack_dir(rec[MCR_U], rec[MCR_U + 1], s, &ux, &uy, &uz);
    double KE = ekin * 1e9;
    double rk = sqrt(KE);
    out[0] = (float)(rec[MCR_X] * 1e-2);
    out[1] = (float)(rec[MCR_X + 1] * 1e-2);
    out[2] = (float)(rec[MCR_X + 2] * 1e-2);
    out[3] = (float)(ux * rk * SE2V);
    out[4] = (float)(uy * rk * SE2V);
    out[5] = (float)(uz * rk * SE2V);
    out[6] = (float)rec[MCR_PX];
    out[7] = (float)rec[MCR_PX + 1];
    out[8] = (float)rec[MCR_PX + 2];
    out[9] = (float)rec[MCR_W];
    out[10] = (float)(rec[MCR_T] * 1e-3);
    out[11] = (float)(mcpl_ctr_base(i, seed) + i);
}

/* write_particle (mcpl_out.py:55-73) with pack_vector (:36-50) and
 * KE_from_v (:52-54).  The reference evaluates the products, the
 * reciprocal and the energy with Python floats (FP64) from the FP32 state
 * and rounds at the write: positions and time scale exactly in FP64, so
 * they round as FP32 products do; the energy is formed in FP64.
 * pack_vector compares signed velocity components and keeps zeros where
 * MCPL would keep a component; `standard` selects MCPL's own packing of
 * the unit direction instead. */
RHD void mcpl_record(const float st[11], int standard, float rec[MCR_N])
{
    float x = st[3], y = st[4], z = st[5];
    rec[MCR_PX] = st[6];
    rec[MCR_PX + 1] = st[7];
    rec[MCR_PX + 2] = st[8];
    rec[MCR_X] = st[0] * 100.0f;
    rec[MCR_X + 1] = st[1] * 100.0f;
    rec[MCR_X + 2] = st[2] * 100.0f;
    double dx = x, dy = y, dz = z;
    double vm2 = (dx * dx + dy * dy) + dz * dz;
    double KE = 5.22703725e-15 * vm2;
    if (standard) {
        double v = sqrt(vm2);
        double p0, p1, s;
        mcpl_pack_dir(dx / v, dy / v, dz / v, &p0, &p1, &s);
        rec[MCR_U] = (float)p0;
        rec[MCR_U + 1] = (float)p1;
        rec[MCR_U + 2] = (float)(s * KE);
    } else {
        double ux = 0.0, uy = 0.0, si = 1.0;
        if (x > y && x > z) {
            ux = 1.0 / dz;
            si = x > 0.0f ? 1.0 : (x < 0.0f ? -1.0 : 0.0);
        } else if (y > x && y > z) {
            ux = dx;
            uy = 1.0 / dz;
            si = y > 0.0f ? 1.0 : (y < 0.0f ? -1.0 : 0.0);
        } else if (z > x && z > y) {
            si = z > 0.0f ? 1.0 : (z < 0.0f ? -1.0 : 0.0);
        }
        rec[MCR_U] = (float)ux;
        rec[MCR_U + 1] = (float)uy;
        rec[MCR_U + 2] = (float)(si * KE);
    }
    rec[MCR_T] = st[10] * 1000.0f;
    rec[MCR_W] = st[9];
}

#ifdef __HIPCC__
DEV void mcpl_in_step(NS &n, const DComp &c, uint32_t seed)
{
    if (n.key == 0xffffffffu) return; /* no history on this lane */
    const DPart &s = c.s;
    const uint32_t i = n.key;
    const bool have = i < s.ni[MCI_N];
    double rec[MCR_N];
#pragma unroll
    for (int k = 0; k < MCR_N; k++) rec[k] = 0.0;
    if (have) {
        /* field-major table: field k of particle i at k * N + i */
        const size_t N = s.ni[MCI_N];
        if (s.ni[MCI_DOUBLE]) {
            const double *r = (const double *)s.tab[MCT_REC] + i;
#pragma unroll
            for (int k = 0; k < MCR_N; k++) rec[k] = r[k * N];
        } else {
            const float *r = s.tab[MCT_REC] + i;
#pragma unroll
            for (int k = 0; k < MCR_N; k++) rec[k] = (double)r[k * N];
        }
    }
    float o[12];
    mcpl_to_state(rec, i, seed, o);
    /* beyond the file's particles every field is zero and the slot dead */
    n.x = have ? o[0] : 0.0f; n.y = have ? o[1] : 0.0f; n.z = have ? o[2] : 0.0f;
    n.vx = have ? o[3] : 0.0f; n.vy = have ? o[4] : 0.0f; n.vz = have ? o[5] : 0.0f;
    n.px = have ? o[6] : 0.0f; n.py = have ? o[7] : 0.0f; n.pz = have ? o[8] : 0.0f;
    n.w = have ? o[9] : 0.0f; n.t = have ? o[10] : 0.0f; n.ctr = have ? o[11] : 0.0f;
    n.sc = n.sd = n.se = 0.0f;
    n.dead = !have;
}

DEV void mcpl_out_step(const NS &n, const DComp &c, float *dump, uint32_t *tags, uint64_t cap,
                       uint32_t tag)
{
    if (n.dead || n.key == 0xffffffffu) return;
    const uint64_t s = c.s.ni[MCO_SLOT];
    const float st[11] = {n.x, n.y, n.z, n.vx, n.vy, n.vz, n.px, n.py, n.pz, n.w, n.t};
    float rec[MCR_N];
    mcpl_record(st, (int)c.s.ni[MCO_STANDARD], rec);
    /* field-major, so that the lanes of a wave store contiguously */
    float *o = dump + (size_t)s * MCR_N * cap + n.key;
    for (int k = 0; k < MCR_N; k++) o[(size_t)k * cap] = rec[k];
    tags[s * cap + n.key] = tag;
}
/* SRescal (scat/rescal.cl:5-77): a scattering point on the path through
 * the sample, a final direction in the cone of half-angle acos(d[0])
 * about the target (rotations about the unnormalised axes d[1..3] =
 * target x target_perp and the target itself, as the reference does), a
 * final energy uniform in E0 +- dE; the event (ki - kf, energy transfer,
 * weight) goes field-major to the component's buffer (DPart.aux) at the
 * history's index.  phi = rand*2*M_PI promotes to double (rescal.cl:40). */
template <class HC>
DEV void s_rescal(NS &n, const DComp &c, HC &hc)
{
    const DPart &s = c.s;
    rf3 p1 = ipos1(n);
    rf3 path = v3_sub(ipos2(n), p1);
    setpos(n, v3_add(p1, v3_scale(drw(n), path)));
    rf3 ki = v3_scale(RC_V2K, nvel(n));
    float theta = rm_acos(1.0f - drw(n) * (1.0f - s.d[0]));
    float phi = (float)((double)(drw(n) * 2.0f) * RC_M_PI_D);
    rf3 tgt = rf3_make(s.p[RS_T], s.p[RS_T + 1], s.p[RS_T + 2]);
    rf3 tf = v3_rotate_axis(theta, rf3_make(s.d[1], s.d[2], s.d[3]), tgt);
    tf = v3_norm(v3_rotate_axis(phi, tgt, tf));
    float Ef = s.p[RS_E0] + s.p[RS_DE] * (1.0f - 2.0f * drw(n));
    float vf = RC_SE2V * rm_sqrt(Ef);
    setvel(n, v3_mulv(tf, vf));
    rf3 kf = v3_scale(RC_V2K, nvel(n));
    float dE = RC_VS2E * rm_pow(v3_len(v3_scale(RC_K2V, ki)), 2.0f) - Ef;
    if (hc.enabled) {
        const uint64_t cap = (uint64_t)s.ni[RSI_CAP_LO] | ((uint64_t)s.ni[RSI_CAP_HI] << 32);
        float *e = (float *)s.aux + n.key;
        e[0] = ki.x - kf.x;
        e[cap] = ki.y - kf.y;
        e[2 * cap] = ki.z - kf.z;
        e[3 * cap] = dE;
        e[4 * cap] = n.w;
    }
    s_done(n, c.comp_idx);
}
#endif

#endif
