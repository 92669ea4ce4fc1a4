/* dev_components.h - per-neutron device functions of every component type,
 * executed in registers inside the fused segment kernels.
 *
 * Semantics are those of the reference kernels (cited per function); the
 * differences are of implementation only:
 *   - the neutron lives in registers for a whole segment instead of being
 *     re-read from a 100-byte AoS record by five launches per component;
 *   - rotation matrices and parameter-only expressions come precomputed
 *     from the plan (same bits as per-neutron evaluation);
 *   - double-precision expression sites of the reference (double literals
 *     and M_PI promote float operands) are evaluated in FP64 as the
 *     reference evaluates them; parameter-only ones once on the host, and
 *     quotients by a parameter-only divisor from its host-computed
 *     reciprocal with one correction step (rd_div_rc, exact).
 */
#ifndef RAMP_DEV_COMPONENTS_H
#define RAMP_DEV_COMPONENTS_H

#include "rmath.h"
#include "rvec.h"
#include "rrng.h"
#include "rconst.h"
#include "rslots.h"
#include "dev_plan.h"
#include "../../include/ramp.h"

#define DEV __device__ static inline __attribute__((always_inline))

/* The plan (components, steps, source, detectors) is written by the host
 * before a launch and only read by the kernels.  Reading it through the
 * constant address space tells the compiler so, and it then loads the
 * wave-uniform plan fields with scalar loads into SGPRs instead of
 * per-lane vector loads into VGPRs. */
template <class T>
DEV const T &plan_at(const T *p, size_t i)
{
    return *(const T *)((const __attribute__((address_space(4))) T *)p + i);
}

/* a pointer known to address global memory (HBM) or the workgroup's LDS:
 * the compiler then emits global or LDS instructions instead of flat ones
 * (which also wait on the other path) */
template <class T>
DEV T *in_global(T *p)
{
    return (T *)(__attribute__((address_space(1))) T *)p;
}

template <class T>
DEV T *in_lds(T *p)
{
    return (T *)(__attribute__((address_space(3))) T *)p;
}

/* a plan table (written before the launch, never by a kernel): loads at
 * wave-uniform indices, as in loops over a line or CDF table, become
 * scalar loads */
template <class T>
DEV const T *in_const(const T *p)
{
    return (const T *)(const __attribute__((address_space(4))) T *)p;
}

/* feature sets compiled into a segment-kernel instance: component types
 * outside the set are compiled out, which keeps register use (occupancy)
 * at what the instrument actually needs */
#define FEAT_MESH 1u    /* GMesh */
#define FEAT_SAMPLE 2u  /* SIsotropic, SPowder1, SPowderN */
#define FEAT_MONO 4u    /* SMonochromator */
#define FEAT_ISIS 8u    /* MISIS */
#define FEAT_NLIN 16u   /* non-linear blocks: their two-sweep step (mcramp.py:167-217)
                         * holds a second copy of every component's code, and
                         * only the full instance carries it */
#define FEAT_ALL 31u
#define DEVNI __device__ static __attribute__((noinline))

struct NS {
    float x, y, z, vx, vy, vz, px, py, pz, w, t, ctr, sc, sd, se;
    float is[8];
    uint32_t iidx, key;
    uint32_t aux;  /* state of a component split at a segment boundary */
    bool dead;
    bool pend;     /* the first half of a split component is done, the second is due */
};

DEV float drw(NS &n) { return rr_draw(&n.ctr, n.key); }

DEV void isect_reset(NS &n)
{
    n.is[0] = 0.0f; n.is[1] = 0.0f; n.is[2] = 0.0f; n.is[3] = RC_SENTINEL;
    n.is[4] = 0.0f; n.is[5] = 0.0f; n.is[6] = 0.0f; n.is[7] = RC_SENTINEL;
}

DEV rf3 npos(const NS &n) { return rf3_make(n.x, n.y, n.z); }
DEV rf3 nvel(const NS &n) { return rf3_make(n.vx, n.vy, n.vz); }
DEV void setpos(NS &n, rf3 v) { n.x = v.x; n.y = v.y; n.z = v.z; }
DEV void setvel(NS &n, rf3 v) { n.vx = v.x; n.vy = v.y; n.vz = v.z; }
DEV rf3 ipos1(const NS &n) { return rf3_make(n.is[0], n.is[1], n.is[2]); }
DEV rf3 ipos2(const NS &n) { return rf3_make(n.is[4], n.is[5], n.is[6]); }
DEV void setp1(NS &n, rf3 v) { n.is[0] = v.x; n.is[1] = v.y; n.is[2] = v.z; }
DEV void setp2(NS &n, rf3 v) { n.is[4] = v.x; n.is[5] = v.y; n.is[6] = v.z; }

/* ---------------- frame transforms (frametransform.cl:4-43) ------------ */
DEV rf3 mat_derot(const float *R, rf3 v)
{
    return rf3_make((R[0] * v.x + R[3] * v.y) + R[6] * v.z,
                    (R[1] * v.x + R[4] * v.y) + R[7] * v.z,
                    (R[2] * v.x + R[5] * v.y) + R[8] * v.z);
}
DEV rf3 mat_rot(const float *R, rf3 v)
{
    return rf3_make((R[0] * v.x + R[1] * v.y) + R[2] * v.z,
                    (R[3] * v.x + R[4] * v.y) + R[5] * v.z,
                    (R[6] * v.x + R[7] * v.y) + R[8] * v.z);
}

/* For the matrix of zero Euler angles (1 on the diagonal, signed zeros
 * elsewhere) every product of an off-diagonal entry is a signed zero, so
 * (x*1 + y*0) + z*0 equals x exactly whenever x, y, z are finite and x is
 * nonzero.  When all six components of a wave's neutrons satisfy this the
 * matrix product is skipped; otherwise it is evaluated in full. */
DEV bool fin_nz(float v) { return v != 0.0f && (v - v) == 0.0f; }
DEV bool ident_ok(const NS &n, float x, float y, float z)
{
    return n.dead || (fin_nz(x) && fin_nz(y) && fin_nz(z) && fin_nz(n.vx) && fin_nz(n.vy) &&
                      fin_nz(n.vz));
}

template <bool FULL>
DEV void tr_in(NS &n, const DComp &c)
{
    rf3 p = rf3_make(n.x - c.pos[0], n.y - c.pos[1], n.z - c.pos[2]);
    if (!FULL && c.pad == 1u && __all(ident_ok(n, p.x, p.y, p.z))) {
        if (n.dead) return;
        setpos(n, p);
        return;
    }
    if (n.dead) return;
    setpos(n, mat_derot(c.R, p));
    setvel(n, mat_derot(c.R, nvel(n)));
    if (FULL) {
        rf3 q = mat_derot(c.R, rf3_make(n.px, n.py, n.pz));
        n.px = q.x; n.py = q.y; n.pz = q.z;
    }
}

template <bool FULL>
DEV void tr_out(NS &n, const DComp &c)
{
    if (!FULL && c.pad == 1u && __all(ident_ok(n, n.x, n.y, n.z))) {
        if (n.dead) return;
        n.x = n.x + c.pos[0]; n.y = n.y + c.pos[1]; n.z = n.z + c.pos[2];
        return;
    }
    if (n.dead) return;
    rf3 p = mat_rot(c.R, npos(n));
    n.x = p.x + c.pos[0]; n.y = p.y + c.pos[1]; n.z = p.z + c.pos[2];
    setvel(n, mat_rot(c.R, nvel(n)));
    if (FULL) {
        rf3 q = mat_rot(c.R, rf3_make(n.px, n.py, n.pz));
        n.px = q.x; n.py = q.y; n.pz = q.z;
    }
}

/* terminator.cl:3-22 */
DEV void terminate(NS &n, uint32_t restore)
{
    if (n.dead) return;
    /* length(pos2) == 0 exactly when its squared norm is 0 */
    if (v3_dot(ipos2(n), ipos2(n)) == 0.0f && restore == 0) n.dead = true;
}

/* ---------------- sources ---------------------------------------------- */
DEV void src_clear(NS &n, int seed, float sb_old)
{
    n.x = n.y = n.z = n.vx = n.vy = n.vz = n.px = n.py = n.pz = 0.0f;
    n.w = n.t = n.sc = n.sd = n.se = 0.0f;
    n.ctr = (float)seed + sb_old;
    n.dead = false;
    n.pend = false;
    n.aux = 0u;
}

/* position on the face and direction to a point of the target window
 * (flat_mod.cl:28-36) */
DEV void src_posdir(NS &n, const DPart &s, float vel)
{
    n.x = s.p[SP_MOD_W] * (0.5f - drw(n));
    n.y = s.p[SP_MOD_H] * (0.5f - drw(n));
    n.z = 0.0f;
    float Dx = s.p[SP_TGT_W] * (0.5f - drw(n)) - n.x;
    float Dy = s.p[SP_TGT_H] * (0.5f - drw(n)) - n.y;
    rf3 dir = v3_norm(rf3_make(Dx, Dy, s.p[SP_TGT_D]));
    setvel(n, v3_scale(vel, dir));
}

/* polarisation on the unit sphere (flat_mod.cl:39-43); the two draws are
 * always consumed, the vector is formed only in full-state mode */
template <bool FULL>
DEV void src_pol(NS &n, const DPart &s)
{
    float u1 = drw(n);
    float u2 = drw(n);
    if (FULL) {
        float theta = (float)(2 * RC_M_PI_D * (double)u1);
        float phi = rm_acos(1 - 2 * u2);
        n.px = rm_sin(phi) * rm_cos(theta);
        n.py = rm_sin(phi) * rm_sin(theta);
        n.pz = rm_cos(phi);
    }
}

/* maxwell_energy_distn (cont_mod.cl:4-9) for a per-T constant set:
 * overKT = d[i], pow(overKT,1.5) = d[3+i]; the reference's double chain
 * 2.0f*rootE/sqrt(M_PI)*pow(..)*exp(..) is evaluated in double, the
 * quotient exactly from the host's sqrt(M_PI) and its reciprocal */
DEV float maxwell_dev(float E, float overKT, float pf, double sqpi, double rsqpi)
{
    float rootE = rm_sqrt(E);
    double r = rd_div_rc((double)(2.0f * rootE), sqpi, rsqpi);
    r = r * (double)pf;
    r = r * (double)rm_exp(-E * overKT);
    return (float)r;
}

template <bool FULL, uint32_t F>
DEV void src_generate(NS &n, const DPart &s, int num_sim, int seed, float sb_old)
{
    src_clear(n, seed, sb_old);
    switch (s.type) {
    case RAMP_SRC_FLAT: {
        float E_range = s.p[SP_EMAX] - s.p[SP_EMIN];
        float E_val = s.p[SP_EMIN] + E_range * drw(n);
        n.w = 1.0f / (float)num_sim;
        src_posdir(n, s, RC_SE2V * rm_sqrt(E_val));
        src_pol<FULL>(n, s);
        break;
    }
    case RAMP_SRC_CONT: {
        float E_range = s.p[SP_EMAX] - s.p[SP_EMIN];
        float E_val = s.p[SP_EMIN] + E_range * drw(n);
        float E_J = RC_MEV2J * E_val;
        const double sqpi = dp_f64(s.d, 18), rsqpi = dp_f64(s.d, 20);
        float term1;
        if (s.p[SP_TAIL_JOIN] > 0.0f && E_J > s.d[9])
            term1 = s.p[SP_I1] * (s.d[10] * (s.d[9] / E_J)) / s.d[6];
        else
            term1 = s.p[SP_I1] * maxwell_dev(E_J, s.d[0], s.d[3], sqpi, rsqpi) / s.d[6];
        float M = term1;
        /* d[16]/d[17]: the term has I = 0 and a finite Maxwellian value, so
         * adding it leaves M unchanged and it is not evaluated */
        if (s.d[16] == 0.0f) M += s.p[SP_I2] * maxwell_dev(E_J, s.d[1], s.d[4], sqpi, rsqpi) / s.d[7];
        if (s.d[17] == 0.0f) M += s.p[SP_I3] * maxwell_dev(E_J, s.d[2], s.d[5], sqpi, rsqpi) / s.d[8];
        n.w = M / (float)num_sim;
        src_posdir(n, s, RC_SE2V * rm_sqrt(E_val));
        src_pol<FULL>(n, s);
        break;
    }
    case RAMP_SRC_ISIS: if constexpr ((F & FEAT_ISIS) != 0) {
        /* isis_mod.cl:15-105 (scans bounded by the table, index clamped) */
        const float *flux = in_const(s.tab[ST_ISIS_FLUX]), *tb = in_const(s.tab[ST_ISIS_TBINS]);
        const float *eb = in_const(s.tab[ST_ISIS_EBINS]), *Eint = in_const(s.tab[ST_ISIS_EINT]);
        int nt = (int)s.ni[SI_ISIS_NT], ne = (int)s.ni[SI_ISIS_NE];
        int nflux = (int)s.tab_len[ST_ISIS_FLUX];
        float deviate = s.p[SP_ISIS_TOTAL] * drw(n);
        int lim = nt * ne < nflux ? nt * ne : nflux;
        int idx = 0;
        for (int i = 0; i < lim; i++)
            if ((flux[i] - deviate) > 0.0f) { idx = i - 1; break; }
        int Epnt = 0;
        for (int i = 0; i < ne; i++)
            if ((Eint[i] - deviate) > 0.0f) { Epnt = i - 1; break; }
        int Tpnt = (int)rm_fmod((float)idx, (float)nt);
        if (Tpnt >= nt - 1) Tpnt -= 1;
        if (Tpnt < 0) Tpnt = 0;
        float time_val = tb[Tpnt];
        float time_range = tb[Tpnt + 1] - tb[Tpnt];
        float time_spread = flux[Epnt * nt + Tpnt + 1] - flux[Epnt * nt + Tpnt];
        float R = deviate - flux[Epnt * nt + Tpnt];
        R /= time_spread;
        time_val += time_range * R;
        float ener_spread = Eint[Epnt + 1] - Eint[Epnt];
        int is = (Epnt > 3) ? Epnt - 3 : 0;
        int ie = ((ne - Epnt - 1) > 3) ? Epnt + 3 : (ne - 1);
        deviate = Eint[Epnt] + ener_spread * drw(n);
        float acc = 0.0f;
        for (int i = is; i <= ie; i++) {
            float Pj = 1.0f;
            for (int j = is; j <= ie; j++) {
                if (j != i) Pj *= (deviate - Eint[j]) / (Eint[i] - Eint[j]);
                else Pj *= 1;
            }
            Pj *= eb[i];
            acc += Pj;
        }
        src_posdir(n, s, RC_SE2V * rm_sqrt(acc));
        src_pol<FULL>(n, s);
        n.w = s.p[SP_ISIS_STR_AREA] * s.p[SP_ISIS_TOTAL] * RC_ISIS_WSCALE / (float)num_sim;
        n.t = time_val;
        break;
    } else { break; }
    case RAMP_SRC_LAMBDA: {
        float lam = s.p[SP_LMIN] + (s.p[SP_LMAX] - s.p[SP_LMIN]) * drw(n);
        n.w = (1.0f / lam) * s.p[SP_LNORM] / (float)num_sim;
        src_posdir(n, s, RC_K2V * (6.28318530717958647692f / lam));
        src_pol<FULL>(n, s);
        break;
    }
    case RAMP_SRC_PARALLEL: {
        float lam = s.p[SP_LMIN] + (s.p[SP_LMAX] - s.p[SP_LMIN]) * drw(n);
        n.w = 1.0f / (float)num_sim;
        n.x = s.p[SP_MOD_W] * (0.5f - drw(n));
        n.y = s.p[SP_MOD_H] * (0.5f - drw(n));
        n.z = 0.0f;
        n.vx = 0.0f;
        n.vy = 0.0f;
        n.vz = RC_K2V * (6.28318530717958647692f / lam);
        src_pol<FULL>(n, s);
        break;
    }
    default:
        n.dead = true;
        break;
    }
    isect_reset(n);
}

/* ---------------- geometry --------------------------------------------- */
DEV float ocl_min(float x, float y) { return (y < x) ? y : x; }
DEV float ocl_max(float x, float y) { return (x < y) ? y : x; }

/* geom/plane.cl:3-53 */
DEV void g_plane(NS &n, const DPart &g, uint32_t ci)
{
    float t = 0.0f, x = 0.0f, y = 0.0f;
    uint32_t o = g.ni[GI_ORIENT];
    if (o == 0) {
        t = (-n.z) / n.vz;
        x = n.x + t * n.vx;
        y = n.y + t * n.vy;
    } else if (o == 1) {
        t = (-n.x) / n.vx;
        x = n.z + t * n.vz;
        y = n.y + t * n.vy;
    }
    rf3 vel = nvel(n);
    if ((rm_fabs(x) < g.d[0]) && (rm_fabs(y) < g.d[1]) && t < n.is[3] && t < n.is[7] &&
        t > 0.0f && v3_dot(vel, rf3_make(0.0f, 0.0f, 1.0f)) > 0.0f) {
        rf3 h = v3_add(npos(n), v3_scale(t, vel));
        setp1(n, h);
        setp2(n, h);
        n.is[3] = t;
        n.is[7] = t;
        n.iidx = ci;
    }
}

/* geom/box.cl:3-54 (the kernel writes its private intersection record,
 * zeros on a miss) */
DEV void g_box(NS &n, const DPart &g, uint32_t ci)
{
    rf3 pos = npos(n), vel = nvel(n);
    float t1 = (g.d[0] - pos.x) / vel.x, t2 = (g.d[3] - pos.x) / vel.x;
    float t3 = (g.d[1] - pos.y) / vel.y, t4 = (g.d[4] - pos.y) / vel.y;
    float t5 = (g.d[2] - pos.z) / vel.z, t6 = (g.d[5] - pos.z) / vel.z;
    float tmin = ocl_max(ocl_max(ocl_min(t1, t2), ocl_min(t3, t4)), ocl_min(t5, t6));
    float tmax = ocl_min(ocl_min(ocl_max(t1, t2), ocl_max(t3, t4)), ocl_max(t5, t6));
    if (tmax < 0.0f || tmin > tmax) {
        for (int i = 0; i < 8; i++) n.is[i] = 0.0f;
    } else {
        setp1(n, v3_add(pos, v3_scale(tmin, vel)));
        n.is[3] = tmin;
        setp2(n, v3_add(pos, v3_scale(tmax, vel)));
        n.is[7] = tmax;
        n.iidx = ci;
    }
}

/* geom/sphere.cl:3-44, geom/sphere_int.cl:3-43 */
DEV void g_sphere(NS &n, const DPart &g, uint32_t ci, bool interior)
{
    rf3 pos = npos(n), vel = nvel(n);
    float r = g.p[GP_R];
    float a = v3_dot(vel, vel);
    float b = 2.0f * v3_dot(vel, pos);
    float c = v3_dot(pos, pos) - r * r;
    float q = b * b - 4.0f * a * c;
    if (q > 0.0f) {
        float sq = rm_sqrt(q);
        float d1 = (-b - sq) / (2.f * a);
        float d2 = (-b + sq) / (2.f * a);
        if (!interior) {
            if (d1 < n.is[3] && d1 > 0.f) {
                setp1(n, v3_add(pos, v3_scale(d1, vel)));
                n.is[3] = d1;
                n.iidx = ci;
                setp2(n, v3_add(pos, v3_scale(d2, vel)));
                n.is[7] = d2;
            }
        } else if (d1 < 0.0f && d2 < n.is[3]) {
            rf3 h = v3_add(pos, v3_scale(d2, vel));
            setp1(n, h);
            n.is[3] = d2;
            n.iidx = ci;
            setp2(n, h);
            n.is[7] = d2;
        }
    }
}

/* geom/cylinder_ext.cl:3-53 and geom/banana.cl:3-55.  banana d[0]: 0 the
 * chained angle test is always true, 1 always false, 2 evaluate it */
DEV void g_cyl(NS &n, const DPart &g, uint32_t ci, bool banana)
{
    float r = g.p[GP_R], hgt = g.p[GP_HEIGHT];
    float a = n.vx * n.vx + n.vz * n.vz;
    float b = 2.0f * (n.x * n.vx + n.z * n.vz);
    float c = (n.x * n.x + n.z * n.z) - r * r;
    float q = b * b - 4.0f * a * c;
    float sq = rm_sqrt(q);
    float t1 = (-b - sq) / (2.f * a);
    float t2 = (-b + sq) / (2.f * a);
    float yh = n.y + t2 * n.vy;
    bool cond = (q > 0.0f) && ((-hgt / 2.0f) < yh) && (yh < (hgt / 2));
    if (banana && cond) {
        float mode = g.d[0];
        if (mode == 1.0f) cond = false;
        else if (mode == 2.0f) {
            rf2 e = v2_norm(rf2_make(n.x + t2 * n.vx, n.z + t2 * n.vz));
            float theta2 = rm_acos(v2_dot(e, rf2_make(0.0f, 1.0f)));
            float first = (g.p[GP_MINTH] < theta2) ? 1.0f : 0.0f;
            cond = first < g.p[GP_MAXTH];
        }
    }
    if (!cond) return;
    bool hit = banana ? (t2 < n.is[3] && t1 < 0.0f && t2 > 0.0f)
                      : (t1 < n.is[3] && t1 > 0.0f && t2 > 0.0f);
    if (hit) {
        rf3 pos = npos(n), vel = nvel(n);
        setp1(n, v3_add(pos, v3_scale(t1, vel)));
        n.is[3] = t1;
        n.iidx = ci;
        setp2(n, v3_add(pos, v3_scale(t2, vel)));
        n.is[7] = t2;
    }
}

#include "dev_mesh.h"

template <uint32_t F>
DEV void geom_dispatch(NS &n, const DComp &c)
{
    if (n.dead) return;
    const DPart &g = c.g;
    switch (g.type) {
    case RAMP_G_ARM:
        n.is[4] = 1.0f; n.is[5] = 1.0f; n.is[6] = 1.0f;
        n.iidx = c.comp_idx;
        break;
    case RAMP_G_PLANE: g_plane(n, g, c.comp_idx); break;
    case RAMP_G_BOX: g_box(n, g, c.comp_idx); break;
    case RAMP_G_SPHERE: g_sphere(n, g, c.comp_idx, false); break;
    case RAMP_G_SPHERE_INT: g_sphere(n, g, c.comp_idx, true); break;
    case RAMP_G_CYLINDER: g_cyl(n, g, c.comp_idx, false); break;
    case RAMP_G_BANANA: g_cyl(n, g, c.comp_idx, true); break;
    case RAMP_G_MESH:
        if constexpr ((F & FEAT_MESH) != 0) g_mesh(n, g, c.comp_idx);
        break;
    default: break;
    }
}

/* ---------------- detectors -------------------------------------------- */
/* per-lane event tallies are kept in registers and reduced once per wave
 * at the end of a segment kernel (no same-address atomics on the hot path) */
struct HistCtx {
    static constexpr bool adjoint = false;
    const DDet *dets;
    int mode;       /* RAMP_HIST_* */
    int enabled;    /* 0 in pilot runs */
    uint32_t events, oob, overflow;
    /* the workgroup's LDS copy of the small detectors (null: global
     * atomics only): (w, w^2) per bin, float or fixed point by mode */
    void *acc;
};

DEV void hist_add(HistCtx &hc, int det, uint64_t idx, float w)
{
    if (!hc.enabled) return;
    const DDet &d = plan_at(hc.dets, det);
    if (idx >= d.n) {
        hc.oob++;
        return;
    }
    float w2 = w * w;
    if (hc.mode == RAMP_HIST_DETERMINISTIC) {
        float a = rm_rint(rm_ldexp(w, d.sh_w));
        float b = rm_rint(rm_ldexp(w2, d.sh_w2));
        if (!(rm_fabs(a) < 0x1p62f) || !(rm_fabs(b) < 0x1p62f)) {
            hc.overflow++;
            return;
        }
        if (d.acc >= 0 && hc.acc) {
            unsigned long long *q = in_lds((unsigned long long *)hc.acc + 2 * ((uint64_t)d.acc + idx));
            atomicAdd(q, (unsigned long long)(long long)a);
            atomicAdd(q + 1, (unsigned long long)(long long)b);
        } else {
            atomicAdd(in_global(&d.qw[idx]), (unsigned long long)(long long)a);
            atomicAdd(in_global(&d.qw2[idx]), (unsigned long long)(long long)b);
        }
    } else if (d.acc >= 0 && hc.acc) {
        float *f = in_lds((float *)hc.acc + 2 * ((uint64_t)d.acc + idx));
        atomicAdd(f, w);
        atomicAdd(f + 1, w2);
    } else {
        unsafeAtomicAdd(in_global(&d.w[idx]), w);
        unsafeAtomicAdd(in_global(&d.w2[idx]), w2);
    }
    hc.events++;
}

/* start and end of a workgroup's LDS accumulation (all threads call both):
 * the copy is zeroed, and at the end every nonzero bin is added to HBM */
DEV void hist_acc_begin(HistCtx &hc, void *lds, uint32_t bytes)
{
    for (uint32_t i = threadIdx.x; i < bytes / 4; i += blockDim.x) ((uint32_t *)lds)[i] = 0u;
    __syncthreads();
    hc.acc = lds;
}

DEV void hist_acc_end(const HistCtx &hc, int n_det, uint32_t bins)
{
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < bins; i += blockDim.x) {
        for (int k = 0; k < n_det; k++) {
            const DDet &d = plan_at(hc.dets, k);
            if (d.acc < 0 || i < (uint32_t)d.acc || i >= (uint32_t)d.acc + d.n) continue;
            const uint64_t b = i - (uint32_t)d.acc;
            if (hc.mode == RAMP_HIST_DETERMINISTIC) {
                const unsigned long long *q = in_lds((const unsigned long long *)hc.acc + 2 * (uint64_t)i);
                if (q[0]) atomicAdd(in_global(&d.qw[b]), q[0]);
                if (q[1]) atomicAdd(in_global(&d.qw2[b]), q[1]);
            } else {
                const float *f = in_lds((const float *)hc.acc + 2 * (uint64_t)i);
                if (f[0] != 0.0f) unsafeAtomicAdd(in_global(&d.w[b]), f[0]);
                if (f[1] != 0.0f) unsafeAtomicAdd(in_global(&d.w2[b]), f[1]);
            }
            break;
        }
    }
}

#include "dev_adjoint.h"

/* wavelength variable: 2*M_PI/(V2K*|v|), a double quotient in the
 * reference, evaluated in double (d[0..1]: 2*M_PI as a double) */
DEV float det_wavelength(const NS &n, const DPart &s)
{
    float x = RC_V2K * v3_len(nvel(n));
    return (float)(dp_f64(s.d, 0) / (double)x);
}

DEV void s_done(NS &n, uint32_t ci)
{
    n.iidx = 0;
    n.sc = (float)ci;
    isect_reset(n);
}

DEV void det_finish(NS &n, uint32_t restore)
{
    if (restore == 0) {
        setpos(n, ipos2(n));
        n.t += n.is[7];
        n.dead = true;
    }
}

/* scat/detector1D.cl:20-77 */
template <class HC>
DEV void s_det1d(NS &n, const DComp &c, HC &hc)
{
    const DPart &s = c.s;
    float mn = s.p[D1_MIN], st = s.p[D1_STEP], mx = s.p[D1_MAX];
    uint32_t var = s.ni[D1I_VAR];
    float v = 0.0f;
    if (var == 0) v = RC_VS2E * rm_pow(v3_len(nvel(n)), 2.0f);
    else if (var == 1) v = rm_degrees(rm_atan2(n.is[4], n.is[6]));
    else if (var == 2) v = (1.0e6f) * (n.t + n.is[7]);
    else if (var == 3) v = det_wavelength(n, s);
    if (mn <= v && v <= mx) {
        uint32_t idx = (uint32_t)rm_floor((v - mn) / st);
        hist_add(hc, c.det, idx, n.w);
    }
    det_finish(n, s.ni[D1I_RESTORE]);
    s_done(n, c.comp_idx);
}

DEV float det2d_var(const NS &n, const DPart &s, uint32_t var, int axis)
{
    float x2 = n.is[4], y2 = n.is[5], z2 = n.is[6];
    switch (var) {
    case 0: return x2;
    case 1: return y2;
    case 2: return rm_degrees(rm_atan2(x2, z2));
    case 3: return rm_degrees(rm_atan2(y2, rm_sqrt(x2 * x2 + z2 * z2)));
    case 4: return rm_sign(y2) * rm_degrees(rm_acos(x2 / rm_sqrt(x2 * x2 + y2 * y2)));
    case 5: return (1.0e6f) * (n.t + n.is[7]);
    case 6: return rm_degrees(rm_atan2(n.vx, axis == 1 ? y2 : n.vz));
    case 7: return rm_degrees(rm_atan2(n.vy, axis == 1 ? y2 : n.vz));
    case 8: return det_wavelength(n, s);
    case 9: return RC_VS2E * rm_pow(v3_len(nvel(n)), 2.0f);
    case 10: return rm_sign(x2) * rm_degrees(rm_atan2(rm_sqrt(x2 * x2 + y2 * y2), z2));
    default: return 0.0f;
    }
}

DEV int det2d_idx(float v, float mn, float st, float mx)
{
    if (mn <= v && v <= mx) return (int)rm_round((v - mn) / st);
    return -1;
}

/* scat/detector2D.cl:31-186 */
template <class HC>
DEV void s_det2d(NS &n, const DComp &c, HC &hc)
{
    const DPart &s = c.s;
    float v1 = det2d_var(n, s, s.ni[D2I_VAR1], 1);
    int a1 = det2d_idx(v1, s.p[D2_MIN1], s.p[D2_STEP1], s.p[D2_MAX1]);
    float v2 = det2d_var(n, s, s.ni[D2I_VAR2], 2);
    int a2 = det2d_idx(v2, s.p[D2_MIN2], s.p[D2_STEP2], s.p[D2_MAX2]);
    if (!((a1 == -1) || (a2 == -1))) {
        int flat = a1 * (int)s.ni[D2I_N2] + a2;
        hist_add(hc, c.det, (uint64_t)(uint32_t)flat, n.w);
    }
    det_finish(n, s.ni[D2I_RESTORE]);
    s_done(n, c.comp_idx);
}

/* tube x pixel x TOF bank (no reference counterpart) */
template <class HC>
DEV void s_detpix(NS &n, const DComp &c, HC &hc)
{
    const DPart &s = c.s;
    uint32_t nt = s.ni[DTI_NTUBE], np = s.ni[DTI_NPIX], nf = s.ni[DTI_NTOF];
    float th = rm_degrees(rm_atan2(n.is[4], n.is[6]));
    float tof = (1.0e6f) * (n.t + n.is[7]);
    float ft = (th - s.p[DT_THMIN]) / (s.p[DT_THMAX] - s.p[DT_THMIN]) * (float)nt;
    float fp = (n.is[5] - s.p[DT_YMIN]) / (s.p[DT_YMAX] - s.p[DT_YMIN]) * (float)np;
    float ff = (tof - s.p[DT_TMIN]) / (s.p[DT_TMAX] - s.p[DT_TMIN]) * (float)nf;
    if (ft >= 0.0f && ft < (float)nt && fp >= 0.0f && fp < (float)np && ff >= 0.0f &&
        ff < (float)nf) {
        uint64_t it = (uint64_t)ft, ip = (uint64_t)fp, iff = (uint64_t)ff;
        hist_add(hc, c.det, (it * np + ip) * nf + iff, n.w);
    }
    det_finish(n, s.ni[DTI_RESTORE]);
    s_done(n, c.comp_idx);
}

/* scat/counter.cl:18-49 */
template <class HC>
DEV void s_counter(NS &n, const DComp &c, HC &hc)
{
    hist_add(hc, c.det, 0, n.w);
    n.se = 1.0f;
    s_done(n, c.comp_idx);
}

/* ---------------- beam components -------------------------------------- */

/* scat/chopper.cl:4-59; d[0] = Tg (parameter-only double expression) */
DEV void s_chopper(NS &n, const DComp &c)
{
    const DPart &s = c.s;
    setpos(n, ipos2(n));
    n.t += n.is[7];
    float Tg = s.d[0];
    /* chopper.cl:35 draws even without jitter.  With jitter = 0 the phase
     * error is a signed zero, and no sign of zero can change the result
     * (it is added to a nonzero value or vanishes under the fabs below), so
     * only the draw's counter step is kept: the stream stays draw-for-draw. */
    float pha_error = 0.0f;
    if (s.p[CP_JITTER] != 0.0f) pha_error = s.p[CP_JITTER] * 2.0f * (drw(n) - 0.5f);
    else n.ctr = n.ctr + 1.0f;
    float thi_chop = s.p[CP_FREQ] * (n.t - rm_fabs(s.p[CP_PHASE])) + pha_error;
    float thi_neut = rm_atan2(n.x, n.y + s.p[CP_RADIUS]);
    float thi_diff = rm_fabs(thi_chop - thi_neut);
    float ang_diff = rm_fmod(thi_diff, Tg);
    if (rm_fabs(ang_diff - Tg) < ang_diff) ang_diff = rm_fabs(ang_diff - Tg);
    float width = rm_atan2(s.p[CP_SLIT_W], n.y + s.p[CP_RADIUS]);
    if (ang_diff > width / 2.0f) n.dead = true;
    s_done(n, c.comp_idx);
}

/* reflectivity_func (include/ref.h:3-54) with the parameter-only part
 * precomputed: d[4] m, d[5] W, d[6] alpha, d[7] beta, d[14] Qc (after the
 * m<1 rescale), d[15] m (after it), d[16] 1 when the parameters force R=0.
 * The reference's double chain R0*0.5*(1-tanh)*(poly) is evaluated in
 * double. */
DEV float guide_refl(const DPart &s, float q)
{
    float m = s.d[4], W = s.d[5], alpha = s.d[6], beta = s.d[7];
    float Qc = s.p[GU_QC], R0 = s.p[GU_R0];
    float arg = W > 0 ? (q - m * Qc) / W : 11;
    if (arg > 10 || s.d[16] != 0.0f) return 0.0f;
    float Qc2 = s.d[14];
    if (q <= Qc2) return R0;
    double R = (double)R0 * 0.5 * (double)(1 - rm_tanh(arg));
    return (float)(R * (double)(1 - alpha * (q - Qc2) + beta * (q - Qc2) * (q - Qc2)));
}

/* scat/guide.cl:9-150.  d[0..3] ww, hh, whalf, hhalf; d[8] the float
 * threshold equivalent to refl < 1e-12 (double); d[9] -2*V2K;
 * d[10]/d[11] sqrt(l*l+ww*ww), sqrt(l*l+hh*hh); d[12]/d[13] the nlen2s;
 * d[18..19]/d[20..21] the reciprocals of the two roots, as doubles (guides
 * of positive length) */
DEV void s_guide(NS &n, const DComp &c)
{
    const DPart &s = c.s;
    float l = s.p[GU_L];
    float ww = s.d[0], hh = s.d[1], whalf = s.d[2], hhalf = s.d[3];
    /* wall constants of the two orientations, loaded once (the per-lane
     * choice between them is a select) */
    const float sq_v = s.d[10], sq_h = s.d[11], n2_v = s.d[12], n2_h = s.d[13];
    const double rsq_v = dp_f64(s.d, 18), rsq_h = dp_f64(s.d, 20);
    const double k2 = (double)s.d[9];
    setpos(n, ipos2(n));
    n.t += n.is[7];
    uint32_t max_bounces = s.ni[GUI_MAXB];
    uint32_t i = 0;
    while (i < max_bounces) {
        float av = l * n.vx, bv = ww * n.vz;
        float ah = l * n.vy, bh = hh * n.vz;
        float vdotn_v1 = bv + av, vdotn_v2 = bv - av;
        float vdotn_h1 = bh + ah, vdotn_h2 = bh - ah;
        float cv1 = -whalf * l - n.z * ww, cv2 = n.x * l;
        float ch1 = -hhalf * l - n.z * hh, ch2 = n.y * l;
        float t1 = (l - n.z) / n.vz, t2;
        i = 0;
        /* the reference tests walls 1..4 in order, each when vdotn < 0
         * (guide.cl:67-86).  Of a pair of opposite walls at most one faces
         * the neutron unless the guide narrows, so a pair tests its first
         * facing wall, and a second division runs only for the neutrons
         * facing both walls; the order of the tests is unchanged, and a wave
         * holding both kinds of neutrons runs two divisions per pair, not
         * three. */
        {
            const bool f1 = vdotn_v1 < 0.0f, f2 = vdotn_v2 < 0.0f;
            if (f1 || f2) {
                if ((t2 = (f1 ? cv1 - cv2 : cv1 + cv2) / (f1 ? vdotn_v1 : vdotn_v2)) < t1) {
                    t1 = t2;
                    i = f1 ? 1 : 2;
                }
                if (f1 && f2 && (t2 = (cv1 + cv2) / vdotn_v2) < t1) { t1 = t2; i = 2; }
            }
        }
        {
            const bool f3 = vdotn_h1 < 0.0f, f4 = vdotn_h2 < 0.0f;
            if (f3 || f4) {
                if ((t2 = (f3 ? ch1 - ch2 : ch1 + ch2) / (f3 ? vdotn_h1 : vdotn_h2)) < t1) {
                    t1 = t2;
                    i = f3 ? 3 : 4;
                }
                if (f3 && f4 && (t2 = (ch1 + ch2) / vdotn_h2) < t1) { t1 = t2; i = 4; }
            }
        }
        if (i == 0) break;
        rf3 vel = nvel(n);
        n.x = n.x + t1 * vel.x;
        n.y = n.y + t1 * vel.y;
        n.z = n.z + t1 * vel.z;
        n.t += t1;
        float vdn = (i == 1) ? vdotn_v1 : (i == 2) ? vdotn_v2 : (i == 3) ? vdotn_h1 : vdotn_h2;
        bool vert = (i <= 2);
        /* q = V2K*(-2.0)*vdotn/sqrt(nlen2) is a double expression in the
         * reference: the product is exact in double and the quotient is
         * formed in double from the host's reciprocal of the root, then
         * rounded to float.  d = 2.0*vdotn/nlen2 is a double quotient of two
         * floats, whose float rounding equals the float quotient's. */
        float q = (float)rd_div_rc(k2 * (double)vdn, (double)(vert ? sq_v : sq_h), vert ? rsq_v : rsq_h);
        float d = (2.0f * vdn) / (vert ? n2_v : n2_h);
        if (i == 1) { n.vx = n.vx - d * l; n.vz = n.vz - d * ww; }
        else if (i == 2) { n.vx = n.vx + d * l; n.vz = n.vz - d * ww; }
        else if (i == 3) { n.vy = n.vy - d * l; n.vz = n.vz - d * hh; }
        else { n.vy = n.vy + d * l; n.vz = n.vz - d * hh; }
        float refl = guide_refl(s, q);
        n.w *= refl;
        if (refl < s.d[8] || refl > 1.0f) n.w = 0.0f;
        i++;
    }
    if (i >= max_bounces) n.dead = true;
    s_done(n, c.comp_idx);
}

/* path through a sample (isotropic.cl:127-130, powderN.cl:35-39) */
DEV rf3 sample_path(const NS &n)
{
    if (v3_eq(ipos1(n), ipos2(n))) return v3_sub(ipos2(n), npos(n));
    return v3_sub(ipos2(n), ipos1(n));
}

/* constant-Sigma incoherent sample (no reference counterpart) */
DEV void s_incoherent(NS &n, const DComp &c)
{
    const DPart &s = c.s;
    rf3 path = sample_path(n);
    float L = v3_len(path);
    float v = v3_len(nvel(n));
    float lam = 6.28318530717958647692f / (RC_V2K * v);
    float sig_s = s.p[IN_SIGMA_S] * 100.0f;
    float sig_a = s.p[IN_SIGMA_A] * (lam / s.p[IN_LREF]) * 100.0f;
    float sig = sig_s + sig_a;
    float p_int = 1.0f - rm_exp(-sig * L);
    n.w *= p_int * (sig_s / sig);
    float depth = -rm_log(1.0f - drw(n) * p_int) / sig;
    float frac = depth / L;
    bool inside = v3_eq(ipos1(n), ipos2(n));
    rf3 start = inside ? npos(n) : ipos1(n);
    float t_start = inside ? 0.0f : n.is[3];
    setpos(n, v3_add(start, v3_scale(frac, path)));
    n.t += t_start + depth / v;
    if (s.p[IN_FOCUS_R] > 0.0f) {
        /* uniform in the cone around u (sample frame), weight by its solid angle */
        const float *P = s.p;
        float ct = 1.0f - drw(n) * P[IN_OMC];
        float ph = 6.28318530717958647692f * drw(n);
        float st = rm_sqrt(rm_fmax(0.0f, 1.0f - ct * ct));
        float cp = rm_cos(ph), sp = rm_sin(ph);
        float dx = ct * P[IN_U + 0] + st * (cp * P[IN_E1 + 0] + sp * P[IN_E2 + 0]);
        float dy = ct * P[IN_U + 1] + st * (cp * P[IN_E1 + 1] + sp * P[IN_E2 + 1]);
        float dz = ct * P[IN_U + 2] + st * (cp * P[IN_E1 + 2] + sp * P[IN_E2 + 2]);
        n.vx = v * dx;
        n.vy = v * dy;
        n.vz = v * dz;
        n.w *= P[IN_WF];
    } else {
        float ct = 1.0f - 2.0f * drw(n);
        float ph = 6.28318530717958647692f * drw(n);
        float st = rm_sqrt(rm_fmax(0.0f, 1.0f - ct * ct));
        n.vx = v * (st * rm_cos(ph));
        n.vy = v * (st * rm_sin(ph));
        n.vz = v * ct;
    }
    s_done(n, c.comp_idx);
}

/* Beer-Lambert absorber (no reference counterpart) */
DEV void s_absorber(NS &n, const DComp &c)
{
    rf3 path = sample_path(n);
    float L = v3_len(path);
    n.w *= rm_exp(-(c.s.p[AB_SIGMA_A] * 100.0f) * L);
    setpos(n, ipos2(n));
    n.t += n.is[7];
    s_done(n, c.comp_idx);
}

#include "dev_samples.h"
#include "dev_mcpl.h"

/* scat/polariser.cl:3-32: the polarisation becomes the parameter vector */
DEV void s_polariser(NS &n, const DComp &c)
{
    n.px = c.s.p[PO_P]; n.py = c.s.p[PO_P + 1]; n.pz = c.s.p[PO_P + 2];
    s_done(n, c.comp_idx);
}

/* scat/analyser.cl:1-31: w *= (1 + P.pol)/2, then pol = P */
DEV void s_analyser(NS &n, const DComp &c)
{
    rf3 P = rf3_make(c.s.p[PO_P], c.s.p[PO_P + 1], c.s.p[PO_P + 2]);
    n.w *= (1.0f + v3_dot(P, rf3_make(n.px, n.py, n.pz))) / 2.0f;
    n.px = P.x; n.py = P.y; n.pz = P.z;
    s_done(n, c.comp_idx);
}

/* scat/collimator_lin.cl:3-55: kill above the slope, else attenuate
 * linearly; the vertical test runs even after the horizontal one killed */
DEV void s_collimator(NS &n, const DComp &c)
{
    const DPart &s = c.s;
    float sh = s.p[CO_SLOPE_H], sv = s.p[CO_SLOPE_V], tr = s.p[CO_TRANS];
    if (sh > 0.0f) {
        float phi = rm_fabs(n.vx / n.vz);
        if (phi > sh) n.dead = true;
        else n.w *= tr * (1.0f - phi / sh);
    }
    if (sv > 0.0f) {
        float phi = rm_fabs(n.vy / n.vz);
        if (phi > sv) n.dead = true;
        else n.w *= tr * (1.0f - phi / sv);
    }
    s_done(n, c.comp_idx);
}

template <bool FULL, uint32_t F, class HC>
DEV void scat_dispatch(NS &n, const DComp &c, HC &hc)
{
    if (n.iidx != c.comp_idx) return;
    if (n.dead) return;
    switch (c.s.type) {
    case RAMP_S_ARM:
        n.iidx = c.comp_idx;
        isect_reset(n);
        break;
    case RAMP_S_BEAMSTOP:
        n.dead = true;
        s_done(n, c.comp_idx);
        break;
    case RAMP_S_CHOPPER: s_chopper(n, c); break;
    case RAMP_S_GUIDE: s_guide(n, c); break;
    case RAMP_S_ISOTROPIC:
        if constexpr ((F & FEAT_SAMPLE) != 0) s_isotropic<FULL>(n, c);
        break;
    case RAMP_S_POWDER1:
        if constexpr ((F & FEAT_SAMPLE) != 0) s_powder1(n, c);
        break;
    case RAMP_S_POWDERN:
        if constexpr ((F & FEAT_SAMPLE) != 0) s_powderN(n, c);
        break;
    case RAMP_S_MONO:
        if constexpr ((F & FEAT_MONO) != 0) s_mono(n, c);
        break;
    case RAMP_S_DET1D: s_det1d(n, c, hc); break;
    case RAMP_S_DET2D: s_det2d(n, c, hc); break;
    case RAMP_S_COUNTER: s_counter(n, c, hc); break;
    case RAMP_S_INCOHERENT: s_incoherent(n, c); break;
    case RAMP_S_ABSORBER: s_absorber(n, c); break;
    case RAMP_S_DETPIX: s_detpix(n, c, hc); break;
    case RAMP_S_POWDERDIFF:
        if constexpr ((F & FEAT_SAMPLE) != 0) s_powderdiff(n, c, hc);
        break;
    case RAMP_S_RESCAL:
        /* plans holding SRescal run in full-state mode only */
        if constexpr (FULL) s_rescal(n, c, hc);
        break;
    case RAMP_S_POLARISER: s_polariser(n, c); break;
    case RAMP_S_ANALYSER:
        /* plans holding an analyser run in full-state mode (it reads pol) */
        if constexpr (FULL) s_analyser(n, c);
        break;
    case RAMP_S_COLLIMATOR: s_collimator(n, c); break;
    default: break;
    }
}

#endif
