/* rvec.h - the geometric builtins of the reference's kernels (OpenCL dot,
 * length, normalize, cross on float2/float3) with one fixed evaluation
 * order, shared by the HIP kernels and the CPU oracle.
 *
 *   dot(a,b)      = (a.x*b.x + a.y*b.y) + a.z*b.z       (no fused ops)
 *   length(a)     = sqrt(dot(a,a))                      (IEEE sqrt)
 *   normalize(a)  = a / length(a) componentwise; a zero vector is returned
 *                   unchanged
 *   cross(a,b)    = (a.y*b.z - a.z*b.y, a.z*b.x - a.x*b.z, a.x*b.y - a.y*b.x)
 */
#ifndef RAMP_RVEC_H
#define RAMP_RVEC_H

#include "rmath.h"

RHD rf3 v3_add(rf3 a, rf3 b) { return rf3_make(a.x + b.x, a.y + b.y, a.z + b.z); }
RHD rf3 v3_sub(rf3 a, rf3 b) { return rf3_make(a.x - b.x, a.y - b.y, a.z - b.z); }
RHD rf3 v3_scale(float s, rf3 a) { return rf3_make(s * a.x, s * a.y, s * a.z); }
RHD rf3 v3_mulv(rf3 a, float s) { return rf3_make(a.x * s, a.y * s, a.z * s); }
RHD float v3_dot(rf3 a, rf3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
RHD float v3_len(rf3 a) { return rm_sqrt(v3_dot(a, a)); }
RHD rf3 v3_norm(rf3 a)
{
    float l = v3_len(a);
    if (l == 0.0f) return a;
    return rf3_make(a.x / l, a.y / l, a.z / l);
}
RHD rf3 v3_cross(rf3 a, rf3 b)
{
    return rf3_make(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
RHD int v3_eq(rf3 a, rf3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

RHD float v2_dot(rf2 a, rf2 b) { return a.x * b.x + a.y * b.y; }
RHD float v2_len(rf2 a) { return rm_sqrt(v2_dot(a, a)); }
RHD rf2 v2_norm(rf2 a)
{
    float l = v2_len(a);
    if (l == 0.0f) return a;
    return rf2_make(a.x / l, a.y / l);
}

/* rotate_about_axis (mcramp/include/geom.h:1-19): the Rodrigues matrix of
 * `angle` about `a`, entries formed as the reference forms them, applied
 * to v.  The axis is used as given; the reference does not normalise it. */
RHD rf3 v3_rotate_axis(float angle, rf3 a, rf3 v)
{
    float c = rm_cos(angle), s = rm_sin(angle), t = 1 - c;
    float rxx = c + a.x * a.x * t, rxy = a.x * a.y * t - a.z * s, rxz = a.x * a.z * t + a.y * s;
    float ryx = a.y * a.x * t + a.z * s, ryy = c + a.y * a.y * t, ryz = a.y * a.z * t - a.x * s;
    float rzx = a.z * a.x * t - a.y * s, rzy = a.z * a.y * t + a.x * s, rzz = c + a.z * a.z * t;
    return rf3_make((rxx * v.x + rxy * v.y) + rxz * v.z, (ryx * v.x + ryy * v.y) + ryz * v.z,
                    (rzx * v.x + rzy * v.y) + rzz * v.z);
}

/* Euler-angle frame rotation R = Rz(g) Ry(b) Rx(a) (mcramp/include/geom.h:
 * 21-53).  The nine entries are formed exactly as the reference forms them;
 * rotate applies R, derotate applies R^T (geom.h:55-87). */
typedef struct { float m[9]; } rmat3;

/* sin and cos of an angle; rm_sin(+-0) is +0 and rm_cos(+-0) is 1 exactly
 * (tested), so a zero angle takes no transcendental and gives the same bits */
RHD void rm_sincos_z(float a, float *s, float *c)
{
    if (a == 0.0f) {
        *s = 0.0f;
        *c = 1.0f;
    } else {
        *s = rm_sin(a);
        *c = rm_cos(a);
    }
}

RHD rmat3 rmat3_euler(float xa, float ya, float za)
{
    rmat3 R;
    float cx, sx, cy, sy, cz, sz;
    rm_sincos_z(xa, &sx, &cx);
    rm_sincos_z(ya, &sy, &cy);
    rm_sincos_z(za, &sz, &cz);
    R.m[0] = cy * cz;
    R.m[1] = -cx * sz + sx * sy * cz;
    R.m[2] = sx * sz + cx * sy * cz;
    R.m[3] = cy * sz;
    R.m[4] = cx * cz + sx * sy * sz;
    R.m[5] = -sx * cz + cx * sy * sz;
    R.m[6] = -sy;
    R.m[7] = sx * cy;
    R.m[8] = cx * cy;
    return R;
}

RHD rf3 rmat3_rotate(const rmat3 *R, rf3 v)
{
    return rf3_make(v3_dot(rf3_make(R->m[0], R->m[1], R->m[2]), v),
                    v3_dot(rf3_make(R->m[3], R->m[4], R->m[5]), v),
                    v3_dot(rf3_make(R->m[6], R->m[7], R->m[8]), v));
}

RHD rf3 rmat3_derotate(const rmat3 *R, rf3 v)
{
    return rf3_make(v3_dot(rf3_make(R->m[0], R->m[3], R->m[6]), v),
                    v3_dot(rf3_make(R->m[1], R->m[4], R->m[7]), v),
                    v3_dot(rf3_make(R->m[2], R->m[5], R->m[8]), v));
}

/* Rodrigues rotation of v by angle about a unit axis (geom.h:1-19); the
 * reference re-evaluates cos/sin per matrix entry, which is value-identical
 * to evaluating them once. */
RHD rf3 rot_about_axis(float angle, rf3 ax, rf3 v)
{
    float c = rm_cos(angle), s = rm_sin(angle);
    float omc = 1.0f - c;
    float Rxx = c + ax.x * ax.x * omc;
    float Rxy = ax.x * ax.y * omc - ax.z * s;
    float Rxz = ax.x * ax.z * omc + ax.y * s;
    float Ryx = ax.y * ax.x * omc + ax.z * s;
    float Ryy = c + ax.y * ax.y * omc;
    float Ryz = ax.y * ax.z * omc - ax.x * s;
    float Rzx = ax.z * ax.x * omc - ax.y * s;
    float Rzy = ax.z * ax.y * omc + ax.x * s;
    float Rzz = c + ax.z * ax.z * omc;
    return rf3_make((Rxx * v.x + Rxy * v.y) + Rxz * v.z,
                    (Ryx * v.x + Ryy * v.y) + Ryz * v.z,
                    (Rzx * v.x + Rzy * v.y) + Rzz * v.z);
}

#endif
