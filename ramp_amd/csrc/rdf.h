/* rdf.h - error-free FP32 transformations that reproduce the reference's
 * double-precision expression sites without FP64 arithmetic on the GPU.
 *
 * The reference kernels promote some float expressions to double (M_PI and
 * unsuffixed literals are double on fp64-capable OpenCL devices) and round
 * the result back to float.  With the exact product a*b = hi + lo (one FMA)
 * and a remainder-corrected quotient, these helpers return the float nearest
 * to the (near-)exact value, which equals the reference's double-then-float
 * result except when the exact value lies within about 2^-48 (relative) of
 * a float rounding midpoint.  A double constant C is passed as its split
 * C = hi + lo with hi = (float)C, lo = (float)(C - hi).
 */
#ifndef RAMP_RDF_H
#define RAMP_RDF_H

#include "rmath.h"

typedef struct { float hi, lo; } rdf;

RHD rdf df_split(double c)
{
    rdf r;
    r.hi = (float)c;
    r.lo = (float)(c - (double)r.hi);
    return r;
}

/* exact product of two floats */
RHD rdf df_prod(float a, float b)
{
    rdf r;
    r.hi = a * b;
    r.lo = rm_fma(a, b, -r.hi);
    return r;
}

/* float nearest to u * C */
RHD float df_mulc(float u, rdf c)
{
    float p = c.hi * u;
    float e = rm_fma(c.hi, u, -p);
    return p + (e + c.lo * u);
}

/* float nearest to (a.hi + a.lo) / x */
RHD float df_div(rdf a, float x)
{
    float q = a.hi / x;
    float r = rm_fma(-q, x, a.hi) + a.lo;
    return q + r / x;
}

/* float nearest to s / (d.hi + d.lo) */
RHD float df_rdiv(float s, rdf d)
{
    float q = s / d.hi;
    float r = rm_fma(-q, d.hi, s) - q * d.lo;
    return q + r / d.hi;
}

/* float nearest to a + C */
RHD float df_addc(float a, rdf c)
{
    float s = a + c.hi;
    float bb = s - a;
    float err = (a - (s - bb)) + (c.hi - bb);
    return s + (err + c.lo);
}

/* float nearest to (a.hi + a.lo) * b */
RHD float df_mul_f(rdf a, float b)
{
    float p = a.hi * b;
    float e = rm_fma(a.hi, b, -p);
    return p + (e + a.lo * b);
}

/* double-float product (a.hi + a.lo) * b, kept as a pair */
RHD rdf df_mul_df(rdf a, float b)
{
    rdf r;
    r.hi = a.hi * b;
    r.lo = rm_fma(a.hi, b, -r.hi) + a.lo * b;
    return r;
}

/* u < (q.hi + q.lo) for a float u */
RHD int df_lt(float u, rdf q) { return u < q.hi || (u == q.hi && q.lo > 0.0f); }
/* u > (q.hi + q.lo) for a float u */
RHD int df_gt(float u, rdf q) { return u > q.hi || (u == q.hi && q.lo < 0.0f); }

#endif
