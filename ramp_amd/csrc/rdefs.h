/* rdefs.h - portable definitions shared by the HIP kernels, the host library
 * and the CPU oracle.
 *
 * Everything in the shared headers is written in the common subset of C99
 * and HIP C++ so that one source text is compiled by gcc (oracle, host) and
 * by hipcc (gfx950 device code).  Floating-point expressions are written with
 * explicit evaluation order; every translation unit that includes these
 * headers is compiled with -ffp-contract=off so that no compiler fuses a
 * multiply and an add on one side only.  Fused operations are requested
 * explicitly through rm_fma().
 */
#ifndef RAMP_RDEFS_H
#define RAMP_RDEFS_H

#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define RHD __host__ __device__ static inline __attribute__((always_inline))
#define RAMP_DEVICE_CODE 1
#else
#define RHD static inline
#endif

typedef struct { float x, y, z; } rf3;
typedef struct { float x, y; } rf2;

RHD rf3 rf3_make(float x, float y, float z) { rf3 r; r.x = x; r.y = y; r.z = z; return r; }
RHD rf2 rf2_make(float x, float y) { rf2 r; r.x = x; r.y = y; return r; }

RHD uint32_t rm_f2u(float f) { union { float f; uint32_t u; } v; v.f = f; return v.u; }
RHD float rm_u2f(uint32_t u) { union { float f; uint32_t u; } v; v.u = u; return v.f; }
RHD uint64_t rm_d2u(double f) { union { double f; uint64_t u; } v; v.f = f; return v.u; }

#endif
