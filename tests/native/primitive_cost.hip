/* primitive_cost.hip - one kernel per arithmetic primitive of the shared
 * math library and RNG, compiled to gfx950 assembly by primitive_cost.py
 * to count the VALU instructions each primitive takes on its own (no
 * divergence, no other work).  Never launched. */
#include <hip/hip_runtime.h>

#include "../../ramp_amd/csrc/rmath.h"
#include "../../ramp_amd/csrc/rrng.h"

#define IDX const int i = blockIdx.x * blockDim.x + threadIdx.x

extern "C" __global__ void pc_base1(const float *x, const float *y, float *o) { IDX; o[i] = x[i]; }
extern "C" __global__ void pc_base2(const float *x, const float *y, float *o) { IDX; o[i] = x[i] + y[i]; }
extern "C" __global__ void pc_threefry(const float *x, const float *y, float *o)
{
    IDX;
    o[i] = rr_u01(rr_word0(__float_as_uint(x[i]), __float_as_uint(y[i])));
}
extern "C" __global__ void pc_div(const float *x, const float *y, float *o) { IDX; o[i] = x[i] / y[i]; }
extern "C" __global__ void pc_sqrt(const float *x, const float *y, float *o) { IDX; o[i] = rm_sqrt(x[i]); }
extern "C" __global__ void pc_exp(const float *x, const float *y, float *o) { IDX; o[i] = rm_exp(x[i]); }
extern "C" __global__ void pc_log(const float *x, const float *y, float *o) { IDX; o[i] = rm_log(x[i]); }
extern "C" __global__ void pc_pow(const float *x, const float *y, float *o) { IDX; o[i] = rm_pow(x[i], y[i]); }
extern "C" __global__ void pc_tanh(const float *x, const float *y, float *o) { IDX; o[i] = rm_tanh(x[i]); }
extern "C" __global__ void pc_atan2(const float *x, const float *y, float *o) { IDX; o[i] = rm_atan2(x[i], y[i]); }
extern "C" __global__ void pc_sin(const float *x, const float *y, float *o) { IDX; o[i] = rm_sin(x[i]); }
extern "C" __global__ void pc_cos(const float *x, const float *y, float *o) { IDX; o[i] = rm_cos(x[i]); }
extern "C" __global__ void pc_asin(const float *x, const float *y, float *o) { IDX; o[i] = rm_asin(x[i]); }
extern "C" __global__ void pc_acos(const float *x, const float *y, float *o) { IDX; o[i] = rm_acos(x[i]); }
extern "C" __global__ void pc_fmod(const float *x, const float *y, float *o) { IDX; o[i] = rm_fmod(x[i], y[i]); }
