/* rconst.h - physical constants of the neutron path, with the values the
 * reference kernels compile with (mcramp/include/consts.h:3-37), which are
 * the McStas conversion constants.  The two entries whose reference
 * definition is a double-precision expression are given as such:
 *   MIN2RAD = M_PI/(180.0f*60.0f)          (double on fp64 devices)
 *   kB      = 8.6173*10e-2f = 0.86173      (double; 10x the physical
 *                                           0.086173 meV/K, kept as is)
 */
#ifndef RAMP_RCONST_H
#define RAMP_RCONST_H

#define RC_K_B 1.38066e-23f     /* J/K */
#define RC_V2K 1.58825361e-3f   /* v[m/s] -> k[1/A] */
#define RC_K2V 629.622368f      /* k[1/A] -> v[m/s] */
#define RC_VS2E 5.22703725e-6f  /* v^2 [m^2/s^2] -> E [meV] */
#define RC_SE2V 437.393377f     /* sqrt(E[meV]) -> v [m/s] */
#define RC_E2KS (1.0f / 2.072f)
#define RC_M_PI_D 3.14159265358979323846 /* double M_PI of OpenCL */
#define RC_MIN2RAD_D (RC_M_PI_D / (double)(180.0f * 60.0f))
#define RC_KB_D (8.6173 * (double)10e-2f)
#define RC_SQRT_2PI_D 0x1.40d931ff62705p+1 /* sqrt(2.0f*M_PI) in double (monochromator.cl:7) */
#define RC_MEV2J 1.6021765e-22f /* cont_mod.cl:38 */
#define RC_ISIS_WSCALE 6.2415093e+12f /* isis_mod.cl:94 */
#define RC_SENTINEL 100000.0f  /* intersection sentinel t (e.g. scat/arm.cl:27) */

#endif
