/* rrng.h - the per-neutron counter-based random stream.
 *
 * Reproduces the draw sequence of the reference's rand() (mcramp/include/
 * rand.h:4-20) bit for bit:
 *   - generator: Threefry-4x32 with 20 rounds (Salmon, Moraes, Dror, Shaw,
 *     "Parallel random numbers: as easy as 1, 2, 3", SC'11), written here
 *     from the published round structure: Threefish-style MIX on word pairs
 *     (0,1),(2,3) on even rounds and (0,3),(2,1) on odd rounds, rotation
 *     constants for the 4x32 variant, key injection every 4 rounds with the
 *     Skein parity word 0x1BD11BDA;
 *   - key  = {tid, 0xdecafbad, 0xfacebead, 0x12345678}  (rand.h:7), tid being
 *     the chunk-local index of the neutron;
 *   - ctr  = {(uint)ctr_f, 0xf00dcafe, 0xdeadbeef, 0xbeeff00d} (rand.h:8,15);
 *     ctr_f is a float carried in the neutron state and incremented by 1.0f
 *     after every draw (rand.h:18);
 *   - output = ((w0 & 0x7fffffc0) + (w0 & 0x40)) * 2^-31 of word 0, i.e.
 *     the closed-closed 24-bit fixed-point map of Random123's u01fixedpt.h
 *     (u01fixedpt.h:137-140); 0 and 1 are both possible.
 */
#ifndef RAMP_RRNG_H
#define RAMP_RRNG_H

#include "rdefs.h"

#define RR_KEY1 0xdecafbadu
#define RR_KEY2 0xfacebeadu
#define RR_KEY3 0x12345678u
#define RR_CTR1 0xf00dcafeu
#define RR_CTR2 0xdeadbeefu
#define RR_CTR3 0xbeeff00du
#define RR_PARITY 0x1BD11BDAu

RHD uint32_t rr_rotl(uint32_t x, int n) { return (x << n) | (x >> (32 - n)); }

#define RR_MIX_EVEN(ra, rb) \
    x0 += x1; x1 = rr_rotl(x1, ra); x1 ^= x0; \
    x2 += x3; x3 = rr_rotl(x3, rb); x3 ^= x2;
#define RR_MIX_ODD(ra, rb) \
    x0 += x3; x3 = rr_rotl(x3, ra); x3 ^= x0; \
    x2 += x1; x1 = rr_rotl(x1, rb); x1 ^= x2;
#define RR_INJECT(a, b, c, d, s) \
    x0 += a; x1 += b; x2 += c; x3 += d + (uint32_t)(s);

/* Full 4-word output (used by the golden-vector tests). */
RHD void rr_threefry4x32_20(const uint32_t in[4], const uint32_t key[4], uint32_t out[4])
{
    uint32_t k0 = key[0], k1 = key[1], k2 = key[2], k3 = key[3];
    uint32_t k4 = RR_PARITY ^ k0 ^ k1 ^ k2 ^ k3;
    uint32_t x0 = in[0] + k0, x1 = in[1] + k1, x2 = in[2] + k2, x3 = in[3] + k3;
    RR_MIX_EVEN(10, 26) RR_MIX_ODD(11, 21) RR_MIX_EVEN(13, 27) RR_MIX_ODD(23, 5)
    RR_INJECT(k1, k2, k3, k4, 1)
    RR_MIX_EVEN(6, 20) RR_MIX_ODD(17, 11) RR_MIX_EVEN(25, 10) RR_MIX_ODD(18, 20)
    RR_INJECT(k2, k3, k4, k0, 2)
    RR_MIX_EVEN(10, 26) RR_MIX_ODD(11, 21) RR_MIX_EVEN(13, 27) RR_MIX_ODD(23, 5)
    RR_INJECT(k3, k4, k0, k1, 3)
    RR_MIX_EVEN(6, 20) RR_MIX_ODD(17, 11) RR_MIX_EVEN(25, 10) RR_MIX_ODD(18, 20)
    RR_INJECT(k4, k0, k1, k2, 4)
    RR_MIX_EVEN(10, 26) RR_MIX_ODD(11, 21) RR_MIX_EVEN(13, 27) RR_MIX_ODD(23, 5)
    RR_INJECT(k0, k1, k2, k3, 5)
    out[0] = x0; out[1] = x1; out[2] = x2; out[3] = x3;
}

/* Word 0 of Threefry-4x32-20 for the neutron key/counter layout. */
RHD uint32_t rr_word0(uint32_t c0, uint32_t tid)
{
    uint32_t k0 = tid, k1 = RR_KEY1, k2 = RR_KEY2, k3 = RR_KEY3;
    uint32_t k4 = RR_PARITY ^ k0 ^ k1 ^ k2 ^ k3;
    uint32_t x0 = c0 + k0, x1 = RR_CTR1 + k1, x2 = RR_CTR2 + k2, x3 = RR_CTR3 + k3;
    RR_MIX_EVEN(10, 26) RR_MIX_ODD(11, 21) RR_MIX_EVEN(13, 27) RR_MIX_ODD(23, 5)
    RR_INJECT(k1, k2, k3, k4, 1)
    RR_MIX_EVEN(6, 20) RR_MIX_ODD(17, 11) RR_MIX_EVEN(25, 10) RR_MIX_ODD(18, 20)
    RR_INJECT(k2, k3, k4, k0, 2)
    RR_MIX_EVEN(10, 26) RR_MIX_ODD(11, 21) RR_MIX_EVEN(13, 27) RR_MIX_ODD(23, 5)
    RR_INJECT(k3, k4, k0, k1, 3)
    RR_MIX_EVEN(6, 20) RR_MIX_ODD(17, 11) RR_MIX_EVEN(25, 10) RR_MIX_ODD(18, 20)
    RR_INJECT(k4, k0, k1, k2, 4)
    RR_MIX_EVEN(10, 26) RR_MIX_ODD(11, 21) RR_MIX_EVEN(13, 27) RR_MIX_ODD(23, 5)
    x0 += k0;
    (void)x1; (void)x2; (void)x3;
    return x0;
}

/* Closed-closed 24-bit fixed-point map of a 32-bit word onto [0, 1]. */
RHD float rr_u01(uint32_t w)
{
    return (float)((w & 0x7fffffc0u) + (w & 0x40u)) * 0x1p-31f;
}

/* One draw: counter word 0 is the float counter converted to uint
 * (the reference adds the float to a uint 0), then the counter advances. */
RHD float rr_draw(float *ctr, uint32_t tid)
{
    uint32_t c0 = (uint32_t)(*ctr);
    float u = rr_u01(rr_word0(c0, tid));
    *ctr = *ctr + 1.0f;
    return u;
}

#endif
