/* Golden-vector driver for the reference's vendored Random123 generator.
 *
 * Compiled against the reference's own headers where they lie
 * (mcramp/include/Random123/threefry.h and u01fixedpt.h, reached through
 * -I <reference>/mcramp/include); nothing from the reference is copied.
 * Prints, one line per case, the 4 output words of threefry4x32 (20 rounds,
 * the default used by rand.h:1,14) and the float bits of
 * u01fixedpt_closed_closed_32_24 of word 0, the mapping used by rand.h:19.
 * Cases: the neutron key/counter layout of rand.h:7-8 for a grid of thread
 * ids and counter values, plus keys and counters from a fixed LCG.
 */
#include <stdio.h>
#include <stdint.h>
#include <string.h>
#include <Random123/threefry.h>
#include <Random123/u01fixedpt.h>

static uint32_t lcg_state = 12345u;
static uint32_t lcg(void) { lcg_state = lcg_state * 1664525u + 1013904223u; return lcg_state; }

static void emit(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                 uint32_t k0, uint32_t k1, uint32_t k2, uint32_t k3)
{
    threefry4x32_ctr_t c = {{c0, c1, c2, c3}};
    threefry4x32_key_t k = {{k0, k1, k2, k3}};
    threefry4x32_ctr_t r = threefry4x32(c, k);
    float u = u01fixedpt_closed_closed_32_24(r.v[0]);
    uint32_t ub;
    memcpy(&ub, &u, 4);
    printf("%u %u %u %u %u %u %u %u %u %u %u %u %u\n", c0, c1, c2, c3, k0, k1, k2, k3,
           r.v[0], r.v[1], r.v[2], r.v[3], ub);
}

int main(void)
{
    const uint32_t tids[] = {0u, 1u, 2u, 63u, 64u, 1000u, 65535u, 1000000u, 99999999u, 4294967295u};
    const uint32_t ctrs[] = {0u, 1u, 2u, 7u, 100u, 4096u, 4103u, 16777215u, 16777216u};
    for (unsigned i = 0; i < sizeof(tids) / sizeof(tids[0]); i++)
        for (unsigned j = 0; j < sizeof(ctrs) / sizeof(ctrs[0]); j++)
            emit(ctrs[j], 0xf00dcafeu, 0xdeadbeefu, 0xbeeff00du,
                 tids[i], 0xdecafbadu, 0xfacebeadu, 0x12345678u);
    for (int i = 0; i < 64; i++) {
        uint32_t a = lcg(), b = lcg(), c = lcg(), d = lcg();
        uint32_t e = lcg(), f = lcg(), g = lcg(), h = lcg();
        emit(a, b, c, d, e, f, g, h);
    }
    /* u01 edge words: exact 0 and exact 1 are reachable */
    return 0;
}
