// rd_div_rc (ramp_amd/csrc/rmath.h) against IEEE double division, run by
// tests/test_div_rc.py: random float numerators and divisors over the whole
// float range, exact products of two floats (48-bit numerators) over float
// divisors and over roots of floats, and divisors and numerators next to
// powers of two.  Also checks that the float quotient of two floats equals
// the float rounding of their double quotient (the reference's d =
// 2.0*vdotn/nlen2, computed as a float division) and of rd_div_rc's, and,
// where it is a normal float, the float rounding of the numerator times the
// divisor's double reciprocal (the monochromator's GAUSS exponents).
// Prints the number of mismatches.
// Usage: div_rc_check <cases> <seed>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include "../../ramp_amd/csrc/rmath.h"

static uint64_t st;
static uint32_t next32()
{
    st ^= st << 13; st ^= st >> 7; st ^= st << 17;
    return (uint32_t)(st >> 16);
}
static float rand_float(int emin, int emax)
{
    uint32_t m = next32() & 0x7fffffu;
    int e = emin + (int)(next32() % (uint32_t)(emax - emin + 1));
    float f = std::ldexp(1.0f + (float)m / 8388608.0f, e);
    return (next32() & 1u) ? -f : f;
}
static float near_pow2(int e)
{
    uint32_t k = next32() % 64u;
    float f = std::ldexp(1.0f, e);
    return (next32() & 1u) ? std::nextafter(f, 0.0f) * (1.0f - (float)k * 0x1p-24f)
                           : f * (1.0f + (float)k * 0x1p-23f);
}

int main(int argc, char **argv)
{
    if (argc != 3) return 2;
    const long n = std::atol(argv[1]);
    st = 0x9e3779b97f4a7c15ull ^ (uint64_t)std::atoll(argv[2]);
    long bad = 0, bad_f = 0, zero = 0;
    for (long i = 0; i < n; i++) {
        double a, c;
        switch (i % 5) {
        case 0: a = rand_float(-126, 127); c = std::fabs(rand_float(-126, 127)); break;
        case 1: a = (double)rand_float(-60, 60) * (double)rand_float(-60, 60); c = std::fabs(rand_float(-60, 60)); break;
        case 2: a = (double)rand_float(-60, 60) * (double)rand_float(-60, 60);
                c = (double)std::sqrt(std::fabs(rand_float(-60, 60))); break;
        case 3: a = near_pow2((int)(next32() % 100u) - 50); c = near_pow2((int)(next32() % 100u) - 50); break;
        default: a = (double)rand_float(-149, -127); c = std::fabs(rand_float(-20, 20)); break;
        }
        if (c == 0.0) { zero++; continue; }
        const double rc = 1.0 / c;
        const double q = rd_div_rc(a, c, rc), want = a / c;
        if (std::memcmp(&q, &want, sizeof q) != 0) {
            if (bad < 5) std::printf("mismatch a=%a c=%a got %a want %a\n", a, c, q, want);
            bad++;
        }
        if (i % 5 == 0 || i % 5 == 3 || i % 5 == 4) {
            const float fa = (float)a, fc = (float)c;
            const float fq = (float)rd_div_rc((double)fa, (double)fc, 1.0 / (double)fc), fw = fa / fc;
            const float fd = (float)((double)fa / (double)fc);
            const float fm = (float)((double)fa * (1.0 / (double)fc));
            const bool normal = std::fabs(fw) >= 0x1p-126f || fw == 0.0f;
            if (std::memcmp(&fq, &fw, sizeof fq) != 0 || std::memcmp(&fd, &fw, sizeof fd) != 0 ||
                (normal && std::memcmp(&fm, &fw, sizeof fm) != 0)) {
                if (bad_f < 5) std::printf("float mismatch a=%a c=%a got %a want %a\n", fa, fc, fq, fw);
                bad_f++;
            }
        }
    }
    std::printf("cases %ld mismatches %ld float_mismatches %ld skipped %ld\n", n, bad, bad_f, zero);
    return bad || bad_f ? 1 : 0;
}
