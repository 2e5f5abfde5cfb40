/* Host check of the division used by the device helpers (div_rn in
 * fairsim-vigor_amd/csrc/simmath.h): q = a * ib with ib = RN(1 / b), the
 * residual r = fma(-q, b, a), then fma(r, ib, q), which Markstein's theorem
 * makes the correctly rounded a / b; a or q that is zero, tiny, huge, NaN or
 * infinite takes the plain division.  Compared bit for bit with IEEE
 * division over random operands of every exponent, the divisors the
 * kernels use, operands near multiples of the divisor, and special values.
 * Usage: div_rn_check COUNT; exit status 0 when every quotient matches. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint64_t st = 88172645463325252ull;
static uint64_t next(void) {
    st ^= st << 13;
    st ^= st >> 7;
    st ^= st << 17;
    return st;
}

static double div_rn(double a, double b, double ib) {
    const double q = a * ib;
    const double m = fabs(a), mq = fabs(q);
    if (!(m >= 0x1p-900 && m <= 0x1p900 && mq >= 0x1p-900 && mq <= 0x1p900)) return a / b;
    const double r = fma(-q, b, a);
    return fma(r, ib, q);
}

static int same(double x, double y) {
    if (isnan(x) && isnan(y)) return 1;
    return memcmp(&x, &y, sizeof x) == 0;
}

static double from_bits(uint64_t u) {
    double d;
    memcpy(&d, &u, sizeof d);
    return d;
}

int main(int argc, char **argv) {
    const long count = argc > 1 ? atol(argv[1]) : 1000000;
    long bad = 0;
    const double sp[] = {0.0, -0.0, INFINITY, -INFINITY, NAN, 1e-310, -1e-320, 1e300, 1e-300, 2.0, 0.5, 3.0,
                         M_PI / 2, 512.0};
    const int ns = sizeof sp / sizeof sp[0];
    for (int i = 0; i < ns; i++)
        for (int j = 0; j < ns; j++)
            if (!same(sp[i] / sp[j], div_rn(sp[i], sp[j], 1.0 / sp[j]))) bad++;
    /* divisors of the kernels: pi / 2 (sine range reduction), OTF sample
     * spacings in cycles/micron, powers of two (phase / N), ring radii */
    const double divs[] = {M_PI / 2, 0.0201, 0.0244140625, 1.0 / (512 * 0.08), 1.0 / (1024 * 0.065), 512.0,
                           2048.0, 137.28493, 3.0};
    const int nd = sizeof divs / sizeof divs[0];
    for (long k = 0; k < count; k++) {
        double a, b;
        const int mode = (int)(k % 4);
        if (mode == 0) {
            a = from_bits((next() & 0x800FFFFFFFFFFFFFull) | ((next() % 2047) << 52));
            b = from_bits((next() & 0x800FFFFFFFFFFFFFull) | ((next() % 2047) << 52));
        } else {
            b = divs[next() % nd];
            if (mode == 1) {
                a = ldexp((double)(next() >> 11) / 9007199254740992.0, (int)(next() % 24) - 8);
            } else if (mode == 2) {
                /* a few ulps around an integer multiple of b */
                uint64_t u;
                double m = (double)(1 + next() % 100000) * b;
                memcpy(&u, &m, sizeof u);
                a = from_bits(u + (next() % 9) - 4);
            } else {
                a = sqrt((double)(next() % 3000000)) * 0.01;
            }
        }
        if (!same(a / b, div_rn(a, b, 1.0 / b))) {
            if (bad < 5) printf("mismatch a=%.17g b=%.17g\n", a, b);
            bad++;
        }
    }
    printf("%ld quotients, %ld mismatches\n", count + (long)ns * ns, bad);
    return bad != 0;
}
