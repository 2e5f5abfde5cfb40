/* Host check of the DC-ring bounds of the batched estimation (eb_ring_pair in
 * fairsim-vigor_amd/csrc/estimate_batch.hip): commonRegion zeroes a pixel
 * when ratio = sqrt(d2) / |k| (rounded as the reference rounds it) is below
 * dist or above 1 - dist, d2 = dx^2 + dy^2 an integer.  The kernel replaces
 * the per-pixel test by d2 < lo || d2 > hi with lo, hi from two binary
 * searches over d2 in [0, n^2 / 2]; this program runs the same searches and
 * compares the two forms for every d2, over sizes 16..4096 and |k| drawn
 * at random, integral (exact ratio ties) and tiny.
 * Usage: ring_bounds_check CASES; exit status 0 when every d2 agrees. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

static double ratio(int d2, double k) { return sqrt((double)d2) / k; }

int main(int argc, char **argv) {
    const int cases = argc > 1 ? atoi(argv[1]) : 300;
    const double dist = 0.15;
    uint64_t st = 0x9E3779B97F4A7C15ull;
    long bad = 0, checked = 0;
    for (int it = 0; it < cases; it++) {
        st ^= st << 13;
        st ^= st >> 7;
        st ^= st << 17;
        const int n = 16 << (it % 9);
        double k = (double)(st >> 11) / 9007199254740992.0 * n * 0.7 + 1e-3;
        if (it % 5 == 0) k = floor(k) + 1;    /* integral |k|: ratios hit dist exactly */
        if (it % 11 == 0) k = 1e-3 * (it + 1); /* tiny |k|: almost everything cut */
        const int dmax = 2 * (n / 2) * (n / 2);
        int a = 0, b = dmax + 1;
        while (a < b) {
            const int m = a + (b - a) / 2;
            if (ratio(m, k) < dist) a = m + 1; else b = m;
        }
        const int lo = a;
        a = 0;
        b = dmax + 1;
        while (a < b) {
            const int m = a + (b - a) / 2;
            if (ratio(m, k) > 1 - dist) b = m; else a = m + 1;
        }
        const int hi = a - 1;
        for (int d2 = 0; d2 <= dmax; d2++) {
            const double r = ratio(d2, k);
            const int direct = r < dist || r > 1 - dist, bounds = d2 < lo || d2 > hi;
            checked++;
            if (direct != bounds) {
                if (bad < 5) printf("mismatch n=%d k=%.17g d2=%d\n", n, k, d2);
                bad++;
            }
        }
    }
    printf("%ld squared radii, %ld mismatches\n", checked, bad);
    return bad != 0;
}
