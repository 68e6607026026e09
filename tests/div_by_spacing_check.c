/* Check of the division by a grid spacing used by the small-grid loop of
 * pts_step2d.hip (div_h): for h > 0 and rh = 1/h rounded to nearest,
 * q = x*rh, r = fma(h, q, -x), fma(-r, rh, q) must equal x/h rounded to
 * nearest (Markstein's theorem), signed zeros included.  Numerators: random
 * mantissas and exponents of both signs, and numerators whose quotient lies
 * next to a midpoint between doubles (the hard cases of correct rounding).
 * Usage: div_by_spacing_check N_RANDOM N_MIDPOINT; exit status 0 when no
 * quotient differs. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint64_t s = 88172645463325252ull;
static uint64_t next(void) {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return s;
}
static double div_h(double x, double h, double rh) {
    const double q = x * rh;
    const double r = fma(h, q, -x);
    return fma(-r, rh, q);
}
static int same_bits(double u, double v) {
    uint64_t a, b;
    memcpy(&a, &u, 8);
    memcpy(&b, &v, 8);
    return a == b;
}
static double from_bits(uint64_t b) {
    double d;
    memcpy(&d, &b, 8);
    return d;
}

int main(int argc, char** argv) {
    const long nrand = argc > 1 ? atol(argv[1]) : 1000000, nmid = argc > 2 ? atol(argv[2]) : 1000000;
    /* the spacings 10/n of the tested and benchmarked grids */
    const int ns[] = {127, 255, 63, 70, 42, 90, 61, 33, 100, 6, 5, 2047, 511, 8191, 31, 47};
    long bad = 0, n = 0;
    for (int k = 0; k < (int)(sizeof ns / sizeof ns[0]); ++k) {
        const double h = 10.0 / ns[k], rh = 1.0 / h;
        for (long i = 0; i < nrand; ++i) {
            const uint64_t m = next() & ((1ull << 52) - 1), e = 1023 - 60 + next() % 81, sg = next() & 1;
            const double x = from_bits((sg << 63) | (e << 52) | m);
            ++n;
            if (!same_bits(div_h(x, h, rh), x / h)) {
                if (bad < 5) printf("h=%a x=%a: %a against %a\n", h, x, div_h(x, h, rh), x / h);
                ++bad;
            }
        }
        for (long i = 0; i < nmid; ++i) {
            const uint64_t m = next() & ((1ull << 52) - 1), e = 1023 - 40 + next() % 61;
            const double q = from_bits((e << 52) | m), half = (nextafter(q, INFINITY) - q) * 0.5;
            const double x0 = h * q + h * half; /* next to h times the midpoint */
            for (int d = -2; d <= 2; ++d) {
                double x = x0;
                for (int t = 0; t < abs(d); ++t) x = nextafter(x, d < 0 ? -INFINITY : INFINITY);
                for (int sg = 0; sg < 2; ++sg) {
                    const double xs = sg ? -x : x;
                    ++n;
                    if (!same_bits(div_h(xs, h, rh), xs / h)) {
                        if (bad < 5) printf("h=%a x=%a: %a against %a\n", h, xs, div_h(xs, h, rh), xs / h);
                        ++bad;
                    }
                }
            }
        }
        for (int sg = 0; sg < 2; ++sg) { /* signed zeros */
            const double x = sg ? -0.0 : 0.0;
            ++n;
            if (!same_bits(div_h(x, h, rh), x / h)) ++bad;
        }
    }
    printf("%ld of %ld quotients differ\n", bad, n);
    return bad != 0;
}
