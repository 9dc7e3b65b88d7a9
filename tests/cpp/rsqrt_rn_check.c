/* Exhaustive CPU check of the march's reciprocal square root (test infrastructure).
 *
 * rsqrt_rn in blackholerender_amd/csrc/lawS_device.hpp computes 1/sqrt(x) correctly rounded to
 * float from the hardware estimate y0 = v_rsq_f32(x) (documented within 1 ulp) and one refinement
 * in double; oracle_rsqrtf in oracle/bhs_oracle.c computes it as (float)(1.0 / sqrt((double)x)).
 * This program checks, for every float x in [1, 4) (both exponent parities; every other positive
 * normal float is one of these times 4^k, and every step of both sequences is exact under that
 * scaling):
 *   1. the exact correctly rounded value c(x), decided by integer arithmetic: c is the float whose
 *      two rounding midpoints m- < 1/sqrt(x) < m+ satisfy x m-^2 < 1 < x m+^2 (128-bit products);
 *   2. oracle_rsqrtf's formula equals c(x);
 *   3. the device refinement, restated below operation for operation, returns c(x) from every
 *      estimate y0 within 2 ulps of c(x) (the five floats c - 2 ulp .. c + 2 ulp);
 * and, for every positive denormal x, that the device's scaled path (refinement of x 2^64 from
 * every estimate within 2 ulps of its value, result times 2^32; the hardware estimate flushes
 * denormal inputs) returns c(x); plus the special values and the ends of the normal range. It also
 * reports the smallest distance of 1/sqrt(x) from a float midpoint in double ulps (long double
 * arithmetic).
 * Exit status 0 when every check holds.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static float fb(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t bf(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

/* sign of x m^2 - 1 for m = (2 MY + d) 2^(ey - 25) (d = +-1: the midpoints above and below y when y
 * is not a power of two) or m = (4 MY - 1) 2^(ey - 26) (the midpoint below a power of two). */
static int sign_xm2_minus_1(float x, float y, int which) {
    int ex, ey;
    const double mx = frexp((double)x, &ex), my = frexp((double)y, &ey);
    const uint64_t MX = (uint64_t)ldexp(mx, 24), MY = (uint64_t)ldexp(my, 24);
    __int128 M;
    int e;
    if (which == 0) { M = (__int128)(2 * (int64_t)MY + 1); e = ex - 24 + 2 * (ey - 25); }
    else if (which == 1) { M = (__int128)(2 * (int64_t)MY - 1); e = ex - 24 + 2 * (ey - 25); }
    else { M = (__int128)(4 * (int64_t)MY - 1); e = ex - 24 + 2 * (ey - 26); }
    const __int128 P = (__int128)MX * M * M;            /* x m^2 = P 2^e */
    if (e >= 0) return 1;
    const int s = -e;
    if (s >= 126) return -1;
    const __int128 one = (__int128)1 << s;
    return P < one ? -1 : (P > one ? 1 : 0);
}

/* Is y the correctly rounded 1/sqrt(x) (x a normal float, y a positive normal float)? */
static int is_cr(float x, float y) {
    const int pow2 = (bf(y) & 0x7fffffu) == 0;
    const int hi = sign_xm2_minus_1(x, y, 0);           /* above the upper midpoint? */
    const int lo = sign_xm2_minus_1(x, y, pow2 ? 2 : 1);
    return hi > 0 && lo < 0;
}

/* oracle/bhs_oracle.c: oracle_rsqrtf */
static float oracle_formula(float x) { return (float)(1.0 / sqrt((double)x)); }

/* csrc/lawS_device.hpp: rsqrt_rn_core, with the hardware estimate as an argument. */
static float device_core(float x, float y0) {
    const double xd = (double)x, yd = (double)y0;
    const double t = xd * yd;
    const double e = fma(-t, yd, 1.0);
    const double c = fma(e, 0.375, 0.5);
    const double w = yd * e;
    return (float)fma(w, c, yd);
}

int main(void) {
    long n = 0, bad_oracle = 0, bad_device = 0, bad_denorm = 0, bad_special = 0;
    long double worst = 1e30L;
    float worst_x = 0.0f;
    for (uint32_t b = bf(1.0f); b < bf(4.0f); ++b) {
        const float x = fb(b);
        float c = oracle_formula(x);
        if (!is_cr(x, c)) {
            ++bad_oracle;
            /* find the correctly rounded neighbour for the device check */
            for (int k = -2; k <= 2; ++k)
                if (is_cr(x, fb(bf(c) + k))) { c = fb(bf(c) + k); break; }
        }
        for (int k = -2; k <= 2; ++k)
            if (device_core(x, fb(bf(c) + k)) != c) ++bad_device;
        /* distance of 1/sqrt(x) from the nearest float midpoint, in double ulps */
        const long double z = 1.0L / sqrtl((long double)x);
        int e;
        frexpl(z, &e);
        const long double u = ldexpl(1.0L, e - 24);
        const long double mid = (floorl(z / u) + 0.5L) * u;
        const long double m = fabsl(z - mid) / ldexpl(1.0L, e - 53);
        if (m < worst) { worst = m; worst_x = x; }
        ++n;
    }
    /* positive denormals: the device refines x 2^64 (exact, normal) and scales the result by
     * 2^32 (exact); the oracle's formula is checked against the exact decision on x 2^64 */
    for (uint32_t b = 1; b < 0x800000u; ++b) {
        const float x = fb(b);
        const float c = oracle_formula(x);
        const float xs = x * 0x1p+64f;
        const float cs = oracle_formula(xs);
        if (!is_cr(xs, cs) || c != cs * 0x1p+32f) ++bad_denorm;
        for (int k = -2; k <= 2; ++k)
            if (device_core(xs, fb(bf(cs) + k)) * 0x1p+32f != c) ++bad_denorm;
    }
    /* special values, set by the device's branch: +-0 -> +-inf, +inf -> +0 */
    if (oracle_formula(0.0f) != INFINITY || oracle_formula(-0.0f) != -INFINITY) ++bad_special;
    if (oracle_formula(INFINITY) != 0.0f || signbit(oracle_formula(INFINITY))) ++bad_special;
    if (!isnan(oracle_formula(-1.0f)) || !isnan(oracle_formula(NAN))) ++bad_special;
    /* the extremes of the normal range and both sides of every exponent parity boundary */
    const float ends[] = {0x1p-126f, 0x1.fffffep+127f, 0x1p-125f, 0x1.fffffep-126f, 0x1p+126f};
    for (unsigned i = 0; i < sizeof ends / sizeof ends[0]; ++i) {
        const float x = ends[i], c = oracle_formula(x);
        for (int k = -2; k <= 2; ++k)
            if (device_core(x, fb(bf(c) + k)) != c) ++bad_special;
    }
    printf("floats in [1,4): %ld; oracle formula not correctly rounded: %ld; device refinement "
           "wrong for some estimate within 2 ulps: %ld; denormal path: %ld; specials: %ld; "
           "closest approach to a float midpoint: %.3Lf double ulps (x = %a)\n",
           n, bad_oracle, bad_device, bad_denorm, bad_special, worst, (double)worst_x);
    return (bad_oracle || bad_device || bad_denorm || bad_special || worst < 1.0L) ? 1 : 0;
}
