/* Exhaustive CPU check of the star layers' division by the lattice size (sky.glsl:88,
 * q = (vec3(c) + u2) / N): for N = 200 and N = 48 and every dividend the layer can form from a
 * finite unit direction (c = floor(d N) + {-1, 0, 1} in [-N - 2, N + 2], u2 = k / 65536 for every
 * 16-bit k), the quotient by the sequence
 *     q0 = a * RN(1/N);  r = fma(-q0, N, a);  q = fma(r, RN(1/N), q0)
 * equals the IEEE quotient a / N. The sum a = c + u2 is exact in float (at most 8 integer and 16
 * fraction bits). Prints the number of dividends checked and of mismatches; exit status 1 on any
 * mismatch. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>

static long check(float N, long* n_checked) {
    const float rN = 1.0f / N;
    const float inv16 = 1.0f / 65536.0f;
    const int lim = (int)N + 2;
    long bad = 0;
    for (int c = -lim; c <= lim; ++c)
        for (uint32_t k = 0; k < 65536u; ++k) {
            const float a = (float)c + (float)k * inv16;
            const float ref = a / N;
            const float q0 = a * rN;
            const float r = fmaf(-q0, N, a);
            const float q = fmaf(r, rN, q0);
            union { float f; uint32_t u; } x = {ref}, y = {q};
            if (x.u != y.u) {
                if (bad < 5) printf("mismatch N=%g a=%.9g ref=%.9g got=%.9g\n", N, a, ref, q);
                ++bad;
            }
            ++*n_checked;
        }
    return bad;
}

int main(void) {
    long n = 0, bad = 0;
    bad += check(200.0f, &n);
    bad += check(48.0f, &n);
    printf("dividends checked: %ld\nmismatches: %ld\n", n, bad);
    return bad != 0;
}
