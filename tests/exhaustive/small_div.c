/* Exhaustive check of the binary distance's quotient on the GPU:
 *   y = v_rcp_f32(b)            (within 1 ulp of 1/b)
 *   q0 = a * y;  r = fma(-q0, b, a);  q = fma(r, y, q0)
 * equals the correctly rounded a / b for every pair of integers
 * 0 <= a <= b <= B, for each of the three floats within one ulp of the
 * correctly rounded reciprocal (a superset of what a 1-ulp reciprocal can
 * return).  Prints the number of failures (0 expected) and exits non-zero
 * on any.  usage: small_div B */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

int main(int argc, char **argv)
{
    const long B = argc > 1 ? atol(argv[1]) : 131072;
    long fails = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : fails)
    for (long b = 1; b <= B; ++b) {
        const float fb = (float)b;
        const float yr = 1.0f / fb;
        const float ys[3] = {nextafterf(yr, 0.0f), yr, nextafterf(yr, 2.0f)};
        for (int t = 0; t < 3; ++t) {
            const float y = ys[t];
            for (long a = 0; a <= b; ++a) {
                const float fa = (float)a;
                const float q0 = fa * y;
                const float r = fmaf(-q0, fb, fa);
                const float q = fmaf(r, y, q0);
                const float want = (float)((double)a / (double)b);
                if (q != want) ++fails;
            }
        }
    }
    printf("B=%ld failures=%ld\n", B, fails);
    return fails != 0;
}
