/* Checks the weight formula of the packed CSR entries: with y = RN(1 / d),
 * q = RN(c y), r = RN(c - q d) (an FMA, exact) and w = RN(q + r y) (an FMA),
 * w equals the correctly rounded FP32 quotient c / d for every
 * 0 <= c <= d <= dmax.  Prints the count of pairs and mismatches; exit
 * status 1 on a mismatch. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
int main(int argc, char** argv) {
    unsigned dmax = (unsigned)atoi(argv[1]);
    unsigned long long bad = 0, n = 0;
    for (unsigned d = 1; d <= dmax; ++d) {
        const float fd = (float)d;
        const float y = 1.0f / fd;
        for (unsigned c = 0; c <= d; ++c) {
            const float fc = (float)c;
            const float ref = fc / fd;
            const float q = fc * y;
            const float r = fmaf(-q, fd, fc);
            const float w = fmaf(r, y, q);
            ++n;
            if (w != ref) { if (bad < 5) printf("mismatch c=%u d=%u ref=%a got=%a\n", c, d, ref, w); ++bad; }
        }
    }
    printf("dmax=%u pairs=%llu mismatches=%llu\n", dmax, n, bad);
    return bad != 0;
}
