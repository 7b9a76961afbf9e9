// hostutil.cpp -- ran2-based seed tables for the Hybrid Taus generator.
#include "hostutil.h"

#include <cstring>

namespace sop {

namespace {

// L'Ecuyer's combined generator with a Bays-Durham shuffle table, as
// Numerical Recipes' ran2 writes it.  The table is only filled for a
// non-positive seed; with a positive seed it stays zero, which is what the
// reference's static state gives (ran2.cpp:54-74).
struct Ran2 {
    static constexpr int IM1 = 2147483563, IM2 = 2147483399, IMM1 = IM1 - 1;
    static constexpr int IA1 = 40014, IA2 = 40692, IQ1 = 53668, IQ2 = 52774, IR1 = 12211, IR2 = 3791;
    static constexpr int NTAB = 32, NDIV = 1 + IMM1 / NTAB;
    int idum2 = 123456789, iy = 0, iv[NTAB] = {0};

    double next(int* idum) {
        int j, k;
        if (*idum <= 0) {
            *idum = (-(*idum) < 1) ? 1 : -(*idum);
            idum2 = *idum;
            for (j = NTAB + 7; j >= 0; j--) {
                k = (*idum) / IQ1;
                *idum = IA1 * (*idum - k * IQ1) - k * IR1;
                if (*idum < 0) *idum += IM1;
                if (j < NTAB) iv[j] = *idum;
            }
            iy = iv[0];
        }
        k = (*idum) / IQ1;
        *idum = IA1 * (*idum - k * IQ1) - k * IR1;
        if (*idum < 0) *idum += IM1;
        k = idum2 / IQ2;
        idum2 = IA2 * (idum2 - k * IQ2) - k * IR2;
        if (idum2 < 0) idum2 += IM2;
        j = iy / NDIV;
        iy = iv[j] - idum2;
        iv[j] = *idum;
        if (iy < 1) iy += IMM1;
        const double am = 1.0 / IM1, rnmx = 1.0 - 1.2e-7;
        double t = am * iy;
        return t > rnmx ? rnmx : t;
    }
};

}  // namespace

void generate_seeds(int seed, int n, uint32_t* out4) {
    Ran2 r;
    for (int i = 0; i < n; ++i)
        for (int c = 0; c < 4; ++c) {
            uint32_t v;
            do {
                v = (uint32_t)(r.next(&seed) * 4294967295.0);
            } while (v < 128u);
            out4[4 * i + c] = v;
        }
}

}  // namespace sop
