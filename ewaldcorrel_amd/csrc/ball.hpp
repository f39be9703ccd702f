// Enumeration of the reciprocal-lattice ball |m| <= max_m and its folding
// onto the half-complex grid (sfac.cpp:213-292).
//
// m = (i iL00, i iL10 + j iL11, i iL20 + j iL21 + k iL22); the i, j, k
// bounds are computed with the same floating-point expressions as the
// reference so that points on the sphere are included or excluded alike.
// Each m maps to (i mod K0, j mod K1, k mod K2); when k mod K2 >= ldim the
// point is stored at its conjugate ((-i) mod K0, (-j) mod K1, K2 - k mod K2).
// Every image of a grid point inside the ball is visited.
#pragma once

#include <math.h>

namespace spme {

template <typename F>
void enumerate_ball(const double iL[3][3], const int K[3], int ldim, double max_m, F &&visit) {
    const double R2 = max_m * max_m;
    const int i0 = (int)(max_m / iL[0][0]);
    for (int i = -i0; i <= i0; ++i) {
        const double mx = i * iL[0][0];
        const double m20 = mx * mx;
        const double dy = sqrt(R2 - m20);
        const double my0 = i * iL[1][0];
        const int j0 = (int)ceil((-dy - my0) / iL[1][1]);
        const int j1 = (int)floor((dy - my0) / iL[1][1]);
        const int n0 = ((i % K[0]) + K[0]) % K[0];
        const int n0c = ((-i % K[0]) + K[0]) % K[0];
        for (int j = j0; j <= j1; ++j) {
            const double my = my0 + j * iL[1][1];
            const double m21 = m20 + my * my;
            const double dz = sqrt(R2 - m21);
            const double mz0 = i * iL[2][0] + j * iL[2][1];
            const int k0 = (int)ceil((-dz - mz0) / iL[2][2]);
            const int k1 = (int)floor((dz - mz0) / iL[2][2]);
            const int n1 = ((j % K[1]) + K[1]) % K[1];
            const int n1c = ((-j % K[1]) + K[1]) % K[1];
            for (int k = k0; k <= k1; ++k) {
                const double mz = mz0 + k * iL[2][2];
                const double m2 = m21 + mz * mz;
                const int n2 = ((k % K[2]) + K[2]) % K[2];
                long idx;
                if (n2 >= ldim)
                    idx = ((long)n0c * K[1] + n1c) * ldim + (K[2] - n2);
                else
                    idx = ((long)n0 * K[1] + n1) * ldim + n2;
                visit(idx, m2, mx, my, mz);
            }
        }
    }
}

}  // namespace spme
