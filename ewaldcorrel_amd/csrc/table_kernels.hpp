// Influence table of the closed-form Ewald kernel built on the device
// (set_A for f(m^2) = exp(-pi^2 m^2/eta^2)/(pi m^2), sfac.cpp:213-292,
// ewald.py:124-132).
//
// One thread per half-complex point h = (kx, ky, kz).  The reference visits
// every lattice point (i, j, k) of its ball and adds it to
//   (i mod K0, j mod K1, k mod K2)                      if k mod K2 < ldim,
//   ((-i) mod K0, (-j) mod K1, K2 - k mod K2)           otherwise.
// So h receives the "direct" images (i, j, k) = (kx, ky, kz) (mod K) and, for
// 1 <= kz <= K2 - ldim, the "conjugate" images (i, j, k) = (-kx, -ky, -kz)
// (mod K).  Each thread walks the residue classes inside the ball bounds,
// which are computed with the host enumeration's floating-point expressions
// (contraction off) so that points on the sphere are included alike.  No
// atomics: the table is deterministic.
//
// The same thread evaluates the closed form of the point (principal image
// plus the Nyquist images) and reports |table - closed form| and the table
// magnitudes; the host selects the on-the-fly closed form only when they
// agree to 1e-13 of the mean magnitude.
#pragma once

#include "common.hpp"
#include "recip_kernels.hpp"

namespace spme {

struct BallArgs {
    int K0, K1, K2, ldim;
    CellArgs cell;   // iL of the box the table is built for
    double R2;       // max_m^2
    double fac;      // -(pi/eta)^2
    int i0;          // (int)(max_m / iL00)
};

// visits the ball's lattice points in the residue class (rx, ry, rz) mod K
__device__ inline void ewald_ball_class(const BallArgs &a, int rx, int ry, int rz, double &A,
                                        double *d) {
#pragma clang fp contract(off)
    const CellArgs &c = a.cell;
    // first i >= -i0 with i = rx (mod K0)
    int i = -a.i0 + imod(rx + a.i0, a.K0);
    for (; i <= a.i0; i += a.K0) {
        const double mx = i * c.iL00;
        const double m20 = mx * mx;
        const double dy = sqrt(a.R2 - m20);
        const double my0 = i * c.iL10;
        const int j0 = (int)ceil((-dy - my0) / c.iL11);
        const int j1 = (int)floor((dy - my0) / c.iL11);
        if (j1 < j0) continue;
        for (int j = j0 + imod(ry - j0, a.K1); j <= j1; j += a.K1) {
            const double my = my0 + j * c.iL11;
            const double m21 = m20 + my * my;
            const double dz = sqrt(a.R2 - m21);
            const double mz0 = i * c.iL20 + j * c.iL21;
            const int k0 = (int)ceil((-dz - mz0) / c.iL22);
            const int k1 = (int)floor((dz - mz0) / c.iL22);
            if (k1 < k0) continue;
            for (int k = k0 + imod(rz - k0, a.K2); k <= k1; k += a.K2) {
                const double mz = mz0 + k * c.iL22;
                const double m2 = m21 + mz * mz;
                if (m2 < 1e-10) continue;  // f = f' = 0 at m = 0 (ewald.py:125-127)
                const double r = 1.0 / m2;
                const double f = exp(m2 * a.fac) * r * (1.0 / M_PI);
                const double fp = (a.fac - r) * f;
                A += 0.5 * f;
                d[0] += fp * mx * mx;
                d[1] += fp * my * my;
                d[2] += fp * mz * mz;
                d[3] += fp * my * mx;
                d[4] += fp * mz * mx;
                d[5] += fp * mz * my;
            }
        }
    }
}

// stats[b*4 + {0,1,2,3}] = per-workgroup sum |A|, max |A - closed form|,
// sum max_q |dA_q|, max max_q |dA_q - closed form|
__global__ void __launch_bounds__(kBlock) ewald_table_kernel(BallArgs a, double *__restrict__ A,
                                                              double *__restrict__ dA,
                                                              double *__restrict__ stats) {
    __shared__ double red[kBlock / 64][4];
    const long H = (long)a.K0 * a.K1 * a.ldim;
    const long h = (long)blockIdx.x * kBlock + threadIdx.x;
    double sA = 0.0, eA = 0.0, sD = 0.0, eD = 0.0;
    if (h < H) {
        const int kz = (int)(h % a.ldim);
        const long r = h / a.ldim;
        const int ky = (int)(r % a.K1);
        const int kx = (int)(r / a.K1);
        double t = 0.0, d[6] = {0, 0, 0, 0, 0, 0};
        ewald_ball_class(a, kx, ky, kz, t, d);
        if (kz >= 1 && a.K2 - kz >= a.ldim)
            ewald_ball_class(a, imod(-kx, a.K0), imod(-ky, a.K1), a.K2 - kz, t, d);
        A[h] = t;
#pragma unroll
        for (int q = 0; q < 6; ++q) dA[6 * h + q] = d[q];
        double ca, cd[6];
        ewald_point_impl(a.K0, a.K1, a.K2, a.cell, a.fac, a.R2, kx, ky, kz, ca, cd);
        sA = fabs(t);
        eA = fabs(ca - t);
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            sD = fmax(sD, fabs(d[q]));
            eD = fmax(eD, fabs(cd[q] - d[q]));
        }
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int off = 32; off > 0; off >>= 1) {
        sA += __shfl_xor(sA, off);
        sD += __shfl_xor(sD, off);
        eA = fmax(eA, __shfl_xor(eA, off));
        eD = fmax(eD, __shfl_xor(eD, off));
    }
    if (lane == 0) {
        red[wv][0] = sA;
        red[wv][1] = eA;
        red[wv][2] = sD;
        red[wv][3] = eD;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double v[4] = {0, 0, 0, 0};
        for (int q = 0; q < kBlock / 64; ++q) {
            v[0] += red[q][0];
            v[1] = fmax(v[1], red[q][1]);
            v[2] += red[q][2];
            v[3] = fmax(v[3], red[q][3]);
        }
        for (int q = 0; q < 4; ++q) stats[blockIdx.x * 4 + q] = v[q];
    }
}

}  // namespace spme
