// Real-space Ewald sum and its corrections (notes.tex:66-79, test.py:83,
// 126-132): with the reciprocal part f(z) = exp(-pi^2 z/eta^2)/(pi z) they
// give the total Coulomb energy of the periodic system, independent of eta.
//   E_pair = 1/2 sum_i sum_{j,n}' w_i w_j erfc(eta r)/r  over images r < rc
//   E_self = -eta/sqrt(pi) sum w^2,  E_bg = -pi (sum w)^2 / (2 V eta^2)
// A brute-force FP64 pair sum (one thread per atom i, atoms j staged through
// LDS in tiles, every image within rc): meant for checks and moderate N,
// O(N^2) work.  Deterministic: fixed j order per thread, fixed-order
// reduction of the workgroup partials.
#pragma once

#include "common.hpp"

namespace spme {

struct RealArgs {
    CellArgs cell;                 // iL, 1/V
    double L00, L10, L11, L20, L21, L22;  // box rows (r = L^T s)
    double eta, rc2;
    int e0, e1, e2;                // images n_d in [-e_d, e_d]
    int n;
};

constexpr int kRealTile = 256;

// partials[b * 10 + c]: c = 0 E_pair, 1..6 virial (xx,yy,zz,yx,zx,zy) times
// -V, 7 sum w^2, 8 sum w
__global__ void __launch_bounds__(kBlock) real_space_kernel(RealArgs a,
                                                            const double *__restrict__ w,
                                                            const double *__restrict__ x,
                                                            double *__restrict__ dx,
                                                            double *__restrict__ phi,
                                                            double *__restrict__ partials) {
    __shared__ double sx[kRealTile], sy[kRealTile], sz[kRealTile], sw[kRealTile];
    __shared__ double red[kBlock / 64][10];
    const CellArgs &c = a.cell;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const bool valid = i < a.n;
    double xi = 0, yi = 0, zi = 0, wi = 0;
    if (valid) {
        xi = x[3 * i];
        yi = x[3 * i + 1];
        zi = x[3 * i + 2];
        wi = w[i];
    }
    const double c2 = 2.0 * a.eta / sqrt(M_PI);
    double ph = 0, gx = 0, gy = 0, gz = 0, e = 0;
    double v[6] = {0, 0, 0, 0, 0, 0};
    for (int j0 = 0; j0 < a.n; j0 += kRealTile) {
        __syncthreads();
        for (int t = threadIdx.x; t < kRealTile; t += kBlock) {
            const int j = j0 + t;
            sx[t] = j < a.n ? x[3 * j] : 0.0;
            sy[t] = j < a.n ? x[3 * j + 1] : 0.0;
            sz[t] = j < a.n ? x[3 * j + 2] : 0.0;
            sw[t] = j < a.n ? w[j] : 0.0;
        }
        __syncthreads();
        if (!valid) continue;
        const int nt = min(kRealTile, a.n - j0);
        for (int t = 0; t < nt; ++t) {
            const double dx0 = xi - sx[t], dy0 = yi - sy[t], dz0 = zi - sz[t];
            double s0 = dx0 * c.iL00 + dy0 * c.iL10 + dz0 * c.iL20;
            double s1 = dy0 * c.iL11 + dz0 * c.iL21;
            double s2 = dz0 * c.iL22;
            s0 -= rint(s0);
            s1 -= rint(s1);
            s2 -= rint(s2);
            const double wj = sw[t];
            for (int n0 = -a.e0; n0 <= a.e0; ++n0)
                for (int n1 = -a.e1; n1 <= a.e1; ++n1)
                    for (int n2 = -a.e2; n2 <= a.e2; ++n2) {
                        const double t0 = s0 + n0, t1 = s1 + n1, t2 = s2 + n2;
                        const double rx = t0 * a.L00 + t1 * a.L10 + t2 * a.L20;
                        const double ry = t1 * a.L11 + t2 * a.L21;
                        const double rz = t2 * a.L22;
                        const double r2 = rx * rx + ry * ry + rz * rz;
                        if (r2 >= a.rc2 || r2 <= 1e-20) continue;
                        const double r = sqrt(r2);
                        const double ir = 1.0 / r;
                        const double g = erfc(a.eta * r) * ir;
                        ph += wj * g;
                        const double ww = wi * wj;
                        e += 0.5 * ww * g;
                        // U'(r)/r
                        const double up = ww * (-g - c2 * exp(-a.eta * a.eta * r2)) * ir * ir;
                        gx += up * rx;
                        gy += up * ry;
                        gz += up * rz;
                        v[0] += 0.5 * up * rx * rx;
                        v[1] += 0.5 * up * ry * ry;
                        v[2] += 0.5 * up * rz * rz;
                        v[3] += 0.5 * up * ry * rx;
                        v[4] += 0.5 * up * rz * rx;
                        v[5] += 0.5 * up * rz * ry;
                    }
        }
    }
    if (valid) {
        if (dx) {
            dx[3 * i] = gx;
            dx[3 * i + 1] = gy;
            dx[3 * i + 2] = gz;
        }
        if (phi) phi[i] = ph - c2 * wi;  // the background term is added by the finish
    }
    double acc[10] = {e, v[0], v[1], v[2], v[3], v[4], v[5], wi * wi, wi, 0.0};
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        double s = acc[q];
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
        if (lane == 0) red[wv][q] = s;
    }
    __syncthreads();
    if (threadIdx.x < 9) {
        double s = 0.0;
        for (int q = 0; q < kBlock / 64; ++q) s += red[q][threadIdx.x];
        partials[blockIdx.x * 10 + threadIdx.x] = s;
    }
}

// One workgroup: sums the partials in a fixed order and writes
// out = [E_pair, vir(6), E_self, E_bg, vir_bg]; adds the background term
// -pi Q / (V eta^2) to phi.
__global__ void __launch_bounds__(kBlock) real_space_finish(RealArgs a, int nparts,
                                                            const double *__restrict__ partials,
                                                            double *__restrict__ phi,
                                                            double *__restrict__ out) {
    __shared__ double tot[9];
    if (threadIdx.x < 9) {
        double s = 0.0;
        for (int b = 0; b < nparts; ++b) s += partials[b * 10 + threadIdx.x];
        tot[threadIdx.x] = s;
    }
    __syncthreads();
    const double iV = a.cell.iV, Q = tot[8];
    const double ebg = -M_PI * Q * Q * iV / (2.0 * a.eta * a.eta);
    if (threadIdx.x == 0) {
        out[0] = tot[0];
        for (int q = 0; q < 6; ++q) out[1 + q] = -tot[1 + q] * iV;
        out[7] = -a.eta / sqrt(M_PI) * tot[7];
        out[8] = ebg;
        out[9] = ebg * iV;
    }
    if (phi) {
        const double pb = -M_PI * Q * iV / (a.eta * a.eta);
        for (int i = threadIdx.x; i < a.n; i += kBlock) phi[i] += pb;
    }
}

}  // namespace spme
