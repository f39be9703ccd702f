// Force and potential gather (de2 / potl, sfac.cpp:409-545).
//
// Per atom: u = scale(x) K, k1 = ceil(u), t = k1 - u, stencil grid points
// (k1 - n/2 + i) mod K with weights M(t+i) and derivatives M'(t+i);
//   dE/ds_d = -w iV K_d sum d_d(prod M) Q,  dx = iL (dE/ds)   (force = -dx)
//   phi     = iV sum prod M Q                                  (= dE/dw)
// n threads per atom, thread i owns the x-plane i of the stencil (n^2 loads,
// n consecutive reals per row); the planes are summed through LDS.  With
// SORTED the workgroups walk the column-ordered records of the last
// spreading pass (coalesced, no re-read of the positions), neighbouring
// atoms share grid lines, and consecutive workgroups are kept on one XCD so
// those lines are shared in its L2.  From a record's cell c and fraction
// f = u - c: ceil(u) = c + 1 and ceil(u) - u = 1 - f for f > 0 (exact in
// FP64), and ceil(u) = c, 0 for f = 0.
#pragma once

#include "common.hpp"
#include "spread_kernels.hpp"

namespace spme {

// four consecutive grid values with element (not vector) alignment
template <typename T> struct Row4;
template <> struct Row4<float> {
    typedef float type __attribute__((ext_vector_type(4), aligned(4)));
};
template <> struct Row4<double> {
    typedef double type __attribute__((ext_vector_type(4), aligned(8)));
};

template <typename T, int ORDER>
__device__ __forceinline__ void spline_wd(const SplineArgs<T> &sp, int n, T x, T *b, T *db) {
#pragma unroll
    for (int j = 0; j < Ord<ORDER>::cap; ++j) {
        if (j >= n) break;
        T d = sp.M[j * n];
        T v = sp.M[j * n] * x + sp.M[j * n + 1];
#pragma unroll
        for (int i = 2; i < Ord<ORDER>::cap; ++i) {
            if (i >= n) break;
            d = d * x + v;
            v = v * x + sp.M[j * n + i];
        }
        b[j] = v;
        db[j] = d;
    }
}

template <typename T, int ORDER, bool SORTED>
__global__ void __launch_bounds__(kBlock) gather_kernel(SPME_FRAME_ARGS, SplineArgs<T> sp) {
    __shared__ double red[kBlock][4];
    int b;
    const int bid = SORTED ? xcd_block((int)blockIdx.x, (int)gridDim.x) : (int)blockIdx.x;
    const FrameDesc &F = frames[find_frame_of(offs, nf, bid, b)];
    const int n = order_of<ORDER>(F.order);
    constexpr int CAP = Ord<ORDER>::cap;
    const int apb = kBlock / n;
    const int slot = threadIdx.x / n;
    const int i = threadIdx.x - slot * n;
    const int p = b * apb + slot;
    const bool valid = slot < apb && p < F.n;
    double d0 = 0, d1 = 0, d2 = 0, pp = 0;
    int a = 0;
    double wa = 0.0;
    if (valid) {
        int k0, k1, k2;
        T t0, t1, t2;
        if (SORTED) {
            const AtomRec<T> r = ((const AtomRec<T> *)F.recs)[p];
            a = r.perm;
            wa = r.w;
            const int c0 = r.cell >> 20, c1 = (r.cell >> 10) & 1023, c2 = r.cell & 1023;
            k0 = r.f[0] > T(0) ? c0 + 1 : c0;
            k1 = r.f[1] > T(0) ? c1 + 1 : c1;
            k2 = r.f[2] > T(0) ? c2 + 1 : c2;
            t0 = r.f[0] > T(0) ? T(1) - r.f[0] : T(0);
            t1 = r.f[1] > T(0) ? T(1) - r.f[1] : T(0);
            t2 = r.f[2] > T(0) ? T(1) - r.f[2] : T(0);
        } else {
            a = p;
            wa = F.dx ? F.w[a] : 0.0;
            double s0, s1, s2;
            scale_wrap(F.cell, F.x[3 * a], F.x[3 * a + 1], F.x[3 * a + 2], s0, s1, s2);
            const double u0 = s0 * F.K0, u1 = s1 * F.K1, u2 = s2 * F.K2;
            const double e0 = ceil(u0), e1 = ceil(u1), e2 = ceil(u2);
            k0 = (int)e0;
            k1 = (int)e1;
            k2 = (int)e2;
            t0 = T(e0 - u0);
            t1 = T(e1 - u1);
            t2 = T(e2 - u2);
        }
        T by[CAP], bz[CAP], dby[CAP], dbz[CAP], bx[CAP], dbx[CAP];
        spline_wd<T, ORDER>(sp, n, t0, bx, dbx);
        spline_wd<T, ORDER>(sp, n, t1, by, dby);
        spline_wd<T, ORDER>(sp, n, t2, bz, dbz);
        const int gx = imod(k0 - n / 2 + i, F.K0);
        const int o1 = imod(k1 - n / 2, F.K1);
        const int o2 = imod(k2 - n / 2, F.K2);
        const T *grid = (const T *)F.grid;
        const long rs = 2L * F.ldim;
        // order 4 with the stencil inside the row: one 16-byte (FP32) or two
        // (FP64) vector loads per row, element-aligned
        const bool vec = ORDER == 4 && o2 + 4 <= F.K2;
        T s_b = 0, s_db = 0, s_bdz = 0;
#pragma unroll
        for (int j = 0; j < CAP; ++j) {
            if (j >= n) break;
            int gy = o1 + j;
            gy = gy >= F.K1 ? gy - F.K1 * (gy / F.K1) : gy;
            const T *row = grid + ((long)gx * F.K1 + gy) * rs;
            T q[CAP];
            if (vec) {
                const typename Row4<T>::type v = *(const typename Row4<T>::type *)(row + o2);
                q[0] = v.x;
                q[1] = v.y;
                q[2] = v.z;
                q[3] = v.w;
            } else {
#pragma unroll
                for (int k = 0; k < CAP; ++k) {
                    if (k >= n) break;
                    int z = o2 + k;
                    z = z >= F.K2 ? z - F.K2 * (z / F.K2) : z;
                    q[k] = row[z];
                }
            }
            T zb = 0, zdb = 0;
#pragma unroll
            for (int k = 0; k < CAP; ++k) {
                if (k >= n) break;
                zb += bz[k] * q[k];
                zdb += dbz[k] * q[k];
            }
            s_b += by[j] * zb;
            s_db += dby[j] * zb;
            s_bdz += by[j] * zdb;
        }
        T bxi = bx[0], dbxi = dbx[0];
#pragma unroll
        for (int q = 1; q < CAP; ++q)
            if (q == i) {
                bxi = bx[q];
                dbxi = dbx[q];
            }
        d0 = (double)(dbxi * s_b);
        d1 = (double)(bxi * s_db);
        d2 = (double)(bxi * s_bdz);
        pp = (double)(bxi * s_b);
    }
    red[threadIdx.x][0] = d0;
    red[threadIdx.x][1] = d1;
    red[threadIdx.x][2] = d2;
    red[threadIdx.x][3] = pp;
    __syncthreads();
    if (!valid || i != 0) return;
    double e0 = 0, e1 = 0, e2 = 0, ph = 0;
    for (int q = 0; q < n; ++q) {
        e0 += red[threadIdx.x + q][0];
        e1 += red[threadIdx.x + q][1];
        e2 += red[threadIdx.x + q][2];
        ph += red[threadIdx.x + q][3];
    }
    const CellArgs &c = F.cell;
    if (F.dx) {
        const double f = -wa * c.iV;
        const double g0 = e0 * f * F.K0, g1 = e1 * f * F.K1, g2 = e2 * f * F.K2;
        F.dx[3 * a + 0] = c.iL00 * g0;
        F.dx[3 * a + 1] = c.iL10 * g0 + c.iL11 * g1;
        F.dx[3 * a + 2] = c.iL20 * g0 + c.iL21 * g1 + c.iL22 * g2;
    }
    if (F.phi) F.phi[a] = ph * c.iV;
}

}  // namespace spme
