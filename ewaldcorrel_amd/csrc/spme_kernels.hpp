// SPME kernels other than the FFT: charge spreading, influence multiply
// with energy/virial reduction, and the force/potential gather.
#pragma once

#include "common.hpp"

namespace spme {

// ---- B-spline weights ------------------------------------------------------
// Kernels are specialised for ORDER = 4, 6, 8; ORDER = 0 is the generic
// kernel with the order read at run time (arrays sized kMaxOrder).
template <int ORDER> struct Ord {
    static constexpr int cap = ORDER > 0 ? ORDER : kMaxOrder;
};
template <int ORDER> __device__ __forceinline__ int order_of(const GridArgs &g) {
    return ORDER > 0 ? ORDER : g.order;
}

// M_n(x + j) for j = 0..n-1 by Horner on the coefficient matrix
// (sfac.hpp:124-131).
template <typename T, int ORDER>
__device__ __forceinline__ void spline_w(const SplineArgs<T> &sp, int n, T x, T *b) {
#pragma unroll
    for (int j = 0; j < Ord<ORDER>::cap; ++j) {
        if (j >= n) break;
        T v = sp.M[j * n];
#pragma unroll
        for (int i = 1; i < Ord<ORDER>::cap; ++i) {
            if (i >= n) break;
            v = v * x + sp.M[j * n + i];
        }
        b[j] = v;
    }
}
// value and derivative (sfac.hpp:146-157)
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

// ---- fixed-point scale -------------------------------------------------------
// max |w| as the bit pattern of a non-negative double (order independent,
// so the scale, and therefore the integer grid, is bit-reproducible).
__global__ void absmax_bits(int n, const double *__restrict__ w,
                            unsigned long long *__restrict__ out) {
    double m = 0.0;
    for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < n; a += gridDim.x * blockDim.x)
        m = fmax(m, fabs(w[a]));
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

// scale = 2^(62 - e) with n*max|w|*Mmax^3 < 2^e, so no partial sum of the
// integer grid can overflow int64.  The bound is clamped below so that the
// scale stays finite when every charge is zero.
__device__ __forceinline__ double fixed_scale(int n, unsigned long long maxbits, double mmax3) {
    const double bound = fmax((double)n * __longlong_as_double((long long)maxbits) * mmax3,
                              1e-200);
    int e;
    frexp(bound, &e);
    return ldexp(1.0, 62 - e);
}

// ---- spreading ------------------------------------------------------------------
// Scatter form of the reference's gather-form accumulation (sfac.cpp:141-168):
// an atom in cell c = (int)u with fraction f = u - c adds
// w * M(f0+i) M(f1+j) M(f2+k) to grid point ((c + n/2 - i) mod K, ...).
// One thread per (atom, i): ORDER^2 updates each, contiguous in z.
// MODE 0: FP64 hardware atomics into the double grid (padded layout).
// MODE 1: int64 fixed point into an unpadded K0*K1*K2 accumulator.
template <typename T, int ORDER, int MODE>
__global__ void __launch_bounds__(256)
spread_kernel(int n, const double *__restrict__ w, const double *__restrict__ x,
              CellArgs cell, GridArgs g, SplineArgs<T> sp, double *__restrict__ gridd,
              unsigned long long *__restrict__ gridfx,
              const unsigned long long *__restrict__ maxbits, double mmax3) {
    const int n_ = order_of<ORDER>(g);
    constexpr int CAP = Ord<ORDER>::cap;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int a = (int)(t / n_);
    const int i = (int)(t - (long)a * n_);
    if (a >= n) return;
    double s0, s1, s2;
    scale_wrap(cell, x[3 * a], x[3 * a + 1], x[3 * a + 2], s0, s1, s2);
    const double u0 = s0 * g.K0, u1 = s1 * g.K1, u2 = s2 * g.K2;
    int c0 = (int)u0, c1 = (int)u1, c2 = (int)u2;
    const T f0 = T(u0 - c0), f1 = T(u1 - c1), f2 = T(u2 - c2);
    // s wrapped to exactly 1.0 -> periodic image (cell 0)
    if (c0 >= g.K0) c0 -= g.K0;
    if (c1 >= g.K1) c1 -= g.K1;
    if (c2 >= g.K2) c2 -= g.K2;
    T bx[CAP], by[CAP], bz[CAP];
    spline_w<T, ORDER>(sp, n_, f0, bx);
    spline_w<T, ORDER>(sp, n_, f1, by);
    spline_w<T, ORDER>(sp, n_, f2, bz);
    const double wi = w[a];
    const int gx = imod(c0 + n_ / 2 - i, g.K0);
    const double wx = wi * (double)bx[i];
    double scale = 0.0;
    if (MODE == 1) scale = fixed_scale(n, *maxbits, mmax3);
    int gzs[CAP];
#pragma unroll
    for (int k = 0; k < CAP; ++k)
        if (k < n_) gzs[k] = imod(c2 + n_ / 2 - k, g.K2);
#pragma unroll
    for (int j = 0; j < CAP; ++j) {
        if (j >= n_) break;
        const int gy = imod(c1 + n_ / 2 - j, g.K1);
        const double wxy = wx * (double)by[j];
        if (MODE == 0) {
            double *row = gridd + ((long)gx * g.K1 + gy) * (2L * g.ldim);
#pragma unroll
            for (int k = 0; k < CAP; ++k) {
                if (k >= n_) break;
                unsafeAtomicAdd(row + gzs[k], wxy * (double)bz[k]);
            }
        } else {
            unsigned long long *row = gridfx + ((long)gx * g.K1 + gy) * g.K2;
#pragma unroll
            for (int k = 0; k < CAP; ++k) {
                if (k >= n_) break;
                const long long q = llrint(wxy * (double)bz[k] * scale);
                atomicAdd(row + gzs[k], (unsigned long long)q);
            }
        }
    }
}

// int64 fixed-point accumulator -> working-precision padded grid.
template <typename T>
__global__ void fixed_to_grid(const unsigned long long *__restrict__ acc, T *__restrict__ grid,
                              GridArgs g, int n, const unsigned long long *__restrict__ maxbits,
                              double mmax3) {
    const double inv = 1.0 / fixed_scale(n, *maxbits, mmax3);
    const long G = (long)g.K0 * g.K1 * g.K2;
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < G;
         p += (long)gridDim.x * blockDim.x) {
        const long row = p / g.K2;
        const int z = (int)(p - row * g.K2);
        grid[row * 2L * g.ldim + z] = T((double)(long long)acc[p] * inv);
    }
}

// ---- influence multiply + energy/virial reduction --------------------------------
// For each half-complex point h = (kx, ky, kz):  s = |S|^2,  E += A s,
// v_ab += s dA_ab; with APPLY, S *= A iB0 iB1 iB2 (x2 on kz = 0 and the
// even-K2 Nyquist plane) (sfac.cpp:304-325, 361-382).  FP64 accumulation in
// registers, wavefront shuffle + LDS tree per block, one partial per block;
// a second kernel sums the partials in a fixed order (deterministic).
constexpr int kRedBlock = 256;
constexpr int kRedMaxBlocks = 1024;

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

template <typename T, bool APPLY>
__global__ void __launch_bounds__(kRedBlock)
influence_kernel(cplx<T> *__restrict__ fq, const double *__restrict__ A,
                 const double *__restrict__ dA, GridArgs g, const double *__restrict__ iB0,
                 const double *__restrict__ iB1, const double *__restrict__ iB2,
                 double *__restrict__ partials) {
    const long H = (long)g.K0 * g.K1 * g.ldim;
    const bool even = (g.K2 % 2) == 0;
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    for (long h = (long)blockIdx.x * blockDim.x + threadIdx.x; h < H;
         h += (long)gridDim.x * blockDim.x) {
        cplx<T> v = fq[h];
        const double re = (double)v.x, im = (double)v.y;
        const double s = re * re + im * im;
        const double a = A[h];
        acc[0] += a * s;
        if (APPLY) {
            const double *d = dA + 6 * h;
            acc[1] += s * d[0];
            acc[2] += s * d[1];
            acc[3] += s * d[2];
            acc[4] += s * d[3];
            acc[5] += s * d[4];
            acc[6] += s * d[5];
            const int kz = (int)(h % g.ldim);
            const long r = h / g.ldim;
            const int ky = (int)(r % g.K1);
            const int kx = (int)(r / g.K1);
            double t = a * iB0[kx] * iB1[ky] * iB2[kz];
            if (kz == 0 || (even && kz == g.ldim - 1)) t *= 2.0;
            v.x = T(re * t);
            v.y = T(im * t);
            fq[h] = v;
        }
    }
    __shared__ double red[kRedBlock / 64][8];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < 7; ++c) {
        const double v = wave_sum(acc[c]);
        if (lane == 0) red[wv][c] = v;
    }
    __syncthreads();
    if (threadIdx.x < 7) {
        double v = 0.0;
        for (int q = 0; q < kRedBlock / 64; ++q) v += red[q][threadIdx.x];
        partials[blockIdx.x * 8 + threadIdx.x] = v;
    }
}

// out[0] = E = iV sum A|S|^2;  out[1..6] = virial (xx,yy,zz,yx,zx,zy)
// = (v_ab + delta_ab E_raw) iV^2 (sfac.cpp:384-395).
__global__ void __launch_bounds__(kRedBlock)
reduce_finalize(const double *__restrict__ partials, int nparts, double iV,
                double *__restrict__ out) {
    __shared__ double red[kRedBlock / 64][8];
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < nparts; b += blockDim.x)
#pragma unroll
        for (int c = 0; c < 7; ++c) acc[c] += partials[b * 8 + c];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < 7; ++c) {
        const double v = wave_sum(acc[c]);
        if (lane == 0) red[wv][c] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double r[7];
        for (int c = 0; c < 7; ++c) {
            r[c] = 0.0;
            for (int q = 0; q < kRedBlock / 64; ++q) r[c] += red[q][c];
        }
        const double e = r[0];
        const double fac = iV * iV;
        out[0] = e * iV;
        out[1] = (r[1] + e) * fac;
        out[2] = (r[2] + e) * fac;
        out[3] = (r[3] + e) * fac;
        out[4] = r[4] * fac;
        out[5] = r[5] * fac;
        out[6] = r[6] * fac;
    }
}

// ---- force / potential gather ----------------------------------------------------
// Per atom: u = scale(x) K, k1 = ceil(u), t = k1 - u, grid (k1 - n/2 + i) mod K
// with M(t+i) (sfac.cpp:428-470, 514-541).
//   dx = iL (dE/ds),  dE/ds_d = -w iV K_d sum d_d(prod M) Q   (de2)
//   phi = iV sum prod M Q                                      (potl)
template <typename T, int ORDER>
__global__ void __launch_bounds__(256)
gather_kernel(int n, const double *__restrict__ w, const double *__restrict__ x,
              CellArgs cell, GridArgs g, SplineArgs<T> sp, const T *__restrict__ grid,
              double *__restrict__ dx, double *__restrict__ phi) {
    const int n_ = order_of<ORDER>(g);
    constexpr int CAP = Ord<ORDER>::cap;
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    double s0, s1, s2;
    scale_wrap(cell, x[3 * a], x[3 * a + 1], x[3 * a + 2], s0, s1, s2);
    const double u0 = s0 * g.K0, u1 = s1 * g.K1, u2 = s2 * g.K2;
    const double k0 = ceil(u0), k1 = ceil(u1), k2 = ceil(u2);
    T bx[CAP], by[CAP], bz[CAP], dbx[CAP], dby[CAP], dbz[CAP];
    spline_wd<T, ORDER>(sp, n_, T(k0 - u0), bx, dbx);
    spline_wd<T, ORDER>(sp, n_, T(k1 - u1), by, dby);
    spline_wd<T, ORDER>(sp, n_, T(k2 - u2), bz, dbz);
    const int o0 = imod((int)k0 - n_ / 2, g.K0);
    const int o1 = imod((int)k1 - n_ / 2, g.K1);
    const int o2 = imod((int)k2 - n_ / 2, g.K2);
    int gz[CAP];
#pragma unroll
    for (int k = 0; k < CAP; ++k) {
        int z = (o2 + k) % g.K2;
        gz[k] = z;
    }
    T d0 = 0, d1 = 0, d2 = 0, p = 0;
#pragma unroll
    for (int i = 0; i < CAP; ++i) {
        if (i >= n_) break;
        const int gx = (o0 + i) % g.K0;
        T s_b = 0, s_db = 0, s_bdz = 0;  // (j,k) sums weighted by by.bz, dby.bz, by.dbz
#pragma unroll
        for (int j = 0; j < CAP; ++j) {
            if (j >= n_) break;
            const int gy = (o1 + j) % g.K1;
            const T *row = grid + ((long)gx * g.K1 + gy) * (2L * g.ldim);
            T zb = 0, zdb = 0;
#pragma unroll
            for (int k = 0; k < CAP; ++k) {
                if (k >= n_) break;
                const T q = row[gz[k]];
                zb += bz[k] * q;
                zdb += dbz[k] * q;
            }
            s_b += by[j] * zb;
            s_db += dby[j] * zb;
            s_bdz += by[j] * zdb;
        }
        d0 += dbx[i] * s_b;
        d1 += bx[i] * s_db;
        d2 += bx[i] * s_bdz;
        p += bx[i] * s_b;
    }
    if (dx) {
        const double f = -w[a] * cell.iV;
        const double e0 = (double)d0 * f * g.K0;
        const double e1 = (double)d1 * f * g.K1;
        const double e2 = (double)d2 * f * g.K2;
        dx[3 * a + 0] = cell.iL00 * e0;
        dx[3 * a + 1] = cell.iL10 * e0 + cell.iL11 * e1;
        dx[3 * a + 2] = cell.iL20 * e0 + cell.iL21 * e1 + cell.iL22 * e2;
    }
    if (phi) phi[a] = (double)p * cell.iV;
}

}  // namespace spme
