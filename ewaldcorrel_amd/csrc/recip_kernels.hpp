// Strided (y, x) FFT passes and the reciprocal-space influence step.
//
// Influence step (sfac.cpp:304-396): for each half-complex point h,
//   s = |S_h|^2,  E_raw += A_h s,  v_ab += s dA_h,ab,
//   S_h <- S_h A_h iB0 iB1 iB2 (x2 on kz = 0 and the even-K2 Nyquist plane),
// then E = E_raw iV and virial = (v_ab + delta_ab E_raw) iV^2.
// The table is either read from HBM (any rad_fn) or, for the closed-form
// Ewald kernel, evaluated from the indices (no table traffic).
// Reductions are FP64: registers -> wavefront shuffles -> LDS -> one
// partial per workgroup; workgroup 0 of the next launch sums the partials
// in a fixed order (deterministic; the kernel boundary orders the stores
// before the loads, so no fences or atomics are needed).
#pragma once

#include "common.hpp"
#include "fft_kernels.hpp"

namespace spme {

// ---- closed-form Ewald influence at one half-complex point -----------------------------
// The images of (kx, ky, kz) taken into account are the principal signed
// indices, with both +-K/2 at an even-K Nyquist index (they alias to the
// same point and have different |m| in a triclinic box).  Interior kz
// points also receive the conjugate images -m (same |m|), hence mult = 2.
// A = mult/2 sum f(m^2),  dA_ab = mult sum f'(m^2) m_a m_b over images with
// 1e-10 <= |m|^2 <= max_m^2; f(z) = exp(fac z)/(pi z), f' = (fac - 1/z) f
// (ewald.py:124-132).  The engine only selects this mode after checking
// on the host that it reproduces the full table.
__host__ __device__ inline void ewald_point_impl(int K0, int K1, int K2, const CellArgs &c,
                                                 double fac, double max_m2, int kx, int ky,
                                                 int kz, double &A, double *dA) {
    int nxs[2], nys[2], nzs[2];
    int cx = 1, cy = 1, cz = 1;
    nxs[0] = kx <= K0 / 2 ? kx : kx - K0;
    if ((K0 % 2) == 0 && kx == K0 / 2 && K0 > 0) { nxs[1] = -kx; cx = 2; }
    nys[0] = ky <= K1 / 2 ? ky : ky - K1;
    if ((K1 % 2) == 0 && ky == K1 / 2 && K1 > 0) { nys[1] = -ky; cy = 2; }
    nzs[0] = kz;
    const bool zny = (K2 % 2) == 0 && kz == K2 / 2;
    if (zny && kz > 0) { nzs[1] = -kz; cz = 2; }
    const double mult = (kz == 0 || zny) ? 1.0 : 2.0;
    double a = 0.0, d0 = 0, d1 = 0, d2 = 0, d3 = 0, d4 = 0, d5 = 0;
    for (int ix = 0; ix < cx; ++ix)
        for (int iy = 0; iy < cy; ++iy)
            for (int iz = 0; iz < cz; ++iz) {
                const double nx = nxs[ix], ny = nys[iy], nz = nzs[iz];
                const double mx = nx * c.iL00;
                const double my = nx * c.iL10 + ny * c.iL11;
                const double mz = nx * c.iL20 + ny * c.iL21 + nz * c.iL22;
                const double m2 = mx * mx + my * my + mz * mz;
                if (m2 < 1e-10 || m2 > max_m2) continue;
                const double f = exp(m2 * fac) / (m2 * M_PI);
                const double fp = (fac - 1.0 / m2) * f;
                a += f;
                d0 += fp * mx * mx;
                d1 += fp * my * my;
                d2 += fp * mz * mz;
                d3 += fp * my * mx;
                d4 += fp * mz * mx;
                d5 += fp * mz * my;
            }
    A = 0.5 * mult * a;
    dA[0] = mult * d0;
    dA[1] = mult * d1;
    dA[2] = mult * d2;
    dA[3] = mult * d3;
    dA[4] = mult * d4;
    dA[5] = mult * d5;
}

__device__ __forceinline__ void table_point(const FrameDesc &F, long h, int kx, int ky, int kz,
                                            double &A, double *dA) {
    if (F.tabmode == kTableAnalytic) {
        ewald_point_impl(F.K0, F.K1, F.K2, F.tcell, F.ewald_fac, F.max_m2, kx, ky, kz, A, dA);
    } else {
        const long H = (long)F.K0 * F.K1 * F.ldim;
        A = F.table[h];
        const double *d = F.table + H + 6 * h;
#pragma unroll
        for (int c = 0; c < 6; ++c) dA[c] = d[c];
    }
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Block-reduce acc[7] into one partial per workgroup: F.partials[b] = sums.
__device__ __forceinline__ void block_partial(const FrameDesc &F, int b, const double *acc) {
    __shared__ double red[kBlock / 64][8];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < 7; ++c) {
        const double v = wave_sum(acc[c]);
        if (lane == 0) red[wv][c] = v;
    }
    __syncthreads();
    if (threadIdx.x < 7) {
        double v = 0.0;
        for (int q = 0; q < kBlock / 64; ++q) v += red[q][threadIdx.x];
        F.partials[b * 8 + threadIdx.x] = v;
    }
}

// Sum the nparts partials of the previous launch in a fixed order
// (deterministic) and write E = iV E_raw and the virial
// (v_ab + delta_ab E_raw) iV^2 (sfac.cpp:384-395).
__device__ __forceinline__ void finalize_reduction(const FrameDesc &F, int nparts) {
    __shared__ double red2[kBlock / 64][8];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    double part[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int q = threadIdx.x; q < nparts; q += blockDim.x)
#pragma unroll
        for (int c = 0; c < 7; ++c) part[c] += F.partials[q * 8 + c];
#pragma unroll
    for (int c = 0; c < 7; ++c) {
        const double v = wave_sum(part[c]);
        if (lane == 0) red2[wv][c] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double r[7];
        for (int c = 0; c < 7; ++c) {
            r[c] = 0.0;
            for (int q = 0; q < kBlock / 64; ++q) r[c] += red2[q][c];
        }
        const double iV = F.cell.iV, e = r[0], fac = iV * iV;
        F.out[0] = e * iV;
        F.out[1] = (r[1] + e) * fac;
        F.out[2] = (r[2] + e) * fac;
        F.out[3] = (r[3] + e) * fac;
        F.out[4] = r[4] * fac;
        F.out[5] = r[5] * fac;
        F.out[6] = r[6] * fac;
    }
}

// One workgroup per frame: finalize the reduction (after `en`).
__global__ void __launch_bounds__(kBlock) finish_kernel(SPME_FRAME_ARGS) {
    const FrameDesc &F = frames[nf <= 1 ? 0 : blockIdx.x];
    finalize_reduction(F, F.nred_i);
}

// ---- standalone influence (en / de1 entry points) ------------------------------------------
// Grid-stride over the H half-complex points with F.nred_i workgroups per frame.
template <typename T, bool APPLY>
__global__ void __launch_bounds__(kBlock) influence_kernel(SPME_FRAME_ARGS) {
    int b;
    const FrameDesc &F = frames[find_frame(offs, nf, b)];
    const long H = (long)F.K0 * F.K1 * F.ldim;
    const bool even = (F.K2 % 2) == 0;
    cplx<T> *fq = (cplx<T> *)F.grid;
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    for (long h = (long)b * blockDim.x + threadIdx.x; h < H; h += (long)F.nred_i * blockDim.x) {
        const int kz = (int)(h % F.ldim);
        const long r = h / F.ldim;
        const int ky = (int)(r % F.K1);
        const int kx = (int)(r / F.K1);
        double A, dA[6];
        table_point(F, h, kx, ky, kz, A, dA);
        cplx<T> v = fq[h];
        const double re = (double)v.x, im = (double)v.y;
        const double s = re * re + im * im;
        acc[0] += A * s;
        if (APPLY) {
#pragma unroll
            for (int c = 0; c < 6; ++c) acc[1 + c] += s * dA[c];
            double t = A * F.iB0[kx] * F.iB1[ky] * F.iB2[kz];
            if (kz == 0 || (even && kz == F.ldim - 1)) t *= 2.0;
            v.x = T(re * t);
            v.y = T(im * t);
            fq[h] = v;
        }
    }
    block_partial(F, b, acc);
}

// ---- strided passes ------------------------------------------------------------------------
// AXIS 1 (y): lines (x = o, kz) along y, o < K0, element stride ldim.
// AXIS 0 (x): lines (y = o, kz) along x, o < K1, element stride K1*ldim.
// A frame's blocks are (o, tile); tile t covers kz in [t ldim/T, (t+1) ldim/T).
// MODE 0: plain transform with `sign`; with `finish`, workgroup 0 of each
//         frame first finalizes the reduction of the previous launch
//         (finish 1: after MODE 2, K1*xT partials; 2: after the standalone
//         influence, nred_i partials).
// MODE 1: forward transform, outputs scaled by iB0 iB1 iB2 (S-hat).
// MODE 2 (x only): forward, /B, influence with reductions, inverse -- the
//         whole reciprocal-space step of an evaluation in one HBM pass.
// LDS: twiddles (N), iB0 and the tile's iB2 (N + C doubles, padded to an
// even count), the lines (C x (N+1)), and a second line buffer for plans
// with a generic radix.
template <typename T>
__host__ __device__ inline size_t strided_lds_bytes(int N, int C, int generic) {
    return fft_lds_bytes<T>(N, C, generic) + (size_t)((N + C + 1) & ~1) * sizeof(double);
}

template <typename T, int AXIS, int MODE>
__global__ void __launch_bounds__(kBlock) fft_strided_kernel(SPME_FRAME_ARGS, int sign,
                                                             int finish) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    int b;
    const FrameDesc &F = frames[find_frame(offs, nf, b)];
    if (finish && b == 0) {
        finalize_reduction(F, finish == 1 ? F.K1 * F.xT : F.nred_i);
        __syncthreads();
    }
    const FftPlanArgs &p = AXIS == 1 ? F.plan1 : F.plan0;
    const void *twg = AXIS == 1 ? F.tw1 : F.tw0;
    const int N = p.N, LS = N + 1, ldim = F.ldim;
    const int C = AXIS == 1 ? F.yC : F.xC;
    const int Tt = AXIS == 1 ? F.yT : F.xT;
    const long ostride = AXIS == 1 ? (long)F.K1 * ldim : (long)ldim;
    const long estride = AXIS == 1 ? (long)ldim : (long)F.K1 * ldim;
    const int o = b / Tt, t = b - o * Tt;
    const int c0 = (int)((long)t * ldim / Tt), c1 = (int)((long)(t + 1) * ldim / Tt);
    const int nc = c1 - c0;
    cplx<T> *twl = reinterpret_cast<cplx<T> *>(smem_raw);
    double *ib0 = reinterpret_cast<double *>(twl + N);
    double *ib2 = ib0 + N;
    cplx<T> *b0 = reinterpret_cast<cplx<T> *>(ib0 + ((N + C + 1) & ~1));
    cplx<T> *b1 = b0 + C * LS;
    cplx<T> *base = (cplx<T> *)F.grid + (long)o * ostride + c0;
    constexpr int EPT = fft_elems<T>() / kBlock;
    const int tot = nc * N;
    {
        // all global loads of the thread first (clamped), then LDS writes
        T vx[EPT], vy[EPT];  // scalar arrays stay in registers
#pragma unroll
        for (int q = 0; q < EPT; ++q) {
            const int l = min((int)threadIdx.x + q * kBlock, tot - 1);
            const int e = l / nc, c = l - e * nc;
            const cplx<T> w = base[(long)e * estride + c];
            vx[q] = w.x;
            vy[q] = w.y;
        }
        load_twiddles<T>(twl, twg, N);
        if (MODE != 0) {
            for (int e = threadIdx.x; e < N; e += blockDim.x) ib0[e] = F.iB0[e];
            for (int c = threadIdx.x; c < nc; c += blockDim.x) ib2[c] = F.iB2[c0 + c];
        }
#pragma unroll
        for (int q = 0; q < EPT; ++q) {
            const int l = threadIdx.x + q * kBlock;
            if (l < tot) {
                const int e = l / nc, c = l - e * nc;
                cplx<T> w;
                w.x = vx[q];
                w.y = vy[q];
                b0[c * LS + e] = w;
            }
        }
    }
    __syncthreads();
    const int s0 = MODE == 0 ? sign : -1;
    cplx<T> *res = fft_lines<T>(b0, b1, nc, LS, p, twl, s0);
    if (MODE == 2) {
        // kx = e, ky = o, kz = c0 + c
        const bool even = (F.K2 % 2) == 0;
        const double iby = F.iB1[o];
        double acc[7] = {0, 0, 0, 0, 0, 0, 0};
        for (int l = threadIdx.x; l < tot; l += blockDim.x) {
            const int c = l / N, e = l - c * N;
            const int kz = c0 + c;
            const double ib = ib0[e] * iby * ib2[c];
            cplx<T> v = res[c * LS + e];
            const double re = (double)v.x * ib, im = (double)v.y * ib;  // S-hat
            const double s = re * re + im * im;
            const long h = ((long)e * F.K1 + o) * ldim + kz;
            double A, dA[6];
            table_point(F, h, e, o, kz, A, dA);
            acc[0] += A * s;
#pragma unroll
            for (int q = 0; q < 6; ++q) acc[1 + q] += s * dA[q];
            double tt = A * ib;
            if (kz == 0 || (even && kz == ldim - 1)) tt *= 2.0;
            v.x = T(re * tt);
            v.y = T(im * tt);
            res[c * LS + e] = v;
        }
        __syncthreads();
        // the inverse runs in the buffer that holds the data
        cplx<T> *other = (res == b0) ? b1 : b0;
        res = fft_lines<T>(res, other, nc, LS, p, twl, +1);
        for (int l = threadIdx.x; l < tot; l += blockDim.x) {
            const int e = l / nc, c = l - e * nc;
            base[(long)e * estride + c] = res[c * LS + e];
        }
        block_partial(F, b, acc);
        return;
    }
    const double iby = (MODE == 1) ? F.iB1[o] : 1.0;
    for (int l = threadIdx.x; l < tot; l += blockDim.x) {
        const int e = l / nc, c = l - e * nc;
        cplx<T> v = res[c * LS + e];
        if (MODE == 1) {
            const T s = T(ib0[e] * iby * ib2[c]);
            v.x *= s;
            v.y *= s;
        }
        base[(long)e * estride + c] = v;
    }
}

}  // namespace spme
