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

#include <type_traits>

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
// Contraction off: the device table builder evaluates the same expressions
// for the same image, so the two agree bitwise where no alias contributes.
__host__ __device__ inline void ewald_image(const CellArgs &c, double fac, double max_m2,
                                            int ix, int iy, int iz, double &a, double *d) {
#pragma clang fp contract(off)
    const double nx = ix, ny = iy, nz = iz;
    const double mx = nx * c.iL00;
    const double my = nx * c.iL10 + ny * c.iL11;
    const double mz = nx * c.iL20 + ny * c.iL21 + nz * c.iL22;
    const double m2 = mx * mx + my * my + mz * mz;
    if (m2 < 1e-10 || m2 > max_m2) return;
    const double r = 1.0 / m2;
    const double f = exp(m2 * fac) * r * (1.0 / M_PI);
    const double fp = (fac - r) * f;
    a += f;
    d[0] += fp * mx * mx;
    d[1] += fp * my * my;
    d[2] += fp * mz * mz;
    d[3] += fp * my * mx;
    d[4] += fp * mz * mx;
    d[5] += fp * mz * my;
}

__host__ __device__ inline void ewald_point_impl(int K0, int K1, int K2, const CellArgs &c,
                                                 double fac, double max_m2, int kx, int ky,
                                                 int kz, double &A, double *dA) {
    const int nx0 = kx <= K0 / 2 ? kx : kx - K0;
    const int ny0 = ky <= K1 / 2 ? ky : ky - K1;
    const bool xny = (K0 % 2) == 0 && kx == K0 / 2 && K0 > 0;
    const bool yny = (K1 % 2) == 0 && ky == K1 / 2 && K1 > 0;
    const bool zny = (K2 % 2) == 0 && kz == K2 / 2;
    const bool zn2 = zny && kz > 0;
    const double mult = (kz == 0 || zny) ? 1.0 : 2.0;
    double a = 0.0, d[6] = {0, 0, 0, 0, 0, 0};
    ewald_image(c, fac, max_m2, nx0, ny0, kz, a, d);
    if (xny || yny || zn2) {  // Nyquist planes: the other signed images
        const int cx = xny ? 2 : 1, cy = yny ? 2 : 1, cz = zn2 ? 2 : 1;
        for (int ix = 0; ix < cx; ++ix)
            for (int iy = 0; iy < cy; ++iy)
                for (int iz = 0; iz < cz; ++iz)
                    if (ix + iy + iz > 0)
                        ewald_image(c, fac, max_m2, ix ? -nx0 : nx0, iy ? -ny0 : ny0,
                                    iz ? -kz : kz, a, d);
    }
    A = 0.5 * mult * a;
#pragma unroll
    for (int q = 0; q < 6; ++q) dA[q] = mult * d[q];
}

// FP32 closed form along one x line (fixed ky, kz) of the fused x pass: the
// line's parts of the wave vector m = n iL are computed once, so a point costs
// three products for m; exp by the hardware exp2.  The ball test |m|^2 <= max_m2
// runs in FP32 and is redone exactly as ewald_image does it (FP64, no
// contraction) within a relative band of 1e-6 around the cutoff, so the set of
// images is that of the FP64 form.  Images and factors as ewald_point_impl.
struct XLineF32 {
    float iL00, iL10, iL20;  // m per unit nx
    float Y1, Y2, Z;         // ny iL11, ny iL21, nz iL22
    float fac, fac2, max_m2f, mult;
    int K0, half0, ny, nz;
    bool even0, yny, zn2;
};

__device__ __forceinline__ void xline_init(XLineF32 &L, const FrameDesc &F, int ky, int kz) {
    const CellArgs &c = F.tcell;
    const int K1 = F.K1, K2 = F.K2;
    L.K0 = F.K0;
    L.half0 = F.K0 / 2;
    L.even0 = (F.K0 % 2) == 0;
    L.ny = ky <= K1 / 2 ? ky : ky - K1;
    L.nz = kz;
    L.yny = (K1 % 2) == 0 && ky == K1 / 2;
    const bool zny = (K2 % 2) == 0 && kz == K2 / 2;
    L.zn2 = zny && kz > 0;
    L.mult = (kz == 0 || zny) ? 1.0f : 2.0f;
    L.iL00 = (float)c.iL00;
    L.iL10 = (float)c.iL10;
    L.iL20 = (float)c.iL20;
    L.Y1 = (float)(L.ny * c.iL11);
    L.Y2 = (float)(L.ny * c.iL21);
    L.Z = (float)(kz * c.iL22);
    L.fac = (float)F.ewald_fac;
    L.fac2 = (float)(F.ewald_fac * 1.4426950408889634);  // log2(e)
    L.max_m2f = (float)F.max_m2;
}

// Point kx of the line, accumulated: returns a = sum over the point's
// images of f = exp(-pi^2 m^2 / beta^2) / (pi m^2) and adds sq fp m_a m_b
// (fp = (pi^2 / beta^2 - 1 / m^2) f, the derivative terms of
// ewald_point_impl) to d0..d5 (xx, yy, zz, xy, xz, yz); the plane's
// multiplicity is left to the caller.  Points on no Nyquist plane have one
// image and skip the image loop.
__device__ __forceinline__ float xline_point_acc(const XLineF32 &L, const FrameDesc &F, int kx,
                                                 float sq, float &d0, float &d1, float &d2,
                                                 float &d3, float &d4, float &d5) {
    const int nx = kx <= L.half0 ? kx : kx - L.K0;
    const bool xny = L.even0 && kx == L.half0;
    const float X0 = (float)nx * L.iL00, X1 = (float)nx * L.iL10, X2 = (float)nx * L.iL20;
    float a = 0.0f;
    auto image = [&](int sx, int sy, int sz) {
        const float fx = (float)sx, fy = (float)sy, fz = (float)sz;
        const float mx = fx * X0;
        const float my = fx * X1 + fy * L.Y1;
        const float mz = fx * X2 + fy * L.Y2 + fz * L.Z;
        const float m2 = mx * mx + my * my + mz * mz;
        bool in = m2 >= 1e-10f && m2 <= L.max_m2f;
        if (fabsf(m2 - L.max_m2f) <= 1e-6f * L.max_m2f) {
#pragma clang fp contract(off)
            const CellArgs &c = F.tcell;
            const double nxd = sx * nx, nyd = sy * L.ny, nzd = sz * L.nz;
            const double mxd = nxd * c.iL00;
            const double myd = nxd * c.iL10 + nyd * c.iL11;
            const double mzd = nxd * c.iL20 + nyd * c.iL21 + nzd * c.iL22;
            const double m2d = mxd * mxd + myd * myd + mzd * mzd;
            in = !(m2d < 1e-10 || m2d > F.max_m2);
        }
        if (in) {
            const float r = __builtin_amdgcn_rcpf(m2);
            const float f = __builtin_amdgcn_exp2f(m2 * L.fac2) * r * 0.318309886183790672f;
            const float sfp = sq * ((L.fac - r) * f);
            const float px = sfp * mx, py = sfp * my;
            a += f;
            d0 += px * mx;
            d1 += py * my;
            d2 += sfp * mz * mz;
            d3 += py * mx;
            d4 += px * mz;
            d5 += py * mz;
        }
    };
    if (!(xny || L.yny || L.zn2)) {
        image(1, 1, 1);
        return a;
    }
    const int nimg = (xny ? 2 : 1) * (L.yny ? 2 : 1) * (L.zn2 ? 2 : 1);
    for (int k = 0; k < nimg; ++k) {
        int kk = k;
        int sx = 1, sy = 1, sz = 1;
        if (xny) {
            sx = (kk & 1) ? -1 : 1;
            kk >>= 1;
        }
        if (L.yny) {
            sy = (kk & 1) ? -1 : 1;
            kk >>= 1;
        }
        if (L.zn2) sz = (kk & 1) ? -1 : 1;
        image(sx, sy, sz);
    }
    return a;
}

// e^x for x <= 0 in FP64 with few registers: Cody-Waite reduction
// x = k ln2 + r, |r| <= ln2/2, with ln2 split into a 32-bit head (k ln2_hi
// exact for |k| < 2^21) and its remainder, then the Taylor series of e^r to
// degree 13 (truncation below 2e-17 relative) by Horner, scaled by 2^k.  A
// result below the FP64 range is flushed to zero.
__device__ __forceinline__ double exp_nonpos(double x) {
    constexpr double kLn2Hi = 0.693147180369123816490;  // ln 2 to 32 bits
    constexpr double kLn2Lo = 1.90821492927058770002e-10;
    constexpr double kInvLn2 = 1.44269504088896338700;
    const double k = rint(x * kInvLn2);
    const double r = (x - k * kLn2Hi) - k * kLn2Lo;
    double p = 1.0 / 6227020800.0;  // 1/13!
    p = fma(p, r, 1.0 / 479001600.0);
    p = fma(p, r, 1.0 / 39916800.0);
    p = fma(p, r, 1.0 / 3628800.0);
    p = fma(p, r, 1.0 / 362880.0);
    p = fma(p, r, 1.0 / 40320.0);
    p = fma(p, r, 1.0 / 5040.0);
    p = fma(p, r, 1.0 / 720.0);
    p = fma(p, r, 1.0 / 120.0);
    p = fma(p, r, 1.0 / 24.0);
    p = fma(p, r, 1.0 / 6.0);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return k < -1074.0 ? 0.0 : ldexp(p, (int)k);
}

// m of image (ix, iy, iz) in ewald_image's arithmetic (same products and sums,
// no contraction), so a ball test on it decides as the table builder does
__device__ __forceinline__ void xline64_m(const CellArgs &c, int ix, int iy, int iz, double &mx,
                                          double &my, double &mz, double &m2) {
#pragma clang fp contract(off)
    const double nx = ix, ny = iy, nz = iz;
    mx = nx * c.iL00;
    my = nx * c.iL10 + ny * c.iL11;
    mz = nx * c.iL20 + ny * c.iL21 + nz * c.iL22;
    m2 = mx * mx + my * my + mz * mz;
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

// Block-reduce acc[7] of the kBlock threads into one partial per workgroup,
// F.partials[b] = sums, in a fixed order: the values go to LDS (`scratch`,
// kPartialScratch bytes that no thread reads any more), then 32 threads per
// component add 8 consecutive values each and a 32-lane shuffle tree adds
// those.  Every thread of the workgroup must call it.
constexpr size_t kPartialScratch = 7 * kBlock * sizeof(double);
__device__ __forceinline__ void block_partial(const FrameDesc &F, int b, const double *acc,
                                              double *scratch) {
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 7; ++c) scratch[c * kBlock + threadIdx.x] = acc[c];
    __syncthreads();
    if (threadIdx.x < 7 * 32) {
        const int c = threadIdx.x >> 5, j = threadIdx.x & 31;
        const double *src = scratch + c * kBlock + j * (kBlock / 32);
        double v = 0.0;
#pragma unroll
        for (int k = 0; k < kBlock / 32; ++k) v += src[k];
#pragma unroll
        for (int off = 16; off > 0; off >>= 1) v += __shfl_xor(v, off, 32);
        if (j == 0) F.partials[b * 8 + c] = v;
    }
}

// Sum the nparts partials of the previous launch in a fixed order
// (deterministic) and write E = iV E_raw and the virial
// (v_ab + delta_ab E_raw) iV^2 (sfac.cpp:384-395).  The first kBlock threads
// of the workgroup do the sums, so the order is the same in workgroups of
// any size >= kBlock; every thread must reach the barrier.
__device__ __forceinline__ void finalize_reduction(const FrameDesc &F, int nparts) {
    __shared__ double red2[kBlock / 64][8];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    double part[7] = {0, 0, 0, 0, 0, 0, 0};
    if (threadIdx.x < kBlock)
        for (int q = threadIdx.x; q < nparts; q += kBlock)
#pragma unroll
            for (int c = 0; c < 7; ++c) part[c] += F.partials[q * 8 + c];
#pragma unroll
    for (int c = 0; c < 7; ++c) {
        const double v = wave_sum(part[c]);
        if (lane == 0 && wv < kBlock / 64) red2[wv][c] = v;
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
    __shared__ double scratch[7 * kBlock];
    block_partial(F, b, acc, scratch);
}

// ---- y lines (used when the y-z slab does not fit in LDS) --------------------------------
// Lines (x = o, kz) along y, o < K0, element stride ldim.  A frame's blocks
// are (o, tile); tile t covers kz in [t ldim/T, (t+1) ldim/T).  finish 1:
// workgroup 0 of each frame first finalizes the reduction of the fused x
// pass that ran before.
template <typename T>
__host__ __device__ inline size_t ylines_lds_bytes(int N, int C, int generic) {
    return fft_lds_bytes<T>(N, C, generic);
}

template <typename T>
__global__ void __launch_bounds__(kBlock) fft_y_kernel(SPME_FRAME_ARGS, int sign, int finish) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    int b;
    const FrameDesc &F = frames[find_frame(offs, nf, b)];
    if (finish && b == 0) {
        finalize_reduction(F, F.xB);
        __syncthreads();
    }
    const FftPlanArgs &p = F.plan1;
    const int N = p.N, LS = N + 1, ldim = F.ldim;
    const int Tt = F.yT;
    const long estride = ldim;
    const int o = b / Tt, t = b - o * Tt;
    const int c0 = (int)((long)t * ldim / Tt), c1 = (int)((long)(t + 1) * ldim / Tt);
    const int nc = c1 - c0;
    cplx<T> *twl = reinterpret_cast<cplx<T> *>(smem_raw);
    cplx<T> *b0 = twl + N;
    cplx<T> *b1 = b0 + F.yC * LS;
    cplx<T> *base = (cplx<T> *)F.grid + (long)o * F.K1 * ldim + c0;
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
        load_twiddles<T>(twl, F.tw1, N);
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
    cplx<T> *res = fft_lines<T>(b0, b1, nc, LS, p, twl, sign);
    for (int l = threadIdx.x; l < tot; l += blockDim.x) {
        const int e = l / nc, c = l - e * nc;
        base[(long)e * estride + c] = res[c * LS + e];
    }
}

// ---- x lines ------------------------------------------------------------------------------
// Lines along x (element stride P = K1 ldim complex) for C = 2^xLog
// consecutive columns f = ky ldim + kz of the (ky, kz) plane: each x element
// of a workgroup is one contiguous C-element run in HBM.  Thread t owns
// column c = t mod C for the whole launch (its ky, kz and 1/B factors are
// computed once) and elements e = t div C + q (256 / C).
// MODE 0: transform with `sign` (finish 2: workgroup 0 of each frame first
//         finalizes the standalone influence reduction).
// MODE 1: forward transform, outputs scaled by iB0 iB1 iB2 (S-hat).
// MODE 2: forward, /B, influence with reductions, inverse -- the whole
//         reciprocal-space step of an evaluation in one HBM pass.
// LDS: twiddles (N), iB0 (N doubles, padded to even), C lines of stride N|1
// (two sets for a plan with a radix other than 2, 3, 4, 5, 8, which runs
// the ping-pong line transform).
constexpr int kXElems = 2048;  // complex elements per x workgroup (8 per thread)

__host__ __device__ inline int xline_stride(int N) { return N | 1; }

template <typename T>
__host__ __device__ inline size_t xlines_lds_bytes(int N, int C, int generic) {
    return ((size_t)N + (size_t)C * xline_stride(N) * (generic ? 2 : 1)) * sizeof(cplx<T>) +
           (size_t)((N + 1) & ~1) * sizeof(double);
}

// x transform of the C lines in LDS by the whole workgroup (plans with a
// generic radix, or fewer than four lines); returns the buffer holding the
// result
template <typename T, bool GENERIC, bool OPQ = false>
__device__ __forceinline__ cplx<T> *x_transform(cplx<T> *buf, cplx<T> *other, int C,
                                                const FastDiv &dC, int LS, const FftPlanArgs &p,
                                                const cplx<T> *tw, int sign) {
    if (GENERIC) return fft_lines<T, kXElems / kBlock>(buf, other, C, LS, p, tw, sign);
    lines_fft<T, kXElems / kBlock, kBlock, OPQ>(buf, C, dC, LS, 1, p, tw, sign);
    return buf;
}

// transform of the x launches: the workgroup's lines with radices 2..8, or
// the ping-pong transform for any radix (a separate instantiation each, so
// that neither holds the other's registers)
enum XPath { kXBlock = 0, kXGeneric = 1 };
__host__ __device__ inline int x_path(const FrameDesc &F) {
    return F.plan0.generic ? kXGeneric : kXBlock;
}

// Wavefront sums of the fused x pass, parked in LDS after the lines (one row
// of 8 doubles per wavefront) between the influence step and the inverse
// transform, so that no sums are held in registers through the inverse.
constexpr size_t kXWaveSums = (kBlock / 64) * 8 * sizeof(double);

// The rest of the fused x pass's LDS, after the wavefront sums, for the
// closed form of x-separable boxes: gx[|nx|] = e^(fac iL00^2 nx^2) for
// |nx| <= N/2 (padded to even) and kXColConsts constants per column (the
// line's class constants, exponentials, 1/B of y and z and multiplicity).
constexpr int kXColConsts = 6;
__host__ __device__ inline size_t xfused_gx_doubles(int N) { return (size_t)((N / 2 + 2) & ~1); }
__host__ __device__ inline size_t xfused_extra_lds(int N, int C) {
    return kXWaveSums + (xfused_gx_doubles(N) + (size_t)kXColConsts * C) * sizeof(double);
}

// The influence step of the fused x pass on the transformed lines in LDS:
// S-hat = S iB0 iB1 iB2, the energy and virial sums of A |S-hat|^2 (thread
// sums in acc[7]) and S-hat A iB (x2 on the kz = 0 and Nyquist planes, which
// stand for their Hermitian partners) written back for the inverse.  Thread
// (c, e0) holds column f = b C + c (ky, kz) and elements e0 + q es.
template <typename T, int INF>
__device__ __forceinline__ void x_influence(const FrameDesc &F, cplx<T> *res,
                                            const double *ib0, int LS, int c, int e0, int es,
                                            int f, bool valid, int ky, int kz, double ibyz,
                                            double acc[7]) {
    constexpr int EPT = kXElems / kBlock;
    const int N = F.plan0.N;
    const bool even = (F.K2 % 2) == 0;
    const double two = (kz == 0 || (even && kz == F.ldim - 1)) ? 2.0 : 1.0;
    if (INF && std::is_same<T, float>::value && F.tabmode == kTableAnalytic) {
        // FP32 engine with the FP32 influence option, closed form: FP32 terms,
        // each thread's sums (at most kXElems / kBlock points) in FP32, the
        // wavefront and block sums in FP64.  With a
        // the sum of f = exp(-pi^2 m^2 / beta^2) / (pi m^2) over the point's
        // images: A = mult a / 2 and dA = mult sum fp m m^T, and since
        // mult x two = 2 on every plane, the output factor A iB two is a iB;
        // mult is applied to the thread's sums at the end.
        const float ibyzf = (float)ibyz;
        XLineF32 L;
        xline_init(L, F, ky, kz);
        float aE = 0.f, d0 = 0.f, d1 = 0.f, d2 = 0.f, d3 = 0.f, d4 = 0.f, d5 = 0.f;
#pragma unroll 1
        for (int q = 0; q < EPT; ++q) {
            const int e = e0 + q * es;
            if (e < N && valid) {
                const float ib = (float)ib0[e] * ibyzf;
                cplx<T> v = res[c * LS + e];
                const float re = (float)v.x * ib, im = (float)v.y * ib;  // S-hat
                const float sq = re * re + im * im;
                const float a = xline_point_acc(L, F, e, sq, d0, d1, d2, d3, d4, d5);
                aE += a * sq;
                const float tt = a * ib;
                v.x = T(re * tt);
                v.y = T(im * tt);
                res[c * LS + e] = v;
            }
        }
        acc[0] = (double)(0.5f * L.mult * aE);
        acc[1] = (double)(L.mult * d0);
        acc[2] = (double)(L.mult * d1);
        acc[3] = (double)(L.mult * d2);
        acc[4] = (double)(L.mult * d3);
        acc[5] = (double)(L.mult * d4);
        acc[6] = (double)(L.mult * d5);
    } else {
        // FP64 /B, table influence and thread sums (each table entry added
        // to the sums as it is read)
        const long H = (long)F.K0 * F.K1 * F.ldim;
#pragma unroll
        for (int k = 0; k < 7; ++k) acc[k] = 0.0;
#pragma unroll 1
        for (int q = 0; q < EPT; ++q) {
            const int e = e0 + q * es;
            if (e < N && valid) {
                const double ib = ib0[e] * ibyz;
                cplx<T> v = res[c * LS + e];
                const double re = (double)v.x * ib, im = (double)v.y * ib;  // S-hat
                const double s = re * re + im * im;
                const long h = (long)e * F.K1 * F.ldim + f;
                const double A = F.table[h];
                const double *d = F.table + H + 6 * h;
                acc[0] += A * s;
#pragma unroll
                for (int k = 0; k < 6; ++k) acc[1 + k] += s * d[k];
                const double tt = A * ib * two;
                v.x = T(re * tt);
                v.y = T(im * tt);
                res[c * LS + e] = v;
            }
        }
    }
}

// FP64 closed-form influence of the thread's x line (its column: fixed ky,
// kz) in the fused x pass, in both engines (the FP32 engine rounds only the
// line values it writes back): /B, A, the energy and the virial in FP64.
//   * An image of point kx is m = nx a + sy u + sz z with a = (iL00, iL10,
//     iL20), u = ny (0, iL11, iL21), z = nz (0, 0, iL22) and signs sy, sz = +1
//     except on the y and z Nyquist lines (+-1 images there).  Its virial
//     term w m_a m_b (w = mult s fp) expands into five sums over the line,
//     W2 = sum w nx^2, Wy = sum w nx sy, Wz = sum w nx sz, W0 = sum w,
//     Wyz = sum w sy sz, folded once per thread:
//       v_ab = W2 a_a a_b + Wy (a_a u_b + a_b u_a) + Wz (a_a z_b + a_b z_a)
//              + W0 (u_a u_b + z_a z_b) + Wyz (u_a z_b + z_a u_b).
//   * m^2 = q00 nx^2 + 2 nx c1 + c0 with the class constants c0 = |b|^2,
//     c1 = a.b of b = sy u + sz z; it is formed with contractions, and within
//     a relative 1e-12 of the cutoff the ball test is redone exactly as
//     ewald_image computes m, so the set of images is the table's.
//   * 1/m^2 by the hardware reciprocal and two Newton steps.
// Boxes whose x axis separates (iL10 = iL20 = 0, x_influence64_xsep) have
// c1 = 0 and e^(fac m^2) = gx[kx] e^(fac |b|^2), gx[kx] = e^(fac iL00^2 nx^2)
// formed once per workgroup in LDS; |b|^2 depends on the class only through
// sy sz, so each thread takes at most two exponentials for its whole line and
// none per point.  Other boxes (x_influence64_gen) take e^(fac m^2) per image.
// The wavefront sums of E and the virial go to wsums[wavefront][0..6].
__device__ __forceinline__ double recip_f64(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(fma(-x, r, 1.0), r, r);
    return fma(fma(-x, r, 1.0), r, r);
}

// Line constants and the per-image accumulation shared by both forms.
struct XLine64 {
    int ny;
    bool yny, zn2;
    double mult;
    double W0 = 0.0, W2 = 0.0, Wy = 0.0, Wz = 0.0, Wyz = 0.0, aE = 0.0;
    __device__ __forceinline__ XLine64(const FrameDesc &F, int ky, int kz) {
        ny = ky <= F.K1 / 2 ? ky : ky - F.K1;
        yny = (F.K1 % 2) == 0 && ky == F.K1 / 2;
        const bool zny = (F.K2 % 2) == 0 && kz == F.K2 / 2;
        zn2 = zny && kz > 0;
        mult = (kz == 0 || zny) ? 1.0 : 2.0;
    }
    // b = sy u + sz z of image class (sy, sz): its y and z components
    __device__ __forceinline__ void bvec(const CellArgs &cl, int kz, int sy, int sz, double &B1,
                                         double &B2) const {
        const double nyd = (double)(sy * ny), nzd = (double)(sz * kz);
        B1 = nyd * cl.iL11;
        B2 = fma(nyd, cl.iL21, nzd * cl.iL22);
    }
    // the ball test of image nx of class (sy, sz), m^2 returned
    __device__ __forceinline__ bool in_ball(const CellArgs &cl, double q00, double max_m2,
                                            int nx, int sy, int sz, int kz, double cc0,
                                            double cc1, double &m2) const {
        const double nxd = nx;
        m2 = fma(fma(q00, nxd, 2.0 * cc1), nxd, cc0);
        bool in = !(m2 < 1e-10 || m2 > max_m2);
        if (fabs(m2 - max_m2) <= 1e-12 * max_m2) {
            double mx, my, mz, m2e;
            xline64_m(cl, nx, sy * ny, sz * kz, mx, my, mz, m2e);
            in = !(m2e < 1e-10 || m2e > max_m2);
        }
        return in;
    }
    // the terms of one image (e^(fac m^2) / pi given): f returned, moments added
    __device__ __forceinline__ double add_image(double fac, int nx, int sy, int sz, double m2,
                                                double ex, double s) {
        const double r = recip_f64(m2);
        const double f = ex * r;
        const double w = mult * s * ((fac - r) * f);
        const double nxd = nx;
        const double t = w * nxd;
        W0 += w;
        W2 = fma(t, nxd, W2);
        Wy = sy > 0 ? Wy + t : Wy - t;
        Wz = sz > 0 ? Wz + t : Wz - t;
        Wyz = sy * sz > 0 ? Wyz + w : Wyz - w;
        return f;
    }
    // image k of a Nyquist point: its signed x index and class
    __device__ __forceinline__ void nyq_image(int k, bool xny, int nx, int &nxs, int &sy,
                                              int &sz) const {
        int kk = k;
        nxs = nx;
        sy = 1;
        sz = 1;
        if (xny) {
            if (kk & 1) nxs = -nx;
            kk >>= 1;
        }
        if (yny) {
            if (kk & 1) sy = -1;
            kk >>= 1;
        }
        if (zn2 && (kk & 1)) sz = -1;
    }
    // fold of the sums into E and the virial, then the wavefront sums:
    // a = (iL00, iL10, iL20), u = (0, uy, uz), z = (0, 0, zz)
    __device__ __forceinline__ void finish(const CellArgs &cl, int kz, double *wsums) const {
        const double uy = (double)ny * cl.iL11, uz = (double)ny * cl.iL21;
        const double zz = (double)kz * cl.iL22;
        const double a0 = cl.iL00, a1 = cl.iL10, a2 = cl.iL20;
        double vv[7];
        vv[0] = 0.5 * mult * aE;
        vv[1] = W2 * a0 * a0;                                                // xx
        vv[2] = W2 * a1 * a1 + 2.0 * Wy * a1 * uy + W0 * uy * uy;            // yy
        vv[3] = W2 * a2 * a2 + 2.0 * (Wy * a2 * uz + Wz * a2 * zz) +
                W0 * (uz * uz + zz * zz) + 2.0 * Wyz * uz * zz;              // zz
        vv[4] = W2 * a1 * a0 + Wy * a0 * uy;                                 // yx
        vv[5] = W2 * a2 * a0 + Wy * a0 * uz + Wz * a0 * zz;                  // zx
        vv[6] = W2 * a2 * a1 + Wy * (a2 * uy + a1 * uz) + Wz * a1 * zz +
                W0 * uy * uz + Wyz * uy * zz;                                // zy
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            const double vs = wave_sum(vv[k]);
            if (lane == 0) wsums[wv * 8 + k] = vs;
        }
    }
};

// x-separable boxes (iL10 = iL20 = 0).  Every image of a point has the
// same m^2 within its group (the classes with sy sz = g, g = +-1: b of one
// is the exact negation of the other's, c1 = 0, and nx -> -nx leaves m^2
// unchanged), so each point evaluates one image per group and adds it with
// its multiplicity and the group's sign sums: images nxs sy sz with nxs in
// {nx} (or {nx, -nx} on the x Nyquist plane, xf = 2) and the line's pairs
// (sy, sz) of the group (np_g pairs, sign sums Sy_g, Sz_g):
//   a += xf np_g f,  W0 += xf np_g w,  W2 += xf np_g w nx^2,
//   Wy += w (sum nxs) Sy_g,  Wz += w (sum nxs) Sz_g,  Wyz += xf g np_g w.
// The column constants of the x-separable form (one thread per column):
// k[0] = c0 = |b|^2 of group +, k[1] = cm of group - (c0 on lines that have
// no group -), k[2], k[3] their e^(fac |b|^2) / pi, k[4] = iB1[ky] iB2[kz],
// k[5] = mult.
__device__ __forceinline__ void xsep_column(const FrameDesc &F, int ky, int kz, double *k) {
    const CellArgs &cl = F.tcell;
    const double fac = F.ewald_fac;
    const XLine64 L(F, ky, kz);
    double B1, B2;
    L.bvec(cl, kz, 1, 1, B1, B2);
    const double c0 = fma(B1, B1, B2 * B2);
    const double el = exp_nonpos(fac * c0) * (1.0 / M_PI);
    double cm = c0, elm = el;
    if (L.yny || L.zn2) {
        L.bvec(cl, kz, -1, 1, B1, B2);
        cm = fma(B1, B1, B2 * B2);
        elm = exp_nonpos(fac * cm) * (1.0 / M_PI);
    }
    k[0] = c0;
    k[1] = cm;
    k[2] = el;
    k[3] = elm;
    k[4] = F.iB1[ky] * F.iB2[kz];
    k[5] = L.mult;
}

// One point kx = e of an x line of the x-separable form: S-hat = v iB (in
// FP64), the images of each group with their multiplicities and sign sums
// added to L, and v A iB written back (A holds the line's mult / 2).
template <typename T>
__device__ __forceinline__ void xsep_point(XLine64 &L, const FrameDesc &F, const double *gx,
                                           const double *k, double q00, bool both, int ng,
                                           int kz, int e, double ib0e, cplx<T> &v) {
    const int K0 = F.K0;
    const CellArgs &cl = F.tcell;
    const double fac = F.ewald_fac, max_m2 = F.max_m2;
    const double ib = ib0e * k[4];
    const double re = (double)v.x * ib, im = (double)v.y * ib;  // S-hat
    const double s = re * re + im * im;
    const int nx = e <= K0 / 2 ? e : e - K0;
    const bool xny = (K0 % 2) == 0 && e == K0 / 2;
    double a = 0.0;
#pragma unroll 1
    for (int gi = 0; gi < ng; ++gi) {
        // group g: the image checked is (nx, sy, sz) = (nx, 1, 1), or for
        // g = -1 (nx, -1, 1) on a y Nyquist line, else (nx, 1, -1)
        const int g = gi ? -1 : 1;
        const int sy = L.yny ? g : 1, sz = L.yny ? 1 : g;
        double m2;
        if (!L.in_ball(cl, q00, max_m2, nx, sy, sz, kz, k[gi], 0.0, m2)) continue;
        const double r = recip_f64(m2);
        const double f = gx[nx < 0 ? -nx : nx] * k[2 + gi] * r;
        const double w = k[5] * s * ((fac - r) * f);
        const double nxd = nx;
        // multiplicity xf np of the group's images, and w times the sums
        // of nxs sy and nxs sz over them
        const int mul = (xny ? 2 : 1) * (both ? 2 : 1);
        const int ssy = (xny || both) ? 0 : nx * sy, ssz = (xny || both) ? 0 : nx * sz;
        const double mw = (double)mul * w;
        L.W0 += mw;
        L.W2 = fma(mw * nxd, nxd, L.W2);
        L.Wy = fma(w, (double)ssy, L.Wy);
        L.Wz = fma(w, (double)ssz, L.Wz);
        L.Wyz = gi ? L.Wyz - mw : L.Wyz + mw;
        a = fma((double)mul, f, a);
    }
    L.aE = fma(a, s, L.aE);
    const double tt = a * ib;
    v.x = T(re * tt);
    v.y = T(im * tt);
}

template <typename T>
__device__ __forceinline__ void x_influence64_xsep(const FrameDesc &F, cplx<T> *res,
                                                   const double *ib0, const double *gx,
                                                   const double *colk, int LS, int c, int e0,
                                                   int es, bool valid, int ky, int kz,
                                                   double *wsums) {
    constexpr int EPT = kXElems / kBlock;
    const int N = F.plan0.N, K0 = F.K0;
    const CellArgs &cl = F.tcell;
    const double fac = F.ewald_fac, max_m2 = F.max_m2;
    const double q00 = cl.iL00 * cl.iL00 + cl.iL10 * cl.iL10 + cl.iL20 * cl.iL20;
    XLine64 L(F, ky, kz);
    // pairs of the groups: no Nyquist line: (+,+); y only: (+,+) | (-,+);
    // z only: (+,+) | (+,-); both: (+,+),(-,-) | (+,-),(-,+)
    const bool both = L.yny && L.zn2;
    const int ng = (L.yny || L.zn2) ? 2 : 1;
    const double *k = colk + kXColConsts * c;
#pragma unroll 1
    for (int q = 0; q < EPT; ++q) {
        const int e = e0 + q * es;
        if (e >= N || !valid) continue;
        cplx<T> v = res[c * LS + e];
        xsep_point<T>(L, F, gx, k, q00, both, ng, kz, e, ib0[e], v);
        res[c * LS + e] = v;
    }
    L.finish(cl, kz, wsums);
}

// any box: e^(fac m^2) per image
template <typename T>
__device__ __forceinline__ void x_influence64_gen(const FrameDesc &F, cplx<T> *res,
                                                  const double *ib0, int LS, int c, int e0,
                                                  int es, bool valid, int ky, int kz,
                                                  double ibyz, double *wsums) {
    constexpr int EPT = kXElems / kBlock;
    const int N = F.plan0.N, K0 = F.K0;
    const CellArgs &cl = F.tcell;
    const double fac = F.ewald_fac, max_m2 = F.max_m2;
    const double q00 = cl.iL00 * cl.iL00 + cl.iL10 * cl.iL10 + cl.iL20 * cl.iL20;
    XLine64 L(F, ky, kz);
    // the line's own class (+, +)
    double c0, c1;
    {
        double B1, B2;
        L.bvec(cl, kz, 1, 1, B1, B2);
        c0 = fma(B1, B1, B2 * B2);
        c1 = fma(cl.iL10, B1, cl.iL20 * B2);
    }
    const bool nyq_line = L.yny || L.zn2;
#pragma unroll 1
    for (int q = 0; q < EPT; ++q) {
        const int e = e0 + q * es;
        if (e >= N || !valid) continue;
        const double ib = ib0[e] * ibyz;
        cplx<T> v = res[c * LS + e];
        const double re = (double)v.x * ib, im = (double)v.y * ib;  // S-hat
        const double s = re * re + im * im;
        const int nx = e <= K0 / 2 ? e : e - K0;
        const bool xny = (K0 % 2) == 0 && e == K0 / 2;
        double a = 0.0;
        if (!(xny || nyq_line)) {
            // one image, the line's own class
            double m2;
            if (L.in_ball(cl, q00, max_m2, nx, 1, 1, kz, c0, c1, m2))
                a = L.add_image(fac, nx, 1, 1, m2, exp_nonpos(m2 * fac) * (1.0 / M_PI), s);
        } else {
            // Nyquist planes: every signed image, its class constants anew
            const int nimg = (xny ? 2 : 1) * (L.yny ? 2 : 1) * (L.zn2 ? 2 : 1);
#pragma unroll 1
            for (int k = 0; k < nimg; ++k) {
                int nxs, sy, sz;
                L.nyq_image(k, xny, nx, nxs, sy, sz);
                double B1, B2, m2;
                L.bvec(cl, kz, sy, sz, B1, B2);
                const double cc0 = fma(B1, B1, B2 * B2);
                const double cc1 = fma(cl.iL10, B1, cl.iL20 * B2);
                if (!L.in_ball(cl, q00, max_m2, nxs, sy, sz, kz, cc0, cc1, m2)) continue;
                a += L.add_image(fac, nxs, sy, sz, m2, exp_nonpos(m2 * fac) * (1.0 / M_PI), s);
            }
        }
        L.aE = fma(a, s, L.aE);
        const double tt = a * ib;
        v.x = T(re * tt);
        v.y = T(im * tt);
        res[c * LS + e] = v;
    }
    L.finish(cl, kz, wsums);
}

// The influence of the fused x pass is a template choice (INF): 0 the FP64
// forms for any frame (table, or the closed form of either box class), 1 the
// FP32 closed form of the FP32 engine's option (table frames still take the
// FP64 table path), 2 the FP64 closed form of x-separable boxes only.  Each is
// its own instantiation, so that none holds another's registers; a frame's
// results do not depend on which of 0 and 2 it runs in (same arithmetic).
// Waves per SIMD the fused x launches are held to (registers): FP32 at 64
// VGPRs (line transforms with per-stage indices) runs eight workgroups per
// CU, with the FP32 influence and with the FP64 x-separable one; the general
// FP64 form (a per-image exponential) needs ~96 VGPRs (at six or more waves
// it spills inside the point loop), five waves.
constexpr int kXWaves64 = 5;
template <typename T, int MODE> constexpr bool x_lean() {
    return MODE == 2 && std::is_same<T, float>::value;
}
template <typename T, int MODE, int INF, int XP> constexpr int x_waves() {
    // the ping-pong transform of generic radices holds a little more
    return x_lean<T, MODE>() ? (INF == 0 ? kXWaves64 : (INF == 2 && XP == 1) ? 7 : 8) : 1;
}

template <typename T, int MODE, int XP, int INF>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu((x_waves<T, MODE, INF, XP>()))))
    fft_x_kernel(SPME_FRAME_ARGS, int sign, int finish) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int EPT = kXElems / kBlock;
    int b;
    const FrameDesc &F = frames[find_frame(offs, nf, b)];
    if (finish && b == 0) {
        finalize_reduction(F, F.nred_i);
        __syncthreads();
    }
    const FftPlanArgs &p = F.plan0;
    const int N = p.N, LS = xline_stride(N), ldim = F.ldim;
    const int C = F.xC, lg = F.xLog;
    const int P = F.K1 * ldim;
    cplx<T> *twl = reinterpret_cast<cplx<T> *>(smem_raw);
    double *ib0 = reinterpret_cast<double *>(twl + N);
    cplx<T> *lines = reinterpret_cast<cplx<T> *>(ib0 + ((N + 1) & ~1));
    cplx<T> *lines2 = lines + C * LS;  // generic plans only
    // HBM <-> LDS mapping: column c = t mod C, elements t div C + q (256 / C)
    const int c = threadIdx.x & (C - 1);
    const int e0 = threadIdx.x >> lg, es = kBlock >> lg;
    const int fc = min(b * C + c, P - 1);  // clamped for loads
    cplx<T> *col = (cplx<T> *)F.grid + fc;
    {
        T vx[EPT], vy[EPT];
#pragma unroll
        for (int q = 0; q < EPT; ++q) {
            const int e = min(e0 + q * es, N - 1);
            const cplx<T> w = col[(long)e * P];
            vx[q] = w.x;
            vy[q] = w.y;
        }
        load_twiddles<T>(twl, F.tw0, N);
        if (MODE != 0)
            for (int e = threadIdx.x; e < N; e += kBlock) ib0[e] = F.iB0[e];
        if (MODE == 2 && INF != 1 && F.tabmode == kTableAnalytic && x_separable(F.tcell)) {
            // gx[|nx|] = e^(fac iL00^2 nx^2) and the column constants, after the
            // lines and the wavefront sums
            double *gx = reinterpret_cast<double *>(
                smem_raw + xlines_lds_bytes<T>(N, C, XP == kXGeneric) + kXWaveSums);
            const double q00 = F.tcell.iL00 * F.tcell.iL00;
            for (int e = threadIdx.x; e <= N / 2; e += kBlock) {
                const double nx = e;
                gx[e] = exp_nonpos(F.ewald_fac * q00 * nx * nx);
            }
            if ((int)threadIdx.x < C) {
                const int fk = min(b * C + (int)threadIdx.x, P - 1);
                const int ky = fdiv(fk, F.dldim);
                xsep_column(F, ky, fk - ky * ldim,
                            gx + xfused_gx_doubles(N) + kXColConsts * threadIdx.x);
            }
        }
#pragma unroll
        for (int q = 0; q < EPT; ++q) {
            const int e = e0 + q * es;
            if (e < N) {
                cplx<T> w;
                w.x = vx[q];
                w.y = vy[q];
                lines[c * LS + e] = w;
            }
        }
    }
    __syncthreads();
    const int s0 = MODE == 0 ? sign : -1;
    // power of two: n div C = n >> lg (fdiv: mulhi(n, 1) = 0; fdiv15: n 2^15 >> (15 + lg))
    const FastDiv dC = {(unsigned)C, 1u, (unsigned)lg, 1u << 15, 15u + (unsigned)lg};
    cplx<T> *res = x_transform<T, XP == kXGeneric, x_lean<T, MODE>()>(lines, lines2, C, dC, LS, p, twl, s0);
    if (MODE == 2) {
        const int f = b * C + c;
        const bool valid = f < P;
        const int fv = valid ? f : P - 1;
        const int ky = fdiv(fv, F.dldim), kz = fv - ky * ldim;
        double *wsums = reinterpret_cast<double *>(
            smem_raw + xlines_lds_bytes<T>(N, C, XP == kXGeneric));
        const double *gx = wsums + 8 * (kBlock / 64);
        const double *colk = gx + xfused_gx_doubles(N);
        if constexpr (INF == 2) {
            x_influence64_xsep<T>(F, res, ib0, gx, colk, LS, c, e0, es, valid, ky, kz, wsums);
        } else if (!INF && F.tabmode == kTableAnalytic) {
            if (x_separable(F.tcell))
                x_influence64_xsep<T>(F, res, ib0, gx, colk, LS, c, e0, es, valid, ky, kz,
                                      wsums);
            else
                x_influence64_gen<T>(F, res, ib0, LS, c, e0, es, valid, ky, kz,
                                     F.iB1[ky] * F.iB2[kz], wsums);
        } else {
            double acc[7];
            x_influence<T, INF>(F, res, ib0, LS, c, e0, es, f, valid, ky, kz,
                                F.iB1[ky] * F.iB2[kz], acc);
            // wavefront sums (a fixed butterfly order) parked in LDS
            const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
            for (int k = 0; k < 7; ++k) {
                const double v = wave_sum(acc[k]);
                if (lane == 0) wsums[wv * 8 + k] = v;
            }
        }
        __syncthreads();
        res = x_transform<T, XP == kXGeneric, x_lean<T, MODE>()>(res, res == lines ? lines2 : lines, C, dC, LS, p,
                                              twl, +1);
        // the store addresses are recomputed here rather than kept (in
        // registers, or spilled) from the loads at the start
        int cs = c, es0 = e0;
        asm volatile("" : "+v"(cs), "+v"(es0));
        const bool vs = b * C + cs < P;
        cplx<T> *cols = (cplx<T> *)F.grid + min(b * C + cs, P - 1);
#pragma unroll
        for (int q = 0; q < EPT; ++q) {
            const int e = es0 + q * es;
            if (e < N && vs) cols[(long)e * P] = res[cs * LS + e];
        }
        // the workgroup's partial: its wavefront sums in order (the inverse
        // transform's barriers made them visible)
        if (threadIdx.x < 7) {
            double v = 0.0;
#pragma unroll
            for (int q = 0; q < kBlock / 64; ++q) v += wsums[q * 8 + threadIdx.x];
            F.partials[b * 8 + threadIdx.x] = v;
        }
        return;
    }
    __syncthreads();
    const bool vs = b * C + c < P;
    const int ky = fdiv(fc, F.dldim), kz = fc - ky * ldim;
    const double ibyz = (MODE == 1) ? F.iB1[ky] * F.iB2[kz] : 1.0;
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
        const int e = e0 + q * es;
        if (e < N && vs) {
            cplx<T> v = res[c * LS + e];
            if (MODE == 1) {
                const T s = T(ib0[e] * ibyz);
                v.x *= s;
                v.y *= s;
            }
            col[(long)e * P] = v;
        }
    }
}

// ---- register x pass: 128-point lines as 16 x 8 --------------------------------------------
// The fused x pass (forward, /B, FP64 influence of x-separable boxes, inverse)
// of FP32 frames whose x plan is radix 16 then 8 (N = 128), with the 16- and
// 8-point transforms of each line in registers and two exchanges through LDS
// in place of the LDS-staged stages (four LDS accesses per element instead of
// twelve, two workgroup barriers instead of ten).  Thread (c, j) = (t mod 16,
// t div 16) of 128 owns column c of the workgroup's 16 consecutive (ky, kz)
// columns (each x element a 128-B run in HBM):
//   forward: x[j + 8 r] (r < 16) loaded from HBM; 16-point DFT over r;
//            twiddle W_128^(j k1); exchange; the thread takes k1 = 2j, 2j+1:
//            8-point DFT over j -> X[k1 + 16 k2] (k2 < 8)
//   influence on its 16 points, per point as the LDS-staged pass
//   inverse: 8-point inverse DFT over k2 (output index j'), twiddle
//            W_128^-(j' k1); exchange; 16-point inverse DFT over k1 ->
//            x[j + 8 r], stored.
// The forward transform performs the operations of the LDS-staged plan
// (16, 8); the inverse is decomposed the other way (k = k1 + 16 k2), so its
// rounding differs from the LDS-staged inverse's.  One partial of E and the
// virial per workgroup, as the LDS-staged pass.
constexpr int kXRegThreads = 128;
constexpr int kXRegN = 128;
constexpr int kXRegC = 16;    // columns per workgroup
constexpr int kXRegLS = 129;  // column stride of the exchange buffer (complex)

// One image group of a point of the register x pass: the terms of xsep_point
// for group gi (its images, multiplicity and sign sums), added when the image
// is in the ball; returns mul f (0 outside the ball).
__device__ __forceinline__ double xreg_group(XLine64 &L, const CellArgs &cl, double q00,
                                             double max_m2, double fac, double cc, double el,
                                             double k5, double gxe, double s, int nx, bool xny,
                                             bool both, int kz, int gi) {
    const int g = gi ? -1 : 1;
    const int sy = L.yny ? g : 1, sz = L.yny ? 1 : g;
    double m2;
    if (!L.in_ball(cl, q00, max_m2, nx, sy, sz, kz, cc, 0.0, m2)) return 0.0;
    const double r = recip_f64(m2);
    const double f = gxe * el * r;
    const double w = k5 * s * ((fac - r) * f);
    const double nxd = nx;
    const int mul = (xny ? 2 : 1) * (both ? 2 : 1);
    const int ssy = (xny || both) ? 0 : nx * sy, ssz = (xny || both) ? 0 : nx * sz;
    const double mw = (double)mul * w;
    L.W0 += mw;
    L.W2 = fma(mw * nxd, nxd, L.W2);
    L.Wy = fma(w, (double)ssy, L.Wy);
    L.Wz = fma(w, (double)ssz, L.Wz);
    L.Wyz = gi ? L.Wyz - mw : L.Wyz + mw;
    return (double)mul * f;
}

// The influence of the 16 points of a thread of the register x pass
// (x-separable boxes): u[h][jj] holds X[e], e = 2 j + h + 16 jj.  The terms,
// images and order of summation of xsep_point, written out point by point
// (every value stays in its register, no loop); the second image group is
// evaluated only in wavefronts that hold a Nyquist column.
template <typename T>
__device__ __forceinline__ void xsep_points_reg(XLine64 &L, const FrameDesc &F,
                                                const double *gx, const double *k, double q00,
                                                bool both, int ng, int kz, int j,
                                                const double *ib0, cplx<T> (&u)[2][8]) {
    const CellArgs &cl = F.tcell;
    const double fac = F.ewald_fac, max_m2 = F.max_m2;
    const bool two = __any(ng == 2);
    // (the second group's constants are read where they are used: held for
    // the whole loop, they cost the registers of the fourth workgroup per CU)
    const double c0 = k[0], el = k[2], k4 = k[4], k5 = k[5];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) {
            const int e = 2 * j + h + 16 * jj;
            // nx = e for e <= 64, e - 128 above (K0 = 128)
            const int nx = jj < 4 ? e : (jj > 4 ? e - kXRegN : (e == kXRegN / 2 ? e : e - kXRegN));
            const bool xny = jj == 4 && e == kXRegN / 2;
            const double ib = ib0[e] * k4;
            cplx<T> &p = u[h][jj];
            const double re = (double)p.x * ib, im = (double)p.y * ib;
            const double s = re * re + im * im;
            const double gxe = gx[nx < 0 ? -nx : nx];
            double a = 0.0;
            a = a + xreg_group(L, cl, q00, max_m2, fac, c0, el, k5, gxe, s, nx, xny, both, kz, 0);
            if (two && ng == 2)
                a = a + xreg_group(L, cl, q00, max_m2, fac, k[1], k[3], k5, gxe, s, nx, xny,
                                   both, kz, 1);
            L.aE = fma(a, s, L.aE);
            const double tt = a * ib;
            p.x = T(re * tt);
            p.y = T(im * tt);
            // one point at a time: the empty asm makes the point's results exist
            // here (left to the compiler, their products sink to the end and
            // every point's terms stay live at once)
            asm volatile("" : "+v"(p.x), "+v"(p.y), "+v"(L.aE));
        }
}
template <typename T> __host__ __device__ inline size_t xreg_lds_bytes() {
    return (size_t)kXRegN * sizeof(cplx<T>) + (size_t)kXRegN * sizeof(double) +
           (size_t)kXRegC * kXRegLS * sizeof(cplx<T>) +
           (size_t)(kXRegThreads / 64) * 8 * sizeof(double) +
           (xfused_gx_doubles(kXRegN) + (size_t)kXColConsts * kXRegC) * sizeof(double);
}

template <typename T>
__global__ void __launch_bounds__(kXRegThreads) __attribute__((amdgpu_waves_per_eu(4)))
fft_x_reg_kernel(SPME_FRAME_ARGS) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int N = kXRegN, C = kXRegC, LS = kXRegLS;
    int b;
    const FrameDesc &F = frames[find_frame(offs, nf, b)];
    const int ldim = F.ldim, P = F.K1 * ldim;
    cplx<T> *twl = reinterpret_cast<cplx<T> *>(smem_raw);
    double *ib0 = reinterpret_cast<double *>(twl + N);
    cplx<T> *buf = reinterpret_cast<cplx<T> *>(ib0 + N);
    double *wsums = reinterpret_cast<double *>(buf + C * LS);
    double *gx = wsums + (kXRegThreads / 64) * 8;
    double *colk = gx + xfused_gx_doubles(N);
    const int t = threadIdx.x, c = t & (C - 1), j = t >> 4;
    const int f = b * C + c;
    const bool valid = f < P;
    const int fc = valid ? f : P - 1;  // clamped for loads
    // element x of the column at 32-bit offset fc + x P (complex) from the
    // frame's grid, a workgroup-uniform global base: one register per
    // address, recomputed for the stores (16 64-bit addresses kept from the
    // loads to the stores were spilled)
    typedef __attribute__((address_space(1))) T GT;
    GT *grid = (GT *)F.grid;
    const unsigned e0 = 2u * ((unsigned)fc + (unsigned)j * (unsigned)P);
    const unsigned es = 16u * (unsigned)P;
    cplx<T> v[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        v[r].x = grid[e0 + (unsigned)r * es];
        v[r].y = grid[e0 + (unsigned)r * es + 1u];
    }
    {
        const cplx<T> *tw = (const cplx<T> *)F.tw0;
        for (int e = t; e < N; e += kXRegThreads) {
            twl[e] = tw[e];
            ib0[e] = F.iB0[e];
        }
        const double q00 = F.tcell.iL00 * F.tcell.iL00;
        for (int e = t; e <= N / 2; e += kXRegThreads) {
            const double nx = e;
            gx[e] = exp_nonpos(F.ewald_fac * q00 * nx * nx);
        }
        if (t < C) {
            const int fk = min(b * C + t, P - 1);
            const int ky = fdiv(fk, F.dldim);
            xsep_column(F, ky, fk - ky * ldim, colk + kXColConsts * t);
        }
    }
    __syncthreads();
    // forward: 16-point DFT over r, twiddles, exchange
    bfly16<T>(v, -1);
#pragma unroll
    for (int k1 = 1; k1 < 16; ++k1) v[k1] = cmul<T>(v[k1], twiddle<T>(twl, j * k1, -1));
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) buf[c * LS + k1 * 8 + j] = v[k1];
    __syncthreads();
    cplx<T> u[2][8];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) u[h][jj] = buf[c * LS + (2 * j + h) * 8 + jj];
    bfly8<T>(u[0], -1);
    bfly8<T>(u[1], -1);
    // influence of the thread's 16 points X[k1 + 16 k2]
    {
        const int ky = fdiv(fc, F.dldim), kz = fc - ky * ldim;
        const CellArgs &cl = F.tcell;
        const double q00 = cl.iL00 * cl.iL00 + cl.iL10 * cl.iL10 + cl.iL20 * cl.iL20;
        XLine64 L(F, ky, kz);
        const bool both = L.yny && L.zn2;
        const int ng = (L.yny || L.zn2) ? 2 : 1;
        const double *k = colk + kXColConsts * c;
        if (valid) xsep_points_reg<T>(L, F, gx, k, q00, both, ng, kz, j, ib0, u);
        L.finish(cl, kz, wsums);
    }
    // inverse: 8-point inverse DFT over k2, twiddles, exchange (each thread
    // writes the places it read: no barrier needed before)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int k1 = 2 * j + h;
        bfly8<T>(u[h], +1);
#pragma unroll
        for (int jj = 1; jj < 8; ++jj) u[h][jj] = cmul<T>(u[h][jj], twiddle<T>(twl, jj * k1, +1));
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) buf[c * LS + k1 * 8 + jj] = u[h][jj];
    }
    __syncthreads();
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) v[k1] = buf[c * LS + k1 * 8 + j];
    bfly16<T>(v, +1);
    unsigned e1 = e0;
    asm volatile("" : "+v"(e1));
    if (valid) {
        if constexpr (std::is_same<T, float>::value) {
            // one 8-byte store per element (with the offset opaque, the
            // compiler cannot merge the two 4-byte stores of an element)
            typedef __attribute__((address_space(1))) unsigned long long GU;
            GU *grid2 = (GU *)F.grid;
            const unsigned h1 = e1 >> 1, hs = es >> 1;
#pragma unroll
            for (int r = 0; r < 16; ++r)
                grid2[h1 + (unsigned)r * hs] =
                    ((unsigned long long)__float_as_uint(v[r].y) << 32) | __float_as_uint(v[r].x);
        } else {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                grid[e1 + (unsigned)r * es] = v[r].x;
                grid[e1 + (unsigned)r * es + 1u] = v[r].y;
            }
        }
    }
    // the workgroup's partial: its wavefront sums in order (made visible by
    // the barrier above)
    if (t < 7) {
        double sum = 0.0;
#pragma unroll
        for (int q = 0; q < kXRegThreads / 64; ++q) sum += wsums[q * 8 + t];
        F.partials[b * 8 + t] = sum;
    }
}



}  // namespace spme
