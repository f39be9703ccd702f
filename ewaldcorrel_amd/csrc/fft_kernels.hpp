// LDS-staged mixed-radix Stockham FFT line kernels (gfx950).
//
// One workgroup owns C lines of length N.  A line is staged once from HBM
// into LDS (coalesced: along the line for the contiguous z axis, across
// adjacent lines for the strided y and x axes), all radix stages run
// LDS->LDS with ping-pong buffers (Stockham autosort, no bit reversal),
// and the result is written back once.  Radix 2, 3, 4, 5 and 8 have
// dedicated butterflies; any other prime factor uses a direct DFT whose
// roots come from the same per-length twiddle table, so every N >= 1 is
// supported (the reference accepts any K >= 1; its test uses K = 13).
//
// Stage s with radix R and span Ns (product of earlier radices), for
// j in [0, N/R):  k = j mod Ns;  v_r = in[j + r N/R] * W^(r k), W =
// exp(sign 2 pi i/(Ns R));  v = DFT_R(v);  out[(j/Ns) Ns R + k + r Ns] = v_r.
//
// Twiddle table: tw[t] = exp(-2 pi i t/N) computed in double on the host
// (rounded once for FP32); the inverse transform conjugates on the fly.
#pragma once

#include "common.hpp"

namespace spme {

struct FftPlanArgs {
    int N;
    int nstages;
    int radix[kMaxRadixStages];
};

template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
    cplx<T> r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}
template <typename T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) {
    cplx<T> r;
    r.x = a.x + b.x;
    r.y = a.y + b.y;
    return r;
}
template <typename T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) {
    cplx<T> r;
    r.x = a.x - b.x;
    r.y = a.y - b.y;
    return r;
}
// multiply by (sign * i)
template <typename T>
__device__ __forceinline__ cplx<T> mul_si(cplx<T> a, int sign) {
    cplx<T> r;
    r.x = sign < 0 ? a.y : -a.y;
    r.y = sign < 0 ? -a.x : a.x;
    return r;
}
template <typename T>
__device__ __forceinline__ cplx<T> twiddle(const cplx<T> *tw, int t, int sign) {
    cplx<T> w = tw[t];
    if (sign > 0) w.y = -w.y;
    return w;
}

// ---- fixed butterflies (in registers) ------------------------------------
template <typename T>
__device__ __forceinline__ void bfly2(cplx<T> *v) {
    cplx<T> a = v[0], b = v[1];
    v[0] = cadd<T>(a, b);
    v[1] = csub<T>(a, b);
}
template <typename T>
__device__ __forceinline__ void bfly4(cplx<T> *v, int sign) {
    cplx<T> s02 = cadd<T>(v[0], v[2]), d02 = csub<T>(v[0], v[2]);
    cplx<T> s13 = cadd<T>(v[1], v[3]), d13 = mul_si<T>(csub<T>(v[1], v[3]), sign);
    v[0] = cadd<T>(s02, s13);
    v[2] = csub<T>(s02, s13);
    v[1] = cadd<T>(d02, d13);
    v[3] = csub<T>(d02, d13);
}
template <typename T>
__device__ __forceinline__ void bfly3(cplx<T> *v, int sign) {
    const T c = T(-0.5);
    const T s = T(0.86602540378443864676);
    cplx<T> t1 = cadd<T>(v[1], v[2]);
    cplx<T> t2;
    t2.x = v[0].x + c * t1.x;
    t2.y = v[0].y + c * t1.y;
    cplx<T> d = csub<T>(v[1], v[2]);
    d.x *= s;
    d.y *= s;
    cplx<T> t3 = mul_si<T>(d, sign);
    v[0] = cadd<T>(v[0], t1);
    v[1] = cadd<T>(t2, t3);
    v[2] = csub<T>(t2, t3);
}
template <typename T>
__device__ __forceinline__ void bfly5(cplx<T> *v, int sign) {
    const T c1 = T(0.30901699437494742410), c2 = T(-0.80901699437494742410);
    const T s1 = T(0.95105651629515357212), s2 = T(0.58778525229247312917);
    cplx<T> a1 = cadd<T>(v[1], v[4]), a2 = cadd<T>(v[2], v[3]);
    cplx<T> b1 = csub<T>(v[1], v[4]), b2 = csub<T>(v[2], v[3]);
    cplx<T> x0 = v[0];
    cplx<T> p1, p2, q1, q2;
    p1.x = x0.x + c1 * a1.x + c2 * a2.x;
    p1.y = x0.y + c1 * a1.y + c2 * a2.y;
    p2.x = x0.x + c2 * a1.x + c1 * a2.x;
    p2.y = x0.y + c2 * a1.y + c1 * a2.y;
    q1.x = s1 * b1.x + s2 * b2.x;
    q1.y = s1 * b1.y + s2 * b2.y;
    q2.x = s2 * b1.x - s1 * b2.x;
    q2.y = s2 * b1.y - s1 * b2.y;
    q1 = mul_si<T>(q1, sign);
    q2 = mul_si<T>(q2, sign);
    v[0] = cadd<T>(x0, cadd<T>(a1, a2));
    v[1] = cadd<T>(p1, q1);
    v[4] = csub<T>(p1, q1);
    v[2] = cadd<T>(p2, q2);
    v[3] = csub<T>(p2, q2);
}
template <typename T>
__device__ __forceinline__ void bfly8(cplx<T> *v, int sign) {
    // radix-2 over two radix-4s (even/odd inputs)
    cplx<T> e[4] = {v[0], v[2], v[4], v[6]};
    cplx<T> o[4] = {v[1], v[3], v[5], v[7]};
    bfly4<T>(e, sign);
    bfly4<T>(o, sign);
    const T h = T(0.70710678118654752440);
    // w8^q, q=1..3, w8 = exp(sign i pi/4)
    cplx<T> o1, o2, o3;
    o1.x = h * (o[1].x - sign * o[1].y);
    o1.y = h * (o[1].y + sign * o[1].x);
    o2 = mul_si<T>(o[2], sign);
    o3.x = h * (-o[3].x - sign * o[3].y);
    o3.y = h * (-o[3].y + sign * o[3].x);
    v[0] = cadd<T>(e[0], o[0]);
    v[4] = csub<T>(e[0], o[0]);
    v[1] = cadd<T>(e[1], o1);
    v[5] = csub<T>(e[1], o1);
    v[2] = cadd<T>(e[2], o2);
    v[6] = csub<T>(e[2], o2);
    v[3] = cadd<T>(e[3], o3);
    v[7] = csub<T>(e[3], o3);
}

// One Stockham stage of fixed radix R over C lines held in LDS.
// src/dst: line c occupies [c*LS, c*LS + N).
template <typename T, int R>
__device__ void stage_fixed(const cplx<T> *__restrict__ src, cplx<T> *__restrict__ dst,
                            int C, int LS, int N, int Ns, const cplx<T> *tw, int sign) {
    const int NR = N / R;
    const int twstep = N / (Ns * R);
    for (int b = threadIdx.x; b < C * NR; b += blockDim.x) {
        const int c = b / NR;
        const int j = b - c * NR;
        const int k = j % Ns;
        const cplx<T> *s = src + c * LS;
        cplx<T> v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = s[j + r * NR];
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) v[r] = cmul<T>(v[r], twiddle<T>(tw, r * k * twstep, sign));
        }
        if (R == 2) bfly2<T>(v);
        else if (R == 3) bfly3<T>(v, sign);
        else if (R == 4) bfly4<T>(v, sign);
        else if (R == 5) bfly5<T>(v, sign);
        else if (R == 8) bfly8<T>(v, sign);
        cplx<T> *d = dst + c * LS + (j / Ns) * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) d[r * Ns] = v[r];
    }
}

// Generic radix (any R dividing N): direct DFT, roots from the N-table.
template <typename T>
__device__ void stage_generic(const cplx<T> *__restrict__ src, cplx<T> *__restrict__ dst,
                              int C, int LS, int N, int Ns, int R, const cplx<T> *tw,
                              int sign) {
    const int NR = N / R;
    const int twstep = N / (Ns * R);
    for (int b = threadIdx.x; b < C * NR; b += blockDim.x) {
        const int c = b / NR;
        const int j = b - c * NR;
        const int k = j % Ns;
        const cplx<T> *s = src + c * LS;
        cplx<T> *d = dst + c * LS + (j / Ns) * Ns * R + k;
        for (int q = 0; q < R; ++q) {
            cplx<T> acc;
            acc.x = T(0);
            acc.y = T(0);
            for (int r = 0; r < R; ++r) {
                cplx<T> v = s[j + r * NR];
                if (Ns > 1 && r > 0) v = cmul<T>(v, twiddle<T>(tw, r * k * twstep, sign));
                const int t = ((r * q) % R) * NR;
                acc = cadd<T>(acc, cmul<T>(v, twiddle<T>(tw, t, sign)));
            }
            d[q * Ns] = acc;
        }
    }
}

// Runs all stages; returns the LDS buffer holding the result.
template <typename T>
__device__ cplx<T> *run_stages(cplx<T> *b0, cplx<T> *b1, int C, int LS,
                               const FftPlanArgs &p, const cplx<T> *tw, int sign) {
    int Ns = 1;
    cplx<T> *src = b0, *dst = b1;
    for (int s = 0; s < p.nstages; ++s) {
        const int R = p.radix[s];
        switch (R) {
            case 2: stage_fixed<T, 2>(src, dst, C, LS, p.N, Ns, tw, sign); break;
            case 3: stage_fixed<T, 3>(src, dst, C, LS, p.N, Ns, tw, sign); break;
            case 4: stage_fixed<T, 4>(src, dst, C, LS, p.N, Ns, tw, sign); break;
            case 5: stage_fixed<T, 5>(src, dst, C, LS, p.N, Ns, tw, sign); break;
            case 8: stage_fixed<T, 8>(src, dst, C, LS, p.N, Ns, tw, sign); break;
            default: stage_generic<T>(src, dst, C, LS, p.N, Ns, R, tw, sign); break;
        }
        __syncthreads();
        Ns *= R;
        cplx<T> *t = src;
        src = dst;
        dst = t;
    }
    return src;
}

// ---------------------------------------------------------------------------
// z axis, real -> half-complex.  rows = K0*K1, each of K2 reals at stride
// 2*ldim reals; output ldim complex in place.  Block: C rows.
template <typename T>
__global__ void __launch_bounds__(256)
fft_z_r2c(T *__restrict__ grid, int nrows, int ldim, FftPlanArgs p,
          const cplx<T> *__restrict__ tw, int C) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cplx<T> *b0 = reinterpret_cast<cplx<T> *>(smem_raw);
    const int N = p.N;
    const int LS = N + 1;
    cplx<T> *b1 = b0 + C * LS;
    const int row0 = blockIdx.x * C;
    const int nr = min(C, nrows - row0);
    const long rs = 2L * ldim;  // row stride in reals
    for (int l = threadIdx.x; l < nr * N; l += blockDim.x) {
        const int c = l / N, z = l - c * N;
        cplx<T> v;
        v.x = grid[(long)(row0 + c) * rs + z];
        v.y = T(0);
        b0[c * LS + z] = v;
    }
    __syncthreads();
    cplx<T> *res = run_stages<T>(b0, b1, nr, LS, p, tw, -1);
    cplx<T> *out = reinterpret_cast<cplx<T> *>(grid);
    for (int l = threadIdx.x; l < nr * ldim; l += blockDim.x) {
        const int c = l / ldim, k = l - c * ldim;
        out[(long)(row0 + c) * ldim + k] = res[c * LS + k];
    }
}

// z axis, half-complex -> real (unnormalised, sign +1).  Hermitian
// completion X[K2-k] = conj(X[k]); the real part of the full complex
// transform equals the c2r result (imaginary parts of the kz=0 and Nyquist
// terms contribute sin(0) = sin(pi z) = 0).
template <typename T>
__global__ void __launch_bounds__(256)
fft_z_c2r(T *__restrict__ grid, int nrows, int ldim, FftPlanArgs p,
          const cplx<T> *__restrict__ tw, int C) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cplx<T> *b0 = reinterpret_cast<cplx<T> *>(smem_raw);
    const int N = p.N;
    const int LS = N + 1;
    cplx<T> *b1 = b0 + C * LS;
    const int row0 = blockIdx.x * C;
    const int nr = min(C, nrows - row0);
    const cplx<T> *in = reinterpret_cast<const cplx<T> *>(grid);
    for (int l = threadIdx.x; l < nr * ldim; l += blockDim.x) {
        const int c = l / ldim, k = l - c * ldim;
        const cplx<T> v = in[(long)(row0 + c) * ldim + k];
        b0[c * LS + k] = v;
        const int km = N - k;
        if (k > 0 && km >= ldim) {
            cplx<T> w;
            w.x = v.x;
            w.y = -v.y;
            b0[c * LS + km] = w;
        }
    }
    __syncthreads();
    cplx<T> *res = run_stages<T>(b0, b1, nr, LS, p, tw, +1);
    const long rs = 2L * ldim;
    for (int l = threadIdx.x; l < nr * N; l += blockDim.x) {
        const int c = l / N, z = l - c * N;
        grid[(long)(row0 + c) * rs + z] = res[c * LS + z].x;
    }
}

// Strided axis (y or x) complex transform over the half-complex array.
// Line (o, c): element e at o*ostride + c + e*estride (complex units),
// c in [0, ninner).  Block = (o, tile of C adjacent c), so each element
// row of the tile is a contiguous run of C complex numbers in HBM.
// Optional per-output scale sc_e[e] * sc_o[o] * sc_c[c] (used to fold the
// B-spline moduli 1/(b0 b1 b2) into the last forward pass).
template <typename T>
__global__ void __launch_bounds__(256)
fft_strided(cplx<T> *__restrict__ data, int ninner, int ntiles, long ostride,
            long estride, FftPlanArgs p, const cplx<T> *__restrict__ tw, int C,
            int sign, const double *__restrict__ sc_e, const double *__restrict__ sc_o,
            const double *__restrict__ sc_c) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cplx<T> *b0 = reinterpret_cast<cplx<T> *>(smem_raw);
    const int N = p.N;
    const int LS = N + 1;
    cplx<T> *b1 = b0 + C * LS;
    const int o = blockIdx.x / ntiles;
    const int c0 = (blockIdx.x - o * ntiles) * C;
    const int nc = min(C, ninner - c0);
    cplx<T> *base = data + (long)o * ostride + c0;
    for (int l = threadIdx.x; l < nc * N; l += blockDim.x) {
        const int e = l / nc, c = l - e * nc;
        b0[c * LS + e] = base[(long)e * estride + c];
    }
    __syncthreads();
    cplx<T> *res = run_stages<T>(b0, b1, nc, LS, p, tw, sign);
    const double so = sc_o ? sc_o[o] : 1.0;
    for (int l = threadIdx.x; l < nc * N; l += blockDim.x) {
        const int e = l / nc, c = l - e * nc;
        cplx<T> v = res[c * LS + e];
        if (sc_e) {
            const T s = T(sc_e[e] * so * sc_c[c0 + c]);
            v.x *= s;
            v.y *= s;
        }
        base[(long)e * estride + c] = v;
    }
}

}  // namespace spme
