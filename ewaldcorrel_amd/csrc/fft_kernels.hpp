// LDS-staged mixed-radix Stockham FFT line kernels (gfx950).
//
// A workgroup owns C lines of length N.  Lines are staged once from HBM into
// one LDS buffer (coalesced: along the line for the contiguous z axis,
// across adjacent kz lines for the strided y and x axes), every radix stage
// runs in place with register staging (each thread reads its butterflies'
// inputs, barrier, writes the outputs), twiddles come from an LDS copy of
// the per-length table, and the result is written back once.  Staging
// loads use clamped addresses with no branches, so every load of a thread
// is in flight before the first one is waited for.
//
// Radix 2, 3, 4, 5, 8 and 16 have dedicated butterflies; any other prime uses a
// direct DFT (two LDS buffers for that stage), so every N >= 1 works (the
// reference accepts any K >= 1; its own test uses K = 13).
//
// Stage with radix R and span Ns (product of earlier radices), j < N/R:
//   k = j mod Ns;  v_r = in[j + r N/R] * W^(r k),  W = exp(sign 2 pi i/(Ns R));
//   v = DFT_R(v);  out[(j/Ns) Ns R + k + r Ns] = v_r     (natural order out)
//
// Real transforms along z pack two real rows into one complex line
// (z = a + i b) and separate the spectra afterwards, so the z passes do half
// the complex work of a zero-imaginary transform.
#pragma once

#include "common.hpp"

namespace spme {

// max complex elements a workgroup stages (bounds the per-thread registers
// of a stage: 8 FP32 or 4 FP64 complex values per thread)
template <typename T> __host__ __device__ constexpr int fft_elems() {
    return sizeof(T) == 8 ? 1024 : 2048;
}

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

// radix 16 as 4 x 4: 4-point DFTs over n1 of x[4 n1 + n2], twiddles
// W16^(n2 k1), 4-point DFTs over n2; outputs in natural order k1 + 4 k2
template <typename T>
__device__ __forceinline__ cplx<T> cmul_c(cplx<T> a, T c, T s) {
    cplx<T> r;
    r.x = a.x * c - a.y * s;
    r.y = a.x * s + a.y * c;
    return r;
}
template <typename T>
__device__ __forceinline__ void bfly16(cplx<T> *v, int sign) {
    const T c1 = T(0.92387953251128675613), s1 = T(0.38268343236508977173);
    const T h = T(0.70710678118654752440);
    const T sg = sign < 0 ? T(-1) : T(1);
    cplx<T> y[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
        cplx<T> a[4] = {v[n2], v[4 + n2], v[8 + n2], v[12 + n2]};
        bfly4<T>(a, sign);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) y[n2][k1] = a[k1];
    }
    // W16^m = (cos(2 pi m/16), sign sin(2 pi m/16)) for m = n2 k1
    y[1][1] = cmul_c<T>(y[1][1], c1, sg * s1);   // m = 1
    y[1][2] = cmul_c<T>(y[1][2], h, sg * h);     // m = 2
    y[1][3] = cmul_c<T>(y[1][3], s1, sg * c1);   // m = 3
    y[2][1] = cmul_c<T>(y[2][1], h, sg * h);     // m = 2
    y[2][2] = mul_si<T>(y[2][2], sign);          // m = 4
    y[2][3] = cmul_c<T>(y[2][3], -h, sg * h);    // m = 6
    y[3][1] = cmul_c<T>(y[3][1], s1, sg * c1);   // m = 3
    y[3][2] = cmul_c<T>(y[3][2], -h, sg * h);    // m = 6
    y[3][3] = cmul_c<T>(y[3][3], -c1, -sg * s1); // m = 9
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        cplx<T> b[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
        bfly4<T>(b, sign);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = b[k2];
    }
}

// In-place stage of fixed radix R over C lines (line c at buf + c*LS):
// every thread loads the inputs of up to MAXQ butterflies into registers,
// transforms them, and after a barrier stores the outputs.  Index arithmetic
// by FastDivs and 24-bit multiplies (every index is an LDS index < 2^15).
template <typename T, int R, int EPT = fft_elems<T>() / kBlock>
__device__ __forceinline__ void stage_ip(cplx<T> *buf, int C, int LS, int N, int Ns,
                                         const FastDiv &dNR, const FastDiv &dNs, int twstep,
                                         const cplx<T> *tw, int sign) {
    constexpr int MAXQ = (EPT + R - 1) / R;
    const int NR = dNR.d;
    const int nb = C * NR;  // wave-uniform products stay on the scalar unit
    cplx<T> v[MAXQ][R];
    int dst[MAXQ];
#pragma unroll
    for (int q = 0; q < MAXQ; ++q) {
        const int bf = threadIdx.x + q * kBlock;
        dst[q] = -1;
        if (bf < nb) {
            const int c = fdiv15(bf, dNR);
            const int j = bf - mul24(c, NR);
            const int jq = fdiv15(j, dNs);
            const int k = j - mul24(jq, Ns);
            const cplx<T> *s = buf + mul24(c, LS) + j;
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = s[r * NR];
            if (Ns > 1) {
                const int kt = mul24(k, twstep);
                int t = 0;
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    t += kt;
                    v[q][r] = cmul<T>(v[q][r], twiddle<T>(tw, t, sign));
                }
            }
            if (R == 2) bfly2<T>(v[q]);
            else if (R == 3) bfly3<T>(v[q], sign);
            else if (R == 4) bfly4<T>(v[q], sign);
            else if (R == 5) bfly5<T>(v[q], sign);
            else if (R == 8) bfly8<T>(v[q], sign);
            else if (R == 16) bfly16<T>(v[q], sign);
            dst[q] = mul24(c, LS) + mul24(jq, Ns * (R - 1)) + j;
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXQ; ++q) {
        if (dst[q] >= 0) {
#pragma unroll
            for (int r = 0; r < R; ++r) buf[dst[q] + r * Ns] = v[q][r];
        }
    }
    __syncthreads();
}

// Generic radix (any R dividing N): direct DFT from src to dst.
template <typename T>
__device__ __forceinline__ void stage_generic(const cplx<T> *__restrict__ src,
                                              cplx<T> *__restrict__ dst, int C, int LS, int N,
                                              int Ns, int R, const cplx<T> *tw, int sign) {
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
    __syncthreads();
}

// All stages of the plan; returns the buffer that holds the result.
// Requires C N <= EPT kBlock.
template <typename T, int EPT = fft_elems<T>() / kBlock>
__device__ __forceinline__ cplx<T> *fft_lines(cplx<T> *b0, cplx<T> *b1, int C, int LS,
                                              const FftPlanArgs &p, const cplx<T> *tw,
                                              int sign) {
    int Ns = 1;
    cplx<T> *cur = b0, *oth = b1;
    for (int s = 0; s < p.nstages; ++s) {
        const int R = p.radix[s];
        switch (R) {
            case 2: stage_ip<T, 2, EPT>(cur, C, LS, p.N, Ns, p.dNR[s], p.dNs[s], p.twstep[s], tw, sign); break;
            case 3: stage_ip<T, 3, EPT>(cur, C, LS, p.N, Ns, p.dNR[s], p.dNs[s], p.twstep[s], tw, sign); break;
            case 4: stage_ip<T, 4, EPT>(cur, C, LS, p.N, Ns, p.dNR[s], p.dNs[s], p.twstep[s], tw, sign); break;
            case 5: stage_ip<T, 5, EPT>(cur, C, LS, p.N, Ns, p.dNR[s], p.dNs[s], p.twstep[s], tw, sign); break;
            case 8: stage_ip<T, 8, EPT>(cur, C, LS, p.N, Ns, p.dNR[s], p.dNs[s], p.twstep[s], tw, sign); break;
            case 16:  // one radix-16 butterfly per thread at most (16 registers of values)
                if constexpr (EPT <= 16)
                    stage_ip<T, 16, EPT>(cur, C, LS, p.N, Ns, p.dNR[s], p.dNs[s], p.twstep[s], tw, sign);
                break;
            default: {
                stage_generic<T>(cur, oth, C, LS, p.N, Ns, R, tw, sign);
                cplx<T> *t = cur;
                cur = oth;
                oth = t;
            }
        }
        Ns *= R;
    }
    return cur;
}

// ---- strided line transforms (slab and x kernels) -----------------------------------------
// C lines in LDS, line c at buf + c LST, element j of a line at + j EST.
// Butterfly b of a stage belongs to line c = b mod C, index j = b div C, so
// adjacent lanes work on adjacent lines: with a line stride that is odd in
// complex elements (or EST = C-contiguous columns) the lanes of a wave hit
// distinct LDS banks whatever the stage's index pattern.  The index
// arithmetic uses FastDivs precomputed per stage.  In place, register staged:
// every thread reads its butterflies' inputs, barrier, writes the outputs.
// Requires C N <= EPT NT.
template <typename T, int R, int EPT, int NT, bool OPQ>
__device__ __forceinline__ void lstage(cplx<T> *buf, int C, const FastDiv &dC, int LST, int EST,
                                       int N, int Ns, const FastDiv &dNs, int twstep,
                                       const cplx<T> *tw, int sign) {
    constexpr int MAXQ = (EPT + R - 1) / R;
    const int NR = N / R;
    const int nb = C * NR;
    cplx<T> v[MAXQ][R];
    int dst[MAXQ];
    // OPQ: the thread's butterfly indices are recomputed in every stage (an
    // opaque thread id) instead of being kept from stage to stage, which
    // costs registers (spilled, or a workgroup less per CU) where registers
    // bound the kernel and a little issue time where they do not
    int tid = (int)threadIdx.x;
    if (OPQ) asm volatile("" : "+v"(tid));
#pragma unroll
    for (int q = 0; q < MAXQ; ++q) {
        const int bf = tid + q * NT;
        dst[q] = -1;
        if (bf < nb) {
            const int j = fdiv(bf, dC);
            const int c = bf - j * C;
            const int jq = fdiv(j, dNs);
            const int k = j - jq * Ns;
            const cplx<T> *s = buf + c * LST + j * EST;
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = s[r * NR * EST];
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r)
                    v[q][r] = cmul<T>(v[q][r], twiddle<T>(tw, r * k * twstep, sign));
            }
            if (R == 2) bfly2<T>(v[q]);
            else if (R == 3) bfly3<T>(v[q], sign);
            else if (R == 4) bfly4<T>(v[q], sign);
            else if (R == 5) bfly5<T>(v[q], sign);
            else if (R == 8) bfly8<T>(v[q], sign);
            else if (R == 16) bfly16<T>(v[q], sign);
            dst[q] = c * LST + (j + jq * Ns * (R - 1)) * EST;
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXQ; ++q) {
        if (dst[q] >= 0) {
#pragma unroll
            for (int r = 0; r < R; ++r) buf[dst[q] + r * Ns * EST] = v[q][r];
        }
    }
    __syncthreads();
}

// All stages of a plan whose radices are 2, 3, 4, 5, 8 and 16 (plan.generic == 0).
template <typename T, int EPT, int NT, bool OPQ = false>
__device__ __forceinline__ void lines_fft(cplx<T> *buf, int C, const FastDiv &dC, int LST,
                                          int EST, const FftPlanArgs &p, const cplx<T> *tw,
                                          int sign) {
    int Ns = 1;
    for (int s = 0; s < p.nstages; ++s) {
        const int R = p.radix[s];
        const FastDiv &dNs = p.dNs[s];
        const int ts = p.twstep[s];
        switch (R) {
            case 2: lstage<T, 2, EPT, NT, OPQ>(buf, C, dC, LST, EST, p.N, Ns, dNs, ts, tw, sign); break;
            case 3: lstage<T, 3, EPT, NT, OPQ>(buf, C, dC, LST, EST, p.N, Ns, dNs, ts, tw, sign); break;
            case 4: lstage<T, 4, EPT, NT, OPQ>(buf, C, dC, LST, EST, p.N, Ns, dNs, ts, tw, sign); break;
            case 5: lstage<T, 5, EPT, NT, OPQ>(buf, C, dC, LST, EST, p.N, Ns, dNs, ts, tw, sign); break;
            case 8: lstage<T, 8, EPT, NT, OPQ>(buf, C, dC, LST, EST, p.N, Ns, dNs, ts, tw, sign); break;
            default:  // 16: variants with at most 16 elements per thread only (the
                      // plan is checked on the host)
                if constexpr (EPT <= 16)
                    lstage<T, 16, EPT, NT, OPQ>(buf, C, dC, LST, EST, p.N, Ns, dNs, ts, tw, sign);
                break;
        }
        Ns *= R;
    }
}

// Twiddles to LDS: all loads first (clamped index), then the LDS writes.
template <typename T>
__device__ __forceinline__ void load_twiddles(cplx<T> *twl, const void *tw, int N) {
    const cplx<T> *g = (const cplx<T> *)tw;
    constexpr int Q = (kMaxK + kBlock - 1) / kBlock;
    T vx[Q], vy[Q];  // scalar arrays stay in registers (vector-typed ones may not)
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const cplx<T> w = g[min((int)threadIdx.x + q * kBlock, N - 1)];
        vx[q] = w.x;
        vy[q] = w.y;
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int t = threadIdx.x + q * kBlock;
        if (t < N) {
            cplx<T> w;
            w.x = vx[q];
            w.y = vy[q];
            twl[t] = w;
        }
    }
}

// LDS bytes of a line workgroup: twiddles + C lines (+ a second buffer for
// plans with a generic radix)
template <typename T>
__host__ __device__ inline size_t fft_lds_bytes(int N, int C, int generic) {
    return ((size_t)N + (size_t)C * (N + 1) * (generic ? 2 : 1)) * sizeof(cplx<T>);
}

// ---------------------------------------------------------------------------
// z axis, real -> half-complex, two real rows per complex line.
// Rows r = 0..K0*K1-1, K2 reals each at stride 2*ldim; the spectra of rows
// 2c and 2c+1 come out of Z = FFT(a + i b) as
//   A_k = (Z_k + conj Z_{N-k})/2,  B_k = (Z_k - conj Z_{N-k})/(2i).
template <typename T>
__global__ void __launch_bounds__(kBlock) fft_z_r2c_kernel(SPME_FRAME_ARGS) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    int b;
    const FrameDesc &F = frames[find_frame(offs, nf, b)];
    const FftPlanArgs &p = F.plan2;
    const int N = p.N, ldim = F.ldim, LS = N + 1;
    const int nrows = F.K0 * F.K1;
    const int zC = F.zC;
    const int row0 = b * 2 * zC;
    const int nl = min(zC, (nrows - row0 + 1) / 2);
    cplx<T> *twl = reinterpret_cast<cplx<T> *>(smem_raw);
    cplx<T> *b0 = twl + N;
    cplx<T> *b1 = b0 + zC * LS;
    load_twiddles<T>(twl, F.tw2, N);
    const T *g = (const T *)F.grid;
    const long rs = 2L * ldim;
    constexpr int EPT = fft_elems<T>() / kBlock;
    const int tot = nl * N;
    T ra[EPT], rb[EPT];
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
        const int l = min((int)threadIdx.x + q * kBlock, tot - 1);
        const int c = l / N, z = l - c * N;
        const int r = row0 + 2 * c;
        ra[q] = g[(long)r * rs + z];
        rb[q] = g[(long)min(r + 1, nrows - 1) * rs + z];
    }
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
        const int l = threadIdx.x + q * kBlock;
        if (l < tot) {
            const int c = l / N, z = l - c * N;
            cplx<T> v;
            v.x = ra[q];
            v.y = (row0 + 2 * c + 1 < nrows) ? rb[q] : T(0);
            b0[c * LS + z] = v;
        }
    }
    __syncthreads();
    const cplx<T> *res = fft_lines<T>(b0, b1, nl, LS, p, twl, -1);
    cplx<T> *out = (cplx<T> *)F.grid;
    for (int l = threadIdx.x; l < nl * ldim; l += blockDim.x) {
        const int c = l / ldim, k = l - c * ldim;
        const cplx<T> zk = res[c * LS + k];
        const cplx<T> zm = res[c * LS + (k == 0 ? 0 : N - k)];
        const int r = row0 + 2 * c;
        cplx<T> a, bb;
        a.x = T(0.5) * (zk.x + zm.x);
        a.y = T(0.5) * (zk.y - zm.y);
        bb.x = T(0.5) * (zk.y + zm.y);
        bb.y = T(0.5) * (zm.x - zk.x);
        out[(long)r * ldim + k] = a;
        if (r + 1 < nrows) out[(long)(r + 1) * ldim + k] = bb;
    }
}

// z axis, half-complex -> real (unnormalised, sign +1), two rows per line:
// Z_k = A_k + i B_k over the Hermitian-completed rows, real part -> row 2c,
// imaginary part -> row 2c+1.  Imaginary parts of the kz = 0 and Nyquist
// terms are dropped, as a c2r transform ignores them.
template <typename T>
__global__ void __launch_bounds__(kBlock) fft_z_c2r_kernel(SPME_FRAME_ARGS) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    int b;
    const FrameDesc &F = frames[find_frame(offs, nf, b)];
    const FftPlanArgs &p = F.plan2;
    const int N = p.N, ldim = F.ldim, LS = N + 1;
    const int nrows = F.K0 * F.K1;
    const int zC = F.zC;
    const int row0 = b * 2 * zC;
    const int nl = min(zC, (nrows - row0 + 1) / 2);
    cplx<T> *twl = reinterpret_cast<cplx<T> *>(smem_raw);
    cplx<T> *b0 = twl + N;
    cplx<T> *b1 = b0 + zC * LS;
    load_twiddles<T>(twl, F.tw2, N);
    const cplx<T> *in = (const cplx<T> *)F.grid;
    const bool even = (N % 2) == 0;
    constexpr int EPT = fft_elems<T>() / kBlock;
    const int tot = nl * ldim;
    T rax[EPT], ray[EPT], rbx[EPT], rby[EPT];
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
        const int l = min((int)threadIdx.x + q * kBlock, tot - 1);
        const int c = l / ldim, k = l - c * ldim;
        const int r = row0 + 2 * c;
        const cplx<T> va = in[(long)r * ldim + k];
        const cplx<T> vb = in[(long)min(r + 1, nrows - 1) * ldim + k];
        rax[q] = va.x;
        ray[q] = va.y;
        rbx[q] = vb.x;
        rby[q] = vb.y;
    }
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
        const int l = threadIdx.x + q * kBlock;
        if (l >= tot) continue;
        const int c = l / ldim, k = l - c * ldim;
        cplx<T> a, bb;
        a.x = rax[q];
        a.y = ray[q];
        bb.x = rbx[q];
        bb.y = rby[q];
        if (row0 + 2 * c + 1 >= nrows) bb.x = bb.y = T(0);
        if (k == 0 || (even && k == N / 2)) {
            a.y = T(0);
            bb.y = T(0);
        }
        cplx<T> z;
        z.x = a.x - bb.y;
        z.y = a.y + bb.x;
        b0[c * LS + k] = z;
        const int km = N - k;
        if (k > 0 && km >= ldim) {
            cplx<T> zc;  // conj(A) + i conj(B)
            zc.x = a.x + bb.y;
            zc.y = bb.x - a.y;
            b0[c * LS + km] = zc;
        }
    }
    __syncthreads();
    const cplx<T> *res = fft_lines<T>(b0, b1, nl, LS, p, twl, +1);
    T *g = (T *)F.grid;
    const long rs = 2L * ldim;
    for (int l = threadIdx.x; l < nl * N; l += blockDim.x) {
        const int c = l / N, z = l - c * N;
        const int r = row0 + 2 * c;
        const cplx<T> v = res[c * LS + z];
        g[(long)r * rs + z] = v.x;
        if (r + 1 < nrows) g[(long)(r + 1) * rs + z] = v.y;
    }
}

}  // namespace spme
