// Spreading fused with the forward y-z slab transform: one workgroup per
// (frame, x plane).  The plane's K1 x (K2 + n - 1) fixed-point sums fill LDS
// (the z stencil unwrapped, as in the spreading tiles); the atoms of the n
// planes whose stencils reach it add their one x weight each; the sums are
// folded and converted in place into the slab's pair layout, and the z and y
// transforms of fft_yz_fwd_kernel follow.  The real grid never goes to HBM:
// the launch reads the sorted records (each by the n workgroups of its
// stencil's planes, neighbours on one XCD) and writes the half-complex plane.
// The reference spreads into the real grid and transforms it with one FFTW
// r2c plan (sfac.cpp:97-128, 274-283).
#pragma once

#include "common.hpp"
#include "slab128_kernels.hpp"
#include "slab_kernels.hpp"
#include "spread_kernels.hpp"

namespace spme {

constexpr int kSpreadSlabThreads = 1024;
// elements per thread of the line transforms: planes of up to 1024 x 9
// (128 x 128: 8,320)
constexpr int kSpreadSlabEPT = 9;

// LDS: twiddles of z and y, then the sums (the slab overlays them)
template <typename T>
__host__ __device__ inline size_t spread_slab_lds_bytes(int K1, int K2, int order) {
    return ((size_t)K1 + K2) * sizeof(cplx<T>) + (size_t)K1 * (K2 + order - 1) * 8;
}

// Whether a grid takes the fused path (with fixed-point sums): a plane of
// sums in the LDS budget with the slab inside it; at least n cells along
// every axis (each atom reaches the plane through one x weight, a stencil
// row wraps at most once in y, and a z halo cell folds onto one cell); the
// line transforms, the fold and the untangling of the pairs within the
// workgroup's elements per thread.
template <typename T>
__host__ inline bool spread_slab_fits(int K0, int K1, int K2, int ldim, int order) {
    const int npair = (K1 + 1) / 2;
    const size_t sums = (size_t)K1 * (K2 + order - 1) * 8;
    const size_t slab = (size_t)slab_rows(K1) * ldim * sizeof(cplx<T>);
    constexpr int NT = kSpreadSlabThreads, EPT = kSpreadSlabEPT;
    return K0 >= order && K1 >= order && K2 >= order && slab <= sums &&
           spread_slab_lds_bytes<T>(K1, K2, order) <= kSlabLds &&
           std::max(npair * K2, ldim * K1) <= NT * EPT && npair * ldim <= NT * (EPT / 2 + 1);
}

// R128: FP32 128 x 128 planes with the plans (16, 8): the transform of the
// register-line slab kernels (slab128_fwd_lines) in place of the LDS-staged
// line passes, the fold writing the z lines straight into registers.
template <typename T, int ORDER, bool R128>
__global__ void __launch_bounds__(kSpreadSlabThreads)
spread_slab_kernel(SPME_FRAME_ARGS, SplineArgs<T> sp) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int NT = kSpreadSlabThreads, EPT = kSpreadSlabEPT;
    constexpr int CAP = Ord<ORDER>::cap;
    __shared__ int seg_begin[CAP], seg_pref[CAP + 1];
    // neighbouring planes (which read the same records) on one XCD
    int x0;
    const FrameDesc &F =
        frames[find_frame_of(offs, nf, xcd_block((int)blockIdx.x, (int)gridDim.x), x0)];
    const int n = order_of<ORDER>(F.order);
    const int K0 = F.K0, K1 = F.K1, K2 = F.K2, RS = F.ldim;
    const int Kp = K2 + n - 1, H = n / 2 - 1;
    const int npair = (K1 + 1) >> 1;
    cplx<T> *twz = reinterpret_cast<cplx<T> *>(smem_raw);
    cplx<T> *twy = twz + K2;
    unsigned long long *acc = reinterpret_cast<unsigned long long *>(twy + K1);
    cplx<T> *slab = reinterpret_cast<cplx<T> *>(acc);
    // segment d: the records of plane x0 - n/2 + d (mod K0), whose atoms reach
    // plane x0 through their x weight d (lane d of the first wavefront; the
    // loads in flight while LDS is cleared)
    int sbeg = 0, slen = 0;
    if ((int)threadIdx.x < n) {
        const int cx = imod(x0 - n / 2 + (int)threadIdx.x, K0);
        sbeg = F.colstart[cx * K1];
        slen = F.colstart[(cx + 1) * K1] - sbeg;
    }
    // the twiddles to registers now, to LDS after the atoms (their region is
    // not touched before): no wait for them at the start
    const cplx<T> *gtz = (const cplx<T> *)F.tw2, *gty = (const cplx<T> *)F.tw1;
    cplx<T> rtz{}, rty{};
    if ((int)threadIdx.x < K2) rtz = gtz[threadIdx.x];
    if ((int)threadIdx.x < K1) rty = gty[threadIdx.x];
    {
        const int cells = K1 * Kp;
        uint4 *z4 = reinterpret_cast<uint4 *>(acc);  // (K1 + K2) * 8 B in: 16-B aligned
        for (int i = threadIdx.x; i < cells / 2; i += NT) z4[i] = make_uint4(0, 0, 0, 0);
        if ((cells & 1) && threadIdx.x == 0) acc[cells - 1] = 0ull;
    }
    if (threadIdx.x < 64) {
        int v = slen;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_up(v, o);
            if ((int)threadIdx.x >= o) v += t;
        }
        if ((int)threadIdx.x < n) {
            seg_begin[threadIdx.x] = sbeg;
            seg_pref[threadIdx.x] = v - slen;
        }
        if ((int)threadIdx.x == n - 1) seg_pref[n] = v;
    }
    __syncthreads();
    const int units = seg_pref[n];
    const double scale = *F.fxscale;
    constexpr int RW = rec_words<T>();
    const uint2 *recw = reinterpret_cast<const uint2 *>(F.recs);
    // FP32: int32 contributions of the pre-scaled charge into int64 sums (the
    // bound of fixed_scale); FP64: int64 of the scaled product
    typedef typename std::conditional<std::is_same<T, float>::value, float, double>::type W;
    // the stencil work of one record (raw words v, x weight d)
    auto atom = [&](const uint2 *v, int d) {
        T f0, f1, f2, w;
        int cell;
        if constexpr (std::is_same<T, float>::value) {
            f0 = __uint_as_float(v[0].x);
            f1 = __uint_as_float(v[0].y);
            f2 = __uint_as_float(v[1].x);
            cell = (int)v[1].y;
            w = __uint_as_float(v[2].x);
        } else {
            f0 = __hiloint2double((int)v[0].y, (int)v[0].x);
            f1 = __hiloint2double((int)v[1].y, (int)v[1].x);
            f2 = __hiloint2double((int)v[2].y, (int)v[2].x);
            w = __hiloint2double((int)v[3].y, (int)v[3].x);
            cell = (int)v[4].x;
        }
        const int cy = (cell >> 10) & 1023, cz = cell & 1023;
        T by[CAP], bz[CAP];
        // the x weight d: every weight from the argument's coefficients and a
        // select (indexed by d, the coefficients would be read from memory)
        T bx = T(0);
#pragma unroll
        for (int i = 0; i < CAP; ++i) {
            if (i >= n) break;
            const T b = spline_one<T, ORDER>(sp, n, f0, i);
            bx = d == i ? b : bx;
        }
        spline_w<T, ORDER>(sp, n, f1, by);
        spline_w<T, ORDER>(sp, n, f2, bz);
        const W wx = W((double)w * scale) * W(bx);
        // row j: y = cy + n/2 - j (mod K1); stencil point k at z cell
        // cz + n - 1 - k of the unwrapped row
#pragma unroll
        for (int j = 0; j < CAP; ++j) {
            if (j >= n) break;
            int ly = cy + n / 2 - j;
            if (ly >= K1) ly -= K1;
            else if (ly < 0) ly += K1;
            unsigned long long *row = acc + ly * Kp + cz;
            const W wxy = wx * W(by[j]);
#pragma unroll
            for (int k = 0; k < CAP; ++k) {
                if (k >= n) break;
                if constexpr (std::is_same<W, float>::value) {
                    const int qv = round_half_up(wxy * bz[k]);
                    atomicAdd(&row[n - 1 - k], (unsigned long long)(long long)qv);
                } else {
                    const long long qv = llrint((double)wxy * (double)bz[k]);
                    atomicAdd(&row[n - 1 - k], (unsigned long long)qv);
                }
            }
        }
    };
    // the records of kPre rounds are loaded before any is used: with one
    // workgroup per CU nothing else hides their latency
    constexpr int kPre = 4;
    constexpr int RWL = std::is_same<T, float>::value ? RW : RW - 1;  // FP64: no padding word
    for (int u0 = 0; u0 < units; u0 += kPre * NT) {
        uint2 v[kPre][RWL];
        int dq[kPre];
#pragma unroll
        for (int q = 0; q < kPre; ++q) {
            const int u = u0 + q * NT + (int)threadIdx.x;
            int d = 0;
#pragma unroll
            for (int i = 1; i < CAP; ++i)
                if (i < n && u >= seg_pref[i]) d = i;
            dq[q] = u < units ? d : -1;
            const long g = u < units ? (long)seg_begin[d] + (u - seg_pref[d]) : 0;
#pragma unroll
            for (int k = 0; k < RWL; ++k) v[q][k] = recw[g * RW + k];
        }
        // the loaded words exist here (left alone, the compiler moves each
        // round's loads into its branch below, one round in flight at a time)
#pragma unroll
        for (int q = 0; q < kPre; ++q)
#pragma unroll
            for (int k = 0; k < RWL; ++k) asm volatile("" : "+v"(v[q][k].x), "+v"(v[q][k].y));
#pragma unroll
        for (int q = 0; q < kPre; ++q)
            if (dq[q] >= 0) atom(v[q], dq[q]);
    }
    // the twiddles held since the start (K1, K2 <= 1024 = NT)
    if ((int)threadIdx.x < K2) twz[threadIdx.x] = rtz;
    if ((int)threadIdx.x < K1) twy[threadIdx.x] = rty;
    __syncthreads();
    const double inv = 1.0 / scale;
    auto cellv = [&](int y, int z) -> T {
        int zz = z + H;
        if (zz >= K2) zz -= K2;
        const unsigned long long *arow = acc + y * Kp;
        unsigned long long v = arow[zz];
        if (zz < n - 1) v += arow[zz + K2];
        return T((double)(long long)v * inv);
    };
    if constexpr (R128 && std::is_same<T, float>::value) {
        // z line c = t / 8 (rows 2c, 2c + 1), j = t mod 8: x[j + 8 r] folded
        // and converted into registers, then the register-line transform
        constexpr int N = kSlab128N;
        float2 v[16];
        const int t = threadIdx.x;
        const int la = t < 8 * (N / 2) ? t >> 3 : 0, ja = t & 7;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            v[r].x = cellv(2 * la, ja + 8 * r);
            v[r].y = cellv(2 * la + 1, ja + 8 * r);
        }
        __syncthreads();
        slab128_fwd_lines<NT>(F, x0, reinterpret_cast<float2 *>(acc), twz, twy, v);
        return;
    }
    // fold the z halo (cell z holds zz = (z + H) mod K2, and zz + K2 for
    // zz < n - 1) and convert into the pair layout: element (c, z) =
    // row_2c[z] + i row_2c+1[z] at slab[c 2 RS + z]; all reads first, as the
    // slab overlays the sums
    {
        constexpr int Q = EPT;  // npair K2 <= NT EPT elements
        const int tot = npair * K2;
        T va[Q], vb[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int l = (int)threadIdx.x + q * NT;
            va[q] = vb[q] = T(0);
            if (l < tot) {
                const int c = fdiv(l, F.dK2), z = l - c * K2;
                va[q] = cellv(2 * c, z);
                if (2 * c + 1 < K1) vb[q] = cellv(2 * c + 1, z);
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int l = (int)threadIdx.x + q * NT;
            if (l < tot) {
                const int c = fdiv(l, F.dK2), z = l - c * K2;
                cplx<T> v;
                v.x = va[q];
                v.y = vb[q];
                slab[c * 2 * RS + z] = v;
            }
        }
        __syncthreads();
    }
    // the forward slab transform (fft_yz_fwd_kernel from its z pass on)
    lines_fft<T, EPT, NT, true>(slab, npair, F.dpair, 2 * RS, 1, F.plan2, twz, -1);
    {
        constexpr int Q = EPT / 2 + 1;
        const int nk = npair * RS;
        cplx<T> za[Q], zb[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int l = min((int)threadIdx.x + q * NT, nk - 1);
            const int c = fdiv(l, F.dldim), k = l - c * RS;
            za[q] = slab[c * 2 * RS + k];
            zb[q] = slab[c * 2 * RS + (k == 0 ? 0 : K2 - k)];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int l = (int)threadIdx.x + q * NT;
            if (l < nk) {
                const int c = fdiv(l, F.dldim), k = l - c * RS;
                const cplx<T> zk = za[q], zm = zb[q];
                cplx<T> a, bb;
                a.x = T(0.5) * (zk.x + zm.x);
                a.y = T(0.5) * (zk.y - zm.y);
                bb.x = T(0.5) * (zk.y + zm.y);
                bb.y = T(0.5) * (zm.x - zk.x);
                slab[(2 * c) * RS + k] = a;
                slab[(2 * c + 1) * RS + k] = bb;
            }
        }
        __syncthreads();
    }
    lines_fft<T, EPT, NT, true>(slab, RS, F.dldim, 1, RS, F.plan1, twy, -1);
    cplx<T> *out = (cplx<T> *)F.grid + (long)x0 * K1 * RS;
    for (int l = threadIdx.x; l < K1 * RS; l += NT) out[l] = slab[l];
}

}  // namespace spme
