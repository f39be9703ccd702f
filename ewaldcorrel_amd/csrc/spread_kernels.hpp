// Charge spreading: column sort + LDS-tiled owner-computes scatter.
//
// The reference spreads in gather form, one grid point at a time, with
// order^3 multimap lookups per point (sfac.cpp:113-124,141-168).  Here:
//   prep     per atom: cell c = (int)(s K), column (cx, cy), rank within the
//            column (one counter atomic), block max |w|
//   scan     per frame: exclusive prefix of the column counts, counters
//            cleared for the next call, fixed-point scale fixed
//   scatter  per atom: one record (fraction, w, cell, original index) in
//            column order
//   spread   per x-y tile of full-z columns: every atom of the
//            (tile + order - 1) neighbourhood adds the part of its stencil
//            that falls in the tile to an LDS accumulator (int64 fixed
//            point or FP64 LDS atomics); the tile is then written once with
//            coalesced stores.  No global atomics, no separate zeroing or
//            conversion pass, and every grid point is written exactly once.
// The stencil convention is the reference's: an atom in cell c with
// fraction f adds w M(f0+i) M(f1+j) M(f2+k) to ((c + n/2 - i) mod K, ...).
#pragma once

#include <type_traits>

#include "common.hpp"

namespace spme {

template <int ORDER> struct Ord {
    static constexpr int cap = ORDER > 0 ? ORDER : kMaxOrder;
};
template <int ORDER> __device__ __forceinline__ int order_of(int runtime_order) {
    return ORDER > 0 ? ORDER : runtime_order;
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
template <typename T, int ORDER>
__device__ __forceinline__ T spline_one(const SplineArgs<T> &sp, int n, T x, int j) {
    T v = sp.M[j * n];
#pragma unroll
    for (int i = 1; i < Ord<ORDER>::cap; ++i) {
        if (i >= n) break;
        v = v * x + sp.M[j * n + i];
    }
    return v;
}

// cell and fraction of one atom (sfac.cpp:97-104); a coordinate wrapped to
// exactly 1.0 is the periodic image in cell 0.
__device__ __forceinline__ void atom_cell(const FrameDesc &F, int a, int c[3], double f[3]) {
    double s0, s1, s2;
    scale_wrap(F.cell, F.x[3 * a], F.x[3 * a + 1], F.x[3 * a + 2], s0, s1, s2);
    const double u0 = s0 * F.K0, u1 = s1 * F.K1, u2 = s2 * F.K2;
    c[0] = (int)u0;
    c[1] = (int)u1;
    c[2] = (int)u2;
    f[0] = u0 - c[0];
    f[1] = u1 - c[1];
    f[2] = u2 - c[2];
    if (c[0] >= F.K0) c[0] -= F.K0;
    if (c[1] >= F.K1) c[1] -= F.K1;
    if (c[2] >= F.K2) c[2] -= F.K2;
}

__device__ __forceinline__ double block_max(double v, double *red) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) red[wv] = v;
    __syncthreads();
    double m = 0.0;
    if (threadIdx.x == 0)
        for (int q = 0; q < kBlock / 64; ++q) m = fmax(m, red[q]);
    return m;
}

// FP64 values: scale = 2^(62 - e) with n*max|w| < 2^e; an atom adds at most
// |w| to any one grid point in total (the spline weights of each axis sum to
// one, also when a grid shorter than the stencil wraps it onto itself), so no
// partial sum of the integer grid can overflow int64.  FP32 values: scale =
// 2^(30 - e) with max|w|*Mmax^3 < 2^e, so every contribution converts to an
// int32 (|q| < 2^30, a quantum 2^-30 of the largest one, far below FP32
// rounding) and a sum of fewer than 2^33 of them fits in int64.  Clamped so
// that it stays finite when every charge is zero.
__device__ __forceinline__ double fixed_scale(int n, unsigned long long maxbits, double mmax3,
                                              int fx32) {
    const double wmax = __longlong_as_double((long long)maxbits);
    const double bound = fmax(fx32 ? wmax * mmax3 : (double)n * wmax, 1e-200);
    int e;
    frexp(bound, &e);
    return ldexp(1.0, (fx32 ? 30 : 62) - e);
}

// ---- prep: one thread per atom -------------------------------------------------------
__global__ void __launch_bounds__(kBlock) prep_kernel(SPME_FRAME_ARGS) {
    __shared__ double red[kBlock / 64];
    int b;
    const FrameDesc &F = frames[find_frame(offs, nf, b)];
    const int a = b * kBlock + threadIdx.x;
    double wa = 0.0;
    if (a < F.n) {
        int c[3];
        double f[3];
        atom_cell(F, a, c, f);
        const int col = c[0] * F.K1 + c[1];
        F.a_col[a] = col;
        F.a_rank[a] = atomicAdd(F.colcount + col, 1);
        wa = fabs(F.w[a]);
    }
    if (F.fixed) {
        const double m = block_max(wa, red);
        if (threadIdx.x == 0 && m > 0.0)
            atomicMax(F.maxbits, (unsigned long long)__double_as_longlong(m));
    }
}

// ---- scan: one workgroup per frame ----------------------------------------------------
// Tiles of 1024 x 16 columns: each thread loads its 16 consecutive counts
// at once (all loads in flight), the 1024 thread sums are scanned in LDS,
// and the tile's offsets continue from the running total of earlier tiles.
__global__ void __launch_bounds__(1024) scan_kernel(SPME_FRAME_ARGS) {
    constexpr int E = 16;
    __shared__ int part[1024];
    const FrameDesc &F = frames[nf <= 1 ? 0 : blockIdx.x];
    const int ncol = F.K0 * F.K1;
    int carry = 0;
    for (int base = 0; base < ncol; base += 1024 * E) {
        const int i0 = base + (int)threadIdx.x * E;
        // colcount is zero-padded to a multiple of 1024*E: four unconditional
        // 16-byte loads per thread, all in flight together
        int v[E];
        const int4 *src = reinterpret_cast<const int4 *>(F.colcount + i0);
#pragma unroll
        for (int e4 = 0; e4 < E / 4; ++e4) {
            const int4 q = src[e4];
            v[4 * e4 + 0] = q.x;
            v[4 * e4 + 1] = q.y;
            v[4 * e4 + 2] = q.z;
            v[4 * e4 + 3] = q.w;
        }
        int s = 0;
#pragma unroll
        for (int e = 0; e < E; ++e) s += v[e];
        part[threadIdx.x] = s;
        __syncthreads();
        // Hillis-Steele inclusive scan over the 1024 thread sums
        for (int off = 1; off < 1024; off <<= 1) {
            const int t = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
            __syncthreads();
            part[threadIdx.x] += t;
            __syncthreads();
        }
        int run = carry + (threadIdx.x > 0 ? part[threadIdx.x - 1] : 0);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            if (i0 + e < ncol) F.colstart[i0 + e] = run;
            run += v[e];
        }
        int4 *dst = reinterpret_cast<int4 *>(F.colcount + i0);
#pragma unroll
        for (int e4 = 0; e4 < E / 4; ++e4) dst[e4] = make_int4(0, 0, 0, 0);
        carry += part[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) F.colstart[ncol] = carry;
    if (threadIdx.x == 0 && F.fixed) {
        *F.fxscale = fixed_scale(F.n, *F.maxbits, F.mmax3, F.fx32);
        *F.maxbits = 0ull;
    }
}

// ---- scatter: one thread per atom -------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kBlock) scatter_kernel(SPME_FRAME_ARGS) {
    int b;
    const FrameDesc &F = frames[find_frame(offs, nf, b)];
    const int a = b * kBlock + threadIdx.x;
    if (a >= F.n) return;
    int c[3];
    double f[3];
    atom_cell(F, a, c, f);
    const int pos = F.colstart[F.a_col[a]] + F.a_rank[a];
    AtomRec<T> r;
    r.f[0] = T(f[0]);
    r.f[1] = T(f[1]);
    r.f[2] = T(f[2]);
    r.w = F.w[a];
    r.cell = (c[0] << 20) | (c[1] << 10) | c[2];
    r.perm = a;
    ((AtomRec<T> *)F.recs)[pos] = r;
}

// ---- spread: one workgroup per x-y tile -------------------------------------------------------
constexpr int kMaxSeg = 64;

template <typename T, int ORDER, bool FIXED>
__global__ void __launch_bounds__(kSpreadBlock)
spread_tile_kernel(SPME_FRAME_ARGS, SplineArgs<T> sp) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ int seg_begin[kMaxSeg];
    __shared__ int seg_cnt[kMaxSeg];
    __shared__ int seg_pref[kMaxSeg + 1];
    __shared__ int seg_ilo[kMaxSeg / 2], seg_np[kMaxSeg / 2];
    __shared__ unsigned seg_inv[kMaxSeg / 2];
    int b;
    const FrameDesc &F = frames[find_frame(offs, nf, b)];
    const int n = order_of<ORDER>(F.order);
    constexpr int CAP = Ord<ORDER>::cap;
    const int K0 = F.K0, K1 = F.K1, K2 = F.K2;
    const int X = b % F.ntx, Y = b / F.ntx;
    const int x0 = X * F.tx, y0 = Y * F.ty;
    const int txr = min(F.tx, K0 - x0), tyr = min(F.ty, K1 - y0);
    const int cells = txr * tyr * K2;

    typedef typename std::conditional<FIXED, unsigned long long, double>::type Acc;
    Acc *acc = reinterpret_cast<Acc *>(smem_raw);
    for (int i = threadIdx.x; i < cells; i += blockDim.x) acc[i] = Acc(0);

    // column segments of the neighbourhood: for each cx, the cy range
    // [y0 - n/2, y0 + tyr + n/2 - 2] (mod K1) is one or two runs of
    // consecutive column ids, i.e. one or two runs of sorted atoms.  An atom
    // in neighbourhood row d (cx = x0 - n/2 + d) reaches the tile with its
    // planes i in [d - txr + 1, d] only; when the neighbourhood wraps onto
    // itself (K0 < txr + n - 1) every plane is tried.
    const int ncx = min(txr + n - 1, K0), ncy = min(tyr + n - 1, K1);
    const bool wrapped = ncx < txr + n - 1;
    if ((int)threadIdx.x < ncx) {  // segments 2d and 2d+1 of column row d
        const int d = threadIdx.x;
        const int cy0 = imod(y0 - n / 2, K1);
        const int len0 = min(ncy, K1 - cy0);
        const int cx = imod(x0 - n / 2 + d, K0);
        const int col = cx * K1 + cy0;
        const int a0 = F.colstart[col], a1 = F.colstart[col + len0];
        int c0 = 0, c1 = 0;
        if (len0 < ncy) {
            c0 = F.colstart[cx * K1];
            c1 = F.colstart[cx * K1 + (ncy - len0)];
        }
        const int ilo = wrapped ? 0 : max(d - txr + 1, 0);
        const int ihi = wrapped ? n - 1 : min(d, n - 1);
        const int np = ihi - ilo + 1;  // >= 1
        seg_begin[2 * d] = a0;
        seg_cnt[2 * d] = (a1 - a0) * np;
        seg_begin[2 * d + 1] = c0;
        seg_cnt[2 * d + 1] = (c1 - c0) * np;
        seg_ilo[d] = ilo;
        seg_np[d] = np;
        seg_inv[d] = np > 1 ? 0xffffffffu / (unsigned)np + 1u : 0u;  // ceil(2^32 / np)
    }
    __syncthreads();
    const int nseg = 2 * ncx;
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int s = 0; s < nseg; ++s) {
            seg_pref[s] = tot;
            tot += seg_cnt[s];
        }
        seg_pref[nseg] = tot;
    }
    __syncthreads();
    const int units = seg_pref[nseg];
    const double scale = FIXED ? *F.fxscale : 1.0;
    const AtomRec<T> *recs = (const AtomRec<T> *)F.recs;

    // one unit = (atom, i): the x-plane of its stencil.  UPT units per thread
    // per round: their records are loaded together, then processed.
    constexpr int UPT = 4;
    for (int base = 0; base < units; base += UPT * (int)blockDim.x) {
        int pos[UPT], iu[UPT], cell[UPT];
        T fx[UPT], fy[UPT], fz[UPT];
        double wv[UPT];
#pragma unroll
        for (int q = 0; q < UPT; ++q) {
            const int u = base + q * (int)blockDim.x + (int)threadIdx.x;
            pos[q] = -1;
            if (u < units) {
                int lo = 0, hi = nseg - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (seg_pref[mid] <= u) lo = mid;
                    else hi = mid - 1;
                }
                // off = a * np + p: atom a of the segment, its plane ilo + p
                const int d = lo >> 1, off = u - seg_pref[lo];
                // multiply-high by ceil(2^32/np) (np >= 2) divides exactly below 2^24
                const int np = seg_np[d];
                const int a = np == 1             ? off
                              : off < (1 << 24) ? (int)__umulhi((unsigned)off, seg_inv[d])
                                                : off / np;
                iu[q] = seg_ilo[d] + (off - a * seg_np[d]);
                pos[q] = seg_begin[lo] + a;
            }
        }
        // unconditional loads (index 0 is always allocated) keep them all in flight
#pragma unroll
        for (int q = 0; q < UPT; ++q) {
            const AtomRec<T> &r = recs[pos[q] >= 0 ? pos[q] : 0];
            cell[q] = r.cell;
            fx[q] = r.f[0];
            fy[q] = r.f[1];
            fz[q] = r.f[2];
            wv[q] = r.w;
        }
#pragma unroll
        for (int q = 0; q < UPT; ++q) {
            if (pos[q] < 0) continue;
            const int i = iu[q];
            const int cx = cell[q] >> 20;
            const int cy = (cell[q] >> 10) & 1023;
            const int cz = cell[q] & 1023;
            const int lx = imod(cx + n / 2 - i, K0) - x0;
            if (lx < 0 || lx >= txr) continue;
            T by[CAP], bz[CAP];
            spline_w<T, ORDER>(sp, n, fy[q], by);
            spline_w<T, ORDER>(sp, n, fz[q], bz);
            // z and y targets once per unit; one conditional wrap suffices when
            // the stencil is no longer than the grid, imod otherwise
            int zt[CAP], yt[CAP];
#pragma unroll
            for (int k = 0; k < CAP; ++k) {
                int z = cz + n / 2 - k;
                if (K2 >= n) z = z < 0 ? z + K2 : (z >= K2 ? z - K2 : z);
                else z = imod(z, K2);
                zt[k] = z;
                int y = cy + n / 2 - k;
                if (K1 >= n) y = y < 0 ? y + K1 : (y >= K1 ? y - K1 : y);
                else y = imod(y, K1);
                yt[k] = y - y0;
            }
            if (std::is_same<T, float>::value && FIXED) {
                // FP32 path: the scaled x factor (|wx| < 2^30, formed in FP64 since
                // the scale alone may exceed the FP32 range), FP32 products, int32
                // contributions
                const float wx =
                    (float)(wv[q] * (double)spline_one<T, ORDER>(sp, n, fx[q], i) * scale);
#pragma unroll
                for (int j = 0; j < CAP; ++j) {
                    if (j >= n) break;
                    const int ly = yt[j];
                    if (ly < 0 || ly >= tyr) continue;
                    const float wxy = wx * (float)by[j];
                    Acc *row = acc + (lx * tyr + ly) * K2;
#pragma unroll
                    for (int k = 0; k < CAP; ++k) {
                        if (k >= n) break;
                        const int qv = __float2int_rn(wxy * (float)bz[k]);
                        atomicAdd((unsigned long long *)&row[zt[k]],
                                  (unsigned long long)(long long)qv);
                    }
                }
            } else {
                const double wx = wv[q] * (double)spline_one<T, ORDER>(sp, n, fx[q], i);
#pragma unroll
                for (int j = 0; j < CAP; ++j) {
                    if (j >= n) break;
                    const int ly = yt[j];
                    if (ly < 0 || ly >= tyr) continue;
                    const double wxy = wx * (double)by[j];
                    Acc *row = acc + (lx * tyr + ly) * K2;
#pragma unroll
                    for (int k = 0; k < CAP; ++k) {
                        if (k >= n) break;
                        if (FIXED) {
                            const long long qv = llrint(wxy * (double)bz[k] * scale);
                            atomicAdd((unsigned long long *)&row[zt[k]], (unsigned long long)qv);
                        } else {
                            atomicAdd((double *)&row[zt[k]], wxy * (double)bz[k]);
                        }
                    }
                }
            }
        }
    }
    __syncthreads();
    // write the tile once (coalesced along z)
    T *grid = (T *)F.grid;
    const double inv = FIXED ? 1.0 / scale : 1.0;
    const long rs = 2L * F.ldim;
    for (int idx = threadIdx.x; idx < cells; idx += blockDim.x) {
        const int r = idx / K2;
        const int z = idx - r * K2;
        const int lx = r / tyr, ly = r - lx * tyr;
        double v;
        if (FIXED) v = (double)(long long)acc[idx] * inv;
        else v = (double)acc[idx];
        grid[((long)(x0 + lx) * K1 + (y0 + ly)) * rs + z] = T(v);
    }
}

}  // namespace spme
