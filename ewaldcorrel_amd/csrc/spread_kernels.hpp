// Charge spreading: two-level counting sort + LDS-tiled owner-computes scatter.
//
// The reference spreads in gather form, one grid point at a time, with
// order^3 multimap lookups per point (sfac.cpp:113-124,141-168).  Here:
//   bin      per atom: cell c = (int)(s K) and its record (fraction, w, cell,
//            original index); each workgroup counting-sorts its own atoms by
//            x plane in LDS and writes them as one contiguous chunk with the
//            plane starts (no global atomics, no capacity limits); block max |w|
//   sort     per (frame, x plane): counting sort of the plane's runs in all
//            chunks by column cy in LDS; writes the column starts and the
//            records in column order (frame-wide: planes in x order, columns
//            in y order); plane 0 also fixes the fixed-point scale
//   spread   per x-y tile of full-z columns: every atom of the
//            (tile + order - 1) neighbourhood adds the part of its stencil
//            that falls in the tile to an LDS accumulator (int64 fixed
//            point or FP64 LDS atomics); the tile is then written once with
//            coalesced stores.  No global atomics, no separate zeroing or
//            conversion pass, and every grid point is written exactly once.
// The stencil convention is the reference's: an atom in cell c with
// fraction f adds w M(f0+i) M(f1+j) M(f2+k) to ((c + n/2 - i) mod K, ...).
#pragma once

#include <cstddef>
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
#pragma clang fp contract(off)
    double s0, s1, s2;
    const gmem<double> *gx = gptr(F.x);
    scale_wrap(F.cell, gx[3 * a], gx[3 * a + 1], gx[3 * a + 2], s0, s1, s2);
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

// FP64 values: scale = 2^(62 - e) with n*max|w| < 2^e; an atom adds at most
// |w| to any one grid point in total (the spline weights of each axis sum to
// one, also when a grid shorter than the stencil wraps it onto itself), so no
// partial sum of the integer grid can overflow int64.  FP32 values: scale =
// 2^(30 - e) with max|w|*Mmax^3 < 2^e, so every contribution converts to an
// int32 (|q| < 2^30, a quantum 2^-30 of the largest one, far below FP32
// rounding) and a sum of fewer than 2^33 of them fits in int64.  Clamped so
// that it stays finite when every charge is zero.
__device__ __forceinline__ double fixed_scale(int n, double wmax, double mmax3, int fx32) {
    const double bound = fmax(fx32 ? wmax * mmax3 : (double)n * wmax, 1e-200);
    int e;
    frexp(bound, &e);
    return ldexp(1.0, (fx32 ? 30 : 62) - e);
}

// the record of atom a (scale_wrap, cell and fraction as atom_cell) as raw
// 8-byte words in the layout of AtomRec<T>
template <typename T> __host__ __device__ constexpr int rec_words() {
    return (int)(sizeof(AtomRec<T>) / sizeof(uint2));
}

template <typename T>
__device__ __forceinline__ void make_rec_raw(const FrameDesc &F, int a, int &cx, uint2 *v) {
    int c[3];
    double f[3];
    atom_cell(F, a, c, f);
    cx = c[0];
    const unsigned cell = (unsigned)((c[0] << 20) | (c[1] << 10) | c[2]);
    if constexpr (std::is_same<T, float>::value) {
        static_assert(offsetof(AtomRec<float>, cell) == 12 && offsetof(AtomRec<float>, w) == 16,
                      "AtomRec<float> layout");
        v[0] = make_uint2(__float_as_uint((float)f[0]), __float_as_uint((float)f[1]));
        v[1] = make_uint2(__float_as_uint((float)f[2]), cell);
        v[2] = make_uint2(__float_as_uint((float)gptr(F.w)[a]), (unsigned)a);
    } else {
        static_assert(offsetof(AtomRec<double>, w) == 24 && offsetof(AtomRec<double>, cell) == 32,
                      "AtomRec<double> layout");
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const unsigned long long fb = (unsigned long long)__double_as_longlong(f[d]);
            v[d] = make_uint2((unsigned)fb, (unsigned)(fb >> 32));
        }
        const unsigned long long wb = (unsigned long long)__double_as_longlong(gptr(F.w)[a]);
        v[3] = make_uint2((unsigned)wb, (unsigned)(wb >> 32));
        v[4] = make_uint2(cell, (unsigned)a);
        v[5] = make_uint2(0u, 0u);
    }
}

// ---- bin: each workgroup's atoms in x-plane order ------------------------------------------
// Workgroup b takes atoms [b BA, b BA + BA) (BA = bin_atoms<T>(), all loads in
// flight together), computes their records, counts them per x plane in LDS,
// places them in plane order in LDS and writes them as one contiguous chunk
// bins[b BA ..); binstart[p][b] (p <= K0, nbin chunks per row) is the start
// of plane p within chunk b (binstart[K0][b] = the chunk's atoms).  No global
// atomics and no capacity limit: a plane may hold any number of atoms.
constexpr int kBinThreads = 1024;
template <typename T> __host__ __device__ constexpr int bin_atoms() {
    // 72 KB (FP32) / 48 KB (FP64) of records in LDS; 3072 FP32 atoms per
    // workgroup: a 16-frame water launch (512 workgroups) in one round of
    // two per CU (2048 measured 1.8 us slower: a second, half-empty round)
    return std::is_same<T, float>::value ? 3072 : 1024;
}
template <typename T> __host__ __device__ constexpr size_t bin_lds_bytes() {
    return (size_t)bin_atoms<T>() * sizeof(AtomRec<T>);
}
__host__ __device__ inline int bin_count(int n, bool fp32) {
    const int ba = fp32 ? bin_atoms<float>() : bin_atoms<double>();
    return (n + ba - 1) / ba;
}

// exclusive prefix of v over the workgroup (blockDim.x threads, <= 1024, in
// thread order); *total receives the sum.  wt: blockDim.x / 64 ints of LDS.
__device__ __forceinline__ int block_exclusive_scan(int v, int *wt, int *total) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    int incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(incl, o);
        if (lane >= o) incl += t;
    }
    if (lane == 63) wt[wv] = incl;
    __syncthreads();
    int off = 0, tot = 0;
    for (int k = 0; k < nw; ++k) {
        const int x = wt[k];
        off += k < wv ? x : 0;
        tot += x;
    }
    __syncthreads();  // wt may be reused once every thread has read it
    *total = tot;
    return off + incl - v;
}

template <typename T>
__global__ void __launch_bounds__(kBinThreads) bin_kernel(SPME_FRAME_ARGS) {
    constexpr int BA = bin_atoms<T>(), A = BA / kBinThreads;
    constexpr int RW = rec_words<T>();
    extern __shared__ __align__(16) unsigned char smem_raw[];  // BA records in plane order
    __shared__ int pl[kMaxK];  // per-plane count, then the plane's start in the chunk
    __shared__ int wt[kBinThreads / 64];
    __shared__ double red[kBinThreads / 64];
    int b;
    const FrameDesc &F = frames[find_frame(offs, nf, b)];
    const int K0 = F.K0;
    for (int p = threadIdx.x; p < K0; p += blockDim.x) pl[p] = 0;
    __syncthreads();
    uint2 r[A][RW];
    int cx[A], lr[A];
    double wm = 0.0;
#pragma unroll
    for (int q = 0; q < A; ++q) {
        const int a = b * BA + q * kBinThreads + (int)threadIdx.x;
        cx[q] = -1;
        if (a < F.n) {
            make_rec_raw<T>(F, a, cx[q], r[q]);
            // the charge as the record holds it (FP32 records: rounded to FP32)
            wm = fmax(wm, fabs((double)(T)gptr(F.w)[a]));
            lr[q] = atomicAdd(&pl[cx[q]], 1);
        }
    }
    __syncthreads();
    // plane starts (K0 <= 1024 = kBinThreads: one plane per thread)
    {
        const int p = threadIdx.x;
        int tot;
        const int st = block_exclusive_scan(p < K0 ? pl[p] : 0, wt, &tot);
        const int nbin = (F.n + BA - 1) / BA;
        if (p < K0) {
            pl[p] = st;
            gptr(F.binstart)[(long)p * nbin + b] = st;
        }
        if (p == 0) gptr(F.binstart)[(long)K0 * nbin + b] = tot;
    }
    __syncthreads();
    uint2 *stage = reinterpret_cast<uint2 *>(smem_raw);
#pragma unroll
    for (int q = 0; q < A; ++q) {
        if (cx[q] < 0) continue;
        const int slot = pl[cx[q]] + lr[q];
#pragma unroll
        for (int w = 0; w < RW; ++w) stage[slot * RW + w] = r[q][w];
    }
    __syncthreads();
    const int nb = min(BA, F.n - b * BA);
    uint2 *dst = reinterpret_cast<uint2 *>(F.bins) + (long)b * BA * RW;
    for (int i = threadIdx.x; i < nb * RW; i += blockDim.x) gstore(dst, i, stage[i]);
    if (F.fixed) {  // one partial maximum per workgroup
        for (int off = 32; off > 0; off >>= 1) wm = fmax(wm, __shfl_xor(wm, off));
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        if (lane == 0) red[wv] = wm;
        __syncthreads();
        if (threadIdx.x == 0) {
            double m = 0.0;
            for (int k = 0; k < kBinThreads / 64; ++k) m = fmax(m, red[k]);
            gptr(F.wmax)[b] = m;
        }
    }
}

// ---- sort: one workgroup per (frame, x plane) ------------------------------------
// Counting sort of the plane's records by column.  The plane's records are
// its runs in the bin chunks; with at most 2 kSortThreads chunks a map in LDS
// gives the chunk index of each record of a window of the plane (the threads
// load consecutive records, all lanes busy).  Staged (the plane fits F.scap
// records and one window): each thread loads its records into registers
// once, counts their columns, and places them in column order in LDS,
// written out contiguously.  Larger planes: window by window, once to count
// and once to place each record straight into its place.  More chunks: each
// run is read by one wavefront, the same two passes.  Column starts (offset by the atoms of all earlier planes)
// are written either way.  The order within a column is the order of the
// cursor atomics; no result depends on it (fixed-point sums are exact
// integers and every atom's gather is its own).
constexpr int kSortThreads = 512;
constexpr int kSortRegs = 4;  // records per thread held in registers (staged sort)

// the packed cell of a record held as raw 8-byte words
template <typename T> __device__ __forceinline__ int raw_cell(const uint2 *v);
template <> __device__ __forceinline__ int raw_cell<float>(const uint2 *v) {
    static_assert(offsetof(AtomRec<float>, cell) == 12, "cell is word 1, high half");
    return (int)v[1].y;
}
template <> __device__ __forceinline__ int raw_cell<double>(const uint2 *v) {
    static_assert(offsetof(AtomRec<double>, cell) == 32, "cell is word 4, low half");
    return (int)v[4].x;
}
// the packed cell of record i of a chunk, loaded alone
template <typename T> __device__ __forceinline__ int rec_cell(const uint2 *chunk, int i) {
    constexpr int RW = rec_words<T>();
    typedef __attribute__((address_space(1))) const int GI;
    GI *wd = (GI *)(chunk + (long)i * RW);
    return std::is_same<T, float>::value ? wd[3] : wd[8];
}

// records of the staged sort that fit in LDS (the launch gives scap records)
template <typename T> __host__ __device__ inline size_t sort_lds_bytes(int scap) {
    return (size_t)scap * sizeof(AtomRec<T>);
}

template <typename T>
__global__ void __launch_bounds__(kSortThreads) sort_kernel(SPME_FRAME_ARGS) {
    constexpr int BA = bin_atoms<T>();
    constexpr int RW = rec_words<T>();
    constexpr int NT = kSortThreads, NW = kSortThreads / 64;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ int cur[kMaxK];  // column counts, then their starts, then cursors
    __shared__ int rmap[NT * kSortRegs];  // bins index of the plane's record i (a window)
    __shared__ int wtot[NW];
    __shared__ int wcnt[NW];
    __shared__ double wred[NW];
    int cx;
    const FrameDesc &F = frames[find_frame(offs, nf, cx)];
    const int K0 = F.K0, K1 = F.K1;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int nbin = (F.n + BA - 1) / BA;
    // the plane's start and length in every chunk: rows cx and cx + 1
    const gmem<int> *row0 = gptr(F.binstart) + (long)cx * nbin, *row1 = row0 + nbin;
    const uint2 *bins = reinterpret_cast<const uint2 *>(F.bins);
    // atoms of the planes before this one (part), this plane's run in chunk
    // k = thread (c0, when there are at most NT chunks) and the runs' wave sums
    // table mode (at most 2 NT chunks): thread t owns the plane's runs in
    // chunks t and t + NT, placed one after the other at ex in the plane's
    // record order (any order of the runs will do: the records are sorted by
    // column afterwards); otherwise a wavefront per run
    const bool table = nbin <= 2 * NT;
    const int ka = threadIdx.x, kb = threadIdx.x + NT;
    int part = 0, ca = 0, cb = 0, sa = 0, sb = 0;
    if (table) {
        if (ka < nbin) {
            sa = row0[ka];
            ca = row1[ka] - sa;
        }
        if (kb < nbin) {
            sb = row0[kb];
            cb = row1[kb] - sb;
        }
        part = sa + sb;
    } else {
        for (int k = threadIdx.x; k < nbin; k += blockDim.x) {
            const int s0 = row0[k];
            part += s0;
            ca += row1[k] - s0;
        }
    }
    const int c0 = ca + cb;
    int incl = c0;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(incl, o);
        if (lane >= o) incl += t;
    }
    for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off);
    if (lane == 63) {
        wtot[wv] = part;
        wcnt[wv] = incl;
    }
    for (int y = threadIdx.x; y < K1; y += blockDim.x) cur[y] = 0;
    __syncthreads();
    int pre = 0, nb = 0, woff = 0;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
        pre += wtot[k];
        nb += wcnt[k];
        woff += k < wv ? wcnt[k] : 0;
    }
    const int ex = woff + incl - c0;
    const int bega = ka * BA + sa, begb = kb * BA + sb;
    constexpr int WIN = NT * kSortRegs;  // records of one window of the map
    // the map of the plane's records [i0, i0 + WIN): the bins index of each
    // (every thread writes the parts of its runs inside the window)
    auto fill_map = [&](int i0) {
        for (int j = max(ex, i0); j < min(ex + ca, i0 + WIN); ++j) rmap[j - i0] = bega + (j - ex);
        for (int j = max(ex + ca, i0); j < min(ex + c0, i0 + WIN); ++j)
            rmap[j - i0] = begb + (j - ex - ca);
    };
    const bool staged = table && nb <= F.scap && nb <= WIN;
    uint2 r[kSortRegs][RW];
    if (staged) {
        fill_map(0);
        __syncthreads();
        // this thread's records into registers (all loads in flight together)
#pragma unroll
        for (int q = 0; q < kSortRegs; ++q) {
            const int i = q * NT + (int)threadIdx.x;
            const long g = i < nb ? rmap[i] : 0;
#pragma unroll
            for (int w = 0; w < RW; ++w) r[q][w] = i < nb ? gload(bins, g * RW + w) : make_uint2(0, 0);
        }
#pragma unroll
        for (int q = 0; q < kSortRegs; ++q)
            if (q * NT + (int)threadIdx.x < nb) atomicAdd(&cur[(raw_cell<T>(r[q]) >> 10) & 1023], 1);
    } else if (table) {
        // column counts, window by window
        for (int i0 = 0; i0 < nb; i0 += WIN) {
            fill_map(i0);
            __syncthreads();
#pragma unroll
            for (int q = 0; q < kSortRegs; ++q) {
                const int i = i0 + q * NT + (int)threadIdx.x;
                if (i < nb) atomicAdd(&cur[(rec_cell<T>(bins, rmap[i - i0]) >> 10) & 1023], 1);
            }
            __syncthreads();  // the map is rewritten for the next window
        }
    } else {
        // column counts, a wavefront per run (chunks wv, wv + NW, ...)
        for (int k = wv; k < nbin; k += NW) {
            const int s0 = row0[k], c = row1[k] - s0;
            const uint2 *run = bins + ((long)k * BA + s0) * RW;
            for (int j = lane; j < c; j += 64)
                atomicAdd(&cur[(rec_cell<T>(run, j) >> 10) & 1023], 1);
        }
    }
    __syncthreads();
    // exclusive scan of cur[0..K1): two consecutive columns per thread
    {
        const int y0 = 2 * (int)threadIdx.x;
        const int v0 = y0 < K1 ? cur[y0] : 0, v1 = y0 + 1 < K1 ? cur[y0 + 1] : 0;
        const int sm = v0 + v1;
        int ci = sm;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_up(ci, o);
            if (lane >= o) ci += t;
        }
        if (lane == 63) wcnt[wv] = ci;
        __syncthreads();
        int coff = 0;
#pragma unroll
        for (int k = 0; k < NW; ++k) coff += k < wv ? wcnt[k] : 0;
        const int e0 = coff + ci - sm;  // start of column y0 within the plane
        if (y0 < K1) {
            cur[y0] = e0;
            gptr(F.colstart)[cx * K1 + y0] = pre + e0;
        }
        if (y0 + 1 < K1) {
            cur[y0 + 1] = e0 + v0;
            gptr(F.colstart)[cx * K1 + y0 + 1] = pre + e0 + v0;
        }
        if (cx == K0 - 1 && threadIdx.x == 0) gptr(F.colstart)[K0 * K1] = pre + nb;
    }
    __syncthreads();
    uint2 *dst = reinterpret_cast<uint2 *>(F.recs) + (long)pre * RW;
    if (staged) {
        uint2 *sbuf = reinterpret_cast<uint2 *>(smem_raw);
#pragma unroll
        for (int q = 0; q < kSortRegs; ++q)
            if (q * NT + (int)threadIdx.x < nb) {
                const int slot = atomicAdd(&cur[(raw_cell<T>(r[q]) >> 10) & 1023], 1);
#pragma unroll
                for (int w = 0; w < RW; ++w) sbuf[slot * RW + w] = r[q][w];
            }
        __syncthreads();  // contiguous write-out, 16 bytes per lane
        for (int i = threadIdx.x; i < nb * RW; i += blockDim.x) gstore(dst, i, sbuf[i]);
    } else if (table) {
        // every record straight to its place, window by window
        for (int i0 = 0; i0 < nb; i0 += WIN) {
            fill_map(i0);
            __syncthreads();
#pragma unroll
            for (int q = 0; q < kSortRegs; ++q) {
                const int i = i0 + q * NT + (int)threadIdx.x;
                if (i < nb) {
                    const long g = rmap[i - i0];
                    uint2 v[RW];
#pragma unroll
                    for (int w = 0; w < RW; ++w) v[w] = gload(bins, g * RW + w);
                    // the cell by a load of its own (as below)
                    const long slot = atomicAdd(&cur[(rec_cell<T>(bins, (int)g) >> 10) & 1023], 1);
#pragma unroll
                    for (int w = 0; w < RW; ++w) gstore(dst, slot * RW + w, v[w]);
                }
            }
            __syncthreads();  // the map is rewritten for the next window
        }
    } else {
        for (int k = wv; k < nbin; k += NW) {
            const int s0 = row0[k], c = row1[k] - s0;
            const uint2 *run = bins + ((long)k * BA + s0) * RW;
            for (int j = lane; j < c; j += 64) {
                uint2 v[RW];
#pragma unroll
                for (int w = 0; w < RW; ++w) v[w] = gload(run, (long)j * RW + w);
                // the cell by a load of its own: taken from the loaded words,
                // the compiler moved the rest of them through scratch memory
                const long slot = atomicAdd(&cur[(rec_cell<T>(run, j) >> 10) & 1023], 1);
#pragma unroll
                for (int w = 0; w < RW; ++w) gstore(dst, slot * RW + w, v[w]);
            }
        }
    }
    if (cx == 0 && F.fixed) {  // the fixed-point scale, from the bin partial maxima
        double m = 0.0;
        for (int i = threadIdx.x; i < nbin; i += blockDim.x) m = fmax(m, gptr(F.wmax)[i]);
        for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
        if (lane == 0) wred[wv] = m;
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int k = 0; k < NW; ++k) m = fmax(m, wred[k]);
            *gptr(F.fxscale) = fixed_scale(F.n, m, F.mmax3, F.fx32);
        }
    }
}

// ---- spread: one workgroup per x-y tile -------------------------------------------------------
// LDS holds the tile's full-z columns as rows of Kp = K2 + n - 1 cells: an atom's z
// stencil z = cz + n/2 - k is stored unwrapped at zz = cz + n - 1 - k, so every
// stencil line is n consecutive cells, and the write-out folds zz back onto
// (zz - n/2 + 1) mod K2.  One thread per atom of the neighbourhood: its record and
// splines are loaded once, then its planes and rows that fall in the tile add
// their lines.
constexpr int kMaxSeg = 64;

// (int)floor(x + 0.5) in one instruction; ties are irrelevant here (any fixed rule
// keeps the sums deterministic)
__device__ __forceinline__ int round_half_up(float x) {
    int q;
    asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(q) : "v"(x));
    return q;
}

template <typename T, int ORDER, bool FIXED, bool ZCH>
__global__ void __launch_bounds__(kSpreadBlock)
spread_tile_kernel(SPME_FRAME_ARGS, SplineArgs<T> sp) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ int seg_begin[kMaxSeg];
    __shared__ int seg_pref[kMaxSeg + 1];
    int b;
    const FrameDesc &F = frames[find_frame(offs, nf, b)];
    const int n = order_of<ORDER>(F.order);
    constexpr int CAP = Ord<ORDER>::cap;
    const int K0 = F.K0, K1 = F.K1, K2 = F.K2;
    // tile (X, Y) of columns and, for long columns, z chunk Zc of F.tz cells
    const int nxy = F.ntx * F.nty;
    const int Zc = b / nxy, bxy = b - Zc * nxy;
    const int X = bxy % F.ntx, Y = bxy / F.ntx;
    const int x0 = X * F.tx, y0 = Y * F.ty;
    const int txr = min(F.tx, K0 - x0), tyr = min(F.ty, K1 - y0);
    const bool zchunk = ZCH && F.ntz > 1;  // ZCH: launches with chunked frames
    const int zc0 = Zc * F.tz;
    const int tzr = zchunk ? min(F.tz, K2 - zc0) : K2;
    // row length: the whole column with the z stencil unwrapped (folded back at
    // write-out), or the chunk's own cells only (stencil points outside dropped)
    const int Kp = zchunk ? tzr : K2 + n - 1;
    const int cells = txr * tyr * Kp;

    typedef typename std::conditional<FIXED, unsigned long long, double>::type Acc;
    Acc *acc = reinterpret_cast<Acc *>(smem_raw);
    {
        uint4 *z4 = reinterpret_cast<uint4 *>(smem_raw);
        for (int i = threadIdx.x; i < cells / 2; i += blockDim.x) z4[i] = make_uint4(0, 0, 0, 0);
        if ((cells & 1) && threadIdx.x == 0) acc[cells - 1] = Acc(0);
    }

    // column segments of the neighbourhood: for each cx-row d (cx = x0 - n/2 + d),
    // the cy range [y0 - n/2, y0 - n/2 + ncy) (mod K1) is one or two runs of
    // consecutive column ids, i.e. one or two runs of sorted atoms (segments 2d,
    // 2d+1).  The first wavefront builds them and scans their sizes.
    const int ncx = min(txr + n - 1, K0), ncy = min(tyr + n - 1, K1);
    const bool wrapped = ncx < txr + n - 1 || ncy < tyr + n - 1;
    const int nseg = 2 * ncx;
    if (threadIdx.x < 64) {
        const int s = threadIdx.x;
        int beg = 0, cnt = 0;
        if (s < nseg) {
            const int d = s >> 1;
            const int cy0 = imod(y0 - n / 2, K1);
            const int len0 = min(ncy, K1 - cy0);
            const int cx = imod(x0 - n / 2 + d, K0);
            if ((s & 1) == 0) {
                beg = F.colstart[cx * K1 + cy0];
                cnt = F.colstart[cx * K1 + cy0 + len0] - beg;
            } else if (len0 < ncy) {
                beg = F.colstart[cx * K1];
                cnt = F.colstart[cx * K1 + (ncy - len0)] - beg;
            }
        }
        int v = cnt;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_up(v, o);
            if (s >= o) v += t;
        }
        if (s < nseg) {
            seg_begin[s] = beg;
            seg_pref[s] = v - cnt;
        }
        if (s == 63) seg_pref[nseg] = v;
    }
    __syncthreads();
    const int units = seg_pref[nseg];
    const double scale = FIXED ? *F.fxscale : 1.0;
    typedef __attribute__((address_space(1))) const AtomRec<T> GRec;
    GRec *recs = (GRec *)F.recs;

    const int H = n / 2 - 1;
    // the stencil work of one atom (record r, neighbourhood row d)
    auto atom = [&](GRec &r, int d) {
        const int cell = r.cell;
        const int cx = cell >> 20;
        const int cy = (cell >> 10) & 1023;
        const int cz = cell & 1023;
        // z chunk: the stencil covers z = cz - H .. cz + n/2 (mod K2), H = n/2 - 1;
        // zoff = its first cell in chunk coordinates when it does not wrap
        int zoff = cz;
        bool zfast = true;
        if (zchunk) {
            zoff = cz - H - zc0;
            zfast = zoff >= 0 && zoff + n - 1 < tzr;
            if (!zfast) zoff = 0;
        }
        T bx[CAP], by[CAP], bz[CAP];
        spline_w<T, ORDER>(sp, n, r.f[0], bx);
        spline_w<T, ORDER>(sp, n, r.f[1], by);
        spline_w<T, ORDER>(sp, n, r.f[2], bz);
        // FP32: int32 contributions of the pre-scaled charge (|w scale| Mmax^3 < 2^30
        // keeps every product in range); FP64 fixed: int64 of the scaled product;
        // FP64 float: the product itself
        typedef typename std::conditional<std::is_same<T, float>::value && FIXED, float,
                                          double>::type W;
        const W ws = W(r.w * scale);
        auto add = [&](Acc *cellp, W wxy, int k) {
            if (std::is_same<W, float>::value) {
                const int qv = round_half_up((float)wxy * (float)bz[k]);
                atomicAdd((unsigned long long *)cellp, (unsigned long long)(long long)qv);
            } else if (FIXED) {
                const long long qv = llrint((double)wxy * (double)bz[k]);
                atomicAdd((unsigned long long *)cellp, (unsigned long long)qv);
            } else {
                atomicAdd((double *)cellp, (double)wxy * (double)bz[k]);
            }
        };
        auto line = [&](Acc *row, W wxy) {
            if (zfast) {
                // row points at the stencil's first cell (zz = cz, or the chunk
                // offset); stencil point k lands at row[n - 1 - k]
#pragma unroll
                for (int k = 0; k < CAP; ++k) {
                    if (k >= n) break;
                    add(&row[n - 1 - k], wxy, k);
                }
            } else {
                // chunk edge or wrap: row points at the chunk's first cell; only
                // the stencil points inside the chunk are added
#pragma unroll
                for (int k = 0; k < CAP; ++k) {
                    if (k >= n) break;
                    int z = cz + n / 2 - k;
                    if (z >= K2) z -= K2;
                    else if (z < 0) z += K2;
                    const int zl = z - zc0;
                    if ((unsigned)zl < (unsigned)tzr) add(&row[zl], wxy, k);
                }
            }
        };
        if (!wrapped) {
            // the column's place in the neighbourhood; plane i lands on lx = d - i,
            // row j on ly = dy - j
            int dy = cy - y0 + n / 2;
            if (dy < 0) dy += K1;
            else if (dy >= K1) dy -= K1;
            // the line of plane i, row j starts (i tyr + j) Kp cells before the
            // line of plane 0, row 0: a workgroup-uniform offset
            Acc *base = acc + ((d * tyr + dy) * Kp + zoff);
            const int sx = tyr * Kp;
#pragma unroll
            for (int i = 0; i < CAP; ++i) {
                if (i >= n) break;
                if ((unsigned)(d - i) >= (unsigned)txr) continue;
                const W wx = ws * W(bx[i]);
#pragma unroll
                for (int j = 0; j < CAP; ++j) {
                    if (j >= n) break;
                    if ((unsigned)(dy - j) >= (unsigned)tyr) continue;
                    const int off = __builtin_amdgcn_readfirstlane(i * sx + j * Kp);
                    line(base - off, wx * W(by[j]));
                }
            }
        } else {
            // the neighbourhood wraps onto itself: every plane and row is tried
#pragma unroll
            for (int i = 0; i < CAP; ++i) {
                if (i >= n) break;
                const int lx = imod(cx + n / 2 - i, K0) - x0;
                if ((unsigned)lx >= (unsigned)txr) continue;
                const W wx = ws * W(bx[i]);
#pragma unroll
                for (int j = 0; j < CAP; ++j) {
                    if (j >= n) break;
                    const int ly = imod(cy + n / 2 - j, K1) - y0;
                    if ((unsigned)ly >= (unsigned)tyr) continue;
                    line(acc + ((lx * tyr + ly) * Kp + zoff), wx * W(by[j]));
                }
            }
        }
    };
    if constexpr (!ZCH) {
        // one atom per thread
        for (int u = threadIdx.x; u < units; u += blockDim.x) {
            int lo = 0, hi = nseg - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (seg_pref[mid] <= u) lo = mid;
                else hi = mid - 1;
            }
            atom(recs[seg_begin[lo] + (u - seg_pref[lo])], lo >> 1);
        }
    } else {
        // One atom per thread and round; in a z chunk, the round's atoms whose
        // stencil reaches the chunk are first compacted into clist, so that
        // the wavefronts that do stencil work are full.
        __shared__ int wsum[kSpreadBlock / 64];
        int *clist = reinterpret_cast<int *>(acc + cells);  // kSpreadBlock ints
        for (int ub = 0; ub < units; ub += blockDim.x) {
            int key = -1;  // record index << 5 | neighbourhood row d
            const int u = ub + (int)threadIdx.x;
            if (u < units) {
                int lo = 0, hi = nseg - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (seg_pref[mid] <= u) lo = mid;
                    else hi = mid - 1;
                }
                const int ridx = seg_begin[lo] + (u - seg_pref[lo]);
                key = (ridx << 5) | (lo >> 1);
                if (zchunk) {  // the stencil z = cz - H .. cz + n/2 (mod K2) misses the chunk?
                    const int l0 = imod((recs[ridx].cell & 1023) - H - zc0, K2);
                    if (l0 >= tzr && l0 + n - 1 < K2) key = -1;
                }
            }
            if (zchunk) {
                const int cnt = block_compact(key, clist, wsum);
                key = (int)threadIdx.x < cnt ? clist[threadIdx.x] : -1;
            }
            if (key >= 0) atom(recs[key >> 5], key & 31);
            if (zchunk) __syncthreads();  // clist is rewritten by the next round
        }
    }
    __syncthreads();
    // write the tile once: one wavefront per row, lanes along z (coalesced), the
    // halo folded back
    typedef __attribute__((address_space(1))) T GT;
    GT *grid = (GT *)F.grid;
    const double inv = FIXED ? 1.0 / scale : 1.0;
    const long rs = 2L * F.ldim;
    const int rows = txr * tyr;
    const int lane = threadIdx.x & 63;
    const int nw = blockDim.x >> 6;
    for (int rr = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); rr < rows; rr += nw) {
        const int lx = rr / tyr, ly = rr - lx * tyr;
        const Acc *arow = acc + rr * Kp;
        GT *grow = grid + ((long)(x0 + lx) * K1 + (y0 + ly)) * rs;
        if (zchunk) {  // the chunk's own cells, nothing to fold
            for (int zl0 = 0; zl0 < tzr; zl0 += 64) {
                const int zl = zl0 + lane;
                if (zl >= tzr) continue;
                const Acc v = arow[zl];
                double o;
                if (FIXED) o = (double)(long long)v * inv;
                else o = (double)v;
                grow[zc0 + zl] = T(o);
            }
            continue;
        }
        for (int z0 = 0; z0 < K2; z0 += 64) {
            const int z = z0 + lane;
            if (z >= K2) continue;
            int zz = z + H;
            Acc v;
            if (K2 >= n) {  // at most one image of the halo per cell
                if (zz >= K2) zz -= K2;
                v = arow[zz];
                if (zz < n - 1) v += arow[zz + K2];
            } else {
                zz %= K2;
                v = arow[zz];
                for (zz += K2; zz < Kp; zz += K2) v += arow[zz];
            }
            double o;
            if (FIXED) o = (double)(long long)v * inv;
            else o = (double)v;
            grow[z] = T(o);
        }
    }
}

}  // namespace spme
