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

// ---- tiled gather: one workgroup per x-y tile of columns ----------------------------
// The atoms of a tile's columns (contiguous runs of the column-sorted records,
// one per cx) read grid rows of the (tx + n) x (ty + n) block around it only.
// That block, each row unwrapped to K2 + n - 1 reals (position P holds z =
// (P - n/2) mod K2), is staged in LDS with coalesced row loads; then one thread
// per atom sums its whole stencil from LDS: the n^2 scattered global reads per
// atom become LDS reads, and each grid row is fetched from L2 once per tile.
// Workgroup size of the tiled gather: 1024 threads (held to 64 VGPRs, so that
// two workgroups fill a CU with 32 waves) for FP32 order 4, whose stencil
// sums fit; 512 threads held to 128 VGPRs otherwise.  Host-side limits use
// the smaller size (rows staged) and the larger one (the compacted atom list).
constexpr int kGatherBlock = 512;
constexpr int kGatherBlockMax = 1024;
template <typename T, int ORDER> constexpr int gather_block() {
    return sizeof(T) == 4 && ORDER == 4 ? 1024 : 512;
}
template <typename T, int ORDER> constexpr int gather_waves() {
    return gather_block<T, ORDER>() == 1024 ? 8 : 4;
}
constexpr int kGatherLoadRegs = 24;     // staged values each lane holds before storing
constexpr int kGatherMaxZ = 4;          // row length K2 + n - 1 <= 64 kGatherMaxZ

// MZ = ceil((K2 + n - 1) / 64) row chunks per lane; ZCH: launches with
// z-chunked frames
template <typename T, int ORDER, int MZ, bool ZCH>
__global__ void __launch_bounds__((gather_block<T, ORDER>()))
    __attribute__((amdgpu_waves_per_eu((gather_waves<T, ORDER>()))))
gather_tile_kernel(SPME_FRAME_ARGS, SplineArgs<T> sp) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    int b;
    const int bid = xcd_block((int)blockIdx.x, (int)gridDim.x);
    const FrameDesc &F = frames[find_frame_of(offs, nf, bid, b)];
    const int n = order_of<ORDER>(F.order), h = n / 2;
    constexpr int CAP = Ord<ORDER>::cap;
    const int K0 = F.K0, K1 = F.K1, K2 = F.K2;
    // tile (X, Y) of columns and, for long columns, z chunk Zc of F.gtz cells:
    // the chunk's atoms (ceil cell k2 in [zc0, zc0 + tzr)) with rows staged
    // from zc0 - n/2 to zc0 + tzr + n/2 - 2
    const int nxy = F.gntx * F.gnty;
    const int Zc = b / nxy, bxy = b - Zc * nxy;
    const int X = bxy % F.gntx, Y = bxy / F.gntx;
    const int x0 = X * F.gtx, y0 = Y * F.gty;
    const int txr = min(F.gtx, K0 - x0), tyr = min(F.gty, K1 - y0);
    const bool zchunk = ZCH && F.gntz > 1;
    const int zc0 = Zc * F.gtz;
    const int tzr = zchunk ? min(F.gtz, K2 - zc0) : K2;
    const int PY = tyr + n, Kp = tzr + n - 1;
    const int rows = (txr + n) * PY;
    T *tile = reinterpret_cast<T *>(smem_raw);

    // the tile's atoms: cx-row d holds the columns [y0, y0 + tyr) of cx = x0 + d,
    // one run of records; the first wavefront scans the run lengths into LDS
    __shared__ int run_begin[64];
    __shared__ int run_pref[65];
    if (threadIdx.x < 64) {
        const int d = threadIdx.x;
        int beg = 0, cnt = 0;
        if (d < txr) {
            const int c = (x0 + d) * K1 + y0;
            beg = F.colstart[c];
            cnt = F.colstart[c + tyr] - beg;
        }
        int v = cnt;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_up(v, o);
            if (d >= o) v += t;
        }
        run_begin[d] = beg;
        run_pref[d] = v - cnt;
        if (d == 63) run_pref[64] = v;
    }

    // stage the rows: wavefronts over rows, lanes along z.  Every load and store
    // is unconditional within a row: a lane past the end of the row repeats its
    // own previous element (the same value to the same address).  Lane q of a
    // wavefront holds the source offset of the wavefront's q-th row, read back
    // with readlane.
    {
        typedef __attribute__((address_space(1))) const T GT;
        const int lane = threadIdx.x & 63;
        const int nw = blockDim.x >> 6;
        const int w0 = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        int zsrc[MZ], zdst[MZ];
#pragma unroll
        for (int m = 0; m < MZ; ++m) {
            int zp = lane + 64 * m;
            if (zp >= Kp) zp = m > 0 ? zdst[m - 1] : Kp - 1;
            int z = zc0 + zp - h;
            if (K2 >= n) z = z < 0 ? z + K2 : (z >= K2 ? z - K2 : z);
            else z = imod(z, K2);
            zsrc[m] = z;
            zdst[m] = zp;
        }
        const int nrw = rows > w0 ? (rows - w0 + nw - 1) / nw : 0;  // rows of this wavefront
        int my_src = 0, my_dst = 0;
        {
            const int r = w0 + lane * nw;
            if (r < rows) {
                const int a = r / PY, c = r - a * PY;
                const int gx = imod(x0 - h + a, K0), gy = imod(y0 - h + c, K1);
                my_src = (gx * K1 + gy) * (2 * F.ldim);
                my_dst = r * Kp;
            }
        }
        GT *grid = (GT *)F.grid;
        constexpr int RIF = kGatherLoadRegs / MZ;  // rows in flight per wavefront
        for (int q0 = 0; q0 < nrw; q0 += RIF) {
            T v[RIF][MZ];
#pragma unroll
            for (int q = 0; q < RIF; ++q) {
                if (q0 + q < nrw) {  // uniform
                    GT *src = grid + __builtin_amdgcn_readlane(my_src, q0 + q);
#pragma unroll
                    for (int m = 0; m < MZ; ++m) v[q][m] = src[zsrc[m]];
                }
            }
#pragma unroll
            for (int q = 0; q < RIF; ++q) {
                if (q0 + q < nrw) {
                    T *dst = tile + __builtin_amdgcn_readlane(my_dst, q0 + q);
#pragma unroll
                    for (int m = 0; m < MZ; ++m) dst[zdst[m]] = v[q][m];
                }
            }
        }
    }
    __syncthreads();

    const int units = run_pref[64];
    typedef __attribute__((address_space(1))) const AtomRec<T> GRec;
    GRec *recs = (GRec *)F.recs;
    // the stencil work of one atom (record r)
    auto atom = [&](GRec &r) {
        const int a = r.perm;
        const double wa = r.w;
        const int cell = r.cell;
        const T f0 = r.f[0], f1 = r.f[1], f2 = r.f[2];
        const int c0 = cell >> 20, c1 = (cell >> 10) & 1023, c2 = cell & 1023;
        // k = ceil(u): c + 1 for f > 0, c for f = 0; its stencil covers the
        // staged planes k - x0 .. k - x0 + n - 1 (and likewise in y and z)
        const int a0 = (f0 > T(0) ? c0 + 1 : c0) - x0;
        const int b0 = (f1 > T(0) ? c1 + 1 : c1) - y0;
        int k2 = f2 > T(0) ? c2 + 1 : c2;
        if (k2 >= K2) k2 -= K2;
        if (zchunk) k2 -= zc0;  // the compaction kept this chunk's atoms only
        const T t0 = f0 > T(0) ? T(1) - f0 : T(0);
        const T t1 = f1 > T(0) ? T(1) - f1 : T(0);
        const T t2 = f2 > T(0) ? T(1) - f2 : T(0);
        T bx[CAP], dbx[CAP], by[CAP], dby[CAP], bz[CAP], dbz[CAP];
        spline_wd<T, ORDER>(sp, n, t0, bx, dbx);
        spline_wd<T, ORDER>(sp, n, t1, by, dby);
        spline_wd<T, ORDER>(sp, n, t2, bz, dbz);
        // row addresses by running sums (the empty asm keeps the compiler from
        // precomputing all n^2 offsets as loop invariants, which it spills)
        int plane = (a0 * PY + b0) * Kp + k2;
        const int step_x = PY * Kp;
        T e0 = 0, e1 = 0, e2 = 0, ph = 0;
        // planes unrolled for FP32 up to order 4; FP64 and wider stencils keep
        // one plane's n^2 loads in flight (a full unroll needs far more than
        // 128 VGPRs)
        constexpr int kPlanesUnrolled = sizeof(T) == 4 && ORDER > 0 && ORDER <= 4 ? 2 : 1;
#pragma unroll kPlanesUnrolled
        for (int i = 0; i < CAP; ++i) {
            if (i >= n) break;
            T pb = 0, pdb = 0, pbdz = 0;
            int row = plane;
#pragma unroll
            for (int j = 0; j < CAP; ++j) {
                if (j >= n) break;
                T zb = 0, zdb = 0;
#pragma unroll
                for (int k = 0; k < CAP; ++k) {
                    if (k >= n) break;
                    const T q = tile[row + k];
                    zb += bz[k] * q;
                    zdb += dbz[k] * q;
                }
                pb += by[j] * zb;
                pdb += dby[j] * zb;
                pbdz += by[j] * zdb;
                row += Kp;
                asm volatile("" : "+v"(row));
            }
            plane += step_x;
            asm volatile("" : "+v"(plane));
            e0 += dbx[i] * pb;
            e1 += bx[i] * pdb;
            e2 += bx[i] * pbdz;
            ph += bx[i] * pb;
        }
        const CellArgs &cl = F.cell;
        if (F.dx) {
            const double f = -wa * cl.iV;
            const double g0 = (double)e0 * f * K0, g1 = (double)e1 * f * K1,
                         g2 = (double)e2 * f * K2;
            F.dx[3 * a + 0] = cl.iL00 * g0;
            F.dx[3 * a + 1] = cl.iL10 * g0 + cl.iL11 * g1;
            F.dx[3 * a + 2] = cl.iL20 * g0 + cl.iL21 * g1 + cl.iL22 * g2;
        }
        if (F.phi) F.phi[a] = (double)ph * cl.iV;
    };
    if constexpr (!ZCH) {
        // one atom per thread
        for (int u = threadIdx.x; u < units; u += blockDim.x) {
            int lo = 0;
            for (int d = 1; d < txr; ++d)
                if (run_pref[d] <= u) lo = d;
            atom(recs[run_begin[lo] + (u - run_pref[lo])]);
        }
    } else {
        // One atom per thread and round; in a z chunk, the round's atoms of the
        // chunk are first compacted into clist, so that the wavefronts that do
        // stencil work are full.
        __shared__ int wsum[gather_block<T, ORDER>() / 64];
        int *clist = reinterpret_cast<int *>(tile + rows * Kp);  // kGatherBlockMax ints
        for (int ub = 0; ub < units; ub += blockDim.x) {
            int ridx = -1;
            const int u = ub + (int)threadIdx.x;
            if (u < units) {
                int lo = 0;
                for (int d = 1; d < txr; ++d)
                    if (run_pref[d] <= u) lo = d;
                ridx = run_begin[lo] + (u - run_pref[lo]);
                if (zchunk) {  // ceil cell in z inside the chunk?
                    const int c2 = recs[ridx].cell & 1023;
                    int k2 = recs[ridx].f[2] > T(0) ? c2 + 1 : c2;
                    if (k2 >= K2) k2 -= K2;
                    if ((unsigned)(k2 - zc0) >= (unsigned)tzr) ridx = -1;
                }
            }
            if (zchunk) {
                const int cnt = block_compact(ridx, clist, wsum);
                ridx = (int)threadIdx.x < cnt ? clist[threadIdx.x] : -1;
            }
            if (ridx >= 0) atom(recs[ridx]);
            if (zchunk) __syncthreads();  // clist is rewritten by the next round
        }
    }
}

}  // namespace spme
