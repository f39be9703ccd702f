// Shared device/host definitions for the SPME reciprocal-space engine.
//
// Grid layout in HBM (identical to the reference's in-place FFTW buffer,
// sfac.cpp:70-74, so get_S is a plain copy): a K0 x K1 x (2*ldim) array of
// reals with K2 of each row used; the same memory viewed as K0 x K1 x ldim
// interleaved complex numbers holds the half-complex spectrum.
//
// Every kernel receives FrameDescs (one per SPME system: grid, box, tables,
// atoms, workspace) in device memory.  A single evaluation passes one; a
// batch passes the array plus, per stage, the prefix sums of the blocks each
// frame needs, and every workgroup finds its frame by binary search.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spme {

constexpr int kMaxOrder = 12;      // largest even B-spline order compiled in
constexpr int kMaxRadixStages = 16;
constexpr int kBlock = 256;        // threads per workgroup (spreading uses kSpreadBlock)
constexpr int kSpreadBlock = 512;
constexpr int kMaxK = 1024;        // largest grid dimension supported

template <typename T> struct cplx_t;
template <> struct cplx_t<float> { using type = float2; };
template <> struct cplx_t<double> { using type = double2; };
template <typename T> using cplx = typename cplx_t<T>::type;

// Global-memory views of generic pointers.  An access through a generic
// pointer compiles to a flat instruction, which counts against lgkmcnt as
// well as vmcnt, so each LDS wait after it waits for that memory access too
// (a loop of LDS reads feeding stores then waits for every store in turn);
// through these views the access is a global instruction (vmcnt only).
// 8- and 16-byte words move as native vectors (the HIP vector types have
// no assignment in another address space).
template <typename T> using gmem = __attribute__((address_space(1))) T;
typedef unsigned u32x2_t __attribute__((ext_vector_type(2)));
typedef unsigned u32x4_t __attribute__((ext_vector_type(4)));
template <typename T> __device__ __forceinline__ gmem<T> *gptr(T *p) { return (gmem<T> *)p; }
template <typename T> __device__ __forceinline__ const gmem<T> *gptr(const T *p) {
    return (const gmem<T> *)p;
}
__device__ __forceinline__ uint2 gload(const uint2 *p, long i) {
    const u32x2_t v = ((const gmem<u32x2_t> *)p)[i];
    return make_uint2(v.x, v.y);
}
__device__ __forceinline__ void gstore(uint2 *p, long i, uint2 v) {
    u32x2_t w;
    w.x = v.x;
    w.y = v.y;
    ((gmem<u32x2_t> *)p)[i] = w;
}

// Box: lower-triangular inverse iL (rows = reciprocal basis) and 1/V.
// L6 = [a_x, b_y, c_z, b_x, c_x, c_y] (sfac.hpp:37-52).
struct CellArgs {
    double iL00, iL10, iL11, iL20, iL21, iL22;
    double iV;
};

// Division by a launch-invariant divisor d >= 1 without the integer divide
// sequence: q = (umulhi(n, m) + n) >> s with s = ceil(log2 d) and
// m = floor(2^32 (2^s - d) / d) + 1, exact for 0 <= n < 2^31.
// Division by a launch-invariant d without a divide instruction.
//   fdiv:   any n < 2^31: (mulhi(n, m) + n) >> s (a quarter-rate 32-bit mulhi)
//   fdiv15: n < 2^15 (every LDS index of the line transforms): with
//           k = 15 + ceil(log2 d) and m15 = ceil(2^k / d) < 2^17, n m15 < 2^32
//           and floor(n m15 / 2^k) = floor(n / d) exactly (the rounding error
//           n (m15 d - 2^k) / (d 2^k) < 2^15 / 2^k <= 1/d cannot carry
//           floor(n / d) over an integer); one full-rate 24-bit multiply.
struct FastDiv {
    unsigned d, m, s;
    unsigned m15, s15;
};
inline FastDiv make_fastdiv(unsigned d) {
    FastDiv f;
    f.d = d;
    f.s = 0;
    while ((1ull << f.s) < d) ++f.s;
    f.m = (unsigned)(((1ull << 32) * ((1ull << f.s) - d)) / d + 1);
    f.s15 = 15 + f.s;
    f.m15 = d < (1u << 15) ? (unsigned)(((1ull << f.s15) + d - 1) / d) : 0u;
    return f;
}
__device__ __forceinline__ int fdiv(int n, const FastDiv &f) {
    return (int)((__umulhi((unsigned)n, f.m) + (unsigned)n) >> f.s);
}
__device__ __forceinline__ int fdiv15(int n, const FastDiv &f) {
    return (int)(__umul24((unsigned)n, f.m15) >> f.s15);
}
// product of two non-negative ints below 2^24 (LDS indices): full-rate multiply
__device__ __forceinline__ int mul24(int a, int b) {
    return (int)__umul24((unsigned)a, (unsigned)b);
}

struct FftPlanArgs {
    int N;
    int nstages;
    int generic;  // a stage with a radix other than 2,3,4,5,8 (needs two buffers)
    int radix[kMaxRadixStages];
    // per stage s: N / radix[s], the span Ns (product of earlier radices) and
    // the twiddle step N / (Ns radix[s])
    FastDiv dNR[kMaxRadixStages];
    FastDiv dNs[kMaxRadixStages];
    int twstep[kMaxRadixStages];
};

// Coefficients of M_n(x+j), x in [0,1), row j, highest power first.
template <typename T> struct SplineArgs {
    T M[kMaxOrder * kMaxOrder];
};

enum TableMode {
    kTableFull = 0,      // A[H] and dA[6H] read from HBM (any rad_fn)
    kTableAnalytic = 1,  // closed-form Ewald A and dA from the indices (no table)
};

// One atom in column order, written once by the sort and read coalesced by
// the spreading and the gathers: fraction f = u - c of u = s K in the working
// precision, cell c packed as cx << 20 | cy << 10 | cz (K <= 1024), charge
// (FP32 records hold it in FP32, as the FP32 engine spreads and gathers it;
// FP64 records in FP64), and the atom's index in the caller's arrays.  Copied
// as 8-byte words (rec_words).
template <typename T> struct AtomRec;
template <> struct __align__(8) AtomRec<float> {
    float f[3];
    int cell;
    float w;
    int perm;
};
template <> struct __align__(16) AtomRec<double> {
    double f[3];
    double w;
    int cell;
    int perm;
    int pad[2];
};
static_assert(sizeof(AtomRec<float>) == 24, "AtomRec<float> is 24 bytes");
static_assert(sizeof(AtomRec<double>) == 48, "AtomRec<double> is 48 bytes");

// a box whose x axis separates from y and z in reciprocal space (iL10 = iL20
// = 0): m_x depends on kx only and m_y, m_z on (ky, kz) only
__host__ __device__ inline bool x_separable(const CellArgs &cl) {
    return cl.iL10 == 0.0 && cl.iL20 == 0.0;
}

struct FrameDesc {
    // geometry
    int K0, K1, K2, ldim, order;
    int n;  // atoms
    CellArgs cell;
    // grid (T*) and FFT data
    void *grid;
    const double *iB0, *iB1, *iB2;
    const void *tw0, *tw1, *tw2;  // cplx<T> twiddles exp(-2 pi i t/K_d)
    FftPlanArgs plan0, plan1, plan2;
    // influence function
    const double *table;  // A[H] then dA[6H] (kTableFull)
    int tabmode;
    int infl32;        // FP32 engine: closed-form influence and thread sums in FP32 (option)
    double ewald_fac;  // -(pi/eta)^2
    double max_m2;
    CellArgs tcell;    // box the influence function was built for (set_L keeps it)
    // atoms (device pointers, caller-owned)
    const double *w, *x;
    double *dx, *phi;
    double *out;  // E, virial[6]
    // sort + spreading workspace
    int *colstart;    // K0*K1 + 1 exclusive prefix
    void *bins;       // AtomRec<T>[n]: each bin workgroup's chunk in x-plane order
    int *binstart;    // [K0 + 1][bin]: plane starts within each chunk
    int scap;         // records of a plane the sort stages in LDS
    void *recs;       // AtomRec<T>[n] in column order
    double *wmax;     // max |w| of each bin workgroup (fixed point), reduced by the sort
    double *fxscale;  // fixed-point scale, written by the sort
    double mmax3;
    int fixed;
    int fx32;        // FP32 spreading: int32 contributions (|q| < 2^30) into int64 sums
    int tx, ty, ntx, nty;  // spreading tiles
    int tz, ntz;           // z chunks of a spreading tile (ntz = 1: the whole column)
    int gtx, gty, gntx, gnty;  // gather tiles (gtx = 0: no tiled gather)
    int gtz, gntz;             // z chunks of a gather tile (gntz = 1: the whole column)
    // FFT line blocking
    int zC;          // complex lines (two real rows each) per z workgroup
    int yC, yT;      // lines per y tile, tiles over ldim
    int xC, xLog;    // x lines per workgroup (2^xLog consecutive (ky, kz) columns)
    int xB;          // x workgroups per frame
    FastDiv dldim;   // ldim
    FastDiv dK2;     // K2
    FastDiv dpair;   // (K1 + 1) / 2 row pairs of a slab
    // reductions (influence)
    double *partials;
    int nred_i;      // workgroups of the standalone influence launch
};

// Kernel arguments of every stage: frames[0..nf) in device memory (read
// only, so the uniform descriptor fields are fetched with scalar loads) and,
// for nf > 1, offs[f] = first workgroup of frame f in this launch.
#define SPME_FRAME_ARGS \
    const FrameDesc *__restrict__ frames, const int *__restrict__ offs, int nf

// Workgroups are handed to the 8 XCDs of the chip round-robin (workgroup i
// to XCD i mod 8), and each XCD has its own L2.  For kernels whose
// neighbouring workgroups read the same data, xcd_block gives XCD x the
// contiguous range of logical workgroups [start_x, start_x + count_x), so
// the shared lines stay in one L2.  A bijection of [0, nb).
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int x = b % kXcds, i = b / kXcds;
    const int q = nb / kXcds, r = nb % kXcds;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
}

// frame of logical workgroup bid; b = workgroup index within the frame
__device__ __forceinline__ int find_frame_of(const int *__restrict__ offs, int nf, int bid,
                                             int &b) {
    if (nf <= 1) {
        b = bid;
        return 0;
    }
    int lo = 0, hi = nf - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (offs[mid] <= bid) lo = mid;
        else hi = mid - 1;
    }
    b = bid - offs[lo];
    return lo;
}
__device__ __forceinline__ int find_frame(const int *__restrict__ offs, int nf, int &b) {
    return find_frame_of(offs, nf, (int)blockIdx.x, b);
}

// Fractional coordinates wrapped to [0,1): s = wrap(iL^T x) (sfac.hpp:56-62).
// No contraction: every kernel (and the host) that computes an atom's cell
// and fraction gets the same bits, whichever of them does it.
__host__ __device__ inline void scale_wrap(const CellArgs &c, double x0, double x1,
                                           double x2, double &s0, double &s1,
                                           double &s2) {
#pragma clang fp contract(off)
    s0 = x0 * c.iL00 + x1 * c.iL10 + x2 * c.iL20;
    s1 = x1 * c.iL11 + x2 * c.iL21;
    s2 = x2 * c.iL22;
    s0 -= floor(s0);
    s1 -= floor(s1);
    s2 -= floor(s2);
}

// Block-wide stream compaction: the keys >= 0 of the workgroup's threads go to
// list[0..total) in thread order; returns total.  wsum holds a count per
// wavefront.  Every thread must call it (two barriers).
__device__ __forceinline__ int block_compact(int key, int *list, int *wsum) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const unsigned long long m = __ballot(key >= 0);
    const int pre = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[w] = __popcll(m);
    __syncthreads();
    int off = 0, tot = 0;
    for (int q = 0; q < nw; ++q) {
        const int c = wsum[q];
        off += q < w ? c : 0;
        tot += c;
    }
    if (key >= 0) list[off + pre] = key;
    __syncthreads();
    return tot;
}

__host__ __device__ inline int imod(int a, int b) {
    int r = a % b;
    return r < 0 ? r + b : r;
}

}  // namespace spme
