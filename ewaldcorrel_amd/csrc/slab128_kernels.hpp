// y-z slab transforms of 128 x 128 planes (FP32) with the line transforms in
// registers: the same pass structure as fft_yz_fwd_kernel / fft_yz_inv_kernel
// (one workgroup per (frame, x plane), the plane in LDS; the reference runs
// these transforms inside its FFTW r2c / c2r plan, sfac.cpp:70-74,274-283),
// but every 128-point line is held by 8 threads, 16 points each, so that the
// butterflies, twiddles and indices are fixed at compile time:
//   forward line: thread j holds x[j + 8 r] (r < 16); 16-point DFT over r;
//     twiddle W_128^(j k1); exchange through LDS; the thread takes k1 = 2j,
//     2j + 1: 8-point DFT over j -> X[k1 + 16 k2] (k2 < 8)
//   inverse line: the same steps backwards (8-point inverse DFT over k2,
//     twiddle W_128^-(j k1), exchange, 16-point inverse DFT over k1)
// which are the operations of the plan (16, 8) of the generic kernels.
// LDS: the z and y twiddles, then one region that holds in turn the z
// lines (row pairs as complex lines, line c at 129 c: an odd stride, so that
// the 64 lines of a wavefront fill every bank twice), the exchange buffers
// (the same layout) and the half-complex slab (128 rows of 65 complex).
// Lanes run along the lines (or along k where a pass works on whole rows)
// so that no LDS access of a wavefront hits a bank more than twice.
#pragma once

#include "common.hpp"
#include "fft_kernels.hpp"
#include "recip_kernels.hpp"
#include "slab_kernels.hpp"

namespace spme {

constexpr int kSlab128N = 128;
// the slab-kernel code of a launch that runs these kernels (below every EPT)
constexpr int kSlabKernel128 = 1;
constexpr int kSlab128RS = kSlab128N / 2 + 1;  // ldim
constexpr int kSlab128LS = 137;                // line stride in LDS (complex), 9 mod 32
constexpr int kSlab128Threads = 576;           // 65 columns x 8 threads, rounded to waves
constexpr size_t kSlab128Region = (size_t)(kSlab128N / 2 + 1) * kSlab128LS * 8;  // >= the slab

__host__ __device__ inline size_t slab128_lds_bytes() {
    return 2 * kSlab128N * sizeof(float2) + kSlab128Region;
}

// the planes this variant takes: 128 x 128 y-z planes, both plans (16, 8)
__host__ inline bool slab128_applies(int K1, int K2, const FftPlanArgs &p1,
                                     const FftPlanArgs &p2) {
    auto p168 = [](const FftPlanArgs &p) {
        return p.N == kSlab128N && p.nstages == 2 && p.radix[0] == 16 && p.radix[1] == 8;
    };
    return K1 == kSlab128N && K2 == kSlab128N && p168(p1) && p168(p2);
}

// barrier for the LDS only
__device__ __forceinline__ void lds_sync() { __syncthreads(); }

// forward 128-point line: in: x[j + 8 r] in v; out: u[h][k2] = X[(2j + h) + 16 k2].
// buf: the line's exchange buffer; the caller synchronizes before (the buffer
// may alias data others still read) and every thread of the workgroup calls
// it.  Idle threads (active false) compute on the buffer they are given but
// write nothing: values that exist on some threads only are kept twice by
// the register allocator, which spills them.
__device__ __forceinline__ void line128_fwd(float2 (&v)[16], float2 (&u)[2][8], float2 *buf,
                                            const float2 *tw, int j, bool active) {
    bfly16<float>(v, -1);
#pragma unroll
    for (int k1 = 1; k1 < 16; ++k1) v[k1] = cmul<float>(v[k1], twiddle<float>(tw, j * k1, -1));
    if (active) {
#pragma unroll
        for (int k1 = 0; k1 < 16; ++k1) buf[k1 * 8 + j] = v[k1];
    }
    lds_sync();
#pragma unroll
    for (int h = 0; h < 2; ++h) {
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) u[h][jj] = buf[(2 * j + h) * 8 + jj];
        bfly8<float>(u[h], -1);
    }
}

// inverse 128-point line: in: u[h][k2] = X[(2j + h) + 16 k2]; out: x[j + 8 r] in v.
// The caller synchronizes before (buf may alias data others still read);
// idle threads as in line128_fwd.
__device__ __forceinline__ void line128_inv(float2 (&u)[2][8], float2 (&v)[16], float2 *buf,
                                            const float2 *tw, int j, bool active) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int k1 = 2 * j + h;
        bfly8<float>(u[h], +1);
#pragma unroll
        for (int jj = 1; jj < 8; ++jj)
            u[h][jj] = cmul<float>(u[h][jj], twiddle<float>(tw, jj * k1, +1));
        if (active) {
#pragma unroll
            for (int jj = 0; jj < 8; ++jj) buf[k1 * 8 + jj] = u[h][jj];
        }
    }
    lds_sync();
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) v[k1] = buf[k1 * 8 + j];
    bfly16<float>(v, +1);
}

// the y-line lane of thread t: columns 0..63 at t mod 64 with j = t / 64
// (t < 512), column 64 with j = t - 512 (t < 520); the others idle
// (column 0, j = 0); returns whether the thread's line is its own
__device__ __forceinline__ bool slab128_ymap(int t, int &col, int &j) {
    if (t < 512) {
        col = t & 63;
        j = t >> 6;
        return true;
    }
    const bool a = t < 520;
    col = a ? 64 : 0;
    j = a ? t - 512 : 0;
    return a;
}

__device__ __forceinline__ void slab128_twiddles(float2 *twz, float2 *twy, const FrameDesc &F) {
    const float2 *gz = (const float2 *)F.tw2, *gy = (const float2 *)F.tw1;
    for (int t = threadIdx.x; t < kSlab128N; t += (int)blockDim.x) {
        twz[t] = gz[t];
        twy[t] = gy[t];
    }
}

// The forward transform of plane x from the z lines in registers to the
// half-complex plane in HBM, for a workgroup of NT >= 576 threads: thread
// t < 512 holds x[j + 8 r] (r < 16) of z line c = t / 8 (rows 2c, 2c + 1 as
// re + i im) at j = t mod 8 in v (lanes along j: 8 lines per wavefront).
// reg: the LDS region (kSlab128Region bytes, 16-B aligned), free to write
// (the caller has synchronized after its last read of what it held);
// twz, twy: the twiddles in LDS.
template <int NT>
__device__ __forceinline__ void slab128_fwd_lines(const FrameDesc &F, int x, float2 *reg,
                                                  const float2 *twz, const float2 *twy,
                                                  float2 (&v)[16]) {
    constexpr int N = kSlab128N, RS = kSlab128RS, LS = kSlab128LS;
    const int t = threadIdx.x;
    float2 *S = reg;  // half-complex slab: rows of RS
    // z lines, first half (lanes along j): the 16-point DFT, the twiddles
    // and the exchange into line c's buffer at reg + c LS
    {
        const bool aa = t < 8 * (N / 2);
        const int la = aa ? t >> 3 : 0, ja = t & 7;
        bfly16<float>(v, -1);
#pragma unroll
        for (int k1 = 1; k1 < 16; ++k1)
            v[k1] = cmul<float>(v[k1], twiddle<float>(twz, ja * k1, -1));
        if (aa) {
#pragma unroll
            for (int k1 = 0; k1 < 16; ++k1) reg[la * LS + k1 * 8 + ja] = v[k1];
        }
    }
    lds_sync();
    // second half: lanes along the lines (64 per wavefront, one j each; the
    // odd stride LS fills every bank twice): the 8-point DFTs, then Z in
    // natural order in the line's buffer; idle threads (t >= 512) take j = 0
    // and write nothing
    float2 u[2][8];
    const bool za = t < 8 * (N / 2);
    const int lz = t & 63, jz = za ? t >> 6 : 0;
    float2 *lb = reg + lz * LS;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) u[h][jj] = lb[(2 * jz + h) * 8 + jj];
        bfly8<float>(u[h], -1);
    }
    lds_sync();
    if (za) {
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int jj = 0; jj < 8; ++jj) lb[(2 * jz + h) + 16 * jj] = u[h][jj];
    }
    lds_sync();
    // untangle, lanes along k: A_k = (Z_k + conj Z_{N-k}) / 2,
    // B_k = (Z_k - conj Z_{N-k}) / (2i), k <= N/2 -> rows 2c, 2c + 1
    {
        constexpr int Q = (N / 2 * RS + NT - 1) / NT;
        float2 A[Q], B[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int l = min(t + q * NT, N / 2 * RS - 1);
            const int c = l / RS, k = l - c * RS;
            const float2 zk = reg[c * LS + k];
            const float2 zm = reg[c * LS + ((N - k) & (N - 1))];
            A[q].x = 0.5f * (zk.x + zm.x);
            A[q].y = 0.5f * (zk.y - zm.y);
            B[q].x = 0.5f * (zk.y + zm.y);
            B[q].y = 0.5f * (zm.x - zk.x);
        }
        lds_sync();
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int l = t + q * NT;
            if (l < N / 2 * RS) {
                const int c = l / RS, k = l - c * RS;
                S[(2 * c) * RS + k] = A[q];
                S[(2 * c + 1) * RS + k] = B[q];
            }
        }
    }
    lds_sync();
    // y lines: columns kz < RS, rows j + 8 r; columns 0..63 as the z lines
    // (consecutive columns across the lanes), column 64 by threads 512..519
    int ly, jy;
    const bool ya = slab128_ymap(t, ly, jy);
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = S[(jy + 8 * r) * RS + ly];
    lds_sync();
    float2 *cb = reg + ly * LS;
    line128_fwd(v, u, cb, twy, jy, ya);
    lds_sync();
    if (ya) {
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int jj = 0; jj < 8; ++jj) S[((2 * jy + h) + 16 * jj) * RS + ly] = u[h][jj];
    }
    lds_sync();
    // the half-complex plane out, 16 B per lane (global stores)
    gmem<u32x4_t> *out = (gmem<u32x4_t> *)((float2 *)F.grid + (long)x * N * RS);
    const u32x4_t *s4 = reinterpret_cast<const u32x4_t *>(S);
    for (int l = t; l < N * RS / 2; l += NT) out[l] = s4[l];
}

__global__ void __launch_bounds__(kSlab128Threads) __attribute__((amdgpu_waves_per_eu(5)))
fft_yz_fwd128_kernel(SPME_FRAME_ARGS) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int N = kSlab128N, RS = kSlab128RS;
    int x;
    const FrameDesc &F = frames[find_frame(offs, nf, x)];
    float2 *twz = reinterpret_cast<float2 *>(smem_raw);
    float2 *twy = twz + N;
    float2 *reg = twy + N;  // the shared region
    const int t = threadIdx.x;
    // the real plane (K1 rows of 2 RS floats, contiguous) straight into LDS
    float *P = reinterpret_cast<float *>(reg);
    lds_stream16((const float *)F.grid + (long)x * N * 2 * RS, P, N * 2 * RS * (int)sizeof(float));
    slab128_twiddles(twz, twy, F);
    __builtin_amdgcn_s_waitcnt(0);
    lds_sync();
    // z line c = t / 8 (rows 2c, 2c + 1), j = t mod 8: lanes along j, 8 lines
    // per wavefront (the row reads fall on 36 banks)
    float2 v[16];
    {
        const int la = t < 8 * (N / 2) ? t >> 3 : 0, ja = t & 7;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            v[r].x = P[(2 * la) * 2 * RS + ja + 8 * r];
            v[r].y = P[(2 * la + 1) * 2 * RS + ja + 8 * r];
        }
    }
    lds_sync();
    slab128_fwd_lines<kSlab128Threads>(F, x, reg, twz, twy, v);
}

// finish 1: workgroup 0 of each frame first finalizes the reduction of the
// fused x pass that ran before.
__global__ void __launch_bounds__(kSlab128Threads) __attribute__((amdgpu_waves_per_eu(5)))
fft_yz_inv128_kernel(SPME_FRAME_ARGS, int finish) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int N = kSlab128N, RS = kSlab128RS, LS = kSlab128LS, NT = kSlab128Threads;
    int x;
    const FrameDesc &F = frames[find_frame(offs, nf, x)];
    if (finish && x == 0) {
        finalize_reduction(F, F.xB);
        __syncthreads();
    }
    float2 *twz = reinterpret_cast<float2 *>(smem_raw);
    float2 *twy = twz + N;
    float2 *reg = twy + N;
    float2 *S = reg;
    const int t = threadIdx.x;
    const float2 *in = (const float2 *)F.grid + (long)x * N * RS;
    lds_stream16(in, S, N * RS * (int)sizeof(float2));
    slab128_twiddles(twz, twy, F);
    __builtin_amdgcn_s_waitcnt(0);
    lds_sync();
    // y lines: column kz holds Y[ky], ky = k1 + 16 k2 (lanes as in the
    // forward kernel; idle threads take column 0 and write nothing)
    int ly, jy;
    const bool ya = slab128_ymap(t, ly, jy);
    float2 v[16], u[2][8];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) u[h][jj] = S[((2 * jy + h) + 16 * jj) * RS + ly];
    lds_sync();
    float2 *cb = reg + ly * LS;
    line128_inv(u, v, cb, twy, jy, ya);
    lds_sync();
    if (ya) {
#pragma unroll
        for (int r = 0; r < 16; ++r) S[(jy + 8 * r) * RS + ly] = v[r];
    }
    lds_sync();
    // z lines, lanes along k: Z_k = A_k + i B_k (rows 2c, 2c + 1, k <= N/2,
    // imaginary parts of k = 0 and N/2 dropped) and Z_{N-k} = conj(A_k) +
    // i conj(B_k) (0 < k < N/2), in natural order into line c's buffer
    {
        constexpr int Q = (N / 2 * RS + NT - 1) / NT;
        float2 A[Q], B[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int l = min(t + q * NT, N / 2 * RS - 1);
            const int c = l / RS, k = l - c * RS;
            A[q] = S[(2 * c) * RS + k];
            B[q] = S[(2 * c + 1) * RS + k];
        }
        lds_sync();
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int l = t + q * NT;
            if (l < N / 2 * RS) {
                const int c = l / RS, k = l - c * RS;
                float2 a = A[q], bb = B[q];
                if (k == 0 || k == N / 2) {
                    a.y = 0.f;
                    bb.y = 0.f;
                }
                reg[c * LS + k] = make_float2(a.x - bb.y, a.y + bb.x);
                if (k > 0 && k < N / 2)
                    reg[c * LS + (N - k)] = make_float2(a.x + bb.y, bb.x - a.y);
            }
        }
    }
    lds_sync();
    const bool za = t < 8 * (N / 2);
    const int lz = t & 63, jz = za ? t >> 6 : 0;
    float2 *lb = reg + lz * LS;
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) u[h][jj] = lb[(2 * jz + h) + 16 * jj];
    lds_sync();
    line128_inv(u, v, lb, twz, jz, za);
    lds_sync();
    // real rows: line c = Re (row 2c) + i Im (row 2c + 1) back into its
    // buffer, then coalesced stores (z fastest) of the K2 cells of each row
    if (za) {
#pragma unroll
        for (int r = 0; r < 16; ++r) lb[jz + 8 * r] = v[r];
    }
    lds_sync();
    gmem<float> *g = gptr((float *)F.grid) + (long)x * N * 2 * RS;  // global stores
    for (int l = t; l < N / 2 * N; l += NT) {
        const int c = l >> 7, z = l & (N - 1);
        const float2 w = reg[c * LS + z];
        g[(2 * c) * 2 * RS + z] = w.x;
        g[(2 * c + 1) * 2 * RS + z] = w.y;
    }
}

}  // namespace spme
