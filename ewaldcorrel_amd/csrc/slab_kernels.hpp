// y-z slab transforms: one workgroup per (frame, x plane), the whole plane
// in LDS, so the z and y passes of a 3-D transform cost one HBM read and one
// HBM write of the grid together (the reference runs them inside one FFTW
// r2c / c2r plan, sfac.cpp:70-74,274-283).
//
// LDS: twiddles of z (K2) and y (K1), then the slab as K1e rows (K1 rounded
// up to even) of RS = ldim complex numbers, i.e. exactly the half-complex
// layout of the plane in HBM.  Row pair (2c, 2c+1) doubles as one complex z
// line of K2 <= 2 RS elements (z = row_2c + i row_2c+1):
//   forward: load pairs -> z FFT -> untangle into rows
//            A_k = (Z_k + conj Z_{K2-k})/2, B_k = (Z_k - conj Z_{K2-k})/(2i)
//            -> y FFT over the ldim columns -> contiguous store
//   inverse: contiguous load -> y FFT -> Z_k = A_k + i B_k with Hermitian
//            completion (imaginary parts of kz = 0 and Nyquist dropped) ->
//            z FFT -> real rows (Re -> row 2c, Im -> row 2c+1)
// Selected when the slab fits in LDS and the y and z plans use radices
// 2, 3, 4, 5 and 8 only; otherwise the separate z and y line kernels run.
#pragma once

#include "common.hpp"
#include "fft_kernels.hpp"
#include "recip_kernels.hpp"

namespace spme {

// Variants by elements per thread (EPT), each compiled separately. FP32:
// planes of up to 768 x 11 elements (128 x 128 needs 8,320) use 768 threads
// held to 80 VGPRs, so that two workgroups (24 waves) share a CU; up to
// 512 x 20 elements 512 threads held to 128 VGPRs; larger planes 512 threads
// with 32 elements each. FP64 planes (one workgroup per CU by LDS) use 1024
// threads.
template <typename T> struct SlabCfg;
template <> struct SlabCfg<float> {
    static constexpr int small = 11, medium = 20, large = 32;
    static constexpr int block_small = 768, block_medium = 512, block_large = 512;
    static constexpr int waves_small = 6, waves_medium = 4, waves_large = 4;
};
template <> struct SlabCfg<double> {  // medium and large are one variant
    static constexpr int small = 10, medium = 20, large = 20;
    static constexpr int block_small = 1024, block_medium = 1024, block_large = 1024;
    static constexpr int waves_small = 4, waves_medium = 4, waves_large = 4;
};
template <typename T, int EPT> constexpr int slab_block() {
    return EPT == SlabCfg<T>::small    ? SlabCfg<T>::block_small
           : EPT == SlabCfg<T>::medium ? SlabCfg<T>::block_medium
                                       : SlabCfg<T>::block_large;
}
template <typename T, int EPT> constexpr int slab_waves() {
    return EPT == SlabCfg<T>::small    ? SlabCfg<T>::waves_small
           : EPT == SlabCfg<T>::medium ? SlabCfg<T>::waves_medium
                                       : SlabCfg<T>::waves_large;
}
// dynamic LDS budget: gfx950 has 160 KB per workgroup, 1 KB of which is
// left to the static arrays of the reductions
constexpr size_t kSlabLds = 159 * 1024;


__host__ __device__ inline int slab_rows(int K1) { return (K1 + 1) & ~1; }

// The z lines (row pairs) and the y lines (columns) are shared by the
// workgroup's threads; the elements per thread decide the variant (small or
// large, compiled separately; 0: the slab does not apply).
template <typename T> __host__ __device__ inline int slab_ept_for(int K1, int K2, int ldim) {
    const int npair = (K1 + 1) / 2;
    const int elems = max(npair * K2, ldim * K1);
    constexpr int Bs = SlabCfg<T>::block_small, Bm = SlabCfg<T>::block_medium,
                  Bl = SlabCfg<T>::block_large;
    if ((elems + Bs - 1) / Bs <= SlabCfg<T>::small) return SlabCfg<T>::small;
    if ((elems + Bm - 1) / Bm <= SlabCfg<T>::medium) return SlabCfg<T>::medium;
    if ((elems + Bl - 1) / Bl <= SlabCfg<T>::large) return SlabCfg<T>::large;
    return 0;
}

template <typename T>
__host__ __device__ inline size_t slab_lds_bytes(int K1, int K2, int ldim) {
    return ((size_t)K1 + K2 + (size_t)slab_rows(K1) * ldim) * sizeof(cplx<T>);
}

template <typename T>
__device__ __forceinline__ void slab_twiddles(cplx<T> *twz, cplx<T> *twy, const FrameDesc &F) {
    const cplx<T> *gz = (const cplx<T> *)F.tw2, *gy = (const cplx<T> *)F.tw1;
    for (int t = threadIdx.x; t < F.K2; t += (int)blockDim.x) twz[t] = gz[t];
    for (int t = threadIdx.x; t < F.K1; t += (int)blockDim.x) twy[t] = gy[t];
}

// bytes (a multiple of 16) from src (16-B aligned) to LDS dst (16-B aligned)
// by direct global -> LDS loads, 16 B per lane: wave w copies the 1 KB chunks
// w, w + nwaves, ...  Issued from asm so that the compiler does not wait for
// them before unrelated LDS reads; the caller waits (vmcnt 0) and then
// synchronizes the workgroup before reading dst.
__device__ __forceinline__ void lds_stream16(const void *src, void *dst, int bytes) {
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int nw = (int)blockDim.x >> 6;
    const char *sp = (const char *)src;
    const unsigned base =
        (unsigned)(size_t)(__attribute__((address_space(3))) char *)((char *)dst);
    for (int off = wv * 1024; off < bytes; off += nw * 1024) {
        const unsigned m0 = __builtin_amdgcn_readfirstlane(base + (unsigned)off);
        if (off + lane * 16 < bytes) {
            const char *gp = sp + off + lane * 16;
            asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, off"
                         :
                         : "s"(m0), "v"(gp)
                         : "memory", "m0");
        }
    }
}

template <typename T, int EPT>
__global__ void __launch_bounds__((slab_block<T, EPT>()))
    __attribute__((amdgpu_waves_per_eu((slab_waves<T, EPT>()))))
    fft_yz_fwd_kernel(SPME_FRAME_ARGS) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int kSlabBlock = slab_block<T, EPT>();
    constexpr int U = 8;  // staging registers per round
    int x;
    const FrameDesc &F = frames[find_frame(offs, nf, x)];
    const int K1 = F.K1, K2 = F.K2, RS = F.ldim;
    const int npair = (K1 + 1) >> 1;
    cplx<T> *twz = reinterpret_cast<cplx<T> *>(smem_raw);
    cplx<T> *twy = twz + K2;
    cplx<T> *slab = twy + K1;
    slab_twiddles<T>(twz, twy, F);
    const T *g = (const T *)F.grid + (long)x * K1 * 2 * RS;
    // packed loads: element (c, z) = row_2c[z] + i row_2c+1[z]
    const int tot = npair * K2;
    for (int base = 0; base < tot; base += U * kSlabBlock) {
        T ra[U], rb[U];
#pragma unroll
        for (int q = 0; q < U; ++q) {
            const int l = min(base + (int)threadIdx.x + q * kSlabBlock, tot - 1);
            const int c = fdiv(l, F.dK2), z = l - c * K2;
            ra[q] = g[(2 * c) * 2 * RS + z];
            rb[q] = g[min(2 * c + 1, K1 - 1) * 2 * RS + z];
        }
#pragma unroll
        for (int q = 0; q < U; ++q) {
            const int l = base + (int)threadIdx.x + q * kSlabBlock;
            if (l < tot) {
                const int c = fdiv(l, F.dK2), z = l - c * K2;
                cplx<T> v;
                v.x = ra[q];
                v.y = (2 * c + 1 < K1) ? rb[q] : T(0);
                slab[c * 2 * RS + z] = v;
            }
        }
    }
    __syncthreads();
    // per-stage butterfly indices where registers bound the kernel (kept
    // across stages they are spilled); the small FP64 variant has registers
    // to spare (one workgroup per CU by LDS)
    constexpr bool lean = std::is_same<T, float>::value || EPT != SlabCfg<T>::small;
    lines_fft<T, EPT, kSlabBlock, lean>(slab, npair, F.dpair, 2 * RS, 1, F.plan2, twz, -1);
    // untangle the pairs into rows 2c and 2c+1 (all reads before the barrier)
    {
        constexpr int Q = EPT / 2 + 1;
        const int nk = npair * RS;
        cplx<T> za[Q], zb[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int l = min((int)threadIdx.x + q * kSlabBlock, nk - 1);
            const int c = fdiv(l, F.dldim), k = l - c * RS;
            za[q] = slab[c * 2 * RS + k];
            zb[q] = slab[c * 2 * RS + (k == 0 ? 0 : K2 - k)];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int l = (int)threadIdx.x + q * kSlabBlock;
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
    // y transform: ldim lines (columns kz), elements at stride RS
    lines_fft<T, EPT, kSlabBlock, lean>(slab, RS, F.dldim, 1, RS, F.plan1, twy, -1);
    cplx<T> *out = (cplx<T> *)F.grid + (long)x * K1 * RS;
    for (int l = threadIdx.x; l < K1 * RS; l += kSlabBlock) out[l] = slab[l];
}

// finish 1: workgroup 0 of each frame first finalizes the reduction of the
// fused x pass that ran before.
template <typename T, int EPT>
__global__ void __launch_bounds__((slab_block<T, EPT>()))
    __attribute__((amdgpu_waves_per_eu((slab_waves<T, EPT>()))))
    fft_yz_inv_kernel(SPME_FRAME_ARGS, int finish) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int kSlabBlock = slab_block<T, EPT>();
    constexpr int U = 8;
    int x;
    const FrameDesc &F = frames[find_frame(offs, nf, x)];
    if (finish && x == 0) {
        finalize_reduction(F, F.xB);
        __syncthreads();
    }
    const int K1 = F.K1, K2 = F.K2, RS = F.ldim;
    const int npair = (K1 + 1) >> 1;
    cplx<T> *twz = reinterpret_cast<cplx<T> *>(smem_raw);
    cplx<T> *twy = twz + K2;
    cplx<T> *slab = twy + K1;
    slab_twiddles<T>(twz, twy, F);
    const cplx<T> *in = (const cplx<T> *)F.grid + (long)x * K1 * RS;
    const int tot = K1 * RS;
    const int bytes = tot * (int)sizeof(cplx<T>);
    if (((size_t)slab & 15) == 0 && ((size_t)in & 15) == 0 && (bytes & 15) == 0) {
        // the whole plane in flight at once (direct global -> LDS loads, no
        // staging registers): 2-3 us less per 16-frame launch than staging
        // through registers in rounds; the forward kernel's paired rows, at
        // 4 B per lane, measured slower and keep the staged loads
        lds_stream16(in, slab, bytes);
        __builtin_amdgcn_s_waitcnt(0);
    } else {
        for (int base = 0; base < tot; base += U * kSlabBlock) {
            T vx[U], vy[U];
#pragma unroll
            for (int q = 0; q < U; ++q) {
                const cplx<T> v = in[min(base + (int)threadIdx.x + q * kSlabBlock, tot - 1)];
                vx[q] = v.x;
                vy[q] = v.y;
            }
#pragma unroll
            for (int q = 0; q < U; ++q) {
                const int l = base + (int)threadIdx.x + q * kSlabBlock;
                if (l < tot) {
                    cplx<T> v;
                    v.x = vx[q];
                    v.y = vy[q];
                    slab[l] = v;
                }
            }
        }
    }
    if (K1 & 1) {  // the missing odd row of the last pair is zero
        for (int k = threadIdx.x; k < RS; k += kSlabBlock) {
            cplx<T> z;
            z.x = z.y = T(0);
            slab[K1 * RS + k] = z;
        }
    }
    __syncthreads();
    // per-stage butterfly indices where registers bound the kernel (kept
    // across stages they are spilled), as in the forward kernel
    constexpr bool lean = std::is_same<T, float>::value || EPT != SlabCfg<T>::small;
    lines_fft<T, EPT, kSlabBlock, lean>(slab, RS, F.dldim, 1, RS, F.plan1, twy, +1);
    // Z_k = A_k + i B_k and the Hermitian completion (all reads before the barrier)
    {
        constexpr int Q = EPT / 2 + 1;
        const int nk = npair * RS;
        const bool even = (K2 % 2) == 0;
        cplx<T> ra[Q], rb[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int l = min((int)threadIdx.x + q * kSlabBlock, nk - 1);
            const int c = fdiv(l, F.dldim), k = l - c * RS;
            ra[q] = slab[(2 * c) * RS + k];
            rb[q] = slab[(2 * c + 1) * RS + k];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int l = (int)threadIdx.x + q * kSlabBlock;
            if (l < nk) {
                const int c = fdiv(l, F.dldim), k = l - c * RS;
                cplx<T> a = ra[q], bb = rb[q];
                if (k == 0 || (even && k == K2 / 2)) {
                    a.y = T(0);
                    bb.y = T(0);
                }
                cplx<T> z;
                z.x = a.x - bb.y;
                z.y = a.y + bb.x;
                slab[c * 2 * RS + k] = z;
                const int km = K2 - k;
                if (k > 0 && km >= RS) {
                    cplx<T> zc;  // conj(A) + i conj(B)
                    zc.x = a.x + bb.y;
                    zc.y = bb.x - a.y;
                    slab[c * 2 * RS + km] = zc;
                }
            }
        }
        __syncthreads();
    }
    lines_fft<T, EPT, kSlabBlock, lean>(slab, npair, F.dpair, 2 * RS, 1, F.plan2, twz, +1);
    T *g = (T *)F.grid + (long)x * K1 * 2 * RS;
    const int nz = npair * K2;
    for (int l = threadIdx.x; l < nz; l += kSlabBlock) {
        const int c = fdiv(l, F.dK2), z = l - c * K2;
        const cplx<T> v = slab[c * 2 * RS + z];
        g[(2 * c) * 2 * RS + z] = v.x;
        if (2 * c + 1 < K1) g[(2 * c + 1) * 2 * RS + z] = v.y;
    }
}

}  // namespace spme
