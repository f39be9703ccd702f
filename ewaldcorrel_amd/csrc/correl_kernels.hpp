// Structure-factor accumulation on the device (the callers of get_S:
// run.py:36-47 sums |S|^2 over frames, gmx_ewald.py:78-107 sums the species
// cross terms S_i conj(S_j), plot.py:9-32 averages over |k| shells).  The
// accumulators are FP64 device arrays owned by the caller, indexed like the
// half-complex grid, so nothing crosses PCIe per frame.
#pragma once

#include "common.hpp"

namespace spme {

// acc[h] += |S_h|^2 over the H half-complex points of S-hat
template <typename T>
__global__ void __launch_bounds__(kBlock) accum_s2_kernel(const cplx<T> *__restrict__ S, long H,
                                                          double *__restrict__ acc) {
    const long h = (long)blockIdx.x * kBlock + threadIdx.x;
    if (h >= H) return;
    const cplx<T> v = S[h];
    const double re = (double)v.x, im = (double)v.y;
    acc[h] += re * re + im * im;
}

// acc[2h], acc[2h+1] += Re, Im of S_a conj(S_b)
template <typename T>
__global__ void __launch_bounds__(kBlock) accum_cross_kernel(const cplx<T> *__restrict__ Sa,
                                                             const cplx<T> *__restrict__ Sb,
                                                             long H, double *__restrict__ acc) {
    const long h = (long)blockIdx.x * kBlock + threadIdx.x;
    if (h >= H) return;
    const cplx<T> a = Sa[h], b = Sb[h];
    const double ar = a.x, ai = a.y, br = b.x, bi = b.y;
    acc[2 * h] += ar * br + ai * bi;
    acc[2 * h + 1] += ai * br - ar * bi;
}

// One thread per |k| bin: the points of bin b are perm[off[b] .. off[b+1]),
// in increasing h (a fixed order, so the sums are reproducible).
__global__ void __launch_bounds__(kBlock) radial_kernel(const double *__restrict__ val,
                                                        const int *__restrict__ perm,
                                                        const int *__restrict__ off, int nbins,
                                                        double *__restrict__ sum,
                                                        double *__restrict__ cnt) {
    const int b = blockIdx.x * kBlock + threadIdx.x;
    if (b >= nbins) return;
    double s = 0.0;
    for (int p = off[b]; p < off[b + 1]; ++p) s += val[perm[p]];
    sum[b] += s;
    cnt[b] += (double)(off[b + 1] - off[b]);
}

}  // namespace spme
