// Shared device/host definitions for the SPME reciprocal-space engine.
//
// Grid layout in HBM (identical to the reference's in-place FFTW buffer,
// sfac.cpp:70-74, so get_S is a plain copy): a K0 x K1 x (2*ldim) array of
// reals with K2 of each row used; the same memory viewed as K0 x K1 x ldim
// interleaved complex numbers holds the half-complex spectrum.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spme {

constexpr int kMaxOrder = 12;      // largest even B-spline order compiled in
constexpr int kMaxRadixStages = 16;

template <typename T> struct cplx_t;
template <> struct cplx_t<float> { using type = float2; };
template <> struct cplx_t<double> { using type = double2; };
template <typename T> using cplx = typename cplx_t<T>::type;

// Box: lower-triangular inverse iL (rows = reciprocal basis) and 1/V.
// L6 = [a_x, b_y, c_z, b_x, c_x, c_y] (sfac.hpp:37-52).
struct CellArgs {
    double iL00, iL10, iL11, iL20, iL21, iL22;
    double iV;
};

struct GridArgs {
    int K0, K1, K2;
    int ldim;      // K2/2 + 1 complex numbers per row
    int order;
};

// Coefficients of M_n(x+j), x in [0,1), row j, highest power first.
template <typename T> struct SplineArgs {
    T M[kMaxOrder * kMaxOrder];
};

// Fractional coordinates wrapped to [0,1): s = wrap(iL^T x) (sfac.hpp:56-62).
__host__ __device__ inline void scale_wrap(const CellArgs &c, double x0, double x1,
                                           double x2, double &s0, double &s1,
                                           double &s2) {
    s0 = x0 * c.iL00 + x1 * c.iL10 + x2 * c.iL20;
    s1 = x1 * c.iL11 + x2 * c.iL21;
    s2 = x2 * c.iL22;
    s0 -= floor(s0);
    s1 -= floor(s1);
    s2 -= floor(s2);
}

__host__ __device__ inline int imod(int a, int b) {
    int r = a % b;
    return r < 0 ? r + b : r;
}

}  // namespace spme
