// sfac.hpp -- C++ and C interface of the MI355X SPME reciprocal-space engine.
//
// Drop-in for the reference header (frobnitzem/EwaldCorrel sfac.hpp:1-211):
// the same macros, Vec4, Cell, BSpline and SFac names and the same ten
// extern "C" entry points of libsfac.so, with FFTW replaced by the engine's
// own gfx950 FFT (fft_plan_t is an opaque handle here; no <fftw3.h>).
// Extensions (device pointers, precision, batching, serving) are declared in
// sfac_ext.h.
#ifndef EWALDCORREL_AMD_SFAC_HPP
#define EWALDCORREL_AMD_SFAC_HPP

#include <math.h>
#include <stdint.h>

#include <map>
#include <vector>

#define MAX_SPL_ORDER (8)

#define SQR(x) ((x) * (x))
#define MOD(a, b) (((a) % (b) + (b)) % (b))
#define WRAP(a, b) (a > b / 2 ? a - b : a)

// The reference's FFTW plan type (sfac.hpp:13) becomes an opaque handle of
// the engine's own FFT plans.
typedef struct sfac_fft_plan *fft_plan_t;
typedef double mm_complex[2];

// 3-vector plus weight (sfac.hpp:16-31).
struct Vec4 {
    union {
        double r[4];
        struct {
            double x, y, z, w;
        };
    };
    Vec4(const double *p, double w_) : x(p[0]), y(p[1]), z(p[2]), w(w_) {}
    Vec4(double x_, double y_, double z_, double w_) : x(x_), y(y_), z(z_), w(w_) {}
    double operator[](int i) const { return r[i]; }
    double &operator[](int i) { return r[i]; }
};

// Box from L6 = [a_x, b_y, c_z, b_x, c_x, c_y]; rows of L are the box
// vectors, iL its inverse, iV = 1/(a_x b_y c_z) (sfac.hpp:33-63).
struct Cell {
    double L[3][3];
    double iL[3][3];
    double iV;
    explicit Cell(const double L6[6]) {
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) L[a][b] = iL[a][b] = 0.0;
        L[0][0] = L6[0];
        L[1][1] = L6[1];
        L[2][2] = L6[2];
        L[1][0] = L6[3];
        L[2][0] = L6[4];
        L[2][1] = L6[5];
        iV = 1.0 / (L6[0] * L6[1] * L6[2]);
        // inverse of a lower-triangular matrix, written out
        iL[0][0] = 1.0 / L6[0];
        iL[1][1] = 1.0 / L6[1];
        iL[2][2] = 1.0 / L6[2];
        iL[1][0] = -L6[3] * L6[2] * iV;
        iL[2][0] = (L6[3] * L6[5] - L6[1] * L6[4]) * iV;
        iL[2][1] = -L6[0] * L6[5] * iV;
    }
    static double wrap(double s) { return s - floor(s); }
    // fractional coordinates s = wrap(iL^T x), in [0,1]
    Vec4 scale(const double *x) const {
        return Vec4(wrap(x[0] * iL[0][0] + x[1] * iL[1][0] + x[2] * iL[2][0]),
                    wrap(x[1] * iL[1][1] + x[2] * iL[2][1]), wrap(x[2] * iL[2][2]), 0.0);
    }
};

// Cardinal B-spline of order n (sfac.hpp:65-158).  Row j of M holds the
// polynomial coefficients of M_n(x + j) for x in [0,1), highest power
// first; the last column is M_n(j).
struct BSpline {
    const int n;
    std::vector<double> M;

    explicit BSpline(int n_) : n(n_), M((size_t)n_ * n_) { mk_bspl_coeffs(); }

    // M_n(y) = 1/(n-1)! sum_k (-1)^k C(n,k) (y-k)_+^(n-1); for y = x + j only
    // k <= j contributes.  Integer numerators, one division per entry.
    void mk_bspl_coeffs() {
        auto binom = [](int a, int b) {
            long double r = 1;
            for (int t = 1; t <= b; ++t) r = r * (a - b + t) / t;
            return r;
        };
        long double fact = 1;
        for (int t = 2; t < n; ++t) fact *= t;
        for (int j = 0; j < n; ++j) {
            for (int p = 0; p < n; ++p) {  // coefficient of x^p
                long double num = 0;
                for (int k = 0; k <= j; ++k) {
                    long double a = 1;
                    for (int t = 0; t < n - 1 - p; ++t) a *= (j - k);
                    num += ((k & 1) ? -1.0L : 1.0L) * binom(n, k) * binom(n - 1, p) * a;
                }
                M[(size_t)j * n + (n - 1 - p)] = (double)(num / fact);
            }
        }
    }
    // M_n(x + j), x in [0,1): Horner
    double bspl_coef(const double x, const int j) const {
        double b = M[(size_t)n * j];
        for (int i = 1; i < n; ++i) b = b * x + M[(size_t)n * j + i];
        return b;
    }
    // value and derivative of M_n(x + j)
    void dbspl_coef(double *bp, double *dbp, const double x, const int j) const {
        double db = M[(size_t)j * n];
        double b = M[(size_t)j * n] * x + M[(size_t)j * n + 1];
        for (int i = 2; i < n; ++i) {
            db = db * x + b;
            b = b * x + M[(size_t)j * n + i];
        }
        *bp = b;
        *dbp = db;
    }
};

namespace spme {
class Engine;
}

// Structure-factor engine (sfac.hpp:160-192).  The grid, tables and FFT
// plans live in HBM inside `engine`; Q/FQ is a host mirror of the grid in
// the reference's padded layout (K0 x K1 x 2*ldim reals = K0 x K1 x ldim
// complex), refreshed by get_S().
struct SFac {
    int order;
    Cell cell;
    int state;  // -1 nothing computed, 0 S-hat, 1 real-space convolution
    int K[3];
    int ldim;
    double iK;
    BSpline bspc;
    std::vector<std::vector<double> > iB;
    fft_plan_t fft_neg, fft_pos;
    double *A;   // host copy of the influence table (H), or NULL
    double *dA;  // host copy of its virial derivatives (6H), or NULL
    union {
        double *Q;
        mm_complex *FQ;
    };
    spme::Engine *engine;

    // the reference's constructor (sfac.hpp:185, same mangled symbol), and
    // one with the precision flags of sfac_ctor_ex
    SFac(double L[6], int32_t K[3], int order);
    SFac(double L[6], int32_t K[3], int order, int flags);
    ~SFac();
    // Cell sort of the reference (sfac.cpp:92-111): cell id -> (frac, w).
    std::multimap<int, Vec4> sort(int n, const double *w, const double *x) const;
    // spread -> r2c -> /B on the device (sfac.cpp:113-139)
    void operator()(int n, const double *w, const double *x);
    // The reference's gather-form spreading of one grid point (sfac.cpp:141-168):
    // adds sum over the sorted atoms of w M(f0+i) M(f1+j) M(f2+k) from cells
    // (i1 - n/2 + i, ...) mod K to Q at (i1, j1, k1).  It acts on the host
    // mirror Q (the grid as last copied by get_S, taken now if there is none),
    // not on the device grid that operator() spreads into.
    void accum_spline(int i1, int j1, int k1, const std::multimap<int, Vec4> &srt);
};

extern "C" {
void *sfac_ctor(double L[6], int K[3], int order);
void sfac_dtor(void *sfac);
void sfac(void *sfac, int n, const double *w, const double *x);
double *get_S(void *sfac);

// rad_fn must return (A, dA) as a function of r2 = R*R; 'info' is passed
// through (sfac.hpp:201-203).
typedef double (*rad_fn)(double r2, double *dA, void *info);

void set_L(void *sfac, double L[6]);
int set_A(void *sfac, rad_fn f, double max_m, void *info);
double en(void *sfac);
double de1(void *sfac, double vir[6]);
void de2(void *sfac, int n, const double *w, const double *x, double *dx);
void potl(void *sfac, int n, const double *x, double *phi);
}

#include "sfac_ext.h"

#endif
