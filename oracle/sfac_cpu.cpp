// CPU restatement of the reference's SPME reciprocal-space path, in C++ with
// OpenMP as the reference has it -- TEST INFRASTRUCTURE AND CPU BASELINE ONLY.
// It is never linked into libsfac.so; only tests/ and bench.py's cpu_baseline
// leg run it (as the checker, or as the timed CPU path beside the GPU).
//
// It follows frobnitzem/EwaldCorrel sfac.cpp step by step:
//   Cell / scale / BSpline Horner    sfac.hpp:33-63, 81-157
//   bspl_dft -> iB                   sfac.cpp:21-35, 57-68
//   sort (multimap cell -> (f, w))   sfac.cpp:92-111
//   spread in gather form, OpenMP    sfac.cpp:113-124 (parallel over grid points),
//     over grid points                sfac.cpp:141-168 (accum_spline)
//   forward r2c FFT, /B              sfac.cpp:125-138
//   set_A (ball, image fold)         sfac.cpp:213-292, with the Ewald kernel of
//                                    ewald.py:124-132 (eta given)
//   de1 (E, virial, A FQ, c2r)       sfac.cpp:339-396, OpenMP reductions
//   de2, potl (serial)               sfac.cpp:409-545
// The one substitute: the reference transforms with FFTW 3 (single-threaded,
// FFTW_PATIENT), which this host does not have; the transforms here are this
// file's own complex mixed-radix Cooley-Tukey FFT (recursive decimation in
// time, radix-2/4 butterflies and a direct DFT for any other factor), r2c by
// transforming two real rows as one complex row, single-threaded as FFTW is
// in the reference.  Unnormalised, sign -1 forward and +1 inverse (FFTW's
// convention).
//
// Usage: sfac_cpu IN OUT [reps]
//   IN  (binary, native endian): int32 K[3], int32 order, int32 n,
//       float64 L6[6], eta, max_m, w[n], x[3n]
//   OUT (binary, or "-" for none): float64 E, vir[6], dx[3n], phi[n] of the
//       last repetition
// Prints one JSON line: the results' E and virial, the thread count, and the
// median seconds of each step over `reps` evaluations (after one warm-up);
// set_A runs once, before the timed evaluations, as in the reference's use.
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <complex>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

typedef std::complex<double> cd;

#define SQR(x) ((x) * (x))
#define MOD(a, b) (((a) % (b) + (b)) % (b))

struct V4 {
    double r[4];
    double &operator[](int i) { return r[i]; }
    double operator[](int i) const { return r[i]; }
};

// box L6 = [a_x, b_y, c_z, b_x, c_x, c_y]; iL lower triangular (sfac.hpp:33-63)
struct Box {
    double iL[3][3];
    double iV;
    explicit Box(const double *L) {
        memset(iL, 0, sizeof(iL));
        iV = 1.0 / (L[0] * L[1] * L[2]);
        iL[0][0] = 1.0 / L[0];
        iL[1][1] = 1.0 / L[1];
        iL[2][2] = 1.0 / L[2];
        iL[1][0] = -L[3] * L[2] * iV;
        iL[2][0] = (L[3] * L[5] - L[1] * L[4]) * iV;
        iL[2][1] = -L[0] * L[5] * iV;
    }
    V4 scale(const double *x) const {
        V4 s;
        s[0] = x[0] * iL[0][0] + x[1] * iL[1][0] + x[2] * iL[2][0];
        s[1] = x[1] * iL[1][1] + x[2] * iL[2][1];
        s[2] = x[2] * iL[2][2];
        for (int d = 0; d < 3; ++d) s[d] -= floor(s[d]);
        s[3] = 0.0;
        return s;
    }
};

// cardinal B-spline coefficients, row j = M_n(x + j), highest power first,
// from M_n(y) = 1/(n-1)! sum_k (-1)^k C(n,k) (y-k)_+^(n-1)
struct Spline {
    int n;
    std::vector<double> M;
    explicit Spline(int n_) : n(n_), M((size_t)n_ * n_, 0.0) {
        auto binom = [](int a, int b) {
            double r = 1.0;
            for (int t = 1; t <= b; ++t) r = r * (a - b + t) / t;
            return r;
        };
        double fact = 1.0;
        for (int t = 2; t < n; ++t) fact *= t;
        for (int j = 0; j < n; ++j)
            for (int p = 0; p < n; ++p) {
                double num = 0.0;
                for (int k = 0; k <= j; ++k) {
                    double a = 1.0;
                    for (int t = 0; t < n - 1 - p; ++t) a *= (j - k);
                    num += ((k & 1) ? -1.0 : 1.0) * binom(n, k) * binom(n - 1, p) * a;
                }
                M[(size_t)j * n + (n - 1 - p)] = num / fact;
            }
    }
    double val(double x, int j) const {
        double b = M[(size_t)n * j];
        for (int i = 1; i < n; ++i) b = b * x + M[(size_t)n * j + i];
        return b;
    }
    void dval(double *b, double *db, double x, int j) const {
        double d = M[(size_t)j * n];
        double v = M[(size_t)j * n] * x + M[(size_t)j * n + 1];
        for (int i = 2; i < n; ++i) {
            d = d * x + v;
            v = v * x + M[(size_t)j * n + i];
        }
        *b = v;
        *db = d;
    }
};

// ---- complex FFT of one length (recursive mixed-radix DIT) ------------------------
struct Fft {
    int N = 1;
    std::vector<int> fac;  // radix, remaining length pairs
    std::vector<cd> tw;    // exp(-2 pi i t / N)
    explicit Fft(int n) : N(n), tw(n) {
        for (int t = 0; t < n; ++t) tw[t] = std::polar(1.0, -2.0 * M_PI * t / n);
        int m = n;
        while (m > 1) {
            int p = (m % 4 == 0) ? 4 : (m % 2 == 0) ? 2 : 0;
            if (!p)
                for (p = 3; m % p; p += 2) {
                }
            m /= p;
            fac.push_back(p);
            fac.push_back(m);
        }
    }
    cd w(long t, int sign) const {
        const cd v = tw[t % N];
        return sign < 0 ? v : std::conj(v);
    }
    void work(cd *out, const cd *in, long fstride, long istride, const int *f, int sign,
              cd *scr) const {
        const int p = f[0], m = f[1];
        cd *beg = out;
        if (m == 1) {
            for (int q = 0; q < p; ++q) out[q] = in[q * fstride * istride];
        } else {
            for (int q = 0; q < p; ++q)
                work(out + (long)q * m, in + q * fstride * istride, fstride * p, istride, f + 2,
                     sign, scr);
        }
        out = beg;
        if (p == 2) {
            for (int u = 0; u < m; ++u) {
                const cd t = out[u + m] * w((long)u * fstride, sign);
                out[u + m] = out[u] - t;
                out[u] += t;
            }
        } else if (p == 4) {
            const cd j = sign < 0 ? cd(0, -1) : cd(0, 1);
            for (int u = 0; u < m; ++u) {
                const cd a0 = out[u];
                const cd a1 = out[u + m] * w((long)u * fstride, sign);
                const cd a2 = out[u + 2 * m] * w(2L * u * fstride, sign);
                const cd a3 = out[u + 3 * m] * w(3L * u * fstride, sign);
                const cd s02 = a0 + a2, d02 = a0 - a2, s13 = a1 + a3, d13 = (a1 - a3) * j;
                out[u] = s02 + s13;
                out[u + m] = d02 + d13;
                out[u + 2 * m] = s02 - s13;
                out[u + 3 * m] = d02 - d13;
            }
        } else {  // direct DFT of radix p
            for (int u = 0; u < m; ++u) {
                for (int q = 0; q < p; ++q) scr[q] = out[u + (long)q * m];
                for (int q = 0; q < p; ++q) {
                    const long k = u + (long)q * m;
                    cd s = scr[0];
                    for (int r = 1; r < p; ++r) s += scr[r] * w((long)r * k * fstride, sign);
                    out[k] = s;
                }
            }
        }
    }
    // out[k] = sum_t in[t istride] e^(sign 2 pi i t k / N)
    void run(cd *out, const cd *in, long istride, int sign, cd *scr) const {
        if (N == 1) {
            out[0] = in[0];
            return;
        }
        work(out, in, 1, istride, fac.data(), sign, scr);
    }
};

// ---- the reference's SFac, restated ---------------------------------------------------
struct Sys {
    int K[3], order, ldim;
    Box cell;
    Spline bs;
    std::vector<double> iB[3];
    std::vector<double> Q;   // K0 x K1 x 2 ldim (in-place r2c layout)
    std::vector<double> A;   // H, then dA 6H
    Fft f0, f1, f2;

    Sys(const int *K_, int order_, const double *L)
        : order(order_), cell(L), bs(order_), f0(K_[0]), f1(K_[1]), f2(K_[2]) {
        for (int d = 0; d < 3; ++d) K[d] = K_[d];
        ldim = K[2] / 2 + 1;
        Q.assign((size_t)K[0] * K[1] * 2 * ldim, 0.0);
        // bspl_dft (sfac.cpp:21-35): iB[i] = 1 / sum_{j=1}^{n-1} cos(2 pi i (j - n/2)/K) M_n(j)
        for (int d = 0; d < 3; ++d) {
            iB[d].resize(K[d]);
            for (int i = 0; i < K[d]; ++i) {
                const double om = -(2.0 * M_PI * i) / K[d];
                double x = 0.0;
                for (int j = 1; j < order; ++j)
                    x += cos(om * (j - order / 2)) * bs.M[(size_t)order * j + order - 1];
                iB[d][i] = 1.0 / x;
            }
        }
    }
    size_t H() const { return (size_t)K[0] * K[1] * ldim; }

    // sort (sfac.cpp:92-111)
    std::multimap<int, V4> sort(int n, const double *w, const double *x) const {
        std::multimap<int, V4> srt;
        for (int a = 0; a < n; ++a) {
            V4 u = cell.scale(x + 3 * a);
            int c[3];
            for (int i = 0; i < 3; ++i) {
                u[i] *= K[i];
                c[i] = (int)u[i];
                u[i] -= c[i];
            }
            u[3] = w[a];
            srt.insert(std::make_pair((c[0] * K[1] + c[1]) * K[2] + c[2], u));
        }
        return srt;
    }
    // accum_spline (sfac.cpp:141-168)
    void accum(int i1, int j1, int k1, const std::multimap<int, V4> &srt) {
        const long NN = ((long)i1 * K[1] + j1) * ldim * 2 + k1;
        const int i0 = i1 + K[0] - order / 2, j0 = j1 + K[1] - order / 2,
                  k0 = k1 + K[2] - order / 2;
        double ans = 0.0;
        for (int i = 0; i < order; ++i) {
            const int xid = ((i0 + i) % K[0]) * K[1];
            for (int j = 0; j < order; ++j) {
                const int yid = (xid + (j0 + j) % K[1]) * K[2];
                for (int k = 0; k < order; ++k) {
                    const int cid = yid + (k0 + k) % K[2];
                    auto range = srt.equal_range(cid);
                    for (auto it = range.first; it != range.second; ++it) {
                        const V4 &u = it->second;
                        ans += bs.val(u[0], i) * bs.val(u[1], j) * bs.val(u[2], k) * u[3];
                    }
                }
            }
        }
        Q[NN] += ans;
    }

    // forward r2c (in place): z rows in pairs, then y, then x lines
    void fft_fwd() {
        const int K0 = K[0], K1 = K[1], K2 = K[2];
        std::vector<cd> a(std::max({K0, K1, K2})), b(a.size()), scr(a.size());
        const long rs = 2L * ldim;
        const int rows = K0 * K1;
        for (int r = 0; r < rows; r += 2) {
            const bool two = r + 1 < rows;
            double *p0 = Q.data() + (long)r * rs, *p1 = two ? p0 + rs : nullptr;
            for (int z = 0; z < K2; ++z) a[z] = cd(p0[z], two ? p1[z] : 0.0);
            f2.run(b.data(), a.data(), 1, -1, scr.data());
            for (int k = 0; k < ldim; ++k) {
                const cd zk = b[k], zc = std::conj(b[(K2 - k) % K2]);
                const cd A0 = 0.5 * (zk + zc), A1 = cd(0, -0.5) * (zk - zc);
                p0[2 * k] = A0.real();
                p0[2 * k + 1] = A0.imag();
                if (two) {
                    p1[2 * k] = A1.real();
                    p1[2 * k + 1] = A1.imag();
                }
            }
        }
        cd *F = reinterpret_cast<cd *>(Q.data());
        for (int x = 0; x < K0; ++x)
            for (int k = 0; k < ldim; ++k) {
                cd *base = F + (long)x * K1 * ldim + k;
                f1.run(b.data(), base, ldim, -1, scr.data());
                for (int y = 0; y < K1; ++y) base[(long)y * ldim] = b[y];
            }
        for (int y = 0; y < K1; ++y)
            for (int k = 0; k < ldim; ++k) {
                cd *base = F + (long)y * ldim + k;
                f0.run(b.data(), base, (long)K1 * ldim, -1, scr.data());
                for (int x = 0; x < K0; ++x) base[(long)x * K1 * ldim] = b[x];
            }
    }
    // inverse c2r (in place), the mirror of fft_fwd
    void fft_inv() {
        const int K0 = K[0], K1 = K[1], K2 = K[2];
        std::vector<cd> a(std::max({K0, K1, K2})), b(a.size()), scr(a.size());
        cd *F = reinterpret_cast<cd *>(Q.data());
        for (int y = 0; y < K1; ++y)
            for (int k = 0; k < ldim; ++k) {
                cd *base = F + (long)y * ldim + k;
                f0.run(b.data(), base, (long)K1 * ldim, +1, scr.data());
                for (int x = 0; x < K0; ++x) base[(long)x * K1 * ldim] = b[x];
            }
        for (int x = 0; x < K0; ++x)
            for (int k = 0; k < ldim; ++k) {
                cd *base = F + (long)x * K1 * ldim + k;
                f1.run(b.data(), base, ldim, +1, scr.data());
                for (int y = 0; y < K1; ++y) base[(long)y * ldim] = b[y];
            }
        const long rs = 2L * ldim;
        const int rows = K0 * K1;
        for (int r = 0; r < rows; r += 2) {
            const bool two = r + 1 < rows;
            double *p0 = Q.data() + (long)r * rs, *p1 = two ? p0 + rs : nullptr;
            // full Hermitian spectra X0, X1 of the two rows; transform X0 + i X1
            for (int k = 0; k < K2; ++k) {
                const int kk = k < ldim ? k : K2 - k;
                cd x0(p0[2 * kk], p0[2 * kk + 1]);
                cd x1 = two ? cd(p1[2 * kk], p1[2 * kk + 1]) : cd(0, 0);
                if (k >= ldim) {
                    x0 = std::conj(x0);
                    x1 = std::conj(x1);
                } else if (k == 0 || 2 * k == K2) {  // self-conjugate: real, as FFTW takes it
                    x0 = cd(x0.real(), 0.0);
                    x1 = cd(x1.real(), 0.0);
                }
                a[k] = x0 + cd(0, 1) * x1;
            }
            f2.run(b.data(), a.data(), 1, +1, scr.data());
            for (int z = 0; z < K2; ++z) {
                p0[z] = b[z].real();
                if (two) p1[z] = b[z].imag();
            }
        }
    }

    // operator() (sfac.cpp:113-139)
    void sfac(int n, const double *w, const double *x, double *t_sort, double *t_spread,
              double *t_fft, double *t_divb) {
        auto now = [] { return std::chrono::steady_clock::now(); };
        auto sec = [](auto a, auto b) { return std::chrono::duration<double>(b - a).count(); };
        auto t0 = now();
        std::fill(Q.begin(), Q.end(), 0.0);
        std::multimap<int, V4> srt = sort(n, w, x);
        auto t1 = now();
#pragma omp parallel for collapse(3)
        for (int i = 0; i < K[0]; ++i)
            for (int j = 0; j < K[1]; ++j)
                for (int k = 0; k < K[2]; ++k) accum(i, j, k, srt);
        auto t2 = now();
        fft_fwd();
        auto t3 = now();
        const long h = (long)H();
#pragma omp parallel for
        for (long cid = 0; cid < h; ++cid) {
            double *Qi = Q.data() + 2 * cid;
            const double t = iB[0][cid / (ldim * K[1])] * iB[1][(cid / ldim) % K[1]] *
                             iB[2][cid % ldim];
            Qi[0] *= t;
            Qi[1] *= t;
        }
        auto t4 = now();
        *t_sort = sec(t0, t1);
        *t_spread = sec(t1, t2);
        *t_fft = sec(t2, t3);
        *t_divb = sec(t3, t4);
    }

    // set_A (sfac.cpp:213-292) with f(r2) = exp(-pi^2 r2/eta^2)/(pi r2),
    // f' = (-pi^2/eta^2 - 1/r2) f, zero below r2 = 1e-10 (ewald.py:124-132)
    void set_A(double eta, double max_m) {
        const long h = (long)H();
        A.assign(7 * h, 0.0);
        double *dA0 = A.data() + h;
        const double R2 = SQR(max_m);
        const double (*iL)[3] = cell.iL;
        const int i0 = (int)(max_m / iL[0][0]);
        const double fac = -SQR(M_PI / eta);
        auto f = [&](double r2, double *d) {
            if (r2 < 1e-10) {
                *d = 0.0;
                return 0.0;
            }
            const double v = exp(fac * r2) / (M_PI * r2);
            *d = (fac - 1.0 / r2) * v;
            return v;
        };
        int nn[3];
        for (int i = -i0; i <= i0; ++i) {
            const double mx = i * iL[0][0];
            const double m20 = SQR(mx);
            const double dy = sqrt(R2 - m20);
            const double my0 = i * iL[1][0];
            const int j0 = (int)ceil((-dy - my0) / iL[1][1]);
            const int j1 = (int)floor((dy - my0) / iL[1][1]);
            nn[0] = MOD(i, K[0]) * K[1];
            for (int j = j0; j <= j1; ++j) {
                const double my = my0 + j * iL[1][1];
                const double m21 = m20 + SQR(my);
                const double dz = sqrt(R2 - m21);
                const double mz0 = i * iL[2][0] + j * iL[2][1];
                const int k0 = (int)ceil((-dz - mz0) / iL[2][2]);
                const int k1 = (int)floor((dz - mz0) / iL[2][2]);
                nn[1] = (MOD(j, K[1]) + nn[0]) * ldim;
                for (int k = k0; k <= k1; ++k) {
                    const double mz = mz0 + k * iL[2][2];
                    const double m2 = m21 + SQR(mz);
                    double s;
                    nn[2] = MOD(k, K[2]);
                    if (nn[2] >= ldim) {  // the conjugate image (sfac.cpp:267-276)
                        nn[2] = K[2] - nn[2];
                        nn[2] += nn[0] == 0 ? 0 : (K[0] * K[1] - nn[0]) * ldim;
                        nn[2] += j % K[1] == 0 ? 0 : (K[1] + nn[0]) * ldim - nn[1];
                    } else {
                        nn[2] += nn[1];
                    }
                    A[nn[2]] += 0.5 * f(m2, &s);
                    double *dA = dA0 + 6L * nn[2];
                    dA[0] += s * mx * mx;
                    dA[1] += s * my * my;
                    dA[2] += s * mz * mz;
                    dA[3] += s * my * mx;
                    dA[4] += s * mz * mx;
                    dA[5] += s * mz * my;
                }
            }
        }
    }

    // de1 (sfac.cpp:339-396)
    double de1(double vir[6]) {
        const long n = (long)H();
        const int even = (K[2] + 1) % 2;
        double *FQ = Q.data();
        const double *dA0 = A.data() + n;
        double en = 0.0, v00 = 0, v11 = 0, v22 = 0, v10 = 0, v20 = 0, v21 = 0;
#pragma omp parallel for reduction(+ : en, v00, v11, v22, v10, v20, v21)
        for (long i = 0; i < n; ++i) {
            const double s = SQR(FQ[2 * i]) + SQR(FQ[2 * i + 1]);
            const double *dA = dA0 + 6 * i;
            double t = A[i] * iB[0][i / (ldim * K[1])] * iB[1][(i / ldim) % K[1]] *
                       iB[2][i % ldim];
            if (i % ldim == 0 || (even && (i + 1) % ldim == 0)) t *= 2.0;
            FQ[2 * i] *= t;
            FQ[2 * i + 1] *= t;
            en += s * A[i];
            v00 += s * dA[0];
            v11 += s * dA[1];
            v22 += s * dA[2];
            v10 += s * dA[3];
            v20 += s * dA[4];
            v21 += s * dA[5];
        }
        const double fac = SQR(cell.iV);
        vir[0] = (v00 + en) * fac;
        vir[1] = (v11 + en) * fac;
        vir[2] = (v22 + en) * fac;
        vir[3] = v10 * fac;
        vir[4] = v20 * fac;
        vir[5] = v21 * fac;
        fft_inv();
        return en * cell.iV;
    }

    // de2 (sfac.cpp:409-483)
    void de2(int n, const double *w, const double *x, double *dx0) const {
        for (int a = 0; a < n; ++a) {
            int k0[3], k1[3];
            double dx[3] = {0, 0, 0};
            V4 u = cell.scale(x + 3 * a);
            for (int d = 0; d < 3; ++d) {
                u[d] *= K[d];
                k1[d] = (int)ceil(u[d]);
                k0[d] = MOD(k1[d] - order / 2, K[d]);
            }
            for (int i = 0; i < order; ++i) {
                const int ni = ((k0[0] + i) % K[0]) * K[1];
                double mpx, dmpx;
                bs.dval(&mpx, &dmpx, k1[0] - u[0], i);
                for (int j = 0; j < order; ++j) {
                    double mpy, dmpy;
                    bs.dval(&mpy, &dmpy, k1[1] - u[1], j);
                    const double fj0 = dmpx * mpy, fj1 = mpx * dmpy, fj2 = mpx * mpy;
                    const long nj = (long)(ni + (k0[1] + j) % K[1]) * ldim * 2;
                    for (int k = 0; k < order; ++k) {
                        double mpz, dmpz;
                        bs.dval(&mpz, &dmpz, k1[2] - u[2], k);
                        const long nk = nj + (k0[2] + k) % K[2];
                        const double fk = mpz * Q[nk];
                        dx[0] += fj0 * fk;
                        dx[1] += fj1 * fk;
                        dx[2] += fj2 * dmpz * Q[nk];
                    }
                }
            }
            for (int d = 0; d < 3; ++d) dx[d] *= (-w[a] * cell.iV) * K[d];
            dx0[3 * a + 0] = cell.iL[0][0] * dx[0];
            dx0[3 * a + 1] = cell.iL[1][0] * dx[0] + cell.iL[1][1] * dx[1];
            dx0[3 * a + 2] = cell.iL[2][0] * dx[0] + cell.iL[2][1] * dx[1] + cell.iL[2][2] * dx[2];
        }
    }

    // potl (sfac.cpp:497-545)
    void potl(int n, const double *x, double *phi) const {
        for (int a = 0; a < n; ++a) {
            int k0[3], k1[3];
            double pot = 0.0;
            V4 u = cell.scale(x + 3 * a);
            for (int d = 0; d < 3; ++d) {
                u[d] *= K[d];
                k1[d] = (int)ceil(u[d]);
                k0[d] = MOD(k1[d] - order / 2, K[d]);
            }
            for (int i = 0; i < order; ++i) {
                const int ni = ((k0[0] + i) % K[0]) * K[1];
                double s1 = 0.0;
                for (int j = 0; j < order; ++j) {
                    const double mpy = bs.val(k1[1] - u[1], j);
                    const long nj = (long)(ni + (k0[1] + j) % K[1]) * ldim * 2;
                    for (int k = 0; k < order; ++k) {
                        const double mpz = bs.val(k1[2] - u[2], k);
                        s1 += mpy * mpz * Q[nj + (k0[2] + k) % K[2]];
                    }
                }
                pot += bs.val(k1[0] - u[0], i) * s1;
            }
            phi[a] = pot * cell.iV;
        }
    }
};

double median(std::vector<double> v) {
    std::sort(v.begin(), v.end());
    return v.empty() ? 0.0 : v[v.size() / 2];
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: %s IN OUT [reps]\n", argv[0]);
        return 2;
    }
    const int reps = argc > 3 ? std::max(1, atoi(argv[3])) : 3;
    FILE *fi = fopen(argv[1], "rb");
    if (!fi) {
        fprintf(stderr, "cannot open %s\n", argv[1]);
        return 2;
    }
    int32_t hdr[5];
    double par[8];
    if (fread(hdr, sizeof(int32_t), 5, fi) != 5 || fread(par, sizeof(double), 8, fi) != 8) {
        fprintf(stderr, "short header\n");
        return 2;
    }
    const int K[3] = {hdr[0], hdr[1], hdr[2]}, order = hdr[3], n = hdr[4];
    if (order < 2 || order % 2 || K[0] < 1 || K[1] < 1 || K[2] < 1 || n < 0) {
        fprintf(stderr, "bad header\n");
        return 2;
    }
    std::vector<double> w(n), x(3L * n), dx(3L * n), phi(n);
    if (fread(w.data(), sizeof(double), n, fi) != (size_t)n ||
        fread(x.data(), sizeof(double), 3L * n, fi) != 3 * (size_t)n) {
        fprintf(stderr, "short atom data\n");
        return 2;
    }
    fclose(fi);
    Sys s(K, order, par);
    auto now = [] { return std::chrono::steady_clock::now(); };
    auto sec = [](auto a, auto b) { return std::chrono::duration<double>(b - a).count(); };
    auto ta = now();
    s.set_A(par[6], par[7]);
    const double t_setA = sec(ta, now());
    std::vector<double> T[8];
    double E = 0.0, vir[6];
    for (int r = 0; r <= reps; ++r) {  // repetition 0 is the warm-up
        double t[4];
        auto t0 = now();
        s.sfac(n, w.data(), x.data(), &t[0], &t[1], &t[2], &t[3]);
        auto t1 = now();
        E = s.de1(vir);
        auto t2 = now();
        s.de2(n, w.data(), x.data(), dx.data());
        auto t3 = now();
        s.potl(n, x.data(), phi.data());
        auto t4 = now();
        if (r == 0) continue;
        for (int k = 0; k < 4; ++k) T[k].push_back(t[k]);
        T[4].push_back(sec(t1, t2));
        T[5].push_back(sec(t2, t3));
        T[6].push_back(sec(t3, t4));
        T[7].push_back(sec(t0, t4));
    }
    if (strcmp(argv[2], "-") != 0) {
        FILE *fo = fopen(argv[2], "wb");
        if (!fo) {
            fprintf(stderr, "cannot write %s\n", argv[2]);
            return 2;
        }
        fwrite(&E, sizeof(double), 1, fo);
        fwrite(vir, sizeof(double), 6, fo);
        fwrite(dx.data(), sizeof(double), 3L * n, fo);
        fwrite(phi.data(), sizeof(double), n, fo);
        fclose(fo);
    }
    printf("{\"E\": %.17g, \"vir\": [%.17g, %.17g, %.17g, %.17g, %.17g, %.17g], "
           "\"threads\": %d, \"reps\": %d, \"set_A_s\": %.6g, \"eval_s\": %.6g, "
           "\"sort_s\": %.6g, \"spread_s\": %.6g, \"fft_fwd_s\": %.6g, \"divB_s\": %.6g, "
           "\"de1_s\": %.6g, \"de2_s\": %.6g, \"potl_s\": %.6g}\n",
           E, vir[0], vir[1], vir[2], vir[3], vir[4], vir[5], omp_get_max_threads(), reps,
           t_setA, median(T[7]), median(T[0]), median(T[1]), median(T[2]), median(T[3]),
           median(T[4]), median(T[5]), median(T[6]));
    return 0;
}
