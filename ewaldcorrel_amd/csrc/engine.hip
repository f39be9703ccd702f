// SPME reciprocal-space engine: buffers, FFT plans and kernel sequencing.
//
// One evaluation on the device (sfac + de1 + de2 + potl):
//   spread (FP64 atomics, or int64 fixed point + convert)
//   fft_z_r2c -> fft_strided(y) -> fft_strided(x, folds 1/(b0 b1 b2))  = S-hat
//   influence_kernel<APPLY> + reduce_finalize                          = E, virial
//   fft_strided(x, +1) -> fft_strided(y, +1) -> fft_z_c2r              = conv grid
//   gather_kernel                                                       = dx, phi
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cstdio>

#include "ball.hpp"
#include "engine.hpp"
#include "fft_kernels.hpp"
#include "sfac.hpp"
#include "spme_kernels.hpp"

namespace spme {

void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        std::string msg = std::string(what) + ": " + hipGetErrorString(e);
        throw std::runtime_error(msg);
    }
}
#define HIPC(x) hip_check((x), #x)
#define KCHECK(name) hip_check(hipGetLastError(), name)

namespace {

constexpr size_t kLdsBudget = 64 * 1024;

std::vector<int> factor_radices(int N) {
    std::vector<int> r;
    int n = N;
    while (n % 8 == 0) { r.push_back(8); n /= 8; }
    while (n % 4 == 0) { r.push_back(4); n /= 4; }
    while (n % 2 == 0) { r.push_back(2); n /= 2; }
    while (n % 3 == 0) { r.push_back(3); n /= 3; }
    while (n % 5 == 0) { r.push_back(5); n /= 5; }
    for (int p = 7; n > 1; p += 2)
        while (n % p == 0) { r.push_back(p); n /= p; }
    return r;
}

FftPlanArgs plan_args(const FftPlan &p) {
    FftPlanArgs a;
    memset(&a, 0, sizeof(a));
    a.N = p.N;
    a.nstages = (int)p.radix.size();
    for (int s = 0; s < a.nstages; ++s) a.radix[s] = p.radix[s];
    return a;
}

template <typename T> const cplx<T> *tw_of(const FftPlan &p);
template <> const float2 *tw_of<float>(const FftPlan &p) { return p.tw_f; }
template <> const double2 *tw_of<double>(const FftPlan &p) { return p.tw_d; }

// lines per workgroup so that two ping-pong line buffers fit the budget
template <typename T> int lines_per_block(int N, int want) {
    const size_t per = 2 * (size_t)(N + 1) * sizeof(cplx<T>);
    int c = (int)std::max<size_t>(1, kLdsBudget / per);
    return std::max(1, std::min(c, want));
}
template <typename T> size_t lds_bytes(int N, int C) {
    return 2 * (size_t)(N + 1) * C * sizeof(cplx<T>);
}

template <typename T> SplineArgs<T> spline_args(const std::vector<double> &M, int order) {
    SplineArgs<T> s;
    memset(&s, 0, sizeof(s));
    for (int i = 0; i < order * order; ++i) s.M[i] = T(M[i]);
    return s;
}

GridArgs grid_args(const Engine &e) {
    GridArgs g;
    g.K0 = e.K[0];
    g.K1 = e.K[1];
    g.K2 = e.K[2];
    g.ldim = e.ldim;
    g.order = e.order;
    return g;
}

__global__ void zero_kernel(unsigned long long *p, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x)
        p[i] = 0ull;
}
__global__ void zero_words(unsigned int *p, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x)
        p[i] = 0u;
}

int grid_blocks(long work, int block, int cap = 4096) {
    long b = (work + block - 1) / block;
    return (int)std::max<long>(1, std::min<long>(b, cap));
}

}  // namespace

// ---------------------------------------------------------------------------
Engine::Engine(const double L6_[6], const int K_[3], int order_, int flags) {
    if (order_ % 2 != 0 || order_ < 2 || order_ > kMaxOrder)
        throw std::invalid_argument("invalid order (even, 2.." + std::to_string(kMaxOrder) +
                                    "): " + std::to_string(order_));
    if (K_[0] < 1 || K_[1] < 1 || K_[2] < 1) throw std::invalid_argument("invalid K");
    for (int d = 0; d < 3; ++d) K[d] = K_[d];
    order = order_;
    ldim = K[2] / 2 + 1;
    fp32 = (flags & kFlagFP32) != 0;
    fixed = fp32 || (flags & kFlagFixedPoint) != 0;
    BSpline bs(order);
    M = bs.M;
    double mmax = 0.0;
    for (int j = 0; j < order; ++j) mmax = std::max(mmax, M[(size_t)j * order + order - 1]);
    mmax3 = mmax * mmax * mmax;
    // 1/b_d: iB[i] = 1 / sum_{j=1}^{n-1} cos(2 pi i (j - n/2)/K) M_n(j) (sfac.cpp:21-35)
    for (int d = 0; d < 3; ++d) {
        iB[d].resize(K[d]);
        for (int i = 0; i < K[d]; ++i) {
            const double om = -(2.0 * M_PI * i) / K[d];
            double s = 0.0;
            for (int j = 1; j < order; ++j)
                s += cos(om * (j - order / 2)) * M[(size_t)j * order + order - 1];
            iB[d][i] = 1.0 / s;
        }
    }
    set_cell(L6_);

    HIPC(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    const size_t tsz = fp32 ? sizeof(float) : sizeof(double);
    alloc_(&grid_, grid_reals() * tsz);
    if (fixed) alloc_((void **)&acc_, G() * sizeof(unsigned long long));
    for (int d = 0; d < 3; ++d) {
        alloc_((void **)&iB_d_[d], K[d] * sizeof(double));
        HIPC(hipMemcpy(iB_d_[d], iB[d].data(), K[d] * sizeof(double), hipMemcpyHostToDevice));
    }
    alloc_((void **)&partials_, kRedMaxBlocks * 8 * sizeof(double));
    alloc_((void **)&result_, 8 * sizeof(double));
    alloc_((void **)&maxbits_, 64);
    // FFT plans (one per axis; shared tables when lengths coincide are not
    // worth the bookkeeping at these sizes)
    for (int d = 0; d < 3; ++d) {
        FftPlan &p = plan_[d];
        p.N = K[d];
        p.radix = factor_radices(K[d]);
        if ((int)p.radix.size() > kMaxRadixStages) throw std::invalid_argument("K too composite");
        std::vector<double2> twd(K[d]);
        std::vector<float2> twf(K[d]);
        for (int t = 0; t < K[d]; ++t) {
            // exact reduction of the angle keeps FP32 twiddles accurate
            const double ang = -2.0 * M_PI * (double)t / (double)K[d];
            twd[t] = make_double2(cos(ang), sin(ang));
            twf[t] = make_float2((float)twd[t].x, (float)twd[t].y);
        }
        alloc_((void **)&p.tw_d, K[d] * sizeof(double2));
        alloc_((void **)&p.tw_f, K[d] * sizeof(float2));
        HIPC(hipMemcpy(p.tw_d, twd.data(), K[d] * sizeof(double2), hipMemcpyHostToDevice));
        HIPC(hipMemcpy(p.tw_f, twf.data(), K[d] * sizeof(float2), hipMemcpyHostToDevice));
    }
    HIPC(hipMemsetAsync(grid_, 0, grid_reals() * tsz, stream_));
    HIPC(hipStreamSynchronize(stream_));
}

Engine::~Engine() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (void *p : allocs_) (void)hipFree(p);
    if (pin_) (void)hipHostFree(pin_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void Engine::alloc_(void **p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw std::runtime_error("hipMalloc(" + std::to_string(bytes) +
                                 " B) failed: " + hipGetErrorString(e));
    }
    allocs_.push_back(*p);
    dev_bytes_ += bytes;
}

void Engine::set_cell(const double L6_[6]) {
    for (int i = 0; i < 6; ++i) L6[i] = L6_[i];
    Cell c(L6);
    cell.iL00 = c.iL[0][0];
    cell.iL10 = c.iL[1][0];
    cell.iL11 = c.iL[1][1];
    cell.iL20 = c.iL[2][0];
    cell.iL21 = c.iL[2][1];
    cell.iL22 = c.iL[2][2];
    cell.iV = c.iV;
}

void Engine::upload_table(const double *A, const double *dA) {
    const size_t h = H();
    if (!table_) alloc_((void **)&table_, 7 * h * sizeof(double));
    HIPC(hipMemcpyAsync(table_, A, h * sizeof(double), hipMemcpyHostToDevice, stream_));
    HIPC(hipMemcpyAsync(table_ + h, dA, 6 * h * sizeof(double), hipMemcpyHostToDevice, stream_));
    HIPC(hipStreamSynchronize(stream_));
    has_table_ = true;
}

void Engine::build_ewald_table(double eta, double max_m) {
    const size_t h = H();
    std::vector<double> tab(7 * h, 0.0);
    double *A = tab.data();
    double *dA = A + h;
    Cell c(L6);
    const double fac = -(M_PI / eta) * (M_PI / eta);
    enumerate_ball(c.iL, K, ldim, max_m,
                   [&](long idx, double m2, double mx, double my, double mz) {
                       double f = 0.0, fp = 0.0;
                       if (!(m2 < 1e-10)) {
                           f = exp(m2 * fac) / (m2 * M_PI);
                           fp = (fac - 1.0 / m2) * f;
                       }
                       A[idx] += 0.5 * f;
                       double *d = dA + 6 * idx;
                       d[0] += fp * mx * mx;
                       d[1] += fp * my * my;
                       d[2] += fp * mz * mz;
                       d[3] += fp * my * mx;
                       d[4] += fp * mz * mx;
                       d[5] += fp * mz * my;
                   });
    upload_table(A, dA);
}

void Engine::ensure_atoms_(int n) {
    if ((size_t)n <= atom_cap_) return;
    size_t cap = std::max<size_t>(1024, (size_t)n + n / 4);
    double *x, *w, *dx, *phi;
    alloc_((void **)&x, 3 * cap * sizeof(double));
    alloc_((void **)&w, cap * sizeof(double));
    alloc_((void **)&dx, 3 * cap * sizeof(double));
    alloc_((void **)&phi, cap * sizeof(double));
    x_d_ = x;
    w_d_ = w;
    dx_d_ = dx;
    phi_d_ = phi;
    atom_cap_ = cap;
    // old staging buffers stay allocated until the engine dies (rare growth)
    const size_t need = 4 * cap * sizeof(double);
    if (need > pin_cap_) {
        if (pin_) HIPC(hipHostFree(pin_));
        HIPC(hipHostMalloc((void **)&pin_, need, hipHostMallocDefault));
        pin_cap_ = need;
    }
}

// ---- spreading ---------------------------------------------------------------
template <typename T>
void Engine::spread_t_(int n, const double *w_d, const double *x_d) {
    GridArgs g = grid_args(*this);
    SplineArgs<T> sp = spline_args<T>(M, order);
    const int blk = 256;
    if (!fixed) {
        ProfScope ps(this, "zero_grid");
        zero_kernel<<<grid_blocks((long)grid_reals(), blk), blk, 0, stream_>>>(
            (unsigned long long *)grid_, (long)grid_reals());
        KCHECK("zero grid");
    } else {
        ProfScope ps(this, "spread_prep");
        zero_kernel<<<grid_blocks((long)G(), blk), blk, 0, stream_>>>(acc_, (long)G());
        zero_kernel<<<1, 64, 0, stream_>>>(maxbits_, 1);
        absmax_bits<<<grid_blocks(n, blk, 1024), blk, 0, stream_>>>(n, w_d, maxbits_);
        KCHECK("absmax");
    }
    if (n > 0) {
        ProfScope ps(this, "spread");
        const long work = (long)n * order;
        const int nb = (int)((work + blk - 1) / blk);
        double *gd = fixed ? nullptr : (double *)grid_;
#define SPREAD_CASE(O)                                                                     \
    case O:                                                                                \
        if (fixed)                                                                         \
            spread_kernel<T, O, 1><<<nb, blk, 0, stream_>>>(n, w_d, x_d, cell, g, sp, gd,  \
                                                            acc_, maxbits_, mmax3);        \
        else                                                                               \
            spread_kernel<T, O, 0><<<nb, blk, 0, stream_>>>(n, w_d, x_d, cell, g, sp, gd,  \
                                                            acc_, maxbits_, mmax3);        \
        break;
#define SPREAD_CASE_GENERIC                                                                \
    if (fixed)                                                                             \
        spread_kernel<T, 0, 1><<<nb, blk, 0, stream_>>>(n, w_d, x_d, cell, g, sp, gd, acc_, \
                                                        maxbits_, mmax3);                  \
    else                                                                                   \
        spread_kernel<T, 0, 0><<<nb, blk, 0, stream_>>>(n, w_d, x_d, cell, g, sp, gd, acc_, \
                                                        maxbits_, mmax3);                  \
    break;
        switch (order) {
            SPREAD_CASE(4)
            SPREAD_CASE(6)
            SPREAD_CASE(8)
            default: SPREAD_CASE_GENERIC
        }
#undef SPREAD_CASE
#undef SPREAD_CASE_GENERIC
        KCHECK("spread_kernel");
    }
    if (fixed) {
        ProfScope ps(this, "fixed_to_grid");
        fixed_to_grid<T><<<grid_blocks((long)G(), blk), blk, 0, stream_>>>(
            acc_, (T *)grid_, g, n, maxbits_, mmax3);
        KCHECK("fixed_to_grid");
    }
}

// ---- FFTs ------------------------------------------------------------------------
template <typename T>
void Engine::fft_forward_t_() {
    T *grid = (T *)grid_;
    cplx<T> *cg = (cplx<T> *)grid_;
    const int K0 = K[0], K1 = K[1], K2 = K[2];
    {   // z: real rows -> half-complex
        ProfScope ps(this, "fft_z_r2c");
        const FftPlan &p = plan_[2];
        const int C = lines_per_block<T>(K2, 16);
        const int nrows = K0 * K1;
        fft_z_r2c<T><<<(nrows + C - 1) / C, 256, lds_bytes<T>(K2, C), stream_>>>(
            grid, nrows, ldim, plan_args(p), tw_of<T>(p), C);
        KCHECK("fft_z_r2c");
    }
    {   // y: lines along y for each (x, kz)
        ProfScope ps(this, "fft_y_fwd");
        const FftPlan &p = plan_[1];
        const int C = lines_per_block<T>(K1, sizeof(T) == 4 ? 32 : 16);
        const int ntiles = (ldim + C - 1) / C;
        fft_strided<T><<<K0 * ntiles, 256, lds_bytes<T>(K1, C), stream_>>>(
            cg, ldim, ntiles, (long)K1 * ldim, (long)ldim, plan_args(p), tw_of<T>(p), C, -1,
            nullptr, nullptr, nullptr);
        KCHECK("fft_y_fwd");
    }
    {   // x: lines along x for each (ky, kz); folds 1/(b0 b1 b2) into the store
        ProfScope ps(this, "fft_x_fwd");
        const FftPlan &p = plan_[0];
        const int C = lines_per_block<T>(K0, sizeof(T) == 4 ? 32 : 16);
        const int ntiles = (ldim + C - 1) / C;
        fft_strided<T><<<K1 * ntiles, 256, lds_bytes<T>(K0, C), stream_>>>(
            cg, ldim, ntiles, (long)ldim, (long)K1 * ldim, plan_args(p), tw_of<T>(p), C, -1,
            iB_d_[0], iB_d_[1], iB_d_[2]);
        KCHECK("fft_x_fwd");
    }
}

template <typename T>
void Engine::fft_inverse_t_() {
    T *grid = (T *)grid_;
    cplx<T> *cg = (cplx<T> *)grid_;
    const int K0 = K[0], K1 = K[1], K2 = K[2];
    {
        ProfScope ps(this, "fft_x_inv");
        const FftPlan &p = plan_[0];
        const int C = lines_per_block<T>(K0, sizeof(T) == 4 ? 32 : 16);
        const int ntiles = (ldim + C - 1) / C;
        fft_strided<T><<<K1 * ntiles, 256, lds_bytes<T>(K0, C), stream_>>>(
            cg, ldim, ntiles, (long)ldim, (long)K1 * ldim, plan_args(p), tw_of<T>(p), C, +1,
            nullptr, nullptr, nullptr);
        KCHECK("fft_x_inv");
    }
    {
        ProfScope ps(this, "fft_y_inv");
        const FftPlan &p = plan_[1];
        const int C = lines_per_block<T>(K1, sizeof(T) == 4 ? 32 : 16);
        const int ntiles = (ldim + C - 1) / C;
        fft_strided<T><<<K0 * ntiles, 256, lds_bytes<T>(K1, C), stream_>>>(
            cg, ldim, ntiles, (long)K1 * ldim, (long)ldim, plan_args(p), tw_of<T>(p), C, +1,
            nullptr, nullptr, nullptr);
        KCHECK("fft_y_inv");
    }
    {
        ProfScope ps(this, "fft_z_c2r");
        const FftPlan &p = plan_[2];
        const int C = lines_per_block<T>(K2, 16);
        const int nrows = K0 * K1;
        fft_z_c2r<T><<<(nrows + C - 1) / C, 256, lds_bytes<T>(K2, C), stream_>>>(
            grid, nrows, ldim, plan_args(p), tw_of<T>(p), C);
        KCHECK("fft_z_c2r");
    }
}

void Engine::fft_forward_() {
    if (fp32) fft_forward_t_<float>();
    else fft_forward_t_<double>();
}
void Engine::fft_inverse_() {
    if (fp32) fft_inverse_t_<float>();
    else fft_inverse_t_<double>();
}

// ---- influence + reductions -----------------------------------------------------
template <typename T>
void Engine::influence_t_(bool apply, double *out_d) {
    GridArgs g = grid_args(*this);
    const int nb = grid_blocks((long)H(), kRedBlock, kRedMaxBlocks);
    cplx<T> *fq = (cplx<T> *)grid_;
    {
    ProfScope ps(this, apply ? "influence" : "energy");
    if (apply)
        influence_kernel<T, true><<<nb, kRedBlock, 0, stream_>>>(
            fq, table_, table_ + H(), g, iB_d_[0], iB_d_[1], iB_d_[2], partials_);
    else
        influence_kernel<T, false><<<nb, kRedBlock, 0, stream_>>>(
            fq, table_, table_ + H(), g, iB_d_[0], iB_d_[1], iB_d_[2], partials_);
    KCHECK("influence_kernel");
    }
    ProfScope ps(this, "reduce_finalize");
    reduce_finalize<<<1, kRedBlock, 0, stream_>>>(partials_, nb, cell.iV, out_d);
    KCHECK("reduce_finalize");
}

// ---- gather ---------------------------------------------------------------------
template <typename T>
void Engine::gather_t_(int n, const double *w_d, const double *x_d, double *dx_d,
                       double *phi_d) {
    if (n <= 0) return;
    GridArgs g = grid_args(*this);
    SplineArgs<T> sp = spline_args<T>(M, order);
    const int blk = 256;
    const int nb = (n + blk - 1) / blk;
    const T *grid = (const T *)grid_;
    ProfScope ps(this, "gather");
#define GATHER_CASE(O)                                                                    \
    case O:                                                                               \
        gather_kernel<T, O><<<nb, blk, 0, stream_>>>(n, w_d, x_d, cell, g, sp, grid, dx_d, \
                                                     phi_d);                              \
        break;
    switch (order) {
        GATHER_CASE(4)
        GATHER_CASE(6)
        GATHER_CASE(8)
        default:
            gather_kernel<T, 0><<<nb, blk, 0, stream_>>>(n, w_d, x_d, cell, g, sp, grid, dx_d,
                                                         phi_d);
    }
#undef GATHER_CASE
    KCHECK("gather_kernel");
}

// ---- public device-pointer operations ---------------------------------------------
void Engine::spread_forward(int n, const double *w_d, const double *x_d) {
    if (fp32) spread_t_<float>(n, w_d, x_d);
    else spread_t_<double>(n, w_d, x_d);
    fft_forward_();
    state = 0;
}

void Engine::energy(double *out_d) {
    if (state != 0) throw UsageError("You must call sfac before this routine.");
    if (!has_table_) throw UsageError("No convolution kernel has been set.");
    if (fp32) influence_t_<float>(false, out_d);
    else influence_t_<double>(false, out_d);
}

void Engine::convolve(double *out_d) {
    if (state != 0) throw UsageError("You must call sfac before this routine.");
    if (!has_table_) throw UsageError("No convolution kernel has been set.");
    if (fp32) influence_t_<float>(true, out_d);
    else influence_t_<double>(true, out_d);
    fft_inverse_();
    state = 1;
}

void Engine::gather(int n, const double *w_d, const double *x_d, double *dx_d, double *phi_d) {
    if (state != 1) throw UsageError("You must call de1 before this routine.");
    if (fp32) gather_t_<float>(n, w_d, x_d, dx_d, phi_d);
    else gather_t_<double>(n, w_d, x_d, dx_d, phi_d);
}

// ---- host-pointer helpers -----------------------------------------------------------
void Engine::sfac_host(int n, const double *w, const double *x) {
    if (n < 0) throw std::invalid_argument("n < 0");
    ensure_atoms_(n);
    // the caller's arrays are borrowed for the call only: copy them to
    // pinned staging, then upload asynchronously
    HIPC(hipStreamSynchronize(stream_));
    memcpy(pin_, x, 3 * (size_t)n * sizeof(double));
    memcpy(pin_ + 3 * (size_t)n, w, (size_t)n * sizeof(double));
    HIPC(hipMemcpyAsync(x_d_, pin_, 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice,
                        stream_));
    HIPC(hipMemcpyAsync(w_d_, pin_ + 3 * (size_t)n, (size_t)n * sizeof(double),
                        hipMemcpyHostToDevice, stream_));
    spread_forward(n, w_d_, x_d_);
}

double Engine::en_host() {
    energy(result_);
    double out[8];
    HIPC(hipMemcpyAsync(out, result_, 8 * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIPC(hipStreamSynchronize(stream_));
    return out[0];
}

double Engine::de1_host(double vir[6]) {
    convolve(result_);
    double out[8];
    HIPC(hipMemcpyAsync(out, result_, 8 * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIPC(hipStreamSynchronize(stream_));
    for (int c = 0; c < 6; ++c) vir[c] = out[1 + c];
    return out[0];
}

void Engine::gather_host(int n, const double *w, const double *x, double *dx, double *phi) {
    if (state != 1) throw UsageError("You must call de1 before this routine.");
    if (n <= 0) return;
    ensure_atoms_(n);
    HIPC(hipStreamSynchronize(stream_));
    memcpy(pin_, x, 3 * (size_t)n * sizeof(double));
    if (w) memcpy(pin_ + 3 * (size_t)n, w, (size_t)n * sizeof(double));
    HIPC(hipMemcpyAsync(x_d_, pin_, 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice,
                        stream_));
    if (w)
        HIPC(hipMemcpyAsync(w_d_, pin_ + 3 * (size_t)n, (size_t)n * sizeof(double),
                            hipMemcpyHostToDevice, stream_));
    gather(n, w ? w_d_ : nullptr, x_d_, dx ? dx_d_ : nullptr, phi ? phi_d_ : nullptr);
    if (dx)
        HIPC(hipMemcpyAsync(dx, dx_d_, 3 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost,
                            stream_));
    if (phi)
        HIPC(hipMemcpyAsync(phi, phi_d_, (size_t)n * sizeof(double), hipMemcpyDeviceToHost,
                            stream_));
    HIPC(hipStreamSynchronize(stream_));
}

double *Engine::get_S_host() {
    const size_t nr = grid_reals();
    host_grid_.resize(nr);
    if (fp32) {
        std::vector<float> tmp(nr);
        HIPC(hipMemcpyAsync(tmp.data(), grid_, nr * sizeof(float), hipMemcpyDeviceToHost,
                            stream_));
        HIPC(hipStreamSynchronize(stream_));
        for (size_t i = 0; i < nr; ++i) host_grid_[i] = tmp[i];
    } else {
        HIPC(hipMemcpyAsync(host_grid_.data(), grid_, nr * sizeof(double),
                            hipMemcpyDeviceToHost, stream_));
        HIPC(hipStreamSynchronize(stream_));
    }
    return host_grid_.data();
}

void Engine::sync() { HIPC(hipStreamSynchronize(stream_)); }

// ---- profiling ---------------------------------------------------------------------------
void Engine::prof_begin_(const char *name) {
    if (!prof_) return;
    ProfRec r;
    r.name = name;
    HIPC(hipEventCreate(&r.a));
    HIPC(hipEventCreate(&r.b));
    HIPC(hipEventRecord(r.a, stream_));
    prof_recs_.push_back(r);
}

void Engine::prof_end_() {
    if (!prof_ || prof_recs_.empty()) return;
    HIPC(hipEventRecord(prof_recs_.back().b, stream_));
}

void Engine::profile_enable(bool on) {
    sync();
    if (!on) {
        for (auto &r : prof_recs_) {
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
        prof_recs_.clear();
    }
    prof_ = on;
}

void Engine::profile_read(std::vector<std::string> &names, std::vector<double> &ms,
                          std::vector<int> &count) {
    sync();
    names.clear();
    ms.clear();
    count.clear();
    for (auto &r : prof_recs_) {
        float t = 0.f;
        HIPC(hipEventElapsedTime(&t, r.a, r.b));
        size_t k = 0;
        while (k < names.size() && names[k] != r.name) ++k;
        if (k == names.size()) {
            names.push_back(r.name);
            ms.push_back(0.0);
            count.push_back(0);
        }
        ms[k] += t;
        count[k] += 1;
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    prof_recs_.clear();
}

}  // namespace spme
