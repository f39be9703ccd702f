// Device engine for one SPME system (one box, one grid, one spline order).
// Host-only interface; the kernels live in engine.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "common.hpp"

namespace spme {

enum EngineFlags {
    kFlagFP32 = 1,        // spreading, FFTs and gather in FP32
    kFlagFixedPoint = 2,  // deterministic int64 fixed-point spreading (always on for FP32)
};

struct FftPlan {
    int N = 0;
    std::vector<int> radix;
    double2 *tw_d = nullptr;  // exp(-2 pi i t/N), t < N
    float2 *tw_f = nullptr;
};

class Engine {
public:
    Engine(const double L6[6], const int K[3], int order, int flags);
    ~Engine();
    Engine(const Engine &) = delete;
    Engine &operator=(const Engine &) = delete;

    // geometry / tables
    void set_cell(const double L6[6]);
    void upload_table(const double *A, const double *dA);  // H and 6H doubles
    void build_ewald_table(double eta, double max_m);      // closed-form table
    bool has_table() const { return has_table_; }

    // device-pointer operations (asynchronous on stream())
    void spread_forward(int n, const double *w_d, const double *x_d);  // -> state 0
    void energy(double *out_d);                                        // out[0]=E
    void convolve(double *out_d);  // out[0]=E, out[1..6]=virial; -> state 1
    void gather(int n, const double *w_d, const double *x_d, double *dx_d, double *phi_d);

    // host-pointer helpers (stage through pinned memory, synchronous results)
    void sfac_host(int n, const double *w, const double *x);
    double en_host();
    double de1_host(double vir[6]);
    void gather_host(int n, const double *w, const double *x, double *dx, double *phi);
    double *get_S_host();  // host mirror of the grid, K0 x K1 x ldim x 2 doubles

    void sync();
    hipStream_t stream() const { return stream_; }
    double *result_dev() const { return result_; }

    // Per-kernel timing with HIP events recorded on stream() around every
    // launch while enabled.  profile_read synchronises and returns, per
    // kernel name, the summed milliseconds and the launch count.
    void profile_enable(bool on);
    void profile_read(std::vector<std::string> &names, std::vector<double> &ms,
                      std::vector<int> &count);
    struct ProfScope {
        Engine *e;
        ProfScope(Engine *e_, const char *name) : e(e_) { e->prof_begin_(name); }
        ~ProfScope() { e->prof_end_(); }
    };

    int state = -1;  // -1 nothing, 0 S-hat, 1 real-space convolution
    int K[3];
    int ldim;
    int order;
    bool fp32;
    bool fixed;
    double L6[6];
    CellArgs cell;
    std::vector<double> M;       // order x order spline coefficients
    std::vector<double> iB[3];   // 1/b_d moduli
    double mmax3;                // (max M_n)^3, bounds the grid for fixed point

    size_t grid_reals() const { return (size_t)K[0] * K[1] * 2 * ldim; }
    size_t H() const { return (size_t)K[0] * K[1] * ldim; }
    size_t G() const { return (size_t)K[0] * K[1] * K[2]; }
    size_t device_bytes() const { return dev_bytes_; }

private:
    void alloc_(void **p, size_t bytes);
    void ensure_atoms_(int n);
    void fft_forward_();
    void fft_inverse_();
    template <typename T> void fft_forward_t_();
    template <typename T> void fft_inverse_t_();
    template <typename T> void spread_t_(int n, const double *w_d, const double *x_d);
    template <typename T> void gather_t_(int n, const double *w_d, const double *x_d,
                                         double *dx_d, double *phi_d);
    template <typename T> void influence_t_(bool apply, double *out_d);

    hipStream_t stream_ = nullptr;
    void *grid_ = nullptr;               // T[K0*K1*2*ldim]
    unsigned long long *acc_ = nullptr;  // int64[K0*K1*K2] fixed-point accumulator
    double *table_ = nullptr;            // A[H] then dA[6H]
    bool has_table_ = false;
    double *iB_d_[3] = {nullptr, nullptr, nullptr};
    double *partials_ = nullptr;
    double *result_ = nullptr;  // 8 doubles
    unsigned long long *maxbits_ = nullptr;
    double *x_d_ = nullptr, *w_d_ = nullptr, *dx_d_ = nullptr, *phi_d_ = nullptr;
    size_t atom_cap_ = 0;
    double *pin_ = nullptr;  // pinned staging
    size_t pin_cap_ = 0;
    std::vector<double> host_grid_;
    FftPlan plan_[3];
    size_t dev_bytes_ = 0;
    std::vector<void *> allocs_;

    struct ProfRec {
        const char *name;
        hipEvent_t a, b;
    };
    void prof_begin_(const char *name);
    void prof_end_();
    bool prof_ = false;
    std::vector<ProfRec> prof_recs_;
};

// thrown for API misuse; carries the reference's message
struct UsageError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void hip_check(hipError_t e, const char *what);

}  // namespace spme
