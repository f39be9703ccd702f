// SPME reciprocal-space engine: buffers, launch geometry and kernel sequencing.
//
// One evaluation (sfac + de1 + de2 + potl), ten launches, no memsets:
//   prep -> scan -> scatter -> spread          column sort + LDS-tiled spreading
//   fft_z_r2c -> fft_y_fwd                      forward transform, z and y
//   fft_x_recip                                 x forward, /B, influence, E and
//                                               virial, x inverse (one pass)
//   fft_y_inv -> fft_z_c2r                      inverse transform, y and z
//   gather_sorted                               forces and potentials
// The split entry points (sfac / en / de1 / de2 / potl) run the same kernels
// with the x pass, influence and inverse x pass as separate launches, so that
// S-hat is in HBM between sfac and de1 as the reference's state machine
// requires.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cstdio>
#include <functional>
#include <mutex>
#include <type_traits>

#include "engine.hpp"
#include "fft_kernels.hpp"
#include "gather_kernels.hpp"
#include "recip_kernels.hpp"
#include "sfac.hpp"
#include "slab_kernels.hpp"
#include "slab128_kernels.hpp"
#include "spread_kernels.hpp"
#include "spread_slab_kernels.hpp"
#include "table_kernels.hpp"
#include "correl_kernels.hpp"
#include "real_kernels.hpp"

namespace spme {

void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        std::string msg = std::string(what) + ": " + hipGetErrorString(e);
        throw std::runtime_error(msg);
    }
}
#define HIPC(x) hip_check((x), #x)

const char *stage_name(Stage s) {
    static const char *names[ST_COUNT] = {
        "bin",       "sort",      "spread",    "fft_z_r2c",
        "fft_y_fwd", "fft_x_fwd", "fft_x_recip", "influence", "energy",
        "fft_x_inv", "fft_y_inv", "fft_z_c2r", "gather",    "gather_sorted",
        "finish",    "fft_yz_fwd", "fft_yz_inv", "gather_tile", "fft_x_recip",
        "spread_fft_yz"};
    return names[s];
}

namespace {

constexpr size_t kLdsBudget = 64 * 1024;

// Default radix order of a length-N transform: radix 16 first (two
// register-staged stages for 128 instead of three), then 8, 4, 2, 3, 5 and
// any other prime; with16 = false leaves out 16 (kernels holding more than
// 16 elements per thread have no radix-16 stage).
std::vector<int> factor_radices(int N, bool with16 = true) {
    std::vector<int> r;
    int n = N;
    while (with16 && n % 16 == 0) { r.push_back(16); n /= 16; }
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
    int Ns = 1;
    for (int s = 0; s < a.nstages; ++s) {
        a.radix[s] = p.radix[s];
        const int R = p.radix[s];
        if (R != 2 && R != 3 && R != 4 && R != 5 && R != 8 && R != 16) a.generic = 1;
        a.dNR[s] = make_fastdiv((unsigned)(p.N / R));
        a.dNs[s] = make_fastdiv((unsigned)Ns);
        a.twstep[s] = p.N / (Ns * R);
        Ns *= R;
    }
    return a;
}

size_t lds_fft(bool fp32, int N, int C, int generic, bool strided) {
    if (strided)
        return fp32 ? ylines_lds_bytes<float>(N, C, generic)
                    : ylines_lds_bytes<double>(N, C, generic);
    return fp32 ? fft_lds_bytes<float>(N, C, generic) : fft_lds_bytes<double>(N, C, generic);
}

// lines per FFT workgroup: up to fft_elems<T>() elements, within the LDS budget
int fft_lines(bool fp32, int N, int avail, int generic, bool strided) {
    const int target = fp32 ? fft_elems<float>() : fft_elems<double>();
    int C = std::max(1, std::min(target / std::max(N, 1), std::max(avail, 1)));
    while (C > 1 && lds_fft(fp32, N, C, generic, strided) > kLdsBudget) --C;
    return C;
}

template <typename T> SplineArgs<T> spline_args(const std::vector<double> &M, int order) {
    SplineArgs<T> s;
    memset(&s, 0, sizeof(s));
    for (int i = 0; i < order * order; ++i) s.M[i] = T(M[i]);
    return s;
}

template <typename T, int MODE, int INF>
void launch_x_inf(const Launch &L, dim3 g, dim3 bl, size_t lds, hipStream_t st, int sign,
                  int finish) {
    if (L.xpath == kXBlock)
        fft_x_kernel<T, MODE, kXBlock, INF><<<g, bl, lds, st>>>(L.frames, L.offs, L.nf, sign,
                                                                finish);
    else
        fft_x_kernel<T, MODE, kXGeneric, INF><<<g, bl, lds, st>>>(L.frames, L.offs, L.nf, sign,
                                                                  finish);
}

template <typename T, int MODE>
void launch_x(const Launch &L, dim3 g, dim3 bl, size_t lds, hipStream_t st, int sign,
              int finish) {
    // only the fused pass (MODE 2) applies the influence, in the form the
    // launch's frames share (XInfluence), each its own instantiation
    if constexpr (MODE == 2) {
        if (std::is_same<T, float>::value && L.xinf == kXInfFp32) {
            launch_x_inf<T, MODE, kXInfFp32>(L, g, bl, lds, st, sign, finish);
            return;
        }
        if (L.xinf == kXInfSeparable) {
            launch_x_inf<T, MODE, kXInfSeparable>(L, g, bl, lds, st, sign, finish);
            return;
        }
    }
    launch_x_inf<T, MODE, kXInfAny>(L, g, bl, lds, st, sign, finish);
}

template <typename T, int ORDER, int MZ>
void gather_tile_launch(const Launch &L, dim3 g, size_t lds, hipStream_t st,
                        const SplineArgs<T> &sp) {
    constexpr int B = gather_block<T, ORDER>();
    if (L.gzchunk)
        gather_tile_kernel<T, ORDER, MZ, true><<<g, B, lds, st>>>(L.frames, L.offs, L.nf, sp);
    else
        gather_tile_kernel<T, ORDER, MZ, false><<<g, B, lds, st>>>(L.frames, L.offs, L.nf, sp);
}

template <typename T, int ORDER>
void launch_t(Stage s, const Launch &L, int blocks, size_t lds,
              hipStream_t st, const SplineArgs<T> &sp, bool fixed, int finish) {
    const dim3 g(blocks), bl(kBlock);
    switch (s) {
        case ST_BIN:
            bin_kernel<T><<<g, kBinThreads, bin_lds_bytes<T>(), st>>>(L.frames, L.offs, L.nf);
            break;
        case ST_SORT: sort_kernel<T><<<g, kSortThreads, lds, st>>>(L.frames, L.offs, L.nf); break;
        case ST_SPREAD: {
            const dim3 sb(kSpreadBlock);
            auto go = [&](auto fixed_c, auto zch_c) {
                spread_tile_kernel<T, ORDER, decltype(fixed_c)::value, decltype(zch_c)::value>
                    <<<g, sb, lds, st>>>(L.frames, L.offs, L.nf, sp);
            };
            using yes = std::true_type;
            using no = std::false_type;
            if (std::is_same<T, float>::value || fixed) {
                if (L.zchunk) go(yes{}, yes{});
                else go(yes{}, no{});
            } else {
                if (L.zchunk) go(no{}, yes{});
                else go(no{}, no{});
            }
            break;
        }
        case ST_SPREAD_YZ:
            if (std::is_same<T, float>::value && L.slab_ept == kSlabKernel128)
                spread_slab_kernel<T, ORDER, true><<<g, kSpreadSlabThreads, lds, st>>>(
                    L.frames, L.offs, L.nf, sp);
            else
                spread_slab_kernel<T, ORDER, false><<<g, kSpreadSlabThreads, lds, st>>>(
                    L.frames, L.offs, L.nf, sp);
            break;
        case ST_ZR2C: fft_z_r2c_kernel<T><<<g, bl, lds, st>>>(L.frames, L.offs, L.nf); break;
        case ST_YF: fft_y_kernel<T><<<g, bl, lds, st>>>(L.frames, L.offs, L.nf, -1, 0); break;
        case ST_XF_SCALE:
            launch_x<T, 1>(L, g, bl, lds, st, -1, 0);
            break;
        case ST_XFUSED:
            launch_x<T, 2>(L, g, bl, lds, st, -1, 0);
            break;
        case ST_XREG:
            if constexpr (std::is_same<T, float>::value)
                fft_x_reg_kernel<T><<<g, kXRegThreads, lds, st>>>(L.frames, L.offs, L.nf);
            break;
        case ST_YZ_FWD:
        case ST_YZ_INV: {
            if constexpr (std::is_same<T, float>::value) {
                if (L.slab_ept == kSlabKernel128) {
                    if (s == ST_YZ_FWD)
                        fft_yz_fwd128_kernel<<<g, kSlab128Threads, lds, st>>>(L.frames, L.offs,
                                                                              L.nf);
                    else
                        fft_yz_inv128_kernel<<<g, kSlab128Threads, lds, st>>>(L.frames, L.offs,
                                                                              L.nf, finish);
                    break;
                }
            }
            auto go = [&](auto ept_c) {
                constexpr int E = decltype(ept_c)::value;
                if (s == ST_YZ_FWD)
                    fft_yz_fwd_kernel<T, E><<<g, slab_block<T, E>(), lds, st>>>(L.frames, L.offs,
                                                                               L.nf);
                else
                    fft_yz_inv_kernel<T, E><<<g, slab_block<T, E>(), lds, st>>>(
                        L.frames, L.offs, L.nf, finish);
            };
            if (L.slab_ept == SlabCfg<T>::small)
                go(std::integral_constant<int, SlabCfg<T>::small>{});
            else if (L.slab_ept == SlabCfg<T>::medium)
                go(std::integral_constant<int, SlabCfg<T>::medium>{});
            else
                go(std::integral_constant<int, SlabCfg<T>::large>{});
            break;
        }
        case ST_INFL: influence_kernel<T, true><<<g, bl, 0, st>>>(L.frames, L.offs, L.nf); break;
        case ST_INFL_E: influence_kernel<T, false><<<g, bl, 0, st>>>(L.frames, L.offs, L.nf); break;
        case ST_XI:
            launch_x<T, 0>(L, g, bl, lds, st, +1, finish);
            break;
        case ST_YI: fft_y_kernel<T><<<g, bl, lds, st>>>(L.frames, L.offs, L.nf, +1, finish); break;
        case ST_FINISH: finish_kernel<<<g, bl, 0, st>>>(L.frames, L.offs, L.nf); break;
        case ST_ZC2R: fft_z_c2r_kernel<T><<<g, bl, lds, st>>>(L.frames, L.offs, L.nf); break;
        case ST_GATHER: gather_kernel<T, ORDER, false><<<g, bl, 0, st>>>(L.frames, L.offs, L.nf,sp); break;
        case ST_GATHER_SORTED:
            gather_kernel<T, ORDER, true><<<g, bl, 0, st>>>(L.frames, L.offs, L.nf, sp);
            break;
        case ST_GATHER_TILE:
            switch (L.gmz) {
                case 1: gather_tile_launch<T, ORDER, 1>(L, g, lds, st, sp); break;
                case 2: gather_tile_launch<T, ORDER, 2>(L, g, lds, st, sp); break;
                case 3: gather_tile_launch<T, ORDER, 3>(L, g, lds, st, sp); break;
                default: gather_tile_launch<T, ORDER, 4>(L, g, lds, st, sp); break;
            }
            break;
        default: break;
    }
    hip_check(hipGetLastError(), stage_name(s));
}

template <typename T>
void launch_order(int order, Stage s, const Launch &L, int blocks,
                  size_t lds, hipStream_t st, const SplineArgs<T> &sp, bool fixed, int finish) {
    switch (order) {
        case 4: launch_t<T, 4>(s, L, blocks, lds, st, sp, fixed, finish); break;
        case 6: launch_t<T, 6>(s, L, blocks, lds, st, sp, fixed, finish); break;
        case 8: launch_t<T, 8>(s, L, blocks, lds, st, sp, fixed, finish); break;
        default: launch_t<T, 0>(s, L, blocks, lds, st, sp, fixed, finish); break;
    }
}

// the reduction of a stage is finalized by the next launch of the sequence:
// 1 after the fused x pass, 2 after the standalone influence launch
int finishes(Stage prev, Stage s) {
    if ((prev == ST_XFUSED || prev == ST_XREG) && (s == ST_YI || s == ST_YZ_INV)) return 1;
    if (prev == ST_INFL && s == ST_XI) return 2;
    return 0;
}

const Stage kSfacLines[] = {ST_BIN, ST_SORT, ST_SPREAD, ST_ZR2C, ST_YF, ST_XF_SCALE};
const Stage kSfacSlab[] = {ST_BIN, ST_SORT, ST_SPREAD, ST_SPREAD_YZ, ST_YZ_FWD, ST_XF_SCALE};
const Stage kDe1Lines[] = {ST_INFL, ST_XI, ST_YI, ST_ZC2R};
const Stage kDe1Slab[] = {ST_INFL, ST_XI, ST_YZ_INV};
const Stage kEnStages[] = {ST_INFL_E, ST_FINISH};
const Stage kEvalLines[] = {ST_BIN, ST_SORT, ST_SPREAD, ST_ZR2C,
                            ST_YF,   ST_XFUSED, ST_XREG, ST_YI, ST_ZC2R, ST_GATHER_SORTED};
const Stage kEvalSlab[] = {ST_BIN,    ST_SORT,   ST_SPREAD, ST_SPREAD_YZ,    ST_YZ_FWD,
                           ST_XFUSED, ST_XREG,   ST_YZ_INV, ST_GATHER_SORTED};
const Stage kEvalLinesTile[] = {ST_BIN, ST_SORT, ST_SPREAD, ST_ZR2C,
                                ST_YF,   ST_XFUSED, ST_XREG, ST_YI, ST_ZC2R, ST_GATHER_TILE};
const Stage kEvalSlabTile[] = {ST_BIN,    ST_SORT, ST_SPREAD, ST_SPREAD_YZ,  ST_YZ_FWD,
                               ST_XFUSED, ST_XREG, ST_YZ_INV, ST_GATHER_TILE};

template <size_t NS> StageSeq seq_of(const Stage (&a)[NS]) { return StageSeq{a, (int)NS}; }

// The slab kernels use up to 160 KB of dynamic LDS: usable when the device
// allows that much per workgroup and the kernels accept the attribute.
template <typename T> bool slab_lds_usable(size_t bytes) {
    int dev = 0, maxsh = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&maxsh, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
            hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    if (bytes > (size_t)maxsh) return false;
    static std::once_flag once;
    static bool ok = false;
    std::call_once(once, [] {
        const void *fns[14] = {(const void *)fft_yz_fwd_kernel<T, SlabCfg<T>::small>,
                               (const void *)fft_yz_fwd_kernel<T, SlabCfg<T>::medium>,
                               (const void *)fft_yz_fwd_kernel<T, SlabCfg<T>::large>,
                               (const void *)fft_yz_inv_kernel<T, SlabCfg<T>::small>,
                               (const void *)fft_yz_inv_kernel<T, SlabCfg<T>::medium>,
                               (const void *)fft_yz_inv_kernel<T, SlabCfg<T>::large>,
                               (const void *)spread_slab_kernel<T, 4, false>,
                               (const void *)spread_slab_kernel<T, 6, false>,
                               (const void *)spread_slab_kernel<T, 8, false>,
                               (const void *)spread_slab_kernel<T, 0, false>,
                               (const void *)spread_slab_kernel<T, 4, true>,
                               (const void *)spread_slab_kernel<T, 6, true>,
                               (const void *)spread_slab_kernel<T, 8, true>,
                               (const void *)spread_slab_kernel<T, 0, true>};
        ok = true;
        for (const void *fn : fns)
            ok = ok && hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)kSlabLds) == hipSuccess;
        if (std::is_same<T, float>::value)
            for (const void *fn : {(const void *)fft_yz_fwd128_kernel,
                                   (const void *)fft_yz_inv128_kernel})
                ok = ok && hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                               (int)kSlabLds) == hipSuccess;
        if (!ok) (void)hipGetLastError();
    });
    return ok;
}

// Spreading and gather tiles may use more than the default 64 KB of dynamic
// LDS.  The default tile choices stay within kSpreadLds, so that at least two
// workgroups share one CU (gather tiles: two of 1024 threads; spreading tiles
// are smaller, four 512-thread workgroups per CU on water-sized columns); an
// autotuner may choose tiles up to kTileLdsMax (one workgroup per CU).  The attribute is set once for every
// instantiation; without it the tiles stay within 64 KB.
constexpr size_t kSpreadLds = 76 * 1024;
constexpr size_t kTileLdsMax = kSlabLds;
// LDS of a gather tile: (tx + n) x (ty + n) rows of K2 + n - 1 reals
size_t gather_tile_lds(int tx, int ty, int K2, int n, size_t tsz) {
    return (size_t)(tx + n) * (ty + n) * (K2 + n - 1) * tsz;
}
template <typename T, int ORDER, int MZ> bool gather_attr() {
    return hipFuncSetAttribute((const void *)gather_tile_kernel<T, ORDER, MZ, false>,
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)kTileLdsMax) == hipSuccess &&
           hipFuncSetAttribute((const void *)gather_tile_kernel<T, ORDER, MZ, true>,
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)kTileLdsMax) == hipSuccess;
}
template <typename T, int ORDER> bool spread_attr() {
    bool ok = hipFuncSetAttribute((const void *)spread_tile_kernel<T, ORDER, true, false>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)kTileLdsMax) == hipSuccess &&
              hipFuncSetAttribute((const void *)spread_tile_kernel<T, ORDER, true, true>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)kTileLdsMax) == hipSuccess &&
              gather_attr<T, ORDER, 1>() && gather_attr<T, ORDER, 2>() &&
              gather_attr<T, ORDER, 3>() && gather_attr<T, ORDER, 4>();
    if (!std::is_same<T, float>::value)
        ok = ok &&
             hipFuncSetAttribute((const void *)spread_tile_kernel<T, ORDER, false, false>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)kTileLdsMax) == hipSuccess &&
             hipFuncSetAttribute((const void *)spread_tile_kernel<T, ORDER, false, true>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)kTileLdsMax) == hipSuccess;
    return ok;
}
// the bin workgroup's records in LDS (72 KB for FP32): above the default
// 64 KB of dynamic LDS, set once per precision
template <typename T> bool bin_attr_ok() {
    static std::once_flag once;
    static bool ok = false;
    std::call_once(once, [] {
        ok = hipFuncSetAttribute((const void *)bin_kernel<T>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)bin_lds_bytes<T>()) == hipSuccess;
        if (!ok) (void)hipGetLastError();
    });
    return ok;
}
template <typename T> bool tile_attrs_ok() {
    static std::once_flag once;
    static bool ok = false;
    std::call_once(once, [] {
        int dev = 0, maxsh = 0;
        ok = hipGetDevice(&dev) == hipSuccess &&
             hipDeviceGetAttribute(&maxsh, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) ==
                 hipSuccess &&
             (size_t)maxsh >= kTileLdsMax && spread_attr<T, 4>() && spread_attr<T, 6>() &&
             spread_attr<T, 8>() && spread_attr<T, 0>();
        if (!ok) (void)hipGetLastError();
    });
    return ok;
}
// LDS of the default tile choice, and the most any tile may use
template <typename T> size_t spread_lds_budget() {
    return tile_attrs_ok<T>() ? kSpreadLds : kLdsBudget;
}
template <typename T> size_t spread_lds_max() {
    return tile_attrs_ok<T>() ? kTileLdsMax : kLdsBudget;
}

}  // namespace

StageSeq eval_stages(bool slab, bool tile) {
    if (tile) return slab ? seq_of(kEvalSlabTile) : seq_of(kEvalLinesTile);
    return slab ? seq_of(kEvalSlab) : seq_of(kEvalLines);
}
StageSeq sfac_stages(bool slab) { return slab ? seq_of(kSfacSlab) : seq_of(kSfacLines); }
StageSeq de1_stages(bool slab) { return slab ? seq_of(kDe1Slab) : seq_of(kDe1Lines); }

// ---------------------------------------------------------------------------
Engine::Engine(const double L6_[6], const int K_[3], int order_, int flags) {
    if (order_ % 2 != 0 || order_ < 2 || order_ > kMaxOrder)
        throw std::invalid_argument("invalid order (even, 2.." + std::to_string(kMaxOrder) +
                                    "): " + std::to_string(order_));
    for (int d = 0; d < 3; ++d) {
        if (K_[d] < 1) throw std::invalid_argument("invalid K");
        if (K_[d] > kMaxK)
            throw std::invalid_argument("grid dimension " + std::to_string(K_[d]) +
                                        " above the supported maximum " +
                                        std::to_string(kMaxK));
        K[d] = K_[d];
    }
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
    sp_f_ = spline_args<float>(M, order);
    sp_d_ = spline_args<double>(M, order);
    memset(&desc_, 0, sizeof(desc_));
    set_cell(L6_);
    desc_.tcell = desc_.cell;

    HIPC(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    if (!(fp32 ? bin_attr_ok<float>() : bin_attr_ok<double>()))
        throw std::runtime_error("the bin kernel's LDS (" +
                                 std::to_string(fp32 ? bin_lds_bytes<float>()
                                                     : bin_lds_bytes<double>()) +
                                 " bytes) is not available on this device");
    const size_t tsz = fp32 ? sizeof(float) : sizeof(double);
    alloc_(&grid_, grid_reals() * tsz);
    for (int d = 0; d < 3; ++d) {
        alloc_((void **)&iB_d_[d], K[d] * sizeof(double));
        HIPC(hipMemcpy(iB_d_[d], iB[d].data(), K[d] * sizeof(double), hipMemcpyHostToDevice));
    }
    for (int d = 0; d < 3; ++d) {
        FftPlan &p = plan_[d];
        p.N = K[d];
        p.radix = factor_radices(K[d]);
        if ((int)p.radix.size() > kMaxRadixStages) throw std::invalid_argument("K too composite");
        std::vector<double2> twd(K[d]);
        std::vector<float2> twf(K[d]);
        for (int t = 0; t < K[d]; ++t) {
            const double ang = -2.0 * M_PI * (double)t / (double)K[d];
            twd[t] = make_double2(cos(ang), sin(ang));
            twf[t] = make_float2((float)twd[t].x, (float)twd[t].y);
        }
        alloc_((void **)&p.tw_d, K[d] * sizeof(double2));
        alloc_((void **)&p.tw_f, K[d] * sizeof(float2));
        HIPC(hipMemcpy(p.tw_d, twd.data(), K[d] * sizeof(double2), hipMemcpyHostToDevice));
        HIPC(hipMemcpy(p.tw_f, twf.data(), K[d] * sizeof(float2), hipMemcpyHostToDevice));
    }

    // descriptor: geometry, FFT plans, blocking
    FrameDesc &D = desc_;
    D.K0 = K[0];
    D.K1 = K[1];
    D.K2 = K[2];
    D.ldim = ldim;
    D.order = order;
    D.grid = grid_;
    D.iB0 = iB_d_[0];
    D.iB1 = iB_d_[1];
    D.iB2 = iB_d_[2];
    D.tw0 = fp32 ? (const void *)plan_[0].tw_f : (const void *)plan_[0].tw_d;
    D.tw1 = fp32 ? (const void *)plan_[1].tw_f : (const void *)plan_[1].tw_d;
    D.tw2 = fp32 ? (const void *)plan_[2].tw_f : (const void *)plan_[2].tw_d;
    D.plan0 = plan_args(plan_[0]);
    D.plan1 = plan_args(plan_[1]);
    D.plan2 = plan_args(plan_[2]);
    D.mmax3 = mmax3;
    D.fixed = fixed ? 1 : 0;
    D.fx32 = fp32 ? 1 : 0;
    D.infl32 = (fp32 && (flags & kFlagInfl32)) ? 1 : 0;
    D.tabmode = kTableFull;
    const int nrows = K[0] * K[1];
    D.zC = fft_lines(fp32, K[2], (nrows + 1) / 2, D.plan2.generic, false);
    {
        const int c = fft_lines(fp32, K[1], ldim, D.plan1.generic, true);
        D.yT = (ldim + c - 1) / c;
        D.yC = (ldim + D.yT - 1) / D.yT;
    }
    for (int d : {1, 2}) {
        const FftPlanArgs &pa = d == 1 ? D.plan1 : D.plan2;
        const int C = d == 1 ? D.yC : D.zC;
        const size_t maxel = fp32 ? fft_elems<float>() : fft_elems<double>();
        if (lds_fft(fp32, K[d], C, pa.generic, d != 2) > kLdsBudget || (size_t)C * K[d] > maxel)
            throw std::invalid_argument("grid dimension " + std::to_string(K[d]) +
                                        " does not fit the FFT workgroup");
    }
    {
        // x lines: the largest power of two C <= 64 of (ky, kz) columns with
        // C K0 <= kXElems elements per workgroup
        const int P = K[1] * ldim;
        const int cap = kXElems;
        int lg = 0;
        while ((2 << lg) <= 64 && (2 << lg) * K[0] <= cap && (1 << lg) < P) ++lg;
        if ((1 << lg) * K[0] > cap)
            throw std::invalid_argument("grid dimension " + std::to_string(K[0]) +
                                        " does not fit the FFT workgroup");
        D.xLog = lg;
        D.xC = 1 << lg;
        D.xB = (P + D.xC - 1) / D.xC;
    }
    D.dldim = make_fastdiv((unsigned)ldim);
    D.dK2 = make_fastdiv((unsigned)K[2]);
    D.dpair = make_fastdiv((unsigned)((K[1] + 1) / 2));
    {
        const size_t lds = fp32 ? slab_lds_bytes<float>(K[1], K[2], ldim)
                                : slab_lds_bytes<double>(K[1], K[2], ldim);
        slab_ok = !D.plan1.generic && !D.plan2.generic && lds <= kSlabLds &&
                  slab_ept() > 0 &&
                  (fp32 ? slab_lds_usable<float>(lds) : slab_lds_usable<double>(lds));
        use_slab = slab_ok;
        slab128_ok_ = slab_ok && fp32 && slab128_lds_bytes() <= kSlabLds;
        spread_slab_ok =
            slab_ok && fixed &&
            (fp32 ? spread_slab_fits<float>(K[0], K[1], K[2], ldim, order)
                  : spread_slab_fits<double>(K[0], K[1], K[2], ldim, order));
        // the larger slab variants have no radix-16 stage
        if (use_slab && slab_ept() > 16)
            for (int d = 1; d < 3; ++d) {
                plan_[d].radix = factor_radices(K[d], false);
                (d == 1 ? D.plan1 : D.plan2) = plan_args(plan_[d]);
            }
    }
    // spreading tiles: rows of K2 + order - 1 cells (the z stencil unwrapped),
    // the largest of these that fits; z-chunked tiles are left to the autotuner
    // (on the 192 x 192 x 384 order-6 grid an 8 x 8 x 128 chunked tile measured
    // 3.4x slower than a whole-column 4 x 4 one: in a chunk only about a third
    // of the neighbourhood's atoms reach it, and each round of the compacted
    // loop leaves the other threads idle)
    // (512-thread workgroups on 8 x 4 tiles, four per CU, measured 91 us per
    // 16-frame water launch against 98 for 1024 threads on 8 x 8, two per CU)
    static const int tiles[][2] = {{8, 4}, {4, 4}, {4, 2}, {2, 2}, {2, 1}, {1, 1}};
    const size_t spread_budget =
        fp32 ? spread_lds_budget<float>() : spread_lds_budget<double>();
    D.tz = K[2];
    D.ntz = 1;
    for (auto &t : tiles) {
        const int tx = std::min(t[0], K[0]), ty = std::min(t[1], K[1]);
        if ((size_t)tx * ty * (K[2] + order - 1) * 8 <= spread_budget) {
            D.tx = tx;
            D.ty = ty;
            break;
        }
    }
    if (D.tx == 0) throw std::invalid_argument("K2 too large for the spreading tile");
    D.ntx = (K[0] + D.tx - 1) / D.tx;
    D.nty = (K[1] + D.ty - 1) / D.ty;
    // gather tiles (FP32): the staged block must fit the same LDS budget, its rows
    // must fit the per-lane row loads, and a tile of fewer than 16 columns stages
    // too much halo per atom (the per-atom gather is used then; FP64 rows double
    // the LDS and register footprint, and the per-atom gather is faster there)
    // (columns too long for any of these tiles: z chunks of an 8 x 8 tile)
    D.gtz = K[2];
    D.gntz = 1;
    if (fp32 && K[2] + order - 1 <= 64 * kGatherMaxZ) {
        static const int gtiles[][2] = {{8, 8}, {8, 4}, {4, 4}};
        for (auto &t : gtiles) {
            const int tx = std::min(t[0], K[0]), ty = std::min(t[1], K[1]);
            if (tx * ty >= std::min(16, K[0] * K[1]) &&
                gather_tile_lds(tx, ty, K[2], order, tsz) <= spread_budget) {
                D.gtx = tx;
                D.gty = ty;
                D.gntx = (K[0] + tx - 1) / tx;
                D.gnty = (K[1] + ty - 1) / ty;
                break;
            }
        }
    }
    if (fp32 && D.gtx == 0 && K[2] >= 2 * order) {
        const int tx = std::min(8, K[0]), ty = std::min(8, K[1]);
        const int zlds = (int)((spread_budget - kGatherBlockMax * sizeof(int)) /
                               ((size_t)(tx + order) * (ty + order) * tsz));
        const int zmax = std::min(64 * kGatherMaxZ, zlds) - (order - 1);
        if (tx * ty >= std::min(16, K[0] * K[1]) && zmax >= order) {
            const int nch = (K[2] + zmax - 1) / zmax;
            D.gtz = (K[2] + nch - 1) / nch;
            D.gntz = (K[2] + D.gtz - 1) / D.gtz;
            D.gtx = tx;
            D.gty = ty;
            D.gntx = (K[0] + tx - 1) / tx;
            D.gnty = (K[1] + ty - 1) / ty;
        }
    }


    nred_infl_ = (int)std::min<size_t>((H() + kBlock - 1) / kBlock, 1024);
    D.nred_i = nred_infl_;
    const int nred_max = std::max(nred_infl_, D.xB);
    alloc_((void **)&d_desc_, sizeof(FrameDesc));
    alloc_((void **)&d_desc_na_, sizeof(FrameDesc));
    HIPC(hipGetDevice(&dev_));
    memset(&desc_uploaded_, 0xff, sizeof(FrameDesc));
    alloc_((void **)&D.partials, (size_t)nred_max * 8 * sizeof(double));
    alloc_((void **)&D.colstart, ((size_t)nrows + 1) * sizeof(int));
    alloc_((void **)&D.fxscale, 64);
    alloc_((void **)&result_, 8 * sizeof(double));
    HIPC(hipMemsetAsync(grid_, 0, grid_reals() * tsz, stream_));
    HIPC(hipMemsetAsync(result_, 0, 8 * sizeof(double), stream_));
    D.out = result_;
    ensure_atoms(1024);
    HIPC(hipStreamSynchronize(stream_));
    // launch shapes tuned earlier in this process for this grid, order and
    // precision (a record that no longer applies is ignored)
    TuneCfg tc;
    if (tune_recall(*this, tc)) {
        try {
            apply_tune(tc);
        } catch (const std::invalid_argument &) {
        }
    }
}

Engine::~Engine() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    drop_graph_();
    for (auto &r : prof_recs_) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (auto &a : allocs_) (void)hipFree(a.first);
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
    allocs_.push_back({*p, bytes});
    dev_bytes_ += bytes;
}

void Engine::free_(void *p) {
    if (!p) return;
    for (size_t i = 0; i < allocs_.size(); ++i)
        if (allocs_[i].first == p) {
            HIPC(hipFree(p));
            dev_bytes_ -= allocs_[i].second;
            allocs_.erase(allocs_.begin() + (long)i);
            return;
        }
}

void Engine::check_staging(int n, std::initializer_list<const void *> ptrs) const {
    if (n < 0 || (size_t)n <= atom_cap_) return;
    const void *own[4] = {x_d_, w_d_, dx_d_, phi_d_};
    for (const void *p : ptrs)
        for (const void *o : own)
            if (p && p == o)
                throw std::invalid_argument(
                    "an atom buffer of this handle's own staging holds " +
                    std::to_string(atom_cap_) + " atoms, fewer than n = " + std::to_string(n) +
                    " (grow it first and take its address afterwards)");
}

void Engine::drop_graph_() {
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    graph_exec_ = nullptr;
    graph_ = nullptr;
    graph_n_ = -1;
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
    desc_.cell = cell;
    ++gen_;
    drop_graph_();
}

void Engine::upload_table(const double *A, const double *dA) {
    const size_t h = H();
    if (!table_) alloc_((void **)&table_, 7 * h * sizeof(double));
    HIPC(hipMemcpyAsync(table_, A, h * sizeof(double), hipMemcpyHostToDevice, stream_));
    HIPC(hipMemcpyAsync(table_ + h, dA, 6 * h * sizeof(double), hipMemcpyHostToDevice, stream_));
    HIPC(hipStreamSynchronize(stream_));
    desc_.table = table_;
    desc_.tabmode = kTableFull;
    desc_.tcell = cell;
    has_table_ = true;
    ++gen_;
    drop_graph_();
}

// The exact table (every image of the reference's ball, folded as the
// reference folds it) is built on the device, one thread per half-complex
// point, which also compares it with the closed form evaluated from the
// indices.  If they agree to 1e-13 of the mean magnitude, the kernels
// evaluate the closed form on the fly and no table is read; otherwise they
// read the table.
void Engine::build_ewald_table(double eta, double max_m) {
    const size_t h = H();
    if (!table_) alloc_((void **)&table_, 7 * h * sizeof(double));
    const double fac = -(M_PI / eta) * (M_PI / eta);
    BallArgs a;
    a.K0 = K[0];
    a.K1 = K[1];
    a.K2 = K[2];
    a.ldim = ldim;
    a.cell = cell;
    a.R2 = max_m * max_m;
    a.fac = fac;
    a.i0 = (int)(max_m / cell.iL00);
    const int nb = (int)((h + kBlock - 1) / kBlock);
    if ((size_t)nb * 4 > table_stats_cap_) {
        alloc_((void **)&table_stats_, (size_t)nb * 4 * sizeof(double));
        table_stats_cap_ = (size_t)nb * 4;
    }
    ewald_table_kernel<<<nb, kBlock, 0, stream_>>>(a, table_, table_ + h, table_stats_);
    HIPC(hipGetLastError());
    std::vector<double> st((size_t)nb * 4);
    HIPC(hipMemcpyAsync(st.data(), table_stats_, st.size() * sizeof(double),
                        hipMemcpyDeviceToHost, stream_));
    HIPC(hipStreamSynchronize(stream_));
    double sumA = 0.0, errA = 0.0, sumD = 0.0, errD = 0.0;
    for (int b = 0; b < nb; ++b) {
        sumA += st[4 * b];
        errA = std::max(errA, st[4 * b + 1]);
        sumD += st[4 * b + 2];
        errD = std::max(errD, st[4 * b + 3]);
    }
    const double tol = 1e-13;
    const bool analytic = errA <= tol * sumA / (double)h && errD <= tol * sumD / (double)h;
    desc_.table = table_;
    desc_.tabmode = analytic ? kTableAnalytic : kTableFull;
    desc_.ewald_fac = fac;
    desc_.max_m2 = max_m * max_m;
    desc_.tcell = cell;
    has_table_ = true;
    ++gen_;
    drop_graph_();
}

void Engine::copy_table(double *A, double *dA) {
    if (!table_ || !has_table_) throw UsageError("No convolution kernel has been set.");
    const size_t h = H();
    if (A) HIPC(hipMemcpyAsync(A, table_, h * sizeof(double), hipMemcpyDeviceToHost, stream_));
    if (dA)
        HIPC(hipMemcpyAsync(dA, table_ + h, 6 * h * sizeof(double), hipMemcpyDeviceToHost,
                            stream_));
    HIPC(hipStreamSynchronize(stream_));
}

void Engine::ensure_atoms(int n) {
    if ((size_t)n <= atom_cap_) return;
    sync();
    drop_graph_();
    const size_t cap = std::max<size_t>(1024, (size_t)n + n / 4);
    // staging for the host-pointer entry points and the sort workspace: new
    // buffers first, then the old ones are released (the stream is idle);
    // batches holding the old workspace addresses see the generation change
    void *old[5] = {x_d_, w_d_, dx_d_, phi_d_, ws_};
    ++gen_;
    alloc_((void **)&x_d_, 3 * cap * sizeof(double));
    alloc_((void **)&w_d_, cap * sizeof(double));
    alloc_((void **)&dx_d_, 3 * cap * sizeof(double));
    alloc_((void **)&phi_d_, cap * sizeof(double));
    alloc_(&ws_, workspace_bytes(cap));
    bind_workspace(desc_, ws_, cap);
    for (void *p : old) free_(p);
    atom_cap_ = cap;
    const size_t need = 5 * cap * sizeof(double);
    if (need > pin_cap_) {
        if (pin_) HIPC(hipHostFree(pin_));
        HIPC(hipHostMalloc((void **)&pin_, need, hipHostMallocDefault));
        pin_cap_ = need;
    }
}

// The per-atom sort workspace of n atoms in one block: the bin chunks (n
// records in x-plane order per bin workgroup), the plane starts of each chunk
// (K0 + 1 ints per bin workgroup), the column-ordered records and one |w|
// maximum per bin workgroup, each part 256-byte aligned.  Nothing in it needs
// to start at zero.
namespace {
size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
}  // namespace

// Records of one plane the sort stages in LDS: 1.5 times a uniform plane's
// share plus 64 (a plane beyond it is sorted straight into place), within
// 64 KB.
int Engine::sort_cap(size_t n) const {
    const size_t rec = fp32 ? sizeof(AtomRec<float>) : sizeof(AtomRec<double>);
    const size_t want = (3 * n + 2 * (size_t)K[0] - 1) / (2 * (size_t)K[0]) + 64;
    return (int)std::min(want, (size_t)65536 / rec);
}

size_t Engine::workspace_bytes(size_t n) const {
    const size_t rec = fp32 ? sizeof(AtomRec<float>) : sizeof(AtomRec<double>);
    const size_t nbin = (size_t)bin_count((int)n, fp32);
    return align256(n * rec) + align256(nbin * ((size_t)K[0] + 1) * sizeof(int)) +
           align256(n * rec) + align256(nbin * sizeof(double));
}

void Engine::bind_workspace(FrameDesc &D, void *base, size_t n) const {
    const size_t rec = fp32 ? sizeof(AtomRec<float>) : sizeof(AtomRec<double>);
    const size_t nbin = (size_t)bin_count((int)n, fp32);
    char *p = static_cast<char *>(base);
    D.scap = sort_cap(n);
    D.bins = p;
    p += align256(n * rec);
    D.binstart = reinterpret_cast<int *>(p);
    p += align256(nbin * ((size_t)K[0] + 1) * sizeof(int));
    D.recs = p;
    p += align256(n * rec);
    D.wmax = reinterpret_cast<double *>(p);
}

void Engine::set_atoms_(int n, const double *w, const double *x, double *dx, double *phi,
                        double *out) {
    desc_.n = n;
    desc_.w = w;
    desc_.x = x;
    desc_.dx = dx;
    desc_.phi = phi;
    desc_.out = out;
}

int Engine::stage_blocks(Stage s) const { return stage_blocks_n(s, desc_.n); }

bool Engine::uses_stage(Stage s) const {
    switch (s) {
        case ST_SPREAD: return !spread_slab();
        case ST_SPREAD_YZ: return spread_slab();
        case ST_YZ_FWD: return slab() && !spread_slab();
        case ST_YZ_INV: return slab();
        case ST_ZR2C:
        case ST_YF:
        case ST_YI:
        case ST_ZC2R: return !slab();
        case ST_GATHER_TILE: return gtile();
        case ST_GATHER_SORTED: return !gtile();
        case ST_XFUSED: return !xreg();
        case ST_XREG: return xreg();
        default: return true;
    }
}

int Engine::stage_blocks_n(Stage s, int natoms) const {
    const FrameDesc &D = desc_;
    const int nrows = K[0] * K[1];
    switch (s) {
        case ST_BIN: return bin_count(natoms, fp32);
        case ST_SORT: return K[0];
        case ST_SPREAD: return spread_slab() ? 0 : D.ntx * D.nty * D.ntz;
        case ST_SPREAD_YZ: return spread_slab() ? K[0] : 0;
        case ST_ZR2C:
        case ST_ZC2R: return (nrows + 2 * D.zC - 1) / (2 * D.zC);
        case ST_YF:
        case ST_YI: return K[0] * D.yT;
        case ST_XFUSED: return xreg() ? 0 : D.xB;
        case ST_XREG: return xreg() ? D.xB : 0;
        case ST_XF_SCALE:
        case ST_XI: return D.xB;
        case ST_YZ_FWD: return spread_slab() ? 0 : K[0];
        case ST_YZ_INV: return K[0];
        case ST_INFL:
        case ST_INFL_E: return nred_infl_;
        case ST_FINISH: return 1;
        case ST_GATHER_SORTED:
        case ST_GATHER: {
            const int apb = kBlock / order;
            return (natoms + apb - 1) / apb;
        }
        case ST_GATHER_TILE: return D.gntx * D.gnty * D.gntz;
        default: return 0;
    }
}

size_t Engine::stage_lds(Stage s, int xpath) const {
    const FrameDesc &D = desc_;
    // the x launches of a batch run the most general path of its frames: a
    // frame with a plain plan in the generic kernel needs its second line set
    const int xgen = std::max(xpath, x_path(D)) == kXGeneric;
    switch (s) {
        case ST_SORT:  // the staged plane
            return fp32 ? sort_lds_bytes<float>(D.scap) : sort_lds_bytes<double>(D.scap);
        case ST_SPREAD:  // chunked tiles: the compacted atom list after the cells
            return (size_t)D.tx * D.ty * (D.ntz > 1 ? D.tz : K[2] + order - 1) * 8 +
                   (D.ntz > 1 ? kSpreadBlock * sizeof(int) : 0);
        case ST_SPREAD_YZ:
            return fp32 ? spread_slab_lds_bytes<float>(K[1], K[2], order)
                        : spread_slab_lds_bytes<double>(K[1], K[2], order);
        case ST_GATHER_TILE:  // chunked tiles: the compacted atom list after the rows
            return gather_tile_lds(D.gtx, D.gty, D.gntz > 1 ? D.gtz : K[2], order,
                                   fp32 ? 4 : 8) +
                   (D.gntz > 1 ? kGatherBlockMax * sizeof(int) : 0);

        case ST_ZR2C:
        case ST_ZC2R: return lds_fft(fp32, K[2], D.zC, D.plan2.generic, false);
        case ST_YF:
        case ST_YI: return lds_fft(fp32, K[1], D.yC, D.plan1.generic, true);
        case ST_XF_SCALE:
        case ST_XI:
            return fp32 ? xlines_lds_bytes<float>(K[0], D.xC, xgen)
                        : xlines_lds_bytes<double>(K[0], D.xC, xgen);
        case ST_XREG: return xreg_lds_bytes<float>();
        case ST_XFUSED:  // the lines, the wavefront sums, the closed form's x factors
            return (fp32 ? xlines_lds_bytes<float>(K[0], D.xC, xgen)
                         : xlines_lds_bytes<double>(K[0], D.xC, xgen)) +
                   xfused_extra_lds(K[0], D.xC);
        case ST_YZ_FWD:
        case ST_YZ_INV:
            if (slab_kernel() == kSlabKernel128) return slab128_lds_bytes();
            return fp32 ? slab_lds_bytes<float>(K[1], K[2], ldim)
                        : slab_lds_bytes<double>(K[1], K[2], ldim);
        default: return 0;
    }
}

void Engine::launch(Stage s, const Launch &L, int blocks, size_t lds,
                    hipStream_t st, int finish) {
    if (blocks <= 0) return;
    if (fp32) launch_order<float>(order, s, L, blocks, lds, st, sp_f_, fixed, finish);
    else launch_order<double>(order, s, L, blocks, lds, st, sp_d_, fixed, finish);
}

int Engine::slab_kernel() const {
    if (fp32 && slab() && slab128_ok_ && slab128_applies(K[1], K[2], desc_.plan1, desc_.plan2))
        return kSlabKernel128;
    return slab_ept();
}

int Engine::slab_ept() const {
    return fp32 ? slab_ept_for<float>(K[1], K[2], ldim) : slab_ept_for<double>(K[1], K[2], ldim);
}

void Engine::sync_desc_() {
    if (memcmp(&desc_, &desc_uploaded_, sizeof(FrameDesc)) == 0) return;
    // pageable source: the call returns once the bytes are staged
    HIPC(hipMemcpyAsync(d_desc_, &desc_, sizeof(FrameDesc), hipMemcpyHostToDevice, stream_));
    desc_uploaded_ = desc_;
}

void Engine::run_(Stage s, int finish) {
    const int nb = stage_blocks(s);
    if (nb <= 0) return;
    sync_desc_();
    ProfScope ps(this, stage_name(s));
    launch(s, Launch{d_desc_, nullptr, 1, slab_kernel(), x_path(desc_), gmz(), desc_.ntz > 1,
                     desc_.gntz > 1, x_influence_form(desc_)}, nb,
           stage_lds(s), stream_,
           finish);
}

void Engine::run_seq_(StageSeq seq) {
    Stage prev = ST_COUNT;
    for (int i = 0; i < seq.n; ++i) {
        run_(seq.s[i], finishes(prev, seq.s[i]));
        prev = seq.s[i];
    }
}

// ---- public device-pointer operations -----------------------------------------------
void Engine::spread_forward(int n, const double *w_d, const double *x_d) {
    check_staging(n, {w_d, x_d});
    ensure_atoms(n);
    set_atoms_(n, w_d, x_d, desc_.dx, desc_.phi, desc_.out);
    run_seq_(sfac_stages(slab()));
    state = 0;
}

void Engine::energy(double *out_d) {
    if (state != 0) throw UsageError("You must call sfac before this routine.");
    if (!has_table_) throw UsageError("No convolution kernel has been set.");
    desc_.out = out_d;
    run_seq_(seq_of(kEnStages));
}

void Engine::convolve(double *out_d) {
    if (state != 0) throw UsageError("You must call sfac before this routine.");
    if (!has_table_) throw UsageError("No convolution kernel has been set.");
    desc_.out = out_d;
    run_seq_(de1_stages(slab()));
    state = 1;
}

void Engine::gather(int n, const double *w_d, const double *x_d, double *dx_d, double *phi_d) {
    if (state != 1) throw UsageError("You must call de1 before this routine.");
    check_staging(n, {w_d, x_d, dx_d, phi_d});
    ensure_atoms(n);
    set_atoms_(n, w_d, x_d, dx_d, phi_d, desc_.out);
    run_(ST_GATHER, 0);
}

void Engine::evaluate(int n, const double *w_d, const double *x_d, double *out_d,
                      double *dx_d, double *phi_d) {
    if (!has_table_) throw UsageError("No convolution kernel has been set.");
    check_staging(n, {w_d, x_d, dx_d, phi_d});
    ensure_atoms(n);
    set_atoms_(n, w_d, x_d, dx_d, phi_d, out_d);
    const void *key[6] = {w_d, x_d, out_d, dx_d, phi_d, desc_.recs};
    const bool graph = use_graph && !prof_;
    if (graph) {
        const bool same = graph_exec_ && graph_n_ == n && graph_variant_ == variant() &&
                          memcmp(key, graph_key_, sizeof(key)) == 0;
        // the descriptor upload stays outside the graph, and runs before every
        // replay: the split entry points (sfac/en/de1/de2/potl) upload their
        // own descriptors between evaluations (a no-op when nothing changed)
        sync_desc_();
        if (!same) {
            drop_graph_();
            HIPC(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
            try {
                run_seq_(eval_stages(slab(), gtile()));
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(stream_, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            HIPC(hipStreamEndCapture(stream_, &graph_));
            HIPC(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
            graph_n_ = n;
            graph_variant_ = variant();
            memcpy(graph_key_, key, sizeof(key));
        }
        HIPC(hipGraphLaunch(graph_exec_, stream_));
    } else {
        run_seq_(eval_stages(slab(), gtile()));
    }
    state = 1;
}

// ---- evaluation without device allocation ------------------------------------------------
// The per-atom workspace of an allocation-free evaluation lives in one static
// device array per GPU (held until the process ends).  Its device address is
// looked up once per device with that device current; host code passes that
// address, never the array's name, to kernels and copies.  Nothing in it needs
// to start at zero (the column counters, which must, belong to the handle and
// are cleared by the scan of every evaluation).
namespace {
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

void *scratch_ptr(int dev) {
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        int cur = -1;
        HIPC(hipGetDevice(&cur));
        if (cur != dev) HIPC(hipSetDevice(dev));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        hip_check(e, "hipGetSymbolAddress(scratch)");
        if (!p) throw std::runtime_error("hipGetSymbolAddress(scratch) returned NULL");
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}
}  // namespace

void Engine::evaluate_noalloc(int n, const double *w_d, const double *x_d, double *out_d,
                              double *dx_d, double *phi_d) {
    if (!has_table_) throw UsageError("No convolution kernel has been set.");
    if (n < 0) throw std::invalid_argument("n < 0");
    const size_t need = workspace_bytes((size_t)n);
    if (need > kNoAllocScratchBytes)
        throw std::invalid_argument(
            "the allocation-free evaluation holds " + std::to_string(kNoAllocScratchBytes) +
            " bytes of scratch; n = " + std::to_string(n) + " atoms need " +
            std::to_string(need) + " (use sfac_eval_dev)");
    void *base = scratch_ptr(dev_);
    desc_na_ = desc_;
    desc_na_.n = n;
    desc_na_.w = w_d;
    desc_na_.x = x_d;
    desc_na_.out = out_d;
    desc_na_.dx = dx_d;
    desc_na_.phi = phi_d;
    bind_workspace(desc_na_, base, (size_t)n);
    // pageable source: the call returns once the bytes are staged
    HIPC(hipMemcpyAsync(d_desc_na_, &desc_na_, sizeof(FrameDesc), hipMemcpyHostToDevice,
                        stream_));
    const Launch L{d_desc_na_, nullptr, 1, slab_kernel(), x_path(desc_), gmz(), desc_.ntz > 1,
                   desc_.gntz > 1, x_influence_form(desc_)};
    const StageSeq seq = eval_stages(slab(), gtile());
    Stage prev = ST_COUNT;
    for (int i = 0; i < seq.n; ++i) {
        const Stage st = seq.s[i];
        // the sort's LDS follows the staging capacity of this call's workspace
        const size_t lds = st != ST_SORT ? stage_lds(st)
                           : fp32        ? sort_lds_bytes<float>(desc_na_.scap)
                                         : sort_lds_bytes<double>(desc_na_.scap);
        launch(st, L, stage_blocks_n(st, n), lds, stream_, finishes(prev, st));
        prev = st;
    }
    state = 1;
}

// ---- host-pointer helpers ----------------------------------------------------------------
void Engine::sfac_host(int n, const double *w, const double *x) {
    if (n < 0) throw std::invalid_argument("n < 0");
    ensure_atoms(n);
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
    ensure_atoms(n);
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

double Engine::eval_host(int n, const double *w, const double *x, double vir[6], double *dx,
                         double *phi) {
    if (n < 0) throw std::invalid_argument("n < 0");
    ensure_atoms(n);
    HIPC(hipStreamSynchronize(stream_));
    memcpy(pin_, x, 3 * (size_t)n * sizeof(double));
    memcpy(pin_ + 3 * (size_t)n, w, (size_t)n * sizeof(double));
    HIPC(hipMemcpyAsync(x_d_, pin_, 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice,
                        stream_));
    HIPC(hipMemcpyAsync(w_d_, pin_ + 3 * (size_t)n, (size_t)n * sizeof(double),
                        hipMemcpyHostToDevice, stream_));
    evaluate(n, w_d_, x_d_, result_, dx ? dx_d_ : nullptr, phi ? phi_d_ : nullptr);
    double out[8];
    HIPC(hipMemcpyAsync(out, result_, 8 * sizeof(double), hipMemcpyDeviceToHost, stream_));
    if (dx && n)
        HIPC(hipMemcpyAsync(dx, dx_d_, 3 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost,
                            stream_));
    if (phi && n)
        HIPC(hipMemcpyAsync(phi, phi_d_, (size_t)n * sizeof(double), hipMemcpyDeviceToHost,
                            stream_));
    HIPC(hipStreamSynchronize(stream_));
    if (vir)
        for (int c = 0; c < 6; ++c) vir[c] = out[1 + c];
    return out[0];
}

void Engine::upload_atoms(int n, const double *w, const double *x, bool pinned) {
    if (n < 0) throw std::invalid_argument("n < 0");
    ensure_atoms(n);
    if (n == 0) return;
    const size_t nb = (size_t)n * sizeof(double);
    const double *xs = x, *ws = w;
    if (!pinned) {
        HIPC(hipStreamSynchronize(stream_));  // the staging buffer may still be in flight
        memcpy(pin_, x, 3 * nb);
        memcpy(pin_ + 3 * (size_t)n, w, nb);
        xs = pin_;
        ws = pin_ + 3 * (size_t)n;
    }
    HIPC(hipMemcpyAsync(x_d_, xs, 3 * nb, hipMemcpyHostToDevice, stream_));
    HIPC(hipMemcpyAsync(w_d_, ws, nb, hipMemcpyHostToDevice, stream_));
}

void Engine::evaluate_staged(int n, bool want_dx, bool want_phi) {
    evaluate(n, w_d_, x_d_, result_, want_dx ? dx_d_ : nullptr, want_phi ? phi_d_ : nullptr);
}

void Engine::download_results(int n, double *out, double *dx, double *phi, hipStream_t st) {
    if (out) HIPC(hipMemcpyAsync(out, result_, 7 * sizeof(double), hipMemcpyDeviceToHost, st));
    if (dx && n)
        HIPC(hipMemcpyAsync(dx, dx_d_, 3 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost,
                            st));
    if (phi && n)
        HIPC(hipMemcpyAsync(phi, phi_d_, (size_t)n * sizeof(double), hipMemcpyDeviceToHost,
                            st));
}

bool Engine::xreg() const {
    const std::vector<int> &r = plan_[0].radix;
    return fp32 && K[0] == kXRegN && r.size() == 2 && r[0] == 16 && r[1] == 8 &&
           desc_.xC == kXRegC && x_influence_form(desc_) == kXInfSeparable;
}

unsigned long long Engine::variant() const {
    unsigned long long v = 0;
    v |= (unsigned long long)(fp32 ? 1 : 0);
    v |= (unsigned long long)(fixed ? 1 : 0) << 1;
    v |= (unsigned long long)(slab() ? 1 : 0) << 2;
    v |= (unsigned long long)(gtile() ? 1 : 0) << 3;
    v |= (unsigned long long)(x_path(desc_) & 0xf) << 4;
    v |= (unsigned long long)(order & 0xff) << 8;
    v |= (unsigned long long)(slab_kernel() & 0xffff) << 16;
    v |= (unsigned long long)(gmz() & 0xffff) << 32;
    v |= (unsigned long long)(desc_.ntz > 1 ? 1 : 0) << 48;
    v |= (unsigned long long)(desc_.gntz > 1 ? 1 : 0) << 49;
    v |= (unsigned long long)(x_influence_form(desc_) & 3) << 50;
    v |= (unsigned long long)(xreg() ? 1 : 0) << 52;
    v |= (unsigned long long)(spread_slab() ? 1 : 0) << 53;
    return v;
}

// ---- launch shapes for the autotuner -----------------------------------------------------
namespace {
bool plain_radix(int r) { return r == 2 || r == 3 || r == 4 || r == 5 || r == 8 || r == 16; }
}  // namespace

TuneCfg Engine::tune_cfg() const {
    TuneCfg c;
    // an axis whose plan has a generic radix is not tuned (nrad 0: keep)
    for (int d = 0; d < 3; ++d) {
        bool plain = true;
        for (int r : plan_[d].radix) plain = plain && plain_radix(r);
        c.nrad[d] = plain ? (int)plan_[d].radix.size() : 0;
        for (int s = 0; s < c.nrad[d]; ++s) c.radix[d][s] = plan_[d].radix[s];
    }
    c.tx = desc_.tx;
    c.ty = desc_.ty;
    c.tz = desc_.ntz > 1 ? desc_.tz : 0;
    c.gtx = desc_.gtx;
    c.gty = desc_.gty;
    c.gtz = desc_.gntz > 1 ? desc_.gtz : 0;
    c.slab = !use_slab ? 0 : use_spread_slab ? 2 : 1;
    return c;
}

namespace {
// a z chunk (tz > 0) must hold a whole stencil, and the column at least two
bool z_chunk_ok(int tz, const int K[3], int order) {
    return tz == 0 || (tz >= order && tz < K[2] && K[2] >= 2 * order);
}

bool spread_tile_fits(int tx, int ty, int tz, const int K[3], int order, size_t lds_max) {
    if (tx < 1 || ty < 1 || tx > K[0] || ty > K[1]) return false;
    if (std::min(tx + order - 1, K[0]) > kMaxSeg / 2) return false;  // column segments
    if (!z_chunk_ok(tz, K, order)) return false;
    return (size_t)tx * ty * (tz ? tz : K[2] + order - 1) * 8 +
               (tz ? kSpreadBlock * sizeof(int) : 0) <= lds_max;
}

bool gather_tile_fits(int tx, int ty, int tz, const int K[3], int order, size_t lds_max,
                      size_t tsz) {
    if (tx < 1 || ty < 1 || tx > K[0] || ty > K[1] || tx > 64) return false;
    if (tx * ty < std::min(16, K[0] * K[1])) return false;
    if (!z_chunk_ok(tz, K, order)) return false;
    const int zl = tz ? tz : K[2];
    if (zl + order - 1 > 64 * kGatherMaxZ) return false;
    if ((tx + order) * (ty + order) > 64 * (kGatherBlock / 64)) return false;  // staged rows
    return gather_tile_lds(tx, ty, zl, order, tsz) + (tz ? kGatherBlockMax * sizeof(int) : 0) <=
           lds_max;
}
}  // namespace

void Engine::apply_tune(const TuneCfg &c) {
    // validate everything before changing anything
    for (int d = 0; d < 3; ++d) {
        if (c.nrad[d] == 0) continue;
        if (c.nrad[d] < 0 || c.nrad[d] > kMaxRadixStages)
            throw std::invalid_argument("radix stage count");
        long long prod = 1;
        for (int s = 0; s < c.nrad[d]; ++s) {
            if (!plain_radix(c.radix[d][s])) throw std::invalid_argument("radix not in 2,3,4,5,8,16");
            prod *= c.radix[d][s];
        }
        if (prod != K[d]) throw std::invalid_argument("radices do not multiply to K");
        for (int r : plan_[d].radix)
            if (!plain_radix(r)) throw std::invalid_argument("axis has a generic radix");
    }
    const size_t lmax = fp32 ? spread_lds_max<float>() : spread_lds_max<double>();
    if (c.tx || c.ty)
        if (!spread_tile_fits(c.tx, c.ty, c.tz, K, order, lmax))
            throw std::invalid_argument("spreading tile does not fit");
    if (c.gtx || c.gty)
        if (!gather_tile_fits(c.gtx, c.gty, c.gtz, K, order, lmax, fp32 ? 4 : 8))
            throw std::invalid_argument("gather tile does not fit");
    if (c.slab && !slab_ok) throw std::invalid_argument("slab transforms do not apply");
    // radix-16 stages run in kernels with at most 16 elements per thread:
    // the line kernels always, the y-z slab kernels in their small variant
    const bool slab_after = c.slab != 0;
    if (slab_after && slab_ept() > 16)
        for (int d = 1; d < 3; ++d) {
            const std::vector<int> &rd = c.nrad[d] ? std::vector<int>(c.radix[d], c.radix[d] + c.nrad[d])
                                                   : plan_[d].radix;
            for (int r : rd)
                if (r == 16)
                    throw std::invalid_argument("radix 16 does not apply to these slab transforms");
        }
    FrameDesc &D = desc_;
    for (int d = 0; d < 3; ++d) {
        if (c.nrad[d] == 0) continue;
        plan_[d].radix.assign(c.radix[d], c.radix[d] + c.nrad[d]);
        const FftPlanArgs a = plan_args(plan_[d]);
        if (d == 0) D.plan0 = a;
        else if (d == 1) D.plan1 = a;
        else D.plan2 = a;
    }
    if (c.tx || c.ty) {
        D.tx = c.tx;
        D.ty = c.ty;
        D.ntx = (K[0] + D.tx - 1) / D.tx;
        D.nty = (K[1] + D.ty - 1) / D.ty;
        D.tz = c.tz ? c.tz : K[2];
        D.ntz = (K[2] + D.tz - 1) / D.tz;
    }
    D.gtx = c.gtx;
    D.gty = c.gty;
    D.gntx = c.gtx ? (K[0] + c.gtx - 1) / c.gtx : 0;
    D.gnty = c.gty ? (K[1] + c.gty - 1) / c.gty : 0;
    D.gtz = c.gtz ? c.gtz : K[2];
    D.gntz = (K[2] + D.gtz - 1) / D.gtz;
    use_slab = c.slab != 0;
    use_spread_slab = c.slab != 1;
    ++gen_;
    drop_graph_();
}

void Engine::set_transform(int mode) {
    if (mode < 0 || mode > 2) throw std::invalid_argument("transform mode must be 0, 1 or 2");
    use_slab = mode != 1;
    use_spread_slab = mode == 0;
    // a batch holding this frame takes its launches anew
    ++gen_;
    drop_graph_();
}

// Every ordered factorisation of K_d into radices 2, 3, 4, 5, 8, 16 with at most
// one stage more than the default plan (the default first); at most 32.
std::vector<std::vector<int>> Engine::radix_options(int d) const {
    const std::vector<int> &def = plan_[d].radix;
    std::vector<std::vector<int>> out{def};
    for (int r : def)
        if (!plain_radix(r)) return out;
    const int maxlen = std::min<int>((int)def.size() + 1, kMaxRadixStages);
    std::vector<std::vector<int>> all;
    std::vector<int> cur;
    const int rs[6] = {16, 8, 4, 2, 3, 5};
    std::function<void(int)> rec = [&](int rem) {
        if (rem == 1) {
            if (cur != def) all.push_back(cur);
            return;
        }
        if ((int)cur.size() >= maxlen) return;
        for (int r : rs)
            if (rem % r == 0) {
                cur.push_back(r);
                rec(rem / r);
                cur.pop_back();
            }
    };
    if (K[d] > 1) rec(K[d]);
    std::stable_sort(all.begin(), all.end(), [](const std::vector<int> &a,
                                                const std::vector<int> &b) {
        return a.size() < b.size();
    });
    for (auto &v : all) {
        if (out.size() >= 32) break;
        out.push_back(v);
    }
    return out;
}

std::vector<std::array<int, 3>> Engine::spread_tile_options() const {
    std::vector<std::array<int, 3>> out;
    const size_t lmax = fp32 ? spread_lds_max<float>() : spread_lds_max<double>();
    const int sz[] = {1, 2, 4, 8, 16, 32};
    const int zs[] = {0, 32, 64, 128, 256};
    for (int a : sz)
        for (int b : sz)
            for (int tz : zs) {
                const int tx = std::min(a, K[0]), ty = std::min(b, K[1]);
                if (tx * ty < std::min(16, K[0] * K[1])) continue;
                if (!spread_tile_fits(tx, ty, tz, K, order, lmax)) continue;
                const std::array<int, 3> t{tx, ty, tz};
                if (std::find(out.begin(), out.end(), t) == out.end()) out.push_back(t);
            }
    return out;
}

std::vector<std::array<int, 3>> Engine::gather_tile_options() const {
    std::vector<std::array<int, 3>> out{{0, 0, 0}};
    const size_t lmax = fp32 ? spread_lds_max<float>() : spread_lds_max<double>();
    const int sz[] = {2, 4, 8, 16, 32};
    const int zs[] = {0, 32, 64, 128, 192};
    for (int a : sz)
        for (int b : sz)
            for (int tz : zs) {
                const int tx = std::min(a, K[0]), ty = std::min(b, K[1]);
                if (!gather_tile_fits(tx, ty, tz, K, order, lmax, fp32 ? 4 : 8)) continue;
                const std::array<int, 3> t{tx, ty, tz};
                if (std::find(out.begin(), out.end(), t) == out.end()) out.push_back(t);
            }
    return out;
}

std::string Engine::tune_key() const {
    return std::to_string(K[0]) + "x" + std::to_string(K[1]) + "x" + std::to_string(K[2]) +
           "/" + std::to_string(order) + "/" + std::to_string((fp32 ? 1 : 0) | (fixed ? 2 : 0));
}

// ---- process-wide record of tuned shapes ------------------------------------------------
namespace {
std::mutex g_tune_mu;
std::vector<std::pair<std::string, TuneCfg>> g_tuned;

std::string cfg_text(const TuneCfg &c) {
    std::string s;
    for (int d = 0; d < 3; ++d) {
        s += d ? " " : "";
        for (int i = 0; i < c.nrad[d]; ++i) s += (i ? "," : "") + std::to_string(c.radix[d][i]);
        if (c.nrad[d] == 0) s += "-";
    }
    s += " " + std::to_string(c.tx) + "x" + std::to_string(c.ty) + "x" + std::to_string(c.tz);
    s += " " + std::to_string(c.gtx) + "x" + std::to_string(c.gty) + "x" +
         std::to_string(c.gtz);
    s += " " + std::to_string(c.slab);
    return s;
}
}  // namespace

void tune_remember(const Engine &e, const TuneCfg &c) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    const std::string k = e.tune_key();
    for (auto &kv : g_tuned)
        if (kv.first == k) {
            kv.second = c;
            return;
        }
    g_tuned.push_back({k, c});
}

bool tune_recall(const Engine &e, TuneCfg &c) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    const std::string k = e.tune_key();
    for (auto &kv : g_tuned)
        if (kv.first == k) {
            c = kv.second;
            return true;
        }
    return false;
}

std::string tune_line(const Engine &e) { return e.tune_key() + " " + cfg_text(e.tune_cfg()); }

void tune_clear() {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned.clear();
}

// one line per key: "K0xK1xK2/order/flags r0 r1 r2 TXxTYxTZ GTXxGTYxGTZ slab",
// radices comma-separated per axis ("-": the default plan), TZ/GTZ = 0 for
// whole columns (an "AxB" tile without z is read as whole columns)
std::string tune_export() {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    std::string s;
    for (auto &kv : g_tuned) s += kv.first + " " + cfg_text(kv.second) + "\n";
    return s;
}

int tune_import(const std::string &text) {
    int nread = 0;
    size_t pos = 0;
    while (pos < text.size()) {
        size_t eol = text.find('\n', pos);
        if (eol == std::string::npos) eol = text.size();
        const std::string line = text.substr(pos, eol - pos);
        pos = eol + 1;
        char key[64], ax[3][128], st[64], gt[64];
        int slab;
        if (sscanf(line.c_str(), "%63s %127s %127s %127s %63s %63s %d", key, ax[0], ax[1],
                   ax[2], st, gt, &slab) != 7)
            continue;
        int tx = 0, ty = 0, tz = 0, gtx = 0, gty = 0, gtz = 0;
        const int ns = sscanf(st, "%dx%dx%d", &tx, &ty, &tz);
        const int ng = sscanf(gt, "%dx%dx%d", &gtx, &gty, &gtz);
        if (ns < 2 || ng < 2) continue;
        TuneCfg c;
        bool ok = true;
        for (int d = 0; d < 3 && ok; ++d) {
            if (strcmp(ax[d], "-") == 0) continue;
            const char *p = ax[d];
            while (*p && ok) {
                char *end = nullptr;
                const long r = strtol(p, &end, 10);
                if (end == p || c.nrad[d] >= kMaxRadixStages || !plain_radix((int)r)) ok = false;
                else c.radix[d][c.nrad[d]++] = (int)r;
                p = (*end == ',') ? end + 1 : end;
                if (*end && *end != ',') ok = false;
            }
        }
        if (!ok) continue;
        c.tx = tx;
        c.ty = ty;
        c.tz = tz;
        c.gtx = gtx;
        c.gty = gty;
        c.gtz = gtz;
        c.slab = slab;
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            bool found = false;
            for (auto &kv : g_tuned)
                if (kv.first == key) {
                    kv.second = c;
                    found = true;
                }
            if (!found) g_tuned.push_back({key, c});
        }
        ++nread;
    }
    return nread;
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

// ---- structure-factor accumulation ------------------------------------------------------
void Engine::accum_S2(double *acc_d) {
    if (state != 0) throw UsageError("You must call sfac before this routine.");
    const long h = (long)H();
    const int nb = (int)((h + kBlock - 1) / kBlock);
    if (fp32) accum_s2_kernel<float><<<nb, kBlock, 0, stream_>>>((const float2 *)grid_, h, acc_d);
    else accum_s2_kernel<double><<<nb, kBlock, 0, stream_>>>((const double2 *)grid_, h, acc_d);
    HIPC(hipGetLastError());
}

void Engine::accum_cross(const Engine &b, double *acc_d) {
    if (state != 0 || b.state != 0) throw UsageError("You must call sfac before this routine.");
    if (b.K[0] != K[0] || b.K[1] != K[1] || b.K[2] != K[2] || b.fp32 != fp32)
        throw std::invalid_argument("cross terms need the same grid and precision");
    // b's S-hat must be complete before this stream reads it
    hipEvent_t ev;
    HIPC(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    HIPC(hipEventRecord(ev, b.stream_));
    HIPC(hipStreamWaitEvent(stream_, ev, 0));
    HIPC(hipEventDestroy(ev));
    const long h = (long)H();
    const int nb = (int)((h + kBlock - 1) / kBlock);
    if (fp32)
        accum_cross_kernel<float><<<nb, kBlock, 0, stream_>>>((const float2 *)grid_,
                                                             (const float2 *)b.grid_, h, acc_d);
    else
        accum_cross_kernel<double><<<nb, kBlock, 0, stream_>>>(
            (const double2 *)grid_, (const double2 *)b.grid_, h, acc_d);
    HIPC(hipGetLastError());
    // and b's next work (the spread of its next frame overwrites its grid)
    // waits for this read
    HIPC(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    HIPC(hipEventRecord(ev, stream_));
    HIPC(hipStreamWaitEvent(b.stream_, ev, 0));
    HIPC(hipEventDestroy(ev));
}

void Engine::radial(const double *val_d, int nbins, double dk, double *sum_d, double *cnt_d) {
    if (nbins < 1 || !(dk > 0.0)) throw std::invalid_argument("radial: nbins >= 1 and dk > 0");
    const std::vector<double> key = {0.0, L6[0], L6[1], L6[2], L6[3], L6[4], L6[5],
                                     (double)nbins, dk};
    if (key != radial_key_) {
        std::vector<int> bin(H());
        for (int kx = 0; kx < K[0]; ++kx)
            for (int ky = 0; ky < K[1]; ++ky)
                for (int kz = 0; kz < ldim; ++kz) {
                    const double nx = kx <= K[0] / 2 ? kx : kx - K[0];
                    const double ny = ky <= K[1] / 2 ? ky : ky - K[1];
                    const double mx = nx * cell.iL00;
                    const double my = nx * cell.iL10 + ny * cell.iL11;
                    const double mz = nx * cell.iL20 + ny * cell.iL21 + kz * cell.iL22;
                    const double k = 2.0 * M_PI * sqrt(mx * mx + my * my + mz * mz);
                    const double bf = floor(k / dk);
                    bin[((size_t)kx * K[1] + ky) * ldim + kz] = bf < (double)nbins ? (int)bf : -1;
                }
        radial_bins_(key, bin, nbins);
    }
    radial_sum_(val_d, nbins, sum_d, cnt_d);
}

void Engine::radial_plot(const double *val_d, double dk, double clip, int nbins, double *sum_d,
                         double *cnt_d) {
#pragma clang fp contract(off)
    if (nbins < 1 || !(dk > 0.0) || !(clip >= 0.0))
        throw std::invalid_argument("radial_plot: nbins >= 1, dk > 0 and clip >= 0");
    const std::vector<double> key = {1.0, L6[0], L6[1], L6[2], (double)nbins, dk, clip};
    if (key != radial_key_) {
        // plot.py's arithmetic: k_d = i * pi * 2 / L_d, |k| = sqrt(kx^2 + ky^2 + kz^2)
        auto kvec = [](int n, double L) {
            std::vector<double> v(n);
            for (int i = 0; i < n; ++i) {
                const int j = i >= n / 2 ? i - n : i;
                v[i] = (double)j * M_PI * 2.0 / L;
            }
            return v;
        };
        const std::vector<double> kx = kvec(K[0], L6[0]), ky = kvec(K[1], L6[1]),
                                  kz = kvec(ldim, L6[2]);
        std::vector<int> bin(H());
        for (int x = 0; x < K[0]; ++x)
            for (int y = 0; y < K[1]; ++y)
                for (int z = 0; z < ldim; ++z) {
                    const double s2 = kx[x] * kx[x] + ky[y] * ky[y] + kz[z] * kz[z];
                    double d = sqrt(s2) / dk;
                    if (d > clip) d = clip;
                    const int b = (int)d;
                    bin[((size_t)x * K[1] + y) * ldim + z] = b < nbins ? b : -1;
                }
        radial_bins_(key, bin, nbins);
    }
    radial_sum_(val_d, nbins, sum_d, cnt_d);
}

// points grouped by bin (a stable counting sort: increasing h within a bin)
void Engine::radial_bins_(const std::vector<double> &key, const std::vector<int> &bin,
                          int nbins) {
    const size_t h = H();
    std::vector<int> off((size_t)nbins + 1, 0);
    for (size_t p = 0; p < h; ++p)
        if (bin[p] >= 0) off[(size_t)bin[p] + 1] += 1;
    for (int b = 0; b < nbins; ++b) off[(size_t)b + 1] += off[b];
    std::vector<int> perm(std::max<size_t>(off[nbins], 1));
    std::vector<int> pos(off.begin(), off.end() - 1);
    for (size_t p = 0; p < h; ++p)
        if (bin[p] >= 0) perm[pos[bin[p]]++] = (int)p;
    if (!radial_perm_) alloc_((void **)&radial_perm_, h * sizeof(int));
    if ((size_t)nbins + 1 > radial_off_cap_) {
        free_(radial_off_);
        alloc_((void **)&radial_off_, ((size_t)nbins + 1) * sizeof(int));
        radial_off_cap_ = (size_t)nbins + 1;
    }
    HIPC(hipMemcpyAsync(radial_perm_, perm.data(), perm.size() * sizeof(int),
                        hipMemcpyHostToDevice, stream_));
    HIPC(hipMemcpyAsync(radial_off_, off.data(), off.size() * sizeof(int),
                        hipMemcpyHostToDevice, stream_));
    HIPC(hipStreamSynchronize(stream_));
    radial_key_ = key;
}

void Engine::radial_sum_(const double *val_d, int nbins, double *sum_d, double *cnt_d) {
    radial_kernel<<<(nbins + kBlock - 1) / kBlock, kBlock, 0, stream_>>>(
        val_d, radial_perm_, radial_off_, nbins, sum_d, cnt_d);
    HIPC(hipGetLastError());
}

// ---- real-space Ewald ----------------------------------------------------------------------
void Engine::real_space(int n, const double *w_d, const double *x_d, double eta, double rc,
                        double *out_d, double *dx_d, double *phi_d) {
    if (n < 0) throw std::invalid_argument("n < 0");
    if (!(eta > 0.0) || !(rc > 0.0)) throw std::invalid_argument("eta and rc must be > 0");
    RealArgs a;
    a.cell = cell;
    a.L00 = L6[0];
    a.L10 = L6[3];
    a.L11 = L6[1];
    a.L20 = L6[4];
    a.L21 = L6[5];
    a.L22 = L6[2];
    a.eta = eta;
    a.rc2 = rc * rc;
    // |s_d + n_d| <= rc |column d of iL| for an image within rc, |s_d| <= 1/2
    const double c0 = sqrt(cell.iL00 * cell.iL00 + cell.iL10 * cell.iL10 + cell.iL20 * cell.iL20);
    const double c1 = sqrt(cell.iL11 * cell.iL11 + cell.iL21 * cell.iL21);
    const double c2 = fabs(cell.iL22);
    a.e0 = (int)ceil(rc * c0 + 0.5);
    a.e1 = (int)ceil(rc * c1 + 0.5);
    a.e2 = (int)ceil(rc * c2 + 0.5);
    a.n = n;
    const int nb = std::max(1, (n + kBlock - 1) / kBlock);
    if ((size_t)nb * 10 > real_parts_cap_) {
        alloc_((void **)&real_parts_, (size_t)nb * 10 * sizeof(double));
        real_parts_cap_ = (size_t)nb * 10;
    }
    real_space_kernel<<<nb, kBlock, 0, stream_>>>(a, w_d, x_d, dx_d, phi_d, real_parts_);
    HIPC(hipGetLastError());
    real_space_finish<<<1, kBlock, 0, stream_>>>(a, nb, real_parts_, phi_d, out_d);
    HIPC(hipGetLastError());
}

void Engine::real_space_host(int n, const double *w, const double *x, double eta, double rc,
                             double out[10], double *dx, double *phi) {
    if (n < 0) throw std::invalid_argument("n < 0");
    ensure_atoms(n);
    if (!real_out_) alloc_((void **)&real_out_, 16 * sizeof(double));
    HIPC(hipStreamSynchronize(stream_));
    memcpy(pin_, x, 3 * (size_t)n * sizeof(double));
    memcpy(pin_ + 3 * (size_t)n, w, (size_t)n * sizeof(double));
    HIPC(hipMemcpyAsync(x_d_, pin_, 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice,
                        stream_));
    HIPC(hipMemcpyAsync(w_d_, pin_ + 3 * (size_t)n, (size_t)n * sizeof(double),
                        hipMemcpyHostToDevice, stream_));
    real_space(n, w_d_, x_d_, eta, rc, real_out_, dx ? dx_d_ : nullptr, phi ? phi_d_ : nullptr);
    HIPC(hipMemcpyAsync(out, real_out_, 10 * sizeof(double), hipMemcpyDeviceToHost, stream_));
    if (dx && n)
        HIPC(hipMemcpyAsync(dx, dx_d_, 3 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost,
                            stream_));
    if (phi && n)
        HIPC(hipMemcpyAsync(phi, phi_d_, (size_t)n * sizeof(double), hipMemcpyDeviceToHost,
                            stream_));
    HIPC(hipStreamSynchronize(stream_));
}

void Engine::sync() { HIPC(hipStreamSynchronize(stream_)); }

// ---- profiling ---------------------------------------------------------------------------
void Engine::prof_begin_on(const char *name, hipStream_t st) {
    if (!prof_) return;
    ProfRec r;
    r.name = name;
    HIPC(hipEventCreate(&r.a));
    HIPC(hipEventCreate(&r.b));
    HIPC(hipEventRecord(r.a, st));
    prof_recs_.push_back(r);
}

void Engine::prof_end_on(hipStream_t st) {
    if (!prof_ || prof_recs_.empty()) return;
    HIPC(hipEventRecord(prof_recs_.back().b, st));
}

void Engine::prof_begin_(const char *name) { prof_begin_on(name, stream_); }

void Engine::prof_end_() { prof_end_on(stream_); }

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

// ---- batches --------------------------------------------------------------------------------
Batch::Batch(const std::vector<Engine *> &frames) { rebind(frames); }

void Batch::rebind(const std::vector<Engine *> &frames) {
    if (frames.empty()) throw std::invalid_argument("empty batch");
    const Engine *e0 = frames[0];
    for (const Engine *e : frames) {
        if (e->fp32 != e0->fp32 || e->fixed != e0->fixed || e->order != e0->order ||
            e->desc().infl32 != e0->desc().infl32)
            throw std::invalid_argument(
                "a batch needs one precision, accumulation mode, influence precision and "
                "spline order");
        if (!e->has_table()) throw UsageError("No convolution kernel has been set.");
    }
    frames_ = frames;
    const int nf = (int)frames_.size();
    descs_.resize(nf);
    atoms_.resize(nf);
    gens_.resize(nf);
    for (int f = 0; f < nf; ++f) {
        const FrameDesc &D = frames_[f]->desc();
        atoms_[f] = Atoms{D.n, D.w, D.x, D.out, D.dx, D.phi};
        apply_atoms_(f);
    }
    refresh_variant_();
    if (nf > cap_) {
        if (d_descs_) HIPC(hipFree(d_descs_));
        if (d_offs_) HIPC(hipFree(d_offs_));
        d_descs_ = nullptr;
        d_offs_ = nullptr;
        HIPC(hipMalloc((void **)&d_descs_, nf * sizeof(FrameDesc)));
        HIPC(hipMalloc((void **)&d_offs_, (size_t)ST_COUNT * (nf + 1) * sizeof(int)));
        cap_ = nf;
    }
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    graph_exec_ = nullptr;
    graph_ = nullptr;
    dirty_ = true;
    if (lanes_ > 1) build_parts_();
}

void Batch::set_lanes(int k) {
    if (k < 1 || k > 4) throw std::invalid_argument("lanes must be 1 to 4");
    lanes_ = k;
    build_parts_();
}

// groups of consecutive frames, as even in frame count as the division
// allows; the group objects (and their device arrays) are kept across
// rebinds of the same or fewer frames
void Batch::build_parts_() {
    const int nf = (int)frames_.size();
    const int k = std::min(lanes_, nf);
    if (k <= 1) {
        parts_.clear();
        part_begin_.clear();
        return;
    }
    part_begin_.resize(k + 1);
    for (int g = 0; g <= k; ++g) part_begin_[g] = (int)((long)g * nf / k);
    parts_.resize(k);
    for (int g = 0; g < k; ++g) {
        const std::vector<Engine *> fr(frames_.begin() + part_begin_[g],
                                       frames_.begin() + part_begin_[g + 1]);
        if (parts_[g]) parts_[g]->rebind(fr);
        else parts_[g].reset(new Batch(fr));
    }
    for (int f = 0; f < nf; ++f) part_atoms_(f);
    if (!fork_) HIPC(hipEventCreateWithFlags(&fork_, hipEventDisableTiming));
    while ((int)join_.size() < k) {
        hipEvent_t e = nullptr;
        HIPC(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        join_.push_back(e);
    }
}

void Batch::part_atoms_(int f) {
    if (parts_.empty()) return;
    int g = 0;
    while (f >= part_begin_[g + 1]) ++g;
    Batch &p = *parts_[g];
    const int i = f - part_begin_[g];
    p.atoms_[i] = atoms_[f];
    p.apply_atoms_(i);
    p.dirty_ = true;
}

void Batch::synchronize() {
    HIPC(hipStreamSynchronize(default_stream()));
    for (const auto &p : parts_) HIPC(hipStreamSynchronize(p->default_stream()));
}

void Batch::apply_atoms_(int f) {
    Engine *e = frames_[f];
    FrameDesc D = e->desc();
    const Atoms &a = atoms_[f];
    D.n = a.n;
    D.w = a.w;
    D.x = a.x;
    D.out = a.out;
    D.dx = a.dx;
    D.phi = a.phi;
    descs_[f] = D;
    gens_[f] = e->desc_gen();
}

// The launch sequence of a batch: every transform and gather path that one
// of its frames takes (a frame has no workgroups in the launches of the other
// paths), so each frame runs the kernels it runs alone.
void Batch::refresh_variant_() {
    // the x-separable form only if every frame takes it (a frame's results
    // are the same in either kernel)
    xinf_ = kXInfSeparable;
    for (const Engine *e : frames_) {
        const int f = x_influence_form(e->desc());
        if (f != kXInfSeparable) xinf_ = f;
    }
    bool any_slab = false, any_lines = false, any_tile = false, any_sorted = false;
    gmz_ = 1;
    slab_ept_ = 0;
    xpath_ = 0;
    zchunk_ = 0;
    gzchunk_ = 0;
    for (const Engine *e : frames_) {
        zchunk_ = zchunk_ || e->desc().ntz > 1;
        gzchunk_ = gzchunk_ || e->desc().gntz > 1;
        if (e->slab()) {
            any_slab = true;
            slab_ept_ = std::max(slab_ept_, e->slab_kernel());
        } else {
            any_lines = true;
        }
        if (e->gtile()) {
            any_tile = true;
            gmz_ = std::max(gmz_, e->gmz());
        } else {
            any_sorted = true;
        }
        xpath_ = std::max(xpath_, x_path(e->desc()));  // the most general path of the frames
    }
    bool any_tiles = false, any_fused = false, any_yzf = false;
    for (const Engine *e : frames_) {
        any_tiles = any_tiles || !e->spread_slab();
        any_fused = any_fused || e->spread_slab();
        any_yzf = any_yzf || (e->slab() && !e->spread_slab());
    }
    seq_.assign({ST_BIN, ST_SORT});
    if (any_tiles) seq_.push_back(ST_SPREAD);
    if (any_fused) seq_.push_back(ST_SPREAD_YZ);
    if (any_yzf) seq_.push_back(ST_YZ_FWD);
    if (any_lines) seq_.insert(seq_.end(), {ST_ZR2C, ST_YF});
    bool any_xreg = false, any_xfused = false;
    for (const Engine *e : frames_) {
        any_xreg = any_xreg || e->xreg();
        any_xfused = any_xfused || !e->xreg();
    }
    if (any_xfused) seq_.push_back(ST_XFUSED);
    if (any_xreg) seq_.push_back(ST_XREG);
    if (any_slab) seq_.push_back(ST_YZ_INV);
    if (any_lines) seq_.insert(seq_.end(), {ST_YI, ST_ZC2R});
    if (any_tile) seq_.push_back(ST_GATHER_TILE);
    if (any_sorted) seq_.push_back(ST_GATHER_SORTED);
}

Batch::~Batch() {
    parts_.clear();
    if (fork_) (void)hipEventDestroy(fork_);
    for (hipEvent_t e : join_) (void)hipEventDestroy(e);
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    (void)hipFree(d_descs_);
    (void)hipFree(d_offs_);
}

void Batch::set_atoms(int f, int n, const double *w_d, const double *x_d, double *out_d,
                      double *dx_d, double *phi_d) {
    if (f < 0 || f >= (int)frames_.size()) throw std::invalid_argument("frame index");
    Engine *e = frames_[f];
    e->check_staging(n, {w_d, x_d, dx_d, phi_d});
    e->ensure_atoms(n);
    // the engine's descriptor keeps the pointers too (a later rebind of the
    // same engines starts from them)
    FrameDesc &D = e->desc_mut();
    D.n = n;
    D.w = w_d;
    D.x = x_d;
    D.out = out_d;
    D.dx = dx_d;
    D.phi = phi_d;
    atoms_[f] = Atoms{n, w_d, x_d, out_d, dx_d, phi_d};
    apply_atoms_(f);
    dirty_ = true;
    part_atoms_(f);
}

void Batch::upload_(hipStream_t st) {
    const int nf = (int)frames_.size();
    offs_.assign((size_t)ST_COUNT * (nf + 1), 0);
    total_.assign(ST_COUNT, 0);
    lds_.assign(ST_COUNT, 0);
    for (int s = 0; s < ST_COUNT; ++s) {
        int acc = 0;
        for (int f = 0; f < nf; ++f) {
            offs_[(size_t)s * (nf + 1) + f] = acc;
            if (!frames_[f]->uses_stage((Stage)s)) continue;  // another path's launch
            acc += frames_[f]->stage_blocks_n((Stage)s, atoms_[f].n);
            lds_[s] = std::max(lds_[s], frames_[f]->stage_lds((Stage)s, xpath_));
        }
        offs_[(size_t)s * (nf + 1) + nf] = acc;
        total_[s] = acc;
    }
    HIPC(hipMemcpyAsync(d_descs_, descs_.data(), nf * sizeof(FrameDesc), hipMemcpyHostToDevice,
                        st));
    HIPC(hipMemcpyAsync(d_offs_, offs_.data(), offs_.size() * sizeof(int),
                        hipMemcpyHostToDevice, st));
    HIPC(hipStreamSynchronize(st));
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    graph_exec_ = nullptr;
    graph_ = nullptr;
    dirty_ = false;
}

void Batch::evaluate(hipStream_t st) {
    if (!st) st = frames_[0]->stream();  // the legacy null stream cannot be captured
    if (!parts_.empty() && !frames_[0]->profiling()) {
        // group 0 on st, group g on its own stream between a fork and a join
        HIPC(hipEventRecord(fork_, st));
        for (size_t g = 1; g < parts_.size(); ++g) {
            const hipStream_t s = parts_[g]->default_stream();
            if (s != st) HIPC(hipStreamWaitEvent(s, fork_, 0));
        }
        for (size_t g = 0; g < parts_.size(); ++g) {
            parts_[g]->use_graph = use_graph;
            parts_[g]->evaluate(g == 0 ? st : parts_[g]->default_stream());
        }
        for (size_t g = 1; g < parts_.size(); ++g) {
            const hipStream_t s = parts_[g]->default_stream();
            if (s == st) continue;
            HIPC(hipEventRecord(join_[g], s));
            HIPC(hipStreamWaitEvent(st, join_[g], 0));
        }
        return;
    }
    // a frame whose engine changed its box, table, launch shapes or workspace
    // since the snapshot (set_L, set_A, autotuning, growth) is taken anew
    bool changed = false;
    for (int f = 0; f < (int)frames_.size(); ++f)
        if (frames_[f]->desc_gen() != gens_[f]) {
            if (!frames_[f]->has_table()) throw UsageError("No convolution kernel has been set.");
            apply_atoms_(f);
            changed = true;
        }
    if (changed) {
        refresh_variant_();
        dirty_ = true;
    }
    if (dirty_) upload_(st);
    const int nf = (int)frames_.size();
    Engine *e0 = frames_[0];
    auto run_all = [&]() {
        for (const Stage s : seq_) {
            const Launch L{d_descs_, d_offs_ + (size_t)s * (nf + 1), nf, slab_ept_, xpath_, gmz_,
                           zchunk_, gzchunk_, xinf_};
            // the first launch after the fused x pass that holds a frame's
            // workgroups finalizes its energy and virial (the slab or the y
            // inverse, whichever path the frame takes)
            const int finish = (s == ST_YZ_INV || s == ST_YI) ? 1 : 0;
            e0->prof_begin_on(stage_name(s), st);
            e0->launch(s, L, total_[s], lds_[s], st, finish);
            e0->prof_end_on(st);
        }
    };
    if (use_graph && !e0->profiling()) {
        if (!graph_exec_) {
            HIPC(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            try {
                run_all();
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(st, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            HIPC(hipStreamEndCapture(st, &graph_));
            HIPC(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
        }
        HIPC(hipGraphLaunch(graph_exec_, st));
    } else {
        run_all();
    }
    for (Engine *e : frames_) e->state = 1;
}

}  // namespace spme
