// Device engine for one SPME system (one box, one grid, one spline order),
// and batches of such systems launched together.  Host-only interface; the
// kernels live in engine.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <array>
#include <initializer_list>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "common.hpp"

namespace spme {

// Scratch of the allocation-free evaluation: one static device array per GPU.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;

enum EngineFlags {
    kFlagFP32 = 1,        // spreading, FFTs and gather in FP32
    kFlagFixedPoint = 2,  // deterministic int64 fixed-point spreading (always on for FP32)
    kFlagInfl32 = 4,      // FP32 engine: closed-form influence and thread sums in FP32
};

// Kernel stages.  An evaluation runs BIN, SORT, SPREAD, YZ_FWD, XFUSED, YZ_INV,
// GATHER_TILE or GATHER_SORTED (or ZR2C, YF in place of YZ_FWD and YI, ZC2R in
// place of YZ_INV when the y-z slab does not fit in LDS); the split entry
// points use XF_SCALE / INFL / INFL_E / XI.
enum Stage {
    ST_BIN, ST_SORT, ST_SPREAD, ST_ZR2C, ST_YF, ST_XF_SCALE, ST_XFUSED,
    ST_INFL, ST_INFL_E, ST_XI, ST_YI, ST_ZC2R, ST_GATHER, ST_GATHER_SORTED, ST_FINISH,
    ST_YZ_FWD, ST_YZ_INV, ST_GATHER_TILE, ST_XREG, ST_SPREAD_YZ, ST_COUNT
};

// stage sequences of the entry points for the slab / line transforms
struct StageSeq {
    const Stage *s;
    int n;
};
StageSeq eval_stages(bool slab, bool tile);
StageSeq sfac_stages(bool slab);
StageSeq de1_stages(bool slab);
const char *stage_name(Stage s);

// Frames of one launch: frames[0..nf) in device memory, offs = per-frame
// first workgroup (nf > 1 only).
struct Launch {
    const FrameDesc *frames;
    const int *offs;
    int nf;
    int slab_ept;  // elements per thread the slab kernels are compiled for
    int xpath;     // transform path of the x launches (XPath)
    int gmz;       // row chunks of 64 per lane of the tiled gather (ceil((K2 + n - 1) / 64))
    int zchunk;    // some frame spreads through z-chunked tiles
    int gzchunk;   // some frame gathers through z-chunked tiles
    int xinf;      // influence form of the fused x pass the frames share (XInfluence)
};

// Influence form of a fused x launch: kXInfAny (FP64: table, or the closed
// form of any box), kXInfFp32 (the FP32 engine's FP32 influence option),
// kXInfSeparable (FP64 closed form, boxes whose x axis separates).
enum XInfluence { kXInfAny = 0, kXInfFp32 = 1, kXInfSeparable = 2 };
inline int x_influence_form(const FrameDesc &D) {
    if (D.infl32) return kXInfFp32;
    return D.tabmode == kTableAnalytic && x_separable(D.tcell) ? kXInfSeparable : kXInfAny;
}

struct FftPlan {
    int N = 0;
    std::vector<int> radix;
    double2 *tw_d = nullptr;  // exp(-2 pi i t/N), t < N
    float2 *tw_f = nullptr;
};

// Launch shapes of one engine that an autotuner may choose: the radix order
// of each axis's FFT plan, the spreading tile, the gather tile (0 x 0: the
// per-atom gather) and the transform path (y-z slabs or separate lines).
// None of them changes the arithmetic of the deterministic modes beyond the
// FFT's rounding order.
struct TuneCfg {
    int nrad[3] = {0, 0, 0};
    int radix[3][kMaxRadixStages] = {};
    int tx = 0, ty = 0, tz = 0;     // spreading tile; tz = 0: whole columns
    int gtx = 0, gty = 0, gtz = 0;  // gather tile; gtz = 0: whole columns
    int slab = 0;
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
    void build_ewald_table(double eta, double max_m);      // closed-form table (device)
    void copy_table(double *A, double *dA);                // table to host (H, 6H)
    bool has_table() const { return has_table_; }
    int table_mode() const { return desc_.tabmode; }

    // device-pointer operations (asynchronous on stream())
    void spread_forward(int n, const double *w_d, const double *x_d);  // -> state 0
    void energy(double *out_d);                                        // out[0]=E
    void convolve(double *out_d);  // out[0]=E, out[1..6]=virial; -> state 1
    void gather(int n, const double *w_d, const double *x_d, double *dx_d, double *phi_d);
    // one evaluation with the fused reciprocal step; replayed from a captured
    // hipGraph when called again with the same arguments and use_graph
    void evaluate(int n, const double *w_d, const double *x_d, double *out_d, double *dx_d,
                  double *phi_d);

    // The same evaluation with no device allocation: the per-atom workspace
    // in the library's static scratch array of this GPU (kNoAllocScratchBytes,
    // shared by every handle on it, so overlapping calls race), kernels
    // launched one by one (no graph).  std::invalid_argument if n atoms need
    // more scratch than that.
    void evaluate_noalloc(int n, const double *w_d, const double *x_d, double *out_d,
                          double *dx_d, double *phi_d);

    // host-pointer helpers (stage through pinned memory, synchronous results)
    void sfac_host(int n, const double *w, const double *x);
    double en_host();
    double de1_host(double vir[6]);
    void gather_host(int n, const double *w, const double *x, double *dx, double *phi);
    double eval_host(int n, const double *w, const double *x, double vir[6], double *dx,
                     double *phi);
    double *get_S_host();  // host mirror of the grid, K0 x K1 x ldim x 2 doubles

    // Asynchronous host <-> device staging of one evaluation through the
    // engine's own atom buffers, on stream().  Pinned host memory is copied
    // directly; pageable input goes through the pinned staging buffer.
    // download_results writes E and the virial (7 doubles) to out and, when
    // not NULL, dx and phi; the caller synchronises before reading them.
    void upload_atoms(int n, const double *w, const double *x, bool pinned);
    void evaluate_staged(int n, bool want_dx, bool want_phi);
    void download_results(int n, double *out, double *dx, double *phi, hipStream_t st);
    const double *x_dev() const { return x_d_; }
    const double *w_dev() const { return w_d_; }
    double *dx_dev() const { return dx_d_; }
    double *phi_dev() const { return phi_d_; }
    size_t atom_capacity() const { return atom_cap_; }
    // Descriptor generation: bumped whenever the box, the influence table,
    // the launch shapes or the workspace buffers change, so that a batch
    // holding a snapshot of the descriptor knows to take a new one.
    unsigned long long desc_gen() const { return gen_; }
    // std::invalid_argument if one of ptrs is this engine's own atom staging
    // buffer (x_dev, w_dev, dx_dev, phi_dev) and n exceeds its capacity: the
    // growth that n needs would release the buffer the caller points at
    void check_staging(int n, std::initializer_list<const void *> ptrs) const;
    // kernel variants an evaluation of this engine launches (precision,
    // accumulation, order, transform and gather paths, slab/gather template
    // sizes); a batch of engines with one signature runs each frame through
    // exactly the kernels it runs alone
    unsigned long long variant() const;
    // the fused x pass runs the register kernel (FP32, K0 = 128 with the x
    // plan 16, 8, the FP64 influence of an x-separable box)
    bool xreg() const;

    // autotuning: the current launch shapes, applying others (validated
    // against this grid; std::invalid_argument if one does not apply), and
    // the candidates a tuner tries
    TuneCfg tune_cfg() const;
    void apply_tune(const TuneCfg &c);
    std::vector<std::vector<int>> radix_options(int d) const;
    // (tx, ty, tz) candidates; tz = 0: whole columns
    std::vector<std::array<int, 3>> spread_tile_options() const;
    std::vector<std::array<int, 3>> gather_tile_options() const;  // includes {0, 0, 0}
    std::string tune_key() const;  // "K0xK1xK2/order/flags"

    // structure-factor accumulation (device arrays, FP64, caller-owned)
    void accum_S2(double *acc_d);                          // acc[H] += |S|^2
    void accum_cross(const Engine &b, double *acc_d);      // acc[2H] += S conj(S_b)
    // sum_d[b] += sum of val_d over the points of |k| bin b, cnt_d[b] += count;
    // k = 2 pi m, m from the signed principal indices, b = floor(|k| / dk)
    void radial(const double *val_d, int nbins, double dk, double *sum_d, double *cnt_d);
    // the same with plot.py's bins (plot.py:9-32): per axis of the stored
    // K0 x K1 x ldim array, index i -> i - n for i >= n/2 (n the axis length,
    // ldim for z), k_d = 2 pi i / L_d with the box's axis lengths, and bin
    // b = (int) min(|k| / dk, clip); points with b >= nbins are left out
    void radial_plot(const double *val_d, double dk, double clip, int nbins, double *sum_d,
                     double *cnt_d);

    // real-space Ewald sum over the images within rc plus the self and
    // background terms, in the current box: out[10] = E_pair, vir(6), E_self,
    // E_bg, vir_bg; dx = dE_pair/dx; phi = dE/dw of the three terms
    void real_space(int n, const double *w_d, const double *x_d, double eta, double rc,
                    double *out_d, double *dx_d, double *phi_d);
    void real_space_host(int n, const double *w, const double *x, double eta, double rc,
                         double out[10], double *dx, double *phi);

    void sync();
    hipStream_t stream() const { return stream_; }
    double *result_dev() const { return result_; }
    bool use_graph = true;
    // y-z slab transforms (one LDS-resident plane per workgroup) when the
    // plane fits; use_slab = false selects the separate z and y line passes
    bool slab_ok = false;
    bool use_slab = false;
    bool slab() const { return slab_ok && use_slab; }
    // spreading fused with the forward slab transform (one workgroup per x
    // plane, the plane's sums in LDS) where that applies; use_spread_slab =
    // false keeps the separate spreading tiles before the slab transform
    bool spread_slab_ok = false;
    bool slab128_ok_ = false;  // the register-line slab kernels fit this device's LDS
    bool use_spread_slab = true;
    bool spread_slab() const { return slab() && spread_slab_ok && use_spread_slab; }
    // 0: slab transforms (spreading fused where it applies), 1: line passes,
    // 2: slab transforms after spreading tiles; std::invalid_argument otherwise
    void set_transform(int mode);
    int slab_ept() const;
    // the slab kernels a launch runs: kSlabKernel128 for the register-line
    // variant of 128 x 128 FP32 planes with the plans (16, 8), else slab_ept()
    // (a batch runs the largest code of its frames: the generic kernels
    // take every plane)
    int slab_kernel() const;
    // the evaluation gathers through LDS-staged tiles when one fits (FrameDesc::gtx)
    bool gtile() const { return desc_.gtx > 0; }
    int gmz() const {
        return ((desc_.gntz > 1 ? desc_.gtz : K[2]) + order - 1 + 63) / 64;
    }

    // Per-kernel timing with HIP events recorded on stream() around every
    // launch while enabled (graph replay is off meanwhile).
    void profile_enable(bool on);
    void profile_read(std::vector<std::string> &names, std::vector<double> &ms,
                      std::vector<int> &count);
    bool profiling() const { return prof_; }
    // bracket a launch on stream st (a batch records into its first frame)
    void prof_begin_on(const char *name, hipStream_t st);
    void prof_end_on(hipStream_t st);

    // descriptor and launch geometry (used by Batch)
    const FrameDesc &desc() const { return desc_; }
    FrameDesc &desc_mut() { return desc_; }
    int stage_blocks(Stage s) const;
    int stage_blocks_n(Stage s, int natoms) const;  // for natoms atoms
    // whether an evaluation of this engine launches stage s (the transform
    // and gather stages belong to one path each)
    bool uses_stage(Stage s) const;
    // per-atom sort workspace: bytes for n atoms, and its parts in D at base
    size_t workspace_bytes(size_t n) const;
    int sort_cap(size_t n) const;  // records of a plane the sort stages in LDS, n atoms
    void bind_workspace(FrameDesc &D, void *base, size_t n) const;
    // xpath: the x path of the launch (kXBlock, or kXGeneric for a batch with
    // a generic frame); the frame's own path when that is more general
    size_t stage_lds(Stage s, int xpath = 0) const;
    void ensure_atoms(int n);
    void launch(Stage s, const Launch &L, int blocks, size_t lds,
                hipStream_t st, int finish = 0);

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

    struct ProfScope {
        Engine *e;
        ProfScope(Engine *e_, const char *name) : e(e_) { e->prof_begin_(name); }
        ~ProfScope() { e->prof_end_(); }
    };

private:
    void alloc_(void **p, size_t bytes);
    void free_(void *p);  // release one alloc_ buffer (stream synchronised)
    void run_(Stage s, int finish);  // single-frame launch of one stage on stream_
    void run_seq_(StageSeq seq);
    unsigned long long graph_variant_ = 0;  // kernels the captured graph was recorded with
    void drop_graph_();
    void set_atoms_(int n, const double *w, const double *x, double *dx, double *phi,
                    double *out);

    void sync_desc_();  // upload desc_ to d_desc_ when it changed

    hipStream_t stream_ = nullptr;
    FrameDesc desc_;
    FrameDesc desc_uploaded_;
    FrameDesc *d_desc_ = nullptr;
    FrameDesc desc_na_;                // descriptor of the allocation-free evaluation
    FrameDesc *d_desc_na_ = nullptr;
    int dev_ = 0;                      // device the engine was made on
    void *grid_ = nullptr;
    double *table_ = nullptr;  // A[H] then dA[6H]
    double *table_stats_ = nullptr;  // per-workgroup table-vs-closed-form statistics
    size_t table_stats_cap_ = 0;
    bool has_table_ = false;
    double *iB_d_[3] = {nullptr, nullptr, nullptr};
    double *result_ = nullptr;  // 8 doubles
    int nred_infl_ = 1;
    double *x_d_ = nullptr, *w_d_ = nullptr, *dx_d_ = nullptr, *phi_d_ = nullptr;
    void *ws_ = nullptr;  // per-atom sort workspace (bind_workspace)
    size_t atom_cap_ = 0;
    size_t work_cap_ = 0;
    double *pin_ = nullptr;  // pinned staging
    size_t pin_cap_ = 0;
    std::vector<double> host_grid_;
    // |k| bins of the half-complex points for radial(): key, points by bin, offsets
    std::vector<double> radial_key_;
    void radial_bins_(const std::vector<double> &key, const std::vector<int> &bin, int nbins);
    void radial_sum_(const double *val_d, int nbins, double *sum_d, double *cnt_d);
    double *real_out_ = nullptr;    // real-space results of the host-pointer form
    double *real_parts_ = nullptr;  // real-space workgroup partials
    size_t real_parts_cap_ = 0;
    int *radial_perm_ = nullptr;
    int *radial_off_ = nullptr;
    size_t radial_off_cap_ = 0;
    FftPlan plan_[3];
    SplineArgs<float> sp_f_;
    SplineArgs<double> sp_d_;
    size_t dev_bytes_ = 0;
    std::vector<std::pair<void *, size_t>> allocs_;
    unsigned long long gen_ = 1;

    // captured evaluation
    hipGraphExec_t graph_exec_ = nullptr;
    hipGraph_t graph_ = nullptr;
    const void *graph_key_[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    int graph_n_ = -1;

    struct ProfRec {
        const char *name;
        hipEvent_t a, b;
    };
    void prof_begin_(const char *name);
    void prof_end_();
    bool prof_ = false;
    std::vector<ProfRec> prof_recs_;
};

// A set of engines with the same precision, accumulation mode and spline
// order, evaluated together: one launch per stage covers every frame.
class Batch {
public:
    explicit Batch(const std::vector<Engine *> &frames);
    ~Batch();
    // replace the frames, keeping the device descriptor arrays when they are
    // large enough (no allocation or free for a batch of at most the largest
    // earlier size)
    void rebind(const std::vector<Engine *> &frames);
    // per-frame device pointers for the next evaluation
    void set_atoms(int f, int n, const double *w_d, const double *x_d, double *out_d,
                   double *dx_d, double *phi_d);
    // all frames, one launch per stage; st == nullptr: frame 0's stream
    void evaluate(hipStream_t st);
    // the frames as k groups of consecutive frames (k = 1..4, at most one
    // group per frame), each replayed on its own stream: the first group on
    // the evaluation's stream, group g on the stream of its first frame's
    // engine, forked from the evaluation's stream and joined back into it by
    // events, so that the kernels of one group fill the CUs that the other
    // leaves idle.  Every frame runs the same kernels as in one group; k = 1:
    // one launch per stage for every frame.  A profiling frame 0 evaluates
    // the frames as one group (its timings cover every frame).
    void set_lanes(int k);
    int lanes() const { return lanes_; }
    // waits for the default stream and the stream of every group
    void synchronize();
    int size() const { return (int)frames_.size(); }
    const std::vector<Engine *> &frames() const { return frames_; }
    hipStream_t default_stream() const { return frames_[0]->stream(); }
    bool use_graph = true;

private:
    // per-frame atom pointers given by set_atoms (kept apart from the
    // engines' descriptors, which an engine's own calls overwrite)
    struct Atoms {
        int n;
        const double *w, *x;
        double *out, *dx, *phi;
    };
    void upload_(hipStream_t st);
    void refresh_variant_();  // launch variant flags from the frames
    void apply_atoms_(int f);  // descs_[f] = engine descriptor + atoms_[f]
    std::vector<Engine *> frames_;
    std::vector<FrameDesc> descs_;
    std::vector<Atoms> atoms_;
    std::vector<unsigned long long> gens_;  // engine descriptor generation of descs_
    FrameDesc *d_descs_ = nullptr;
    int *d_offs_ = nullptr;                 // ST_COUNT x (nf + 1)
    int cap_ = 0;                           // frames the device arrays hold
    std::vector<int> offs_;                 // host copy
    std::vector<int> total_;                // blocks per stage
    std::vector<size_t> lds_;               // max LDS per stage
    bool dirty_ = true;
    std::vector<Stage> seq_;                // launches of an evaluation (every frame's paths)
    int gmz_ = 1;                           // largest row chunk count of the tiled frames
    int slab_ept_ = 0;                      // slab kernel variant (largest slab frame)
    int xpath_ = 0;                         // x transform path valid for every frame
    int zchunk_ = 0;                        // some frame spreads through z-chunked tiles
    int gzchunk_ = 0;                       // some frame gathers through z-chunked tiles
    int xinf_ = 0;                          // influence form of the fused x launches
    hipGraphExec_t graph_exec_ = nullptr;
    hipGraph_t graph_ = nullptr;
    // groups of frames on their own streams (lanes_ > 1)
    void build_parts_();
    void part_atoms_(int f);  // the group holding frame f takes atoms_[f]
    int lanes_ = 1;
    std::vector<std::unique_ptr<Batch>> parts_;
    std::vector<int> part_begin_;  // first frame of each group, then nf
    hipEvent_t fork_ = nullptr;
    std::vector<hipEvent_t> join_;
};

// Process-wide record of tuned launch shapes, keyed by grid, order and
// precision flags: engines made later with the same key start from it.
void tune_remember(const Engine &e, const TuneCfg &c);
bool tune_recall(const Engine &e, TuneCfg &c);
std::string tune_export();              // one line per key
std::string tune_line(const Engine &e);  // the engine's current shapes, one line
int tune_import(const std::string &s);  // entries read (invalid lines skipped)
void tune_clear();

// Times candidate launch shapes of the frames of `b` (evaluations of the
// whole batch on `st`, median of `reps` after warm-up), one knob at a time
// for each distinct (grid, order, flags) class of frames, keeps the fastest
// (a change must beat the current shape by 1 %), applies it to every frame
// of the class and remembers it.  Returns a report (one line per class and
// knob tried).
std::string autotune(const std::vector<Engine *> &frames, Batch &b, hipStream_t st, int reps);

// thrown for API misuse; carries the reference's message
struct UsageError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void hip_check(hipError_t e, const char *what);

}  // namespace spme
