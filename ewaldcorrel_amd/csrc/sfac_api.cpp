// libsfac.so C ABI (sfac.hpp:194-211 of the reference) and the SFac facade.
//
// Every entry point catches C++ exceptions: nothing is thrown across the
// ABI.  Misuse that the reference reports with printf (sfac.cpp:310-317,
// 348-355, 414-417, 502-505) prints the same message and returns 0.0 / void;
// the message is also kept for sfac_last_error().  A failed constructor
// returns NULL (the reference throws, which aborts a ctypes caller).
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

#include "ball.hpp"
#include "engine.hpp"
#include "server.hpp"
#include "sfac.hpp"

namespace {
thread_local std::string g_last_error;

void set_error(const std::string &m) { g_last_error = m; }

SFac *as_sfac(void *h) {
    if (!h) throw std::invalid_argument("NULL sfac handle");
    return reinterpret_cast<SFac *>(h);
}

template <typename F>
int guard(F &&f) {
    try {
        g_last_error.clear();
        f();
        return 0;
    } catch (const spme::UsageError &e) {
        printf("Error: %s\n", e.what());
        fflush(stdout);
        set_error(e.what());
        return SFAC_ERR_USAGE;
    } catch (const std::invalid_argument &e) {
        set_error(e.what());
        return SFAC_ERR_ARG;
    } catch (const std::exception &e) {
        set_error(e.what());
        return SFAC_ERR_DEVICE;
    }
}
}  // namespace

// ---- SFac facade ---------------------------------------------------------------
SFac::SFac(double L[6], int32_t K_[3], int order_, int flags)
    : order(order_), cell(L), state(-1), bspc(order_ > 0 ? order_ : 2), fft_neg(nullptr),
      fft_pos(nullptr), A(nullptr), dA(nullptr), Q(nullptr), engine(nullptr) {
    if (order % 2 != 0) {
        printf("Bad order: %d (must be even)\n", order);
        fflush(stdout);
        throw std::invalid_argument("invalid order");
    }
    if (K_[0] < 1 || K_[1] < 1 || K_[2] < 1) {
        printf("Bad K: %d %d %d\n", K_[0], K_[1], K_[2]);
        fflush(stdout);
        throw std::invalid_argument("invalid K");
    }
    for (int d = 0; d < 3; ++d) K[d] = K_[d];
    ldim = K[2] / 2 + 1;
    iK = 1.0 / ((double)K[0] * K[1] * K[2]);
    int Ki[3] = {K[0], K[1], K[2]};
    engine = new spme::Engine(L, Ki, order, flags);
    for (int d = 0; d < 3; ++d) iB.push_back(engine->iB[d]);
}

SFac::SFac(double L[6], int32_t K_[3], int order_) : SFac(L, K_, order_, 0) {}

SFac::~SFac() {
    delete engine;
    free(A);
}

void SFac::accum_spline(int i1, int j1, int k1, const std::multimap<int, Vec4> &srt) {
    if (!Q) Q = engine->get_S_host();
    const long NN = ((long)i1 * K[1] + j1) * ldim * 2 + k1;
    const int i0 = i1 + K[0] - order / 2, j0 = j1 + K[1] - order / 2, k0 = k1 + K[2] - order / 2;
    double ans = 0.0;
    for (int i = 0; i < order; ++i) {
        const int xid = ((i0 + i) % K[0]) * K[1];
        for (int j = 0; j < order; ++j) {
            const int yid = (xid + (j0 + j) % K[1]) * K[2];
            for (int k = 0; k < order; ++k) {
                auto range = srt.equal_range(yid + (k0 + k) % K[2]);
                for (auto it = range.first; it != range.second; ++it) {
                    const Vec4 &u = it->second;
                    ans += bspc.bspl_coef(u[0], i) * bspc.bspl_coef(u[1], j) *
                           bspc.bspl_coef(u[2], k) * u.w;
                }
            }
        }
    }
    Q[NN] += ans;
}

std::multimap<int, Vec4> SFac::sort(int n, const double *w, const double *x) const {
    std::multimap<int, Vec4> srt;
    for (int a = 0; a < n; ++a) {
        Vec4 u = cell.scale(x + 3 * a);
        int c[3];
        for (int d = 0; d < 3; ++d) {
            u[d] *= K[d];
            c[d] = (int)u[d];
            u[d] -= c[d];
            if (c[d] >= K[d]) c[d] -= K[d];  // s == 1.0 -> periodic image
        }
        u[3] = w[a];
        srt.insert(std::make_pair((c[0] * K[1] + c[1]) * K[2] + c[2], u));
    }
    return srt;
}

void SFac::operator()(int n, const double *w, const double *x) {
    engine->sfac_host(n, w, x);
    state = engine->state;
}

// ---- the reference C ABI -------------------------------------------------------------
extern "C" {

void *sfac_ctor_ex(const double L[6], const int K[3], int order, int flags) {
    SFac *s = nullptr;
    int rc = guard([&] {
        double L6[6];
        int32_t Ki[3] = {K[0], K[1], K[2]};
        for (int i = 0; i < 6; ++i) L6[i] = L[i];
        s = new SFac(L6, Ki, order, flags);
    });
    return rc == 0 ? reinterpret_cast<void *>(s) : nullptr;
}

void *sfac_ctor(double L[6], int K[3], int order) { return sfac_ctor_ex(L, K, order, 0); }

void sfac_dtor(void *h) {
    guard([&] { delete as_sfac(h); });
}

void sfac(void *h, int n, const double *w, const double *x) {
    guard([&] { (*as_sfac(h))(n, w, x); });
}

double *get_S(void *h) {
    double *p = nullptr;
    guard([&] {
        SFac *s = as_sfac(h);
        p = s->engine->get_S_host();
        s->Q = p;
    });
    return p;
}

void set_L(void *h, double L[6]) {
    guard([&] {
        SFac *s = as_sfac(h);
        s->cell = Cell(L);
        s->engine->set_cell(L);
    });
}

// The reference enumerates the ball and calls f once per lattice point
// (sfac.cpp:213-292); the table is assembled on the host in the same order
// and uploaded once.
int set_A(void *h, rad_fn f, double max_m, void *info) {
    int status = 0;
    int rc = guard([&] {
        SFac *s = as_sfac(h);
        const size_t H = (size_t)s->K[0] * s->K[1] * s->ldim;
        if (!s->A) {
            s->A = (double *)malloc(sizeof(double) * H * 7);
            if (!s->A) {
                status = 2;
                return;
            }
            s->dA = s->A + H;
        }
        memset(s->A, 0, sizeof(double) * H * 7);
        double *A = s->A, *dA = s->dA;
        spme::enumerate_ball(s->cell.iL, s->K, s->ldim, max_m,
                             [&](long idx, double m2, double mx, double my, double mz) {
                                 double d = 0.0;
                                 A[idx] += 0.5 * f(m2, &d, info);
                                 double *p = dA + 6 * idx;
                                 p[0] += d * mx * mx;
                                 p[1] += d * my * my;
                                 p[2] += d * mz * mz;
                                 p[3] += d * my * mx;
                                 p[4] += d * mz * mx;
                                 p[5] += d * mz * my;
                             });
        s->engine->upload_table(A, dA);
    });
    if (rc != 0) return rc == SFAC_ERR_DEVICE ? 2 : rc;
    return status;
}

double en(void *h) {
    double e = 0.0;
    guard([&] { e = as_sfac(h)->engine->en_host(); });
    return e;
}

double de1(void *h, double vir[6]) {
    double e = 0.0;
    guard([&] {
        SFac *s = as_sfac(h);
        e = s->engine->de1_host(vir);
        s->state = s->engine->state;
    });
    return e;
}

void de2(void *h, int n, const double *w, const double *x, double *dx) {
    guard([&] { as_sfac(h)->engine->gather_host(n, w, x, dx, nullptr); });
}

void potl(void *h, int n, const double *x, double *phi) {
    guard([&] { as_sfac(h)->engine->gather_host(n, nullptr, x, nullptr, phi); });
}

// ---- extensions (sfac_ext.h) -------------------------------------------------------
const char *sfac_last_error(void) { return g_last_error.c_str(); }

int sfac_set_A_ewald(void *h, double eta, double max_m) {
    return guard([&] {
        if (!(eta > 0.0)) throw std::invalid_argument("eta must be > 0");
        if (!(max_m >= 0.0)) throw std::invalid_argument("max_m must be >= 0");
        SFac *s = as_sfac(h);
        s->engine->build_ewald_table(eta, max_m);
        // the table now lives on the device; a host copy from an earlier
        // callback set_A would be stale (sfac_table() copies the current one)
        free(s->A);
        s->A = nullptr;
        s->dA = nullptr;
    });
}

int sfac_table(void *h, double *A, double *dA) {
    return guard([&] { as_sfac(h)->engine->copy_table(A, dA); });
}

int sfac_accum_S2(void *h, double *acc_d) {
    return guard([&] { as_sfac(h)->engine->accum_S2(acc_d); });
}

int sfac_accum_cross(void *ha, void *hb, double *acc_d) {
    return guard([&] { as_sfac(ha)->engine->accum_cross(*as_sfac(hb)->engine, acc_d); });
}

int sfac_radial_plot(void *h, const double *val_d, double dk, double clip, int nbins,
                     double *sum_d, double *cnt_d) {
    return guard(
        [&] { as_sfac(h)->engine->radial_plot(val_d, dk, clip, nbins, sum_d, cnt_d); });
}

int sfac_radial(void *h, const double *val_d, int nbins, double dk, double *sum_d,
                double *cnt_d) {
    return guard([&] { as_sfac(h)->engine->radial(val_d, nbins, dk, sum_d, cnt_d); });
}

int sfac_real_space(void *h, int n, const double *w, const double *x, double eta, double rc,
                    double out[10], double *dx, double *phi) {
    return guard([&] { as_sfac(h)->engine->real_space_host(n, w, x, eta, rc, out, dx, phi); });
}

int sfac_real_space_dev(void *h, int n, const double *w_d, const double *x_d, double eta,
                        double rc, double *out_d, double *dx_d, double *phi_d) {
    return guard(
        [&] { as_sfac(h)->engine->real_space(n, w_d, x_d, eta, rc, out_d, dx_d, phi_d); });
}

int sfac_dev(void *h, int n, const double *w_d, const double *x_d) {
    return guard([&] {
        if (n < 0) throw std::invalid_argument("n < 0");
        SFac *s = as_sfac(h);
        s->engine->spread_forward(n, w_d, x_d);
        s->state = s->engine->state;
    });
}

int en_dev(void *h, double *out_d) {
    return guard([&] { as_sfac(h)->engine->energy(out_d); });
}

int de1_dev(void *h, double *out_d) {
    return guard([&] {
        SFac *s = as_sfac(h);
        s->engine->convolve(out_d);
        s->state = s->engine->state;
    });
}

int de2_dev(void *h, int n, const double *w_d, const double *x_d, double *dx_d) {
    return guard([&] { as_sfac(h)->engine->gather(n, w_d, x_d, dx_d, nullptr); });
}

int potl_dev(void *h, int n, const double *x_d, double *phi_d) {
    return guard([&] { as_sfac(h)->engine->gather(n, nullptr, x_d, nullptr, phi_d); });
}

int sfac_eval_dev(void *h, int n, const double *w_d, const double *x_d, double *out_d,
                  double *dx_d, double *phi_d) {
    return guard([&] {
        if (n < 0) throw std::invalid_argument("n < 0");
        SFac *s = as_sfac(h);
        spme::Engine *e = s->engine;
        e->evaluate(n, w_d, x_d, out_d ? out_d : e->result_dev(), dx_d, phi_d);
        s->state = e->state;
    });
}

int sfac_eval_dev_noalloc(void *h, int n, const double *w_d, const double *x_d, double *out_d,
                          double *dx_d, double *phi_d) {
    return guard([&] {
        if (n < 0) throw std::invalid_argument("n < 0");
        SFac *s = as_sfac(h);
        spme::Engine *e = s->engine;
        e->evaluate_noalloc(n, w_d, x_d, out_d ? out_d : e->result_dev(), dx_d, phi_d);
        s->state = e->state;
    });
}

double sfac_eval(void *h, int n, const double *w, const double *x, double vir[6], double *dx,
                 double *phi) {
    double E = 0.0;
    guard([&] {
        SFac *s = as_sfac(h);
        E = s->engine->eval_host(n, w, x, vir, dx, phi);
        s->state = s->engine->state;
    });
    return E;
}

int sfac_set_graph(void *h, int enable) {
    return guard([&] { as_sfac(h)->engine->use_graph = enable != 0; });
}

void *sfac_stream(void *h) {
    void *st = nullptr;
    guard([&] { st = reinterpret_cast<void *>(as_sfac(h)->engine->stream()); });
    return st;
}

int sfac_sync(void *h) {
    return guard([&] { as_sfac(h)->engine->sync(); });
}

int sfac_set_transform(void *h, int mode) {
    return guard([&] {
        as_sfac(h)->engine->set_transform(mode);
    });
}

int sfac_info(void *h, long long out[12]) {
    return guard([&] {
        spme::Engine *e = as_sfac(h)->engine;
        out[0] = e->K[0];
        out[1] = e->K[1];
        out[2] = e->K[2];
        out[3] = e->ldim;
        out[4] = e->order;
        out[5] = (e->fp32 ? SFAC_FP32 : 0) | (e->fixed ? SFAC_FIXED : 0) |
                 (e->desc().infl32 ? SFAC_FP32_INFLUENCE : 0);
        out[6] = (long long)e->device_bytes();
        out[7] = e->state;
        out[8] = e->table_mode();
        out[9] = e->use_graph ? 1 : 0;
        out[10] = !e->slab() ? 0 : e->spread_slab() ? 2 : 1;
        out[11] = e->slab_ok ? 1 : 0;
    });
}

int sfac_profile(void *h, int enable) {
    return guard([&] { as_sfac(h)->engine->profile_enable(enable != 0); });
}

int sfac_profile_read(void *h, char *names, int names_len, double *ms, int *counts, int max) {
    int n = 0;
    int rc = guard([&] {
        std::vector<std::string> nm;
        std::vector<double> t;
        std::vector<int> c;
        as_sfac(h)->engine->profile_read(nm, t, c);
        std::string joined;
        for (size_t k = 0; k < nm.size() && (int)k < max; ++k) {
            if (k) joined += "\n";
            joined += nm[k];
            ms[k] = t[k];
            counts[k] = c[k];
            ++n;
        }
        if (names_len > 0) {
            strncpy(names, joined.c_str(), (size_t)names_len - 1);
            names[names_len - 1] = '\0';
        }
    });
    return rc == 0 ? n : rc;
}

// ---- batches ------------------------------------------------------------------------------
void *sfac_batch_create(void **handles, int nframes) {
    spme::Batch *b = nullptr;
    int rc = guard([&] {
        if (nframes < 1) throw std::invalid_argument("nframes < 1");
        std::vector<spme::Engine *> fr;
        for (int f = 0; f < nframes; ++f) fr.push_back(as_sfac(handles[f])->engine);
        b = new spme::Batch(fr);
    });
    return rc == 0 ? reinterpret_cast<void *>(b) : nullptr;
}

int sfac_batch_set(void *b, int f, int n, const double *w_d, const double *x_d, double *out_d,
                   double *dx_d, double *phi_d) {
    return guard([&] {
        if (!b) throw std::invalid_argument("NULL batch");
        if (n < 0) throw std::invalid_argument("n < 0");
        if (!out_d) throw std::invalid_argument("out_d is required");
        reinterpret_cast<spme::Batch *>(b)->set_atoms(f, n, w_d, x_d, out_d, dx_d, phi_d);
    });
}

int sfac_batch_eval(void *b, void *stream) {
    return guard([&] {
        if (!b) throw std::invalid_argument("NULL batch");
        reinterpret_cast<spme::Batch *>(b)->evaluate(reinterpret_cast<hipStream_t>(stream));
    });
}

int sfac_batch_sync(void *b) {
    return guard([&] {
        if (!b) throw std::invalid_argument("NULL batch");
        reinterpret_cast<spme::Batch *>(b)->synchronize();
    });
}

int sfac_batch_set_lanes(void *b, int lanes) {
    return guard([&] {
        if (!b) throw std::invalid_argument("NULL batch");
        reinterpret_cast<spme::Batch *>(b)->set_lanes(lanes);
    });
}

int sfac_batch_lanes(void *b) {
    int k = 0;
    const int rc = guard([&] {
        if (!b) throw std::invalid_argument("NULL batch");
        k = reinterpret_cast<spme::Batch *>(b)->lanes();
    });
    return rc == 0 ? k : rc;
}

int sfac_batch_set_graph(void *b, int enable) {
    return guard([&] {
        if (!b) throw std::invalid_argument("NULL batch");
        reinterpret_cast<spme::Batch *>(b)->use_graph = enable != 0;
    });
}

int sfac_batch_destroy(void *b) {
    return guard([&] { delete reinterpret_cast<spme::Batch *>(b); });
}

// ---- serving loop --------------------------------------------------------------------------
void *sfac_server_create(int slots, double hbm_budget, int max_batch, int batch_atoms) {
    spme::Server *s = nullptr;
    int rc = guard([&] { s = new spme::Server(slots, hbm_budget, max_batch, batch_atoms); });
    return rc == 0 ? reinterpret_cast<void *>(s) : nullptr;
}

long long sfac_server_submit(void *s, const double L6[6], const int K[3], int order, int flags,
                             double eta, double max_m, int n, const double *w, const double *x,
                             double *out, double *dx, double *phi) {
    long long id = 0;
    int rc = guard([&] {
        if (!s) throw std::invalid_argument("NULL server");
        if (!L6 || !K) throw std::invalid_argument("L6 and K are required");
        id = reinterpret_cast<spme::Server *>(s)->submit(L6, K, order, flags, eta, max_m, n, w,
                                                         x, out, dx, phi);
    });
    return rc == 0 ? id : (long long)rc;
}

int sfac_server_wait(void *s, long long id, double timing[3]) {
    int st = 0;
    int rc = guard([&] {
        if (!s) throw std::invalid_argument("NULL server");
        std::string err;
        st = reinterpret_cast<spme::Server *>(s)->wait(id, timing, &err);
        if (st != 0) set_error(err);
    });
    return rc != 0 ? rc : st;
}

int sfac_server_poll(void *s, long long id) {
    int st = 0;
    int rc = guard([&] {
        if (!s) throw std::invalid_argument("NULL server");
        st = reinterpret_cast<spme::Server *>(s)->poll(id);
    });
    return rc != 0 ? rc : st;
}

int sfac_server_stats(void *s, double out[16]) {
    return guard([&] {
        if (!s || !out) throw std::invalid_argument("NULL argument");
        reinterpret_cast<spme::Server *>(s)->stats(out);
    });
}

int sfac_server_reset_stats(void *s) {
    return guard([&] {
        if (!s) throw std::invalid_argument("NULL server");
        reinterpret_cast<spme::Server *>(s)->reset_stats();
    });
}

int sfac_server_destroy(void *s) {
    return guard([&] { delete reinterpret_cast<spme::Server *>(s); });
}

// ---- autotuner -----------------------------------------------------------------------------
namespace {
int put_text(const std::string &t, char *buf, int len) {
    if (buf && len > 0) {
        const size_t k = std::min<size_t>(t.size(), (size_t)len - 1);
        memcpy(buf, t.data(), k);
        buf[k] = 0;
    }
    return (int)t.size() + 1;
}
}  // namespace

int sfac_autotune(void *h, int n, const double *w_d, const double *x_d, int reps, char *report,
                  int report_len) {
    std::string rep;
    int rc = guard([&] {
        SFac *s = as_sfac(h);
        spme::Engine *e = s->engine;
        if (!e->has_table()) throw spme::UsageError("No convolution kernel has been set.");
        if (n < 1) throw std::invalid_argument("n < 1");
        if ((w_d == nullptr) != (x_d == nullptr))
            throw std::invalid_argument("give both w_d and x_d, or neither");
        if (!w_d) {
            // synthetic atoms: uniform in the box, alternating unit charges
            std::vector<double> w(n), x(3 * (size_t)n);
            unsigned long long r = 0x9e3779b97f4a7c15ull;
            auto uni = [&r]() {
                r ^= r << 13;
                r ^= r >> 7;
                r ^= r << 17;
                return (double)(r >> 11) * (1.0 / 9007199254740992.0);
            };
            const double *L = e->L6;
            for (int a = 0; a < n; ++a) {
                const double s0 = uni(), s1 = uni(), s2 = uni();
                x[3 * a + 0] = s0 * L[0] + s1 * L[3] + s2 * L[4];
                x[3 * a + 1] = s1 * L[1] + s2 * L[5];
                x[3 * a + 2] = s2 * L[2];
                w[a] = (a & 1) ? -1.0 : 1.0;
            }
            e->upload_atoms(n, w.data(), x.data(), false);
            e->sync();
            w_d = e->w_dev();
            x_d = e->x_dev();
        }
        // the outputs go to the engine's staging buffers: size them for n first
        // (growing replaces the buffers, so take their addresses afterwards)
        e->ensure_atoms(n);
        spme::Batch b(std::vector<spme::Engine *>{e});
        b.set_atoms(0, n, w_d, x_d, e->result_dev(), e->dx_dev(), e->phi_dev());
        rep = spme::autotune(b.frames(), b, e->stream(), reps);
        e->sync();
        e->state = -1;
    });
    if (rc != 0) return rc;
    put_text(rep, report, report_len);
    return 0;
}

int sfac_batch_autotune(void *bp, int reps, char *report, int report_len) {
    std::string rep;
    int rc = guard([&] {
        if (!bp) throw std::invalid_argument("NULL batch");
        spme::Batch *b = reinterpret_cast<spme::Batch *>(bp);
        rep = spme::autotune(b->frames(), *b, b->default_stream(), reps);
        b->synchronize();
    });
    if (rc != 0) return rc;
    put_text(rep, report, report_len);
    return 0;
}

int sfac_tune_export(char *buf, int len) {
    int need = 0;
    int rc = guard([&] { need = put_text(spme::tune_export(), buf, len); });
    return rc != 0 ? rc : need;
}

int sfac_tune_import(const char *text) {
    int k = 0;
    int rc = guard([&] {
        if (!text) throw std::invalid_argument("NULL text");
        k = spme::tune_import(text);
    });
    return rc != 0 ? rc : k;
}

int sfac_tune_current(void *h, char *buf, int len) {
    int need = 0;
    int rc = guard([&] { need = put_text(spme::tune_line(*as_sfac(h)->engine), buf, len); });
    return rc != 0 ? rc : need;
}

int sfac_tune_clear(void) {
    return guard([&] { spme::tune_clear(); });
}

}  // extern "C"
