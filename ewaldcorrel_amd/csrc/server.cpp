// Serving loop (continuous batching over resident engines); see server.hpp.
#include "server.hpp"

#include <string.h>

#include <algorithm>
#include <stdexcept>

namespace spme {

namespace {

// Device bytes an engine for (K, order, flags) holding n atoms allocates:
// grid, influence table (A and dA) with its check statistics, 1/B moduli,
// twiddles, column counters and offsets, atom staging and sort records.
size_t engine_bytes(const int K[3], int flags, int n) {
    const bool fp32 = (flags & kFlagFP32) != 0;
    const size_t tsz = fp32 ? 4 : 8;
    const size_t ldim = K[2] / 2 + 1;
    const size_t H = (size_t)K[0] * K[1] * ldim;
    const size_t rows = (size_t)K[0] * K[1];
    const size_t cap = std::max<size_t>(1024, (size_t)n + n / 4);
    size_t b = 2 * H * tsz;                                  // grid
    b += 7 * H * 8 + (H / 256 + 1) * 32;                     // table + statistics
    b += (size_t)(K[0] + K[1] + K[2]) * (8 + 16 + 8);        // 1/B, twiddles
    b += (rows + 16383) / 16384 * 16384 * 4 + (rows + 1) * 4;  // column counters, offsets
    b += cap * (8 * 8 + 2 * 4 + (fp32 ? 32 : 48));           // x,w,dx,phi; col, rank; recs
    b += 1 << 20;                                            // descriptors, partials, rounding
    return b;
}

size_t atom_bytes(int flags, size_t cap) {
    return cap * (8 * 8 + 2 * 4 + ((flags & kFlagFP32) ? 32 : 48));
}

bool host_pinned(const void *p) {
    if (!p) return true;
    hipPointerAttribute_t a;
    memset(&a, 0, sizeof(a));
    const hipError_t e = hipPointerGetAttributes(&a, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeHost;
}

double pct(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = q * (double)(v.size() - 1);
    const size_t i = (size_t)pos;
    const double f = pos - (double)i;
    return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
}

constexpr int kMaxBypass = 64;  // later requests admitted ahead of a waiting one, at most

}  // namespace

Server::Server(int slots, double hbm_budget, int max_batch, int batch_atoms) {
    if (slots < 1 || slots > 64) throw std::invalid_argument("slots must be in 1..64");
    if (max_batch < 1) throw std::invalid_argument("max_batch < 1");
    hip_check(hipGetDevice(&device_), "hipGetDevice");
    size_t fr = 0, tot = 0;
    hip_check(hipMemGetInfo(&fr, &tot), "hipMemGetInfo");
    budget_ = hbm_budget > 0 ? (size_t)hbm_budget : (size_t)(0.8 * (double)fr);
    max_batch_ = max_batch;
    batch_atoms_ = batch_atoms > 0 ? batch_atoms : 32768;
    t0_ = std::chrono::steady_clock::now();
    slots_.resize(slots);
    for (Slot &s : slots_) {
        hip_check(hipStreamCreateWithFlags(&s.st, hipStreamNonBlocking), "hipStreamCreate");
        s.ev.resize(max_batch_);
        for (hipEvent_t &e : s.ev)
            hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    }
    for (int i = 0; i < slots; ++i) threads_.emplace_back(&Server::worker_, this, i);
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_work_.notify_all();
    for (std::thread &t : threads_) t.join();
    for (auto &kv : reqs_) delete kv.second;
    for (Resident &r : pool_) delete r.e;
    for (Slot &s : slots_) {
        delete s.batch;
        for (hipEvent_t e : s.ev) (void)hipEventDestroy(e);
        (void)hipStreamDestroy(s.st);
    }
}

double Server::now_() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
}

long long Server::submit(const double L6[6], const int K[3], int order, int flags, double eta,
                         double max_m, int n, const double *w, const double *x, double *out,
                         double *dx, double *phi) {
    if (n < 0) throw std::invalid_argument("n < 0");
    if (n > 0 && (!w || !x)) throw std::invalid_argument("w and x are required");
    if (!out) throw std::invalid_argument("out (7 doubles) is required");
    if (!(eta > 0) || !(max_m > 0)) throw std::invalid_argument("eta and max_m must be > 0");
    if (order % 2 != 0 || order < 2 || order > kMaxOrder)
        throw std::invalid_argument("invalid order");
    for (int d = 0; d < 3; ++d)
        if (K[d] < 1 || K[d] > kMaxK) throw std::invalid_argument("invalid K");
    if (flags & ~(kFlagFP32 | kFlagFixedPoint)) throw std::invalid_argument("invalid flags");
    ServeRequest *r = new ServeRequest;
    memcpy(r->L6, L6, sizeof(r->L6));
    memcpy(r->K, K, sizeof(r->K));
    r->order = order;
    r->flags = flags;
    r->eta = eta;
    r->max_m = max_m;
    r->n = n;
    r->w = w;
    r->x = x;
    r->out = out;
    r->dx = dx;
    r->phi = phi;
    r->pin_in = host_pinned(w) && host_pinned(x);
    r->est_bytes = engine_bytes(K, flags, n);
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (stop_) {
            delete r;
            throw std::runtime_error("server is shutting down");
        }
        r->id = next_id_++;
        r->t_submit = now_();
        reqs_[r->id] = r;
        queue_.push_back(r);
    }
    cv_work_.notify_one();
    return r->id;
}

int Server::wait(long long id, double timing[3], std::string *err) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = reqs_.find(id);
    if (it == reqs_.end()) throw std::invalid_argument("unknown request id");
    ServeRequest *r = it->second;
    cv_done_.wait(lk, [&] { return r->status == 2 || r->status < 0; });
    if (timing) {
        timing[0] = r->t_admit - r->t_submit;
        timing[1] = r->t_done - r->t_admit;
        timing[2] = r->t_done - r->t_submit;
    }
    const int st = r->status < 0 ? r->status : 0;
    if (err) *err = r->error;
    reqs_.erase(it);
    delete r;
    return st;
}

int Server::poll(long long id) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = reqs_.find(id);
    if (it == reqs_.end()) throw std::invalid_argument("unknown request id");
    const int s = it->second->status;
    return s < 0 ? s : (s == 2 ? 1 : 0);
}

void Server::stats(double out[16]) {
    std::lock_guard<std::mutex> lk(mu_);
    size_t held = 0;
    for (const Resident &r : pool_) held += r.bytes;
    double mean = 0.0, mx = 0.0;
    for (double v : lat_) {
        mean += v;
        mx = std::max(mx, v);
    }
    if (!lat_.empty()) mean /= (double)lat_.size();
    out[0] = (double)lat_.size();
    out[1] = (double)failed_;
    out[2] = pct(lat_, 0.50);
    out[3] = pct(lat_, 0.99);
    out[4] = mean;
    out[5] = mx;
    out[6] = pct(wait_, 0.50);
    out[7] = pct(wait_, 0.99);
    out[8] = (double)pool_.size();
    out[9] = (double)held;
    out[10] = (double)budget_;
    out[11] = (double)created_;
    out[12] = (double)evicted_;
    out[13] = (double)batches_;
    out[14] = (double)batch_frames_;
    out[15] = (double)max_running_;
}

void Server::reset_stats() {
    std::lock_guard<std::mutex> lk(mu_);
    lat_.clear();
    wait_.clear();
    failed_ = created_ = evicted_ = batches_ = batch_frames_ = 0;
    max_running_ = running_;
}

void Server::fail_(ServeRequest *r, int code, const std::string &msg) {
    r->status = code;
    r->error = msg;
    r->t_done = now_();
    if (r->t_admit == 0) r->t_admit = r->t_done;
    ++failed_;
}

// an idle resident engine for r (same grid, order and precision), preferring
// one whose atom buffers already hold r's atoms; -1 if none
int Server::find_idle_(const ServeRequest &r) const {
    int best = -1;
    for (int i = 0; i < (int)pool_.size(); ++i) {
        const Resident &p = pool_[i];
        if (p.busy || p.order != r.order || p.flags != r.flags || p.K[0] != r.K[0] ||
            p.K[1] != r.K[1] || p.K[2] != r.K[2])
            continue;
        if (best < 0) best = i;
        if (p.e->atom_capacity() >= (size_t)r.n) return i;
    }
    return best;
}

// evict idle engines, least recently used first, until `need` more bytes fit
bool Server::make_room_(size_t need, std::vector<Engine *> &victims) {
    if (used_ + need <= budget_) return true;
    size_t idle = 0;
    for (const Resident &p : pool_)
        if (!p.busy) idle += p.bytes;
    if (used_ - std::min(used_, idle) + need > budget_) return false;
    while (used_ + need > budget_) {
        int v = -1;
        for (int i = 0; i < (int)pool_.size(); ++i)
            if (!pool_[i].busy && (v < 0 || pool_[i].last_use < pool_[v].last_use)) v = i;
        if (v < 0) return false;
        victims.push_back(pool_[v].e);
        used_ -= std::min(used_, pool_[v].bytes);
        pool_.erase(pool_.begin() + v);
        ++evicted_;
    }
    return true;
}

Server::Resident *Server::resident_(const Engine *e) {
    for (Resident &p : pool_)
        if (p.e == e) return &p;
    return nullptr;
}

// Admit the first queued request that can run now (an idle engine of its
// kind, or room in the budget for a new one), then, if it is small and has
// an engine, the queued small requests whose idle engines launch the same
// kernel variants.  A request passed over kMaxBypass times blocks later ones
// until it fits.
bool Server::admit_(std::vector<Admitted> &grp, std::vector<Engine *> &victims) {
    grp.clear();
    std::vector<ServeRequest *> blocked;  // earlier requests that do not fit yet
    for (auto it = queue_.begin(); it != queue_.end();) {
        ServeRequest *r = *it;
        if (r->est_bytes > budget_) {
            fail_(r, -2, "request needs about " + std::to_string(r->est_bytes) +
                             " bytes of HBM, above the server's budget of " +
                             std::to_string(budget_));
            it = queue_.erase(it);
            cv_done_.notify_all();
            continue;
        }
        const int res = find_idle_(*r);
        Engine *e = nullptr;
        size_t reserve = 0;
        bool ok = true;
        if (res >= 0) {
            e = pool_[res].e;
            if (e->atom_capacity() < (size_t)r->n) {
                const size_t grow = atom_bytes(r->flags, (size_t)r->n + r->n / 4);
                // make_room_ only evicts idle engines; protect this one first
                pool_[res].busy = true;
                ok = make_room_(grow, victims);
                Resident *p = resident_(e);
                if (ok) {
                    used_ += grow;
                    p->bytes += grow;
                } else {
                    p->busy = false;
                }
            }
        } else {
            ok = make_room_(r->est_bytes, victims);
            if (ok) {
                reserve = r->est_bytes;
                used_ += reserve;
            }
        }
        if (!ok) {
            if (r->bypassed >= kMaxBypass) return false;
            blocked.push_back(r);
            ++it;
            continue;
        }
        for (ServeRequest *b : blocked) ++b->bypassed;
        if (e) {
            Resident *p = resident_(e);
            p->busy = true;
            p->last_use = ++tick_;
        }
        r->status = 1;
        r->t_admit = now_();
        grp.push_back({r, e, reserve});
        queue_.erase(it);
        break;
    }
    if (grp.empty()) return false;
    // continuous batching of small requests onto idle engines of one variant
    if (max_batch_ > 1 && grp[0].e && grp[0].r->n <= batch_atoms_) {
        const unsigned long long var = grp[0].e->variant();
        for (auto it = queue_.begin(); it != queue_.end() && (int)grp.size() < max_batch_;) {
            ServeRequest *r = *it;
            const int res = r->n <= batch_atoms_ && r->est_bytes <= budget_ ? find_idle_(*r) : -1;
            if (res >= 0 && pool_[res].e->variant() == var &&
                pool_[res].e->atom_capacity() >= (size_t)r->n) {
                pool_[res].busy = true;
                pool_[res].last_use = ++tick_;
                r->status = 1;
                r->t_admit = now_();
                grp.push_back({r, pool_[res].e, 0});
                it = queue_.erase(it);
            } else {
                ++it;
            }
        }
    }
    running_ += (int)grp.size();
    max_running_ = std::max(max_running_, running_);
    return true;
}

void Server::worker_(int slot) {
    (void)hipSetDevice(device_);
    Slot &S = slots_[slot];
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
        std::vector<Admitted> grp;
        std::vector<Engine *> victims;
        cv_work_.wait(lk, [&] { return stop_ || admit_(grp, victims); });
        if (grp.empty()) break;  // stopping
        lk.unlock();
        for (Engine *v : victims) delete v;
        victims.clear();
        // engines still to be made (outside the lock: allocation and uploads)
        std::vector<std::string> make_err(grp.size());
        for (size_t i = 0; i < grp.size(); ++i) {
            if (grp[i].e) continue;
            ServeRequest *r = grp[i].r;
            try {
                Engine *e = new Engine(r->L6, r->K, r->order, r->flags);
                e->use_graph = false;
                e->ensure_atoms(r->n);
                grp[i].e = e;
            } catch (const std::exception &ex) {
                make_err[i] = ex.what();
            }
        }
        lk.lock();
        std::vector<ServeRequest *> reqs;
        std::vector<Engine *> eng;
        std::vector<Resident> snap;  // per engine: the box its table was built for
        for (size_t i = 0; i < grp.size(); ++i) {
            Admitted &a = grp[i];
            if (a.reserved) {
                used_ -= std::min(used_, a.reserved);
                if (!a.e) {
                    fail_(a.r, -3, make_err[i]);
                    continue;
                }
                Resident p;
                p.e = a.e;
                memcpy(p.K, a.r->K, sizeof(p.K));
                p.order = a.r->order;
                p.flags = a.r->flags;
                p.busy = true;
                p.last_use = ++tick_;
                p.bytes = a.e->device_bytes();
                used_ += p.bytes;
                pool_.push_back(p);
                ++created_;
            }
            reqs.push_back(a.r);
            eng.push_back(a.e);
            snap.push_back(*resident_(a.e));
        }
        lk.unlock();
        std::string err;
        int code = 0;
        try {
            // per-request influence tables, skipped when the engine already
            // holds the table of this box and kernel
            for (size_t i = 0; i < reqs.size(); ++i) {
                ServeRequest *r = reqs[i];
                Resident &p = snap[i];
                const bool same = p.has_box && p.eta == r->eta && p.max_m == r->max_m &&
                                  memcmp(p.L6, r->L6, sizeof(p.L6)) == 0;
                if (!same) {
                    p.has_box = false;
                    p.e->set_cell(r->L6);
                    p.e->build_ewald_table(r->eta, r->max_m);
                    memcpy(p.L6, r->L6, sizeof(p.L6));
                    p.eta = r->eta;
                    p.max_m = r->max_m;
                    p.has_box = true;
                }
            }
            run_group_(S, reqs, eng);
        } catch (const std::exception &ex) {
            err = ex.what();
            code = -3;
            for (Resident &p : snap) p.has_box = false;
        }
        lk.lock();
        const double t = now_();
        for (size_t i = 0; i < reqs.size(); ++i) {
            ServeRequest *r = reqs[i];
            Resident *p = resident_(eng[i]);
            p->busy = false;
            p->has_box = snap[i].has_box;
            memcpy(p->L6, snap[i].L6, sizeof(p->L6));
            p->eta = snap[i].eta;
            p->max_m = snap[i].max_m;
            const size_t b = p->e->device_bytes();
            if (b > p->bytes) {
                used_ += b - p->bytes;
                p->bytes = b;
            }
            if (code) {
                fail_(r, code, err);
            } else {
                r->status = 2;
                r->t_done = t;
                lat_.push_back(r->t_done - r->t_submit);
                wait_.push_back(r->t_admit - r->t_submit);
            }
        }
        if (reqs.size() > 1) {
            ++batches_;
            batch_frames_ += (long long)reqs.size();
        }
        running_ -= (int)grp.size();
        cv_done_.notify_all();
        cv_work_.notify_all();
    }
}

// Upload, evaluate and download one admitted group.  A single request runs
// on its engine's stream; a group uploads on each engine's stream and is
// evaluated as one batch (one launch per stage) on the slot's stream.
void Server::run_group_(Slot &S, const std::vector<ServeRequest *> &reqs,
                        const std::vector<Engine *> &eng) {
    if (reqs.size() == 1) {
        ServeRequest *r = reqs[0];
        Engine *e = eng[0];
        e->upload_atoms(r->n, r->w, r->x, r->pin_in);
        e->evaluate_staged(r->n, r->dx != nullptr, r->phi != nullptr);
        e->download_results(r->n, r->out, r->dx, r->phi, e->stream());
        e->sync();
        return;
    }
    for (size_t i = 0; i < reqs.size(); ++i) {
        eng[i]->upload_atoms(reqs[i]->n, reqs[i]->w, reqs[i]->x, reqs[i]->pin_in);
        hip_check(hipEventRecord(S.ev[i], eng[i]->stream()), "hipEventRecord");
        hip_check(hipStreamWaitEvent(S.st, S.ev[i], 0), "hipStreamWaitEvent");
    }
    if (!S.batch) {
        S.batch = new Batch(eng);
        S.batch->use_graph = false;
    } else {
        S.batch->rebind(eng);
    }
    for (size_t i = 0; i < reqs.size(); ++i) {
        ServeRequest *r = reqs[i];
        Engine *e = eng[i];
        S.batch->set_atoms((int)i, r->n, e->w_dev(), e->x_dev(), e->result_dev(),
                           r->dx ? e->dx_dev() : nullptr, r->phi ? e->phi_dev() : nullptr);
    }
    S.batch->evaluate(S.st);
    for (size_t i = 0; i < reqs.size(); ++i)
        eng[i]->download_results(reqs[i]->n, reqs[i]->out, reqs[i]->dx, reqs[i]->phi, S.st);
    hip_check(hipStreamSynchronize(S.st), "hipStreamSynchronize");
}

}  // namespace spme
