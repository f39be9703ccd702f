// Autotuner of launch shapes: FFT radix order per axis, transform path,
// spreading tile and gather tile, and the batch's groups on streams, timed
// on the device over whole batched evaluations (declared in engine.hpp).
#include <stdio.h>

#include <algorithm>
#include <array>
#include <string>
#include <vector>

#include "engine.hpp"

namespace spme {

namespace {

std::string radix_text(const std::vector<int> &r) {
    std::string s;
    for (size_t i = 0; i < r.size(); ++i) s += (i ? "," : "") + std::to_string(r[i]);
    return s;
}

struct Timer {
    Batch &b;
    const std::vector<Engine *> &frames;
    hipStream_t st;
    int reps;
    std::vector<hipEvent_t> ev;

    Timer(Batch &b_, const std::vector<Engine *> &f, hipStream_t s, int r)
        : b(b_), frames(f), st(s), reps(r), ev(2 * (size_t)r) {
        for (hipEvent_t &e : ev) hip_check(hipEventCreate(&e), "hipEventCreate");
    }
    ~Timer() {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
    // median milliseconds of one evaluation of the whole batch
    double operator()() {
        b.rebind(frames);  // the frames' current descriptors
        for (int w = 0; w < 3; ++w) b.evaluate(st);
        for (int r = 0; r < reps; ++r) {
            hip_check(hipEventRecord(ev[2 * r], st), "hipEventRecord");
            b.evaluate(st);
            hip_check(hipEventRecord(ev[2 * r + 1], st), "hipEventRecord");
        }
        hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
        std::vector<double> t(reps);
        for (int r = 0; r < reps; ++r) {
            float ms = 0.f;
            hip_check(hipEventElapsedTime(&ms, ev[2 * r], ev[2 * r + 1]), "hipEventElapsedTime");
            t[r] = ms;
        }
        std::sort(t.begin(), t.end());
        return t[t.size() / 2];
    }
};

}  // namespace

std::string autotune(const std::vector<Engine *> &frames, Batch &b, hipStream_t st, int reps) {
    if (frames.empty()) throw std::invalid_argument("no frames to tune");
    if (reps < 1) reps = 1;
    Timer time(b, frames, st, reps);
    std::string rep;
    char line[512];
    (void)time();  // clocks and caches warm before the first measurement
    double best = time();
    const double t_start = best;
    snprintf(line, sizeof(line), "start %.4f ms per batch evaluation (%d frames)\n", best,
             (int)frames.size());
    rep += line;

    // classes of frames sharing grid, order and precision
    std::vector<std::string> keys;
    for (Engine *e : frames)
        if (std::find(keys.begin(), keys.end(), e->tune_key()) == keys.end())
            keys.push_back(e->tune_key());

    for (const std::string &key : keys) {
        std::vector<Engine *> cls;
        for (Engine *e : frames)
            if (e->tune_key() == key) cls.push_back(e);
        Engine *e0 = cls[0];
        TuneCfg cur = e0->tune_cfg();
        for (Engine *e : cls) e->apply_tune(cur);  // one shape for the whole class
        // try one candidate: keep it if it beats the best by 1 %, confirmed
        // by measuring the current shapes and the candidate once more (the
        // candidate's slower time against the current shapes' faster one), so
        // that run-to-run noise does not pick a shape
        auto trial = [&](const TuneCfg &c, const std::string &what) {
            try {
                for (Engine *e : cls) e->apply_tune(c);
            } catch (const std::invalid_argument &) {
                for (Engine *e : cls) e->apply_tune(cur);
                return;
            }
            double t = time();
            bool keep = t < 0.99 * best;
            if (keep) {
                for (Engine *e : cls) e->apply_tune(cur);
                best = std::min(best, time());
                for (Engine *e : cls) e->apply_tune(c);
                t = std::max(t, time());
                keep = t < 0.99 * best;
            }
            snprintf(line, sizeof(line), "%s %s: %.4f ms%s\n", key.c_str(), what.c_str(), t,
                     keep ? " (kept)" : "");
            rep += line;
            if (keep) {
                best = t;
                cur = c;
            } else {
                for (Engine *e : cls) e->apply_tune(cur);
            }
        };
        // transform path
        if (e0->slab_ok) {
            TuneCfg c = cur;
            c.slab = cur.slab ? 0 : 2;
            trial(c, std::string("transforms ") + (c.slab ? "slab" : "lines"));
        }
        // spreading fused with the forward slab transform, or in tiles before it
        if (e0->slab_ok && e0->spread_slab_ok && cur.slab != 0) {
            TuneCfg c = cur;
            c.slab = cur.slab == 2 ? 1 : 2;
            trial(c, std::string("spreading ") + (c.slab == 2 ? "fused" : "in tiles"));
        }
        // radix order of each axis (z, y, x: the slab passes first)
        for (int d : {2, 1, 0}) {
            const auto opts = e0->radix_options(d);
            for (size_t i = 1; i < opts.size(); ++i) {
                TuneCfg c = cur;
                c.nrad[d] = (int)opts[i].size();
                for (size_t s = 0; s < opts[i].size(); ++s) c.radix[d][s] = opts[i][s];
                trial(c, "axis " + std::to_string(d) + " radices " + radix_text(opts[i]));
            }
        }
        auto tile_text = [](const std::array<int, 3> &t) {
            return std::to_string(t[0]) + "x" + std::to_string(t[1]) +
                   (t[2] ? "x" + std::to_string(t[2]) : std::string());
        };
        // spreading tile (z chunks of tz cells, or whole columns; none when the
        // spreading runs fused with the slab transform)
        for (const auto &t : e0->spread_tile_options()) {
            if (e0->spread_slab()) break;
            if (t[0] == cur.tx && t[1] == cur.ty && t[2] == cur.tz) continue;
            TuneCfg c = cur;
            c.tx = t[0];
            c.ty = t[1];
            c.tz = t[2];
            trial(c, "spread tile " + tile_text(t));
        }
        // gather tile (0x0: per-atom gather)
        for (const auto &t : e0->gather_tile_options()) {
            if (t[0] == cur.gtx && t[1] == cur.gty && t[2] == cur.gtz) continue;
            TuneCfg c = cur;
            c.gtx = t[0];
            c.gty = t[1];
            c.gtz = t[2];
            trial(c, "gather tile " + tile_text(t));
        }
        for (Engine *e : cls) e->apply_tune(cur);
        tune_remember(*e0, cur);
    }
    // the batch as two groups of frames on two streams, or as one group (a
    // property of this batch, not of the frames' classes), kept as the
    // shapes are
    if (frames.size() >= 2) {
        const int was = b.lanes(), k = was == 1 ? 2 : 1;
        b.set_lanes(k);
        double t = time();
        bool keep = t < 0.99 * best;
        if (keep) {
            b.set_lanes(was);
            best = std::min(best, time());
            b.set_lanes(k);
            t = std::max(t, time());
            keep = t < 0.99 * best;
        }
        snprintf(line, sizeof(line), "batch in %d group%s on %s: %.4f ms%s\n", k,
                 k == 1 ? "" : "s", k == 1 ? "one stream" : "their own streams", t,
                 keep ? " (kept)" : "");
        rep += line;
        if (keep) best = t;
        else b.set_lanes(was);
    }
    b.rebind(frames);
    snprintf(line, sizeof(line), "end %.4f ms per batch evaluation (start %.4f)\n", best,
             t_start);
    rep += line;
    return rep;
}

}  // namespace spme
