// Serving loop: continuous batching of SPME evaluations of mixed atom counts,
// grids, spline orders and precisions on one GPU.
//
// Requests are submitted with caller-owned host buffers (borrowed until
// wait() returns) and queued.  `slots` worker threads each drive their own
// HIP streams; a worker admits the next request as soon as it is free and
// the request's device memory fits the server's HBM budget, so requests that
// arrive while others run start immediately and their work overlaps on the
// device.  Engines (grid, FFT plans, twiddles, influence table, sort
// workspace) stay resident between requests, keyed by grid, order and
// precision; idle engines are evicted least-recently-used when a new one
// does not fit.  A worker that admits a small request also takes queued
// small requests whose idle engines run the same kernel variants, and
// evaluates them together (one launch per stage), so every frame is computed
// exactly as it is alone.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine.hpp"

namespace spme {

struct ServeRequest {
    long long id = 0;
    double L6[6];
    int K[3];
    int order = 4;
    int flags = 0;
    double eta = 0.0, max_m = 0.0;
    int n = 0;
    const double *w = nullptr, *x = nullptr;          // borrowed
    double *out = nullptr, *dx = nullptr, *phi = nullptr;  // written on completion
    bool pin_in = false;
    size_t est_bytes = 0;
    int status = 0;  // 0 queued, 1 running, 2 done, < 0 error code
    std::string error;
    double t_submit = 0, t_admit = 0, t_done = 0;  // seconds since the server started
    int bypassed = 0;  // later requests admitted ahead of this one
};

class Server {
public:
    Server(int slots, double hbm_budget, int max_batch, int batch_atoms);
    ~Server();
    Server(const Server &) = delete;
    Server &operator=(const Server &) = delete;

    long long submit(const double L6[6], const int K[3], int order, int flags, double eta,
                     double max_m, int n, const double *w, const double *x, double *out,
                     double *dx, double *phi);
    // blocks until the request ends; timing[0..2] = queue wait, service, total seconds.
    // Returns 0, or the request's negative error code (message in *err).
    int wait(long long id, double timing[3], std::string *err);
    int poll(long long id);  // 1 done, 0 pending, < 0 unknown id or failed
    // out[0..15]: completed, failed, p50, p99, mean, max latency (s), p50 and p99
    // queue wait (s), resident engines, HBM bytes held by engines, budget,
    // engines created, evictions, batched launches, frames in batched launches,
    // most requests running at once
    void stats(double out[16]);
    void reset_stats();

private:
    struct Resident {
        Engine *e = nullptr;
        int K[3];
        int order, flags;
        bool busy = false;
        unsigned long long last_use = 0;
        size_t bytes = 0;
        // box and kernel of the table the engine holds
        bool has_box = false;
        double L6[6];
        double eta = 0, max_m = 0;
    };
    struct Admitted {
        ServeRequest *r;
        Engine *e;        // resident engine, or NULL for one still to be made
        size_t reserved;  // bytes reserved for an engine still to be made
    };
    struct Slot {
        hipStream_t st = nullptr;
        std::vector<hipEvent_t> ev;
        Batch *batch = nullptr;
    };

    double now_() const;
    bool admit_(std::vector<Admitted> &grp, std::vector<Engine *> &victims);
    int find_idle_(const ServeRequest &r) const;
    Resident *resident_(const Engine *e);
    bool make_room_(size_t need, std::vector<Engine *> &victims);
    void fail_(ServeRequest *r, int code, const std::string &msg);
    void worker_(int slot);
    void run_group_(Slot &slot, const std::vector<ServeRequest *> &reqs,
                    const std::vector<Engine *> &eng);

    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<ServeRequest *> queue_;
    std::unordered_map<long long, ServeRequest *> reqs_;
    std::vector<Resident> pool_;
    std::vector<Slot> slots_;
    std::vector<std::thread> threads_;
    std::chrono::steady_clock::time_point t0_;
    size_t budget_ = 0, used_ = 0;
    int max_batch_ = 1, batch_atoms_ = 0;
    long long next_id_ = 1;
    unsigned long long tick_ = 0;
    bool stop_ = false;
    int device_ = 0;
    int running_ = 0;
    // statistics
    std::vector<double> lat_, wait_;
    long long failed_ = 0, created_ = 0, evicted_ = 0, batches_ = 0, batch_frames_ = 0;
    int max_running_ = 0;
};

}  // namespace spme
