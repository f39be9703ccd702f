#!/usr/bin/env python3
"""SPME reciprocal-space evaluations/s on one MI355X.

One evaluation = sfac + de1 + de2 + potl (spreading, r2c FFT, /B, influence
multiply with energy/virial, c2r FFT, force + potential gather) of a
water_98k frame (32,768 rigid 3-site waters = 98,304 atoms, 100 A cube,
128^3 grid, B-spline order 4, beta 0.32), with positions and charges
already resident in HBM.  Influence-table setup (set_A) is outside the timed
region, as in the reference's usage.

One step = one batched evaluation of --batch independent frames (each its
own coordinates), one kernel launch per stage for all frames, replayed from
a captured hipGraph.  `value` is the FP32 engine's evaluations/s over the
whole job; the FP64 engine, single-frame latency and per-batch p50/p99 are
reported beside it.  Prints ONE JSON line.

    python bench.py [--steps K] [--warmup W] [--batch B]

HBM traffic of the dominant kernel (roofline.traffic) comes from two
rocprofv3 counter passes over `bench.py --counters` (FETCH_SIZE, then
WRITE_SIZE, CSV output), reduced to calibrated bytes per launch by

    python bench.py --traffic FETCH.csv WRITE.csv

which writes profiles/traffic_<workload>_fp32_batch<B>.json.  The counters
are scaled by what they read for the 256 MiB device copy of the same run.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "SPME recip evals/s at 98k atoms, FP32 vs FP64 GPU; p50/p99 batched latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def frames(name, count):
    from ewaldcorrel_amd import workloads as wl
    if name == "water_98k":
        return [wl.water(32768, 100.0, (128, 128, 128), seed=2 + f, name="water_98k")
                for f in range(count)]
    if name == "point_charges_1k":
        return [wl.point_charges_1k(seed=1 + f) for f in range(count)]
    if name == "gaussian_1m":
        return [wl.gaussian_1m(seed=3 + f) for f in range(count)]
    raise SystemExit(f"unknown workload {name}")


def kernel_bytes(kernel, n, K, p):
    """Algorithmic HBM bytes of one frame in one launch: the compulsory reads
    and writes of that kernel.  p = working-precision bytes of the grid;
    positions, charges, forces and potentials are doubles at the ABI."""
    K0, K1, K2 = K
    G = K0 * K1 * K2
    H = K0 * K1 * (K2 // 2 + 1)
    rec = 4 * p + 8  # atom record: fractions and charge in the working precision, cell, index
    table = {
        "bin": n * 4 * 8 + n * rec,            # x, w in; records out (plane-ordered chunks)
        "sort": 2 * n * rec,                   # records in, records out in column order
        "spread": n * rec + G * p,             # column-sorted records in, grid out once
        # spreading fused with the forward y-z transform: records in (each read
        # by the workgroups of its stencil's planes, from L2), half-complex out
        "spread_fft_yz": n * rec + 2 * H * p,
        "fft_z_r2c": G * p + 2 * H * p,
        "fft_y_fwd": 4 * H * p,
        "fft_x_recip": 4 * H * p,              # closed-form influence: no table read
        "fft_y_inv": 4 * H * p,
        "fft_z_c2r": 2 * H * p + G * p,
        "fft_yz_fwd": G * p + 2 * H * p,       # real plane in, half-complex plane out
        "fft_yz_inv": 2 * H * p + G * p,
        "gather_sorted": G * p + n * (4 * 8 + 4) + n * 4 * 8,
        "gather_tile": G * p + n * (4 * 8 + 4) + n * 4 * 8,  # halo rows re-read from L2
    }
    return table.get(kernel)


def eval_bytes(n, K, p, pg):
    """B_eval: spread + fft_fwd + infl + fft_inv + gather (compulsory traffic
    of one evaluation, FP64 influence table at 8 B per half-complex point)."""
    K0, K1, K2 = K
    G = K0 * K1 * K2
    H = K0 * K1 * (K2 // 2 + 1)
    spread = n * 4 * p + G * pg
    fwd = G * p + 2 * H * p
    infl = 2 * H * p + 2 * H * p + 8 * H
    inv = 2 * H * p + G * p
    gather = G * p + n * 4 * p + n * 4 * p
    return spread + fwd + infl + inv + gather


class Inputs:
    """Device-resident inputs and outputs of one frame."""

    def __init__(self, sy, dev):
        import torch
        n = sy.n
        self.n = n
        self.w = torch.tensor(sy.w, dtype=torch.float64, device=dev)
        self.x = torch.tensor(sy.x, dtype=torch.float64, device=dev)
        self.out = torch.zeros(8, dtype=torch.float64, device=dev)
        self.dx = torch.zeros((n, 3), dtype=torch.float64, device=dev)
        self.phi = torch.zeros(n, dtype=torch.float64, device=dev)


SET_A_MS = []  # wall time of each influence-table build (device builder + check)


def make_handles(systems, precision, set_a_ms=None, fp32_influence=False):
    from ewaldcorrel_amd import ewald
    hs = []
    for sy in systems:
        s = ewald.Sfac(sy.L6, np.array(sy.K), sy.order, precision=precision,
                       fp32_influence=fp32_influence)
        t0 = time.perf_counter()
        s.set_A(ewald.ewald_kernel(sy.beta), sy.max_m)
        (SET_A_MS if set_a_ms is None else set_a_ms).append(1e3 * (time.perf_counter() - t0))
        hs.append(s)
    return hs


def time_single(sy, handle, inp, steps, warmup):
    """One frame per step through sfac_eval_dev (graph replay)."""
    import torch
    stream = torch.cuda.ExternalStream(handle.stream)

    def step():
        handle.eval_dev(inp.n, inp.w.data_ptr(), inp.x.data_ptr(), inp.out.data_ptr(),
                        inp.dx.data_ptr(), inp.phi.data_ptr())

    for _ in range(warmup):
        step()
    handle.sync()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        ev[k][0].record(stream)
        step()
        ev[k][1].record(stream)
    handle.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    lat = np.array([a.elapsed_time(b) for a, b in ev])
    return dict(evals_per_s=steps / wall, ms_per_eval=wall * 1e3 / steps,
                p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)))


def time_batch(batch, handle0, nframes, steps, warmup, profile):
    """One step = one evaluation of every frame of the batch."""
    import torch
    stream = torch.cuda.ExternalStream(handle0.stream)
    for _ in range(warmup):
        batch.eval(None)
    batch.sync()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        ev[k][0].record(stream)
        batch.eval(None)
        ev[k][1].record(stream)
    batch.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    lat = np.array([a.elapsed_time(b) for a, b in ev])
    res = dict(evals_per_s=steps * nframes / wall, ms_per_step=wall * 1e3 / steps,
               p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
               mean_event_ms=float(lat.mean()))
    if profile:
        handle0.profile(True)
        for _ in range(steps):
            batch.eval(None)
        prof = handle0.profile_read()
        handle0.profile(False)
        res["kernels"] = {k: {"total_ms": v[0], "launches": v[1],
                              "avg_us": 1e3 * v[0] / max(v[1], 1)} for k, v in prof.items()}
    return res


def time_streams(batches, streams, nframes, steps, warmup):
    """One step = one evaluation of every sub-batch, each on its own HIP
    stream, so that kernels of different sub-batches overlap.  The step's
    latency runs from an event on the first stream (which every other stream
    waits for) to the point where the first stream has waited for all."""
    import torch
    S = [torch.cuda.ExternalStream(s) for s in streams]
    mk = lambda: torch.cuda.Event(enable_timing=True)  # noqa: E731

    def step(k, evs):
        if evs is not None:
            evs[k][0].record(S[0])
            for st in S[1:]:
                st.wait_event(evs[k][0])
        for b in batches:
            b.eval(None)
        if evs is not None:
            for i, st in enumerate(S[1:]):
                evs[k][2 + i].record(st)
                S[0].wait_event(evs[k][2 + i])
            evs[k][1].record(S[0])

    for k in range(warmup):
        step(k, None)
    torch.cuda.synchronize()
    evs = [[mk() for _ in range(1 + len(S))] for _ in range(steps)]
    t0 = time.perf_counter()
    for k in range(steps):
        step(k, evs)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    lat = np.array([e[0].elapsed_time(e[1]) for e in evs])
    return dict(evals_per_s=steps * nframes / wall, ms_per_step=wall * 1e3 / steps,
                p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                mean_event_ms=float(lat.mean()))


def run_precision(systems, precision, steps, warmup, profile, tune=True, split=(2, 4),
                  fp32_influence=False, single=True):
    """Single-frame evaluations with the default launch shapes, then the batch
    of all frames, autotuned first when `tune` (outside the timed region, as
    the reference's FFTW planning is)."""
    import torch
    from ewaldcorrel_amd import ewald
    dev = torch.device("cuda:0")
    handles = make_handles(systems, precision, None if not fp32_influence else [],
                           fp32_influence=fp32_influence)
    inputs = [Inputs(sy, dev) for sy in systems]
    torch.cuda.synchronize()
    single = time_single(systems[0], handles[0], inputs[0], steps, warmup) if single else None
    batch = ewald.SfacBatch(handles)
    for f, inp in enumerate(inputs):
        batch.set(f, inp.n, inp.w.data_ptr(), inp.x.data_ptr(), inp.out.data_ptr(),
                  inp.dx.data_ptr(), inp.phi.data_ptr())
    tuned = None
    if tune:
        lines = batch.autotune(reps=5).strip().splitlines()
        tuned = {"report_first": lines[0], "report_last": lines[-1],
                 "kept": [ln for ln in lines if ln.endswith("(kept)")]}
        log(f"{precision} autotune: {lines[-1]}; kept {tuned['kept']}")
    bat = time_batch(batch, handles[0], len(handles), steps, warmup, profile)
    bat["autotune"] = tuned
    # the same frames as S sub-batches on S streams
    bat["streams"] = {}
    nf = len(handles)
    for S in split:
        if S > nf or nf % S:
            continue
        per = nf // S
        subs = []
        for g in range(S):
            sb = ewald.SfacBatch(handles[g * per:(g + 1) * per])
            for f in range(per):
                inp = inputs[g * per + f]
                sb.set(f, inp.n, inp.w.data_ptr(), inp.x.data_ptr(), inp.out.data_ptr(),
                       inp.dx.data_ptr(), inp.phi.data_ptr())
            subs.append(sb)
        r = time_streams(subs, [handles[g * per].stream for g in range(S)], nf, steps, warmup)
        bat["streams"][str(S)] = r
        log(f"{precision}: {S} streams x {per} frames: {r['evals_per_s']:.0f} evals/s, "
            f"p50 {r['p50_ms']:.3f} ms/step")
        for sb in subs:
            sb.close()
    bat["table_mode"] = handles[0].info()["table_mode"]
    bat["lanes"] = batch.lanes()
    bat["E"] = [float(inp.out[0].item()) for inp in inputs]
    bat["vir"] = [inp.out[1:7].cpu().numpy().copy() for inp in inputs]
    bat["dx0"] = inputs[0].dx.cpu().numpy()
    bat["phi0"] = inputs[0].phi.cpu().numpy()
    batch.close()
    del handles
    return single, bat


def uneven_batch(nframes, steps, warmup, tune=True):
    """256 independent frames of uneven size (atom counts log-uniform in
    500-20,000, cubic boxes at 0.1 atoms/A^3, K = smallest 2^a 3^b 5^c >= L/1A
    clamped to [32, 96], order 4; seed 4) evaluated in FP32 as one batch (one
    launch per stage for all frames, replayed from a captured hipGraph), timed
    before and after the autotuner chooses the launch shapes of each grid."""
    import torch
    from ewaldcorrel_amd import ewald, workloads as wl
    dev = torch.device("cuda:0")
    systems = wl.uneven_frames(nframes, seed=4)
    handles = make_handles(systems, "fp32", [])
    inputs = [Inputs(sy, dev) for sy in systems]
    batch = ewald.SfacBatch(handles)
    for f, inp in enumerate(inputs):
        batch.set(f, inp.n, inp.w.data_ptr(), inp.x.data_ptr(), inp.out.data_ptr(),
                  inp.dx.data_ptr(), inp.phi.data_ptr())
    torch.cuda.synchronize()
    before = time_batch(batch, handles[0], nframes, steps, warmup, profile=False)
    E0 = [float(inp.out[0].item()) for inp in inputs]
    res = {"workload": "uneven_frames", "frames": nframes,
           "atoms": int(sum(sy.n for sy in systems)),
           "grid_points": int(sum(np.prod(sy.K) for sy in systems)),
           "grids": len({sy.K for sy in systems}),
           "untuned": {"evals_per_s": before["evals_per_s"], "ms_per_batch": before["p50_ms"]}}
    if tune:
        t0 = time.perf_counter()
        rep = batch.autotune(reps=5)
        lines = rep.strip().splitlines()
        res["autotune"] = {"seconds": time.perf_counter() - t0, "report_first": lines[0],
                           "report_last": lines[-1],
                           "kept": [ln for ln in lines if ln.endswith("(kept)")]}
    after = time_batch(batch, handles[0], nframes, steps, warmup, profile=False)
    E1 = [float(inp.out[0].item()) for inp in inputs]
    Bb = sum(eval_bytes(sy.n, sy.K, 4, 4) for sy in systems)
    res.update({"evals_per_s": after["evals_per_s"],
                "latency_ms": {"p50": after["p50_ms"], "p99": after["p99_ms"], "per": "batch"},
                "eval_roofline_frac": Bb / (after["mean_event_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                "tuned_vs_untuned_energy_rel_max": max(abs(a - b) / abs(b)
                                                       for a, b in zip(E1, E0))})
    batch.close()
    return res


def large_box(steps, warmup, tune=True):
    """One 1,048,576-charge system (Gaussian charges, 200 x 200 x 400 A,
    192 x 192 x 384 grid, order 6, beta 0.30; seed 3) evaluated alone, FP32
    (int64 fixed-point spreading) and FP64."""
    import torch
    from ewaldcorrel_amd import workloads as wl
    dev = torch.device("cuda:0")
    sy = wl.gaussian_1m()
    inp = Inputs(sy, dev)
    res = {"workload": sy.name, "atoms": sy.n, "grid": list(sy.K), "order": sy.order}
    outs = {}
    for prec in ("fp32", "fp64"):
        p = 4 if prec == "fp32" else 8  # grid bytes (the FP32 grid is written as FP32)
        (h,) = make_handles([sy], prec, [])
        tuned = None
        if tune:
            lines = h.autotune(sy.n, inp.w.data_ptr(), inp.x.data_ptr(), reps=5).strip()
            lines = lines.splitlines()
            tuned = {"report_first": lines[0], "report_last": lines[-1],
                     "kept": [ln for ln in lines if ln.endswith("(kept)")],
                     "shapes": h.launch_shapes()}
            log(f"large box {prec} autotune: {lines[-1]}; {tuned['shapes']}")
        r = time_single(sy, h, inp, steps, warmup)
        outs[prec] = (float(inp.out[0].item()), inp.dx.cpu().numpy().copy(),
                      inp.out[1:7].cpu().numpy().copy(), inp.phi.cpu().numpy().copy())
        h.profile(True)
        for _ in range(3):
            h.eval_dev(inp.n, inp.w.data_ptr(), inp.x.data_ptr(), inp.out.data_ptr(),
                       inp.dx.data_ptr(), inp.phi.data_ptr())
        prof = h.profile_read()
        h.profile(False)
        res[prec] = {"evals_per_s": r["evals_per_s"],
                     "latency_ms": {"p50": r["p50_ms"], "p99": r["p99_ms"]},
                     "eval_roofline_frac": eval_bytes(sy.n, sy.K, p, p) / (r["p50_ms"] * 1e-3)
                     / 1e9 / HBM_PEAK_GBS,
                     "kernels_avg_us": {k: round(1e3 * v[0] / max(v[1], 1), 3)
                                        for k, v in prof.items()},
                     "autotune": tuned}
        del h
    (e32, f32, v32, p32), (e64, f64, v64, p64) = outs["fp32"], outs["fp64"]
    res["fp32_error_vs_fp64_engine"] = {
        "energy_rel": abs(e32 - e64) / abs(e64),
        "virial_rel_max": float(np.max(np.abs(v32 - v64)) / np.max(np.abs(v64))),
        "force_rel_rms": float(np.sqrt(np.mean((f32 - f64) ** 2)) / np.sqrt(np.mean(f64 ** 2))),
        "potential_rel_rms": float(np.sqrt(np.mean((p32 - p64) ** 2)) /
                                   np.sqrt(np.mean(p64 ** 2)))}
    return res


def serve_stream(count, pool_size, rate, slots, max_batch, seed=5):
    """The serving stream: `count` requests with Poisson arrivals at `rate`/s
    submitted to the native serving loop (continuous batching over resident
    engines), then the same number submitted at once (saturation).

    Requests come from a pool of `pool_size` generated ones (atom counts
    log-uniform in 1k-500k, triclinic boxes, K_d from the box, order 4 or 6,
    FP32 or FP64; seed 5) picked at random; each request carries its own box
    (the pool entry's box scaled by a random factor in [1, 1.001)), so every
    request builds its influence table on the device, inside its latency.
    Inputs and outputs are pinned host arrays (the results come back to the
    host in every request's latency)."""
    import torch
    from ewaldcorrel_amd import ewald, workloads as wl
    rng = np.random.default_rng(seed)
    pool = [wl.serving_request(rng, 1000, 500000, f"request{i}") for i in range(pool_size)]
    arrivals = np.cumsum(np.r_[0.0, rng.exponential(1.0 / rate, count - 1)])
    picks = rng.integers(0, pool_size, 2 * count)
    scales = 1.0 + 1e-3 * rng.uniform(0, 1, 2 * count)

    def pinned(a):
        t = torch.empty(a.shape, dtype=torch.float64, pin_memory=True).numpy()
        t[:] = a
        return t

    ins = [(pinned(q.w), pinned(q.x)) for q in pool]
    outs = [(pinned(np.zeros(7)), pinned(np.zeros((q.n, 3))), pinned(np.zeros(q.n)))
            for q in pool]
    user = [None] * pool_size  # request currently writing pool entry p's outputs
    srv = ewald.SfacServer(slots=slots, max_batch=max_batch)

    def submit(p, scale):
        q = pool[p]
        if user[p] is not None:
            srv.wait(user[p])
        r = srv.submit(q.L6 * scale, q.K, q.order, ins[p][0], ins[p][1], q.beta, q.max_m,
                       precision=q.meta["precision"], out=outs[p])
        user[p] = r
        return r

    def drain():
        for p in range(pool_size):
            if user[p] is not None:
                srv.wait(user[p])
                user[p] = None

    # warm-up: every pool entry once (engines made resident)
    for p in range(pool_size):
        submit(p, 1.0)
    drain()
    srv.reset_stats()
    atoms = 0
    t0 = time.perf_counter()
    for i in range(count):
        t = t0 + arrivals[i]
        while True:
            dt = t - time.perf_counter()
            if dt <= 0:
                break
            if dt > 2e-4:
                time.sleep(dt - 1e-4)
        submit(int(picks[i]), scales[i])
        atoms += pool[int(picks[i])].n
    drain()
    wall = time.perf_counter() - t0
    st = srv.stats()
    stream = {"offered_req_per_s": rate, "requests": count,
              "req_per_s": st["completed"] / wall, "atoms_per_s": atoms / wall,
              "latency_ms": {"p50": 1e3 * st["latency_p50_s"], "p99": 1e3 * st["latency_p99_s"],
                             "mean": 1e3 * st["latency_mean_s"], "max": 1e3 * st["latency_max_s"]},
              "queue_wait_ms": {"p50": 1e3 * st["wait_p50_s"], "p99": 1e3 * st["wait_p99_s"]},
              "failed": st["failed"], "max_running": st["max_running"],
              "batched_launches": st["batched_launches"],
              "batched_frames": st["batched_frames"]}
    # saturation: the same number of requests, all queued at once
    srv.reset_stats()
    atoms = 0
    t0 = time.perf_counter()
    for i in range(count, 2 * count):
        submit(int(picks[i]), scales[i])
        atoms += pool[int(picks[i])].n
    drain()
    wall = time.perf_counter() - t0
    st2 = srv.stats()
    sat = {"requests": count, "req_per_s": st2["completed"] / wall, "atoms_per_s": atoms / wall,
           "latency_ms": {"p50": 1e3 * st2["latency_p50_s"], "p99": 1e3 * st2["latency_p99_s"]},
           "failed": st2["failed"], "max_running": st2["max_running"],
           "batched_launches": st2["batched_launches"], "batched_frames": st2["batched_frames"]}
    res = {"workload": "serving_stream", "slots": slots, "max_batch": max_batch,
           "pool": {"requests": pool_size, "atoms": [min(q.n for q in pool),
                                                     max(q.n for q in pool)],
                    "grids": len({q.K for q in pool}),
                    "fp32": sum(q.meta["precision"] == "fp32" for q in pool),
                    "order6": sum(q.order == 6 for q in pool)},
           "resident_engines": st2["resident_engines"],
           "resident_MB": st2["resident_bytes"] / 1e6, "engines_created": st2["engines_created"],
           "poisson": stream, "saturated": sat,
           "latency_includes": "queueing, influence-table build, H2D of x and w, the "
                               "evaluation, D2H of E, virial, forces and potentials"}
    srv.close()
    return res


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sy, E_gpu64=None):
    """The reference's CPU path, restated in C++ (oracle/sfac_cpu.cpp: multimap
    sort, gather-form OpenMP spreading, OpenMP /B and reductions, serial
    de2/potl; the host has no FFTW, so its transforms are the port's own
    single-threaded FFT), timed on this host on one frame: three evaluations
    after a warm-up at the host's thread count (the value), one at a single
    thread, and one of the -O0 build (the reference's Makefile builds -O0).
    set_A is outside the timing, as in the reference's use."""
    from oracle.cpu_port import run_port
    threads = max(1, min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16"))))
    rep, (E, vir, dx, phi) = run_port(sy.L6, sy.K, sy.order, sy.beta, sy.max_m, sy.w, sy.x,
                                      reps=3, threads=threads)
    rep1, _ = run_port(sy.L6, sy.K, sy.order, sy.beta, sy.max_m, sy.w, sy.x, reps=1, threads=1)
    rep0, _ = run_port(sy.L6, sy.K, sy.order, sy.beta, sy.max_m, sy.w, sy.x, reps=1,
                       threads=threads, opt="O0")
    steps = {k[:-2]: rep[k] for k in ("sort_s", "spread_s", "fft_fwd_s", "divB_s", "de1_s",
                                      "de2_s", "potl_s")}
    return {"value": 1.0 / rep["eval_s"], "unit": "evals/s", "cores": threads, "kind": "port",
            "sample": f"{sy.name} frame 0 ({sy.n} atoms, {sy.K[0]}^3, order {sy.order}): "
                      f"median of 3 evaluations after a warm-up at {threads} OpenMP threads "
                      "(C++ -O3 restatement of the reference CPU path, FP64; no FFTW on the "
                      "host: the port's own single-threaded FFT)",
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "seconds_per_eval": rep["eval_s"], "steps_s": steps,
            "single_thread_evals_per_s": 1.0 / rep1["eval_s"],
            "O0_evals_per_s": 1.0 / rep0["eval_s"],
            "energy_rel_vs_gpu_fp64": (abs(E - E_gpu64) / abs(E_gpu64)
                                       if E_gpu64 is not None else None)}


# kernel symbol prefix -> stage name of the batched FP32 evaluation
STAGE_OF_KERNEL = {
    "spme::bin_kernel": "bin",
    "spme::sort_kernel": "sort",
    "spme::spread_tile_kernel": "spread",
    "spme::spread_slab_kernel": "spread_fft_yz",
    "spme::fft_yz_fwd_kernel": "fft_yz_fwd",
    "spme::fft_yz_fwd128_kernel": "fft_yz_fwd",
    "spme::fft_yz_inv128_kernel": "fft_yz_inv",
    "spme::fft_z_r2c_kernel": "fft_z_r2c",
    "spme::fft_y_kernel": "fft_y",
    "spme::fft_x_kernel<float, 2": "fft_x_recip",
    "spme::fft_x_reg_kernel<float": "fft_x_recip",
    "spme::fft_yz_inv_kernel": "fft_yz_inv",
    "spme::fft_z_c2r_kernel": "fft_z_c2r",
    "spme::gather_kernel": "gather_sorted",
    "spme::gather_tile_kernel": "gather_tile",
}


def traffic_from_counters(fetch_csv, write_csv, copy_bytes=256 << 20):
    """Per-stage HBM bytes per launch from rocprofv3 FETCH_SIZE / WRITE_SIZE
    (KiB per dispatch), each scaled by counter-read / true bytes of the
    largest device copy kernel in the same run."""
    import csv
    import collections

    def load(path):
        per = collections.defaultdict(list)
        copy = 0.0
        for r in csv.DictReader(open(path)):
            name = r["Kernel_Name"].split("(")[0].replace("void ", "")
            v = float(r["Counter_Value"]) * 1024.0
            if "copyBuffer" in name:
                copy = max(copy, v)
            for pre, st in STAGE_OF_KERNEL.items():
                if name.startswith(pre) and "<double" not in name:
                    per[st].append(v)
        return per, copy

    fetch, cf = load(fetch_csv)
    write, cw = load(write_csv)
    ff, fw = copy_bytes / cf, copy_bytes / cw
    out = {"_calibration": {"copy_bytes": copy_bytes, "fetch_counter_bytes": cf,
                            "write_counter_bytes": cw, "fetch_factor": ff, "write_factor": fw}}
    for st in fetch:
        if st in write:
            f = ff * sum(fetch[st]) / len(fetch[st])
            w = fw * sum(write[st]) / len(write[st])
            out[st] = f + w
            out["_" + st] = {"fetch_bytes": f, "write_bytes": w, "dispatches": len(fetch[st])}
    return out


def counter_run(systems, steps, tune=True):
    import torch
    from ewaldcorrel_amd import ewald
    dev = torch.device("cuda:0")
    handles = make_handles(systems, "fp32")
    inputs = [Inputs(sy, dev) for sy in systems]
    batch = ewald.SfacBatch(handles)
    for f, inp in enumerate(inputs):
        batch.set(f, inp.n, inp.w.data_ptr(), inp.x.data_ptr(), inp.out.data_ptr(),
                  inp.dx.data_ptr(), inp.phi.data_ptr())
    if tune:
        batch.autotune(reps=5)  # the launch shapes the bench measures
    for _ in range(steps):
        batch.eval(None)
    batch.sync()
    batch.close()
    nbytes = 256 << 20
    a = torch.ones(nbytes // 4, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"counter_run": systems[0].name, "frames": len(systems), "steps": steps,
                      "copy_bytes": nbytes}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--workload", default="water_98k")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default launch shapes (no autotuning)")
    ap.add_argument("--uneven-frames", type=int, default=256,
                    help="frames of the uneven batch (0: skip it)")
    ap.add_argument("--no-large", action="store_true", help="skip the 1M-charge box")
    ap.add_argument("--serve-requests", type=int, default=2000,
                    help="requests of the serving stream (0: skip it)")
    ap.add_argument("--serve-rate", type=float, default=2000.0)
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-batch", type=int, default=8)
    ap.add_argument("--serve-pool", type=int, default=200)
    ap.add_argument("--traffic", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="reduce two rocprofv3 counter passes of --counters to the "
                         "per-stage traffic file read by the roofline")
    ap.add_argument("--counters", action="store_true",
                    help="short run for hardware-counter passes: the batched FP32 "
                         "evaluation only, then one device copy of 256 MiB whose "
                         "FETCH_SIZE / WRITE_SIZE calibrate the byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "frames and requests shard across GPUs with no exchange, which is not "
              "measured here.", file=sys.stderr)
        sys.exit(2)

    if args.traffic:
        out = traffic_from_counters(*args.traffic)
        path = os.path.join(ROOT, "profiles",
                            f"traffic_{args.workload}_fp32_batch{args.batch}.json")
        with open(path, "w") as f:
            json.dump(out, f, indent=1, sort_keys=True)
        log(f"wrote {path}")
        return

    systems = frames(args.workload, args.batch)
    sy = systems[0]
    log(f"workload {sy.name}: n={sy.n} K={sy.K} order={sy.order} beta={sy.beta} "
        f"batch={args.batch}")
    if args.counters:
        counter_run(systems, args.steps, tune=not args.no_tune)
        return
    s64, b64 = run_precision(systems, "fp64", args.steps, args.warmup, profile=True,
                             tune=not args.no_tune)
    log(f"fp64: single {s64['evals_per_s']:.0f} evals/s ({s64['p50_ms']:.3f} ms); "
        f"batch {b64['evals_per_s']:.0f} evals/s, p50 {b64['p50_ms']:.3f} ms/batch")
    s32, b32 = run_precision(systems, "fp32", args.steps, args.warmup, profile=True,
                             tune=not args.no_tune)
    log(f"fp32: single {s32['evals_per_s']:.0f} evals/s ({s32['p50_ms']:.3f} ms); "
        f"batch {b32['evals_per_s']:.0f} evals/s, p50 {b32['p50_ms']:.3f} ms/batch")
    _, b32i = run_precision(systems, "fp32", args.steps, args.warmup, profile=True,
                            tune=not args.no_tune, split=(), fp32_influence=True, single=False)
    log(f"fp32 with the FP32 influence option: batch {b32i['evals_per_s']:.0f} evals/s, "
        f"p50 {b32i['p50_ms']:.3f} ms/batch")

    # frame 0 against the FP64 oracle
    from oracle.sfac_oracle import OracleSfac, ewald_rad
    o = OracleSfac(sy.L6, sy.K, sy.order)
    o.set_A(ewald_rad(sy.beta), sy.max_m)
    E0, v0, dx0, phi0 = o.evaluate(sy.w, sy.x)

    def vs_oracle(b):
        return {"energy_rel": abs(b["E"][0] - E0) / abs(E0),
                "virial_rel_max": float(np.max(np.abs(b["vir"][0] - v0)) / np.max(np.abs(v0))),
                "force_rel_rms": float(np.sqrt(np.mean((b["dx0"] - dx0) ** 2)) /
                                       np.sqrt(np.mean(dx0 ** 2))),
                "potential_rel_rms": float(np.sqrt(np.mean((b["phi0"] - phi0) ** 2)) /
                                           np.sqrt(np.mean(phi0 ** 2)))}

    # FP32 against the FP64 engine on the same inputs
    eE = max(abs(a - b) / abs(b) for a, b in zip(b32["E"], b64["E"]))
    eF = float(np.sqrt(np.mean((b32["dx0"] - b64["dx0"]) ** 2)) /
               np.sqrt(np.mean(b64["dx0"] ** 2)))
    ePhi = float(np.sqrt(np.mean((b32["phi0"] - b64["phi0"]) ** 2)) /
                 np.sqrt(np.mean(b64["phi0"] ** 2)))
    eVir = max(float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
               for a, b in zip(b32["vir"], b64["vir"]))

    # The spreading stage of the batched FP32 engine (bin + sort + spread, one
    # launch each), timed live with HIP events on the batch's stream, against
    # the spreading bytes of one evaluation: N 4p (x, w) + G p (grid written
    # once), p = 4, for every frame of the launch.  Where the spreading runs
    # fused with the forward y-z transform (spread_fft_yz), its grid leaves
    # the launch as the half-complex plane: 2 H p in place of G p.
    B = args.batch
    ks = b32["kernels"]
    stage = [k for k in ("bin", "sort", "spread", "spread_fft_yz") if k in ks]
    stage_us = sum(ks[k]["avg_us"] for k in stage)
    G = int(np.prod(sy.K))
    H2 = 2 * sy.K[0] * sy.K[1] * (sy.K[2] // 2 + 1)
    Bk = (sy.n * 4 * 4 + (H2 if "spread_fft_yz" in stage else G) * 4) * B
    achieved = Bk / (stage_us * 1e-6) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{sy.name}_fp32_batch{B}.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            tr = json.load(f)
        if all(k in tr for k in stage):
            traffic = sum(tr[k] for k in stage)
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                "traffic_source": (os.path.relpath(tfile, ROOT) + " (rocprofv3 FETCH_SIZE + "
                                   "WRITE_SIZE of the same launches, calibrated by a device "
                                   "copy; a stored counter run, not measured in this run)")
                if traffic is not None else None,
                "kernel": "spreading stage: " + " + ".join(stage),
                "algorithmic_bytes": Bk, "avg_us": stage_us,
                "kernels_avg_us": {k: ks[k]["avg_us"] for k in stage},
                "frames_per_launch": B}
    Be32 = eval_bytes(sy.n, sy.K, 4, 4) * B  # the FP32 grid is written once, as FP32
    Be64 = eval_bytes(sy.n, sy.K, 8, 8) * B

    uneven = None
    if args.uneven_frames > 0:
        uneven = uneven_batch(args.uneven_frames, max(5, args.steps // 2), args.warmup,
                              tune=not args.no_tune)
        log("uneven batch: %.0f evals/s tuned (%.0f untuned), p50 %.3f ms/batch" % (
            uneven["evals_per_s"], uneven["untuned"]["evals_per_s"],
            uneven["latency_ms"]["p50"]))
    large = None
    if not args.no_large:
        large = large_box(max(5, args.steps // 2), args.warmup, tune=not args.no_tune)
        log("large box: fp32 %.0f evals/s, fp64 %.0f evals/s" % (
            large["fp32"]["evals_per_s"], large["fp64"]["evals_per_s"]))

    serving = None
    if args.serve_requests > 0:
        serving = serve_stream(args.serve_requests, args.serve_pool, args.serve_rate,
                               args.serve_slots, args.serve_batch)
        log("serving: poisson %.0f req/s p50 %.3f ms p99 %.3f ms; saturated %.0f req/s" % (
            serving["poisson"]["req_per_s"], serving["poisson"]["latency_ms"]["p50"],
            serving["poisson"]["latency_ms"]["p99"], serving["saturated"]["req_per_s"]))

    cpu = None if args.no_cpu else cpu_baseline(sy, b64["E"][0])
    if cpu:
        log(f"cpu baseline: {cpu['value']:.3f} evals/s at {cpu['cores']} threads "
            f"({cpu['single_thread_evals_per_s']:.3f} at 1, -O0 {cpu['O0_evals_per_s']:.3f}); "
            f"{cpu['cpu_model']}")

    line = {
        "metric": METRIC,
        "value": b32["evals_per_s"],
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": b32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": sy.name, "atoms": sy.n, "grid": list(sy.K), "order": sy.order,
                   "beta": sy.beta, "max_m": sy.max_m, "frames_per_step": B,
                   "precision": "fp32 spreading, FFTs and gather; FP64 coordinates, /B, "
                                "influence, energy/virial sums and reductions; int64 "
                                "fixed-point grid accumulation",
                   "influence": ("closed form from the indices (no table read)"
                                 if b32["table_mode"] == 1 else "table read from HBM"),
                   "batch_groups": b32["lanes"]},
        "set_A_ms": {"median": float(np.median(SET_A_MS)), "max": float(np.max(SET_A_MS)),
                     "what": "influence table built on the device, checked against "
                             "the closed form; per frame"},
        "latency_ms": {"p50": b32["p50_ms"], "p99": b32["p99_ms"], "per": "batch"},
        "autotune": {"fp32": b32["autotune"], "fp64": b64["autotune"]},
        "sub_batches_on_streams": {
            prec: {S: {"evals_per_s": r["evals_per_s"], "p50_ms": r["p50_ms"],
                       "p99_ms": r["p99_ms"]} for S, r in b["streams"].items()}
            for prec, b in (("fp32", b32), ("fp64", b64))},
        "single_frame": {"fp32": {"evals_per_s": s32["evals_per_s"],
                                  "latency_ms": {"p50": s32["p50_ms"], "p99": s32["p99_ms"]}},
                         "fp64": {"evals_per_s": s64["evals_per_s"],
                                  "latency_ms": {"p50": s64["p50_ms"], "p99": s64["p99_ms"]}}},
        "fp64": {"value": b64["evals_per_s"], "ms_per_step": b64["ms_per_step"],
                 "latency_ms": {"p50": b64["p50_ms"], "p99": b64["p99_ms"], "per": "batch"},
                 "dtype": "f64",
                 "eval_roofline_frac": Be64 / (b64["mean_event_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "speedup_fp32_vs_fp64": b32["evals_per_s"] / b64["evals_per_s"],
        "speedup_fp32_batch_vs_fp64_single": b32["evals_per_s"] / s64["evals_per_s"],
        "fp32_error_vs_fp64_oracle": dict(vs_oracle(b32), over="frame 0"),
        "fp64_error_vs_fp64_oracle": dict(vs_oracle(b64), over="frame 0"),
        "fp32_error_vs_fp64_engine": {"energy_rel_max": eE, "virial_rel_max": eVir,
                                      "force_rel_rms": eF, "potential_rel_rms": ePhi,
                                      "over": "every frame's energy and virial (max "
                                              "component error over the largest), frame "
                                              "0's forces and potentials"},
        "fp32_influence_option": {
            "value": b32i["evals_per_s"], "ms_per_step": b32i["ms_per_step"],
            "latency_ms": {"p50": b32i["p50_ms"], "p99": b32i["p99_ms"], "per": "batch"},
            "what": "SFAC_FP32_INFLUENCE: the closed-form influence and the per-thread "
                    "energy/virial sums in FP32 (the rest as the FP32 engine)",
            "error_vs_fp64_oracle": dict(vs_oracle(b32i), over="frame 0"),
            "kernels_avg_us": {k: round(v["avg_us"], 3) for k, v in b32i["kernels"].items()}},
        "eval_roofline": {"algorithmic_bytes_per_step": Be32,
                          "achieved_GBs": Be32 / (b32["mean_event_ms"] * 1e-3) / 1e9,
                          "frac": Be32 / (b32["mean_event_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "roofline": roofline,
        "kernels_fp32_batch_avg_us": {k: round(v["avg_us"], 3) for k, v in ks.items()},
        "kernels_fp64_batch_avg_us": {k: round(v["avg_us"], 3)
                                      for k, v in b64["kernels"].items()},
        "uneven_batch": uneven,
        "large_box": large,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
