#!/usr/bin/env python3
"""Benchmark of the GLST force step on one MI355X.

Workload: the 288,000-charge rigid-water box (96,000 SPC/E-geometry waters,
seed 1, L = 142.0 A, rcut = 10 A, tol = 1e-4), inputs resident in HBM.  A step
is what the reference driver times per iteration (test/full_example.cu:135-210):
clear the accumulators, then calc_ener_force (assign -> S(k) -> shell sums ->
back-projection -> near field -> comm), which ends with a stream
synchronisation.

Prints ONE JSON line: atom-steps/s of the mixed FP32/FP64 path (value), the
plain FP64 path beside it, the relative force error of both against the
direct FP64 all-pairs sum, the roofline of the dominant kernel (HIP events on
the handle's stream), a CPU baseline (the oracle's FP64 restatement of the
reference path timed on a bounded sample on the host cores) and the served
p50/p99 latency of the continuous-batching server under Poisson arrivals.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "gauss-legendre-spherical-t_amd"))
sys.path.insert(0, REPO)

NWATER, BOX, RCUT, TOL, SEED = 96000, 142.0, 10.0, 1e-4, 1
PEAK_FP32_VALU_TFLOPS = 157.3   # MI355X published FP32 vector peak
PEAK_HBM_GBS = 8000.0
STAGES = ["bin", "sort", "sf", "rmt", "lr", "lr_reduce", "sr"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rel_force_err(got, ref):
    import numpy as np
    num = sum(float(((g - r) ** 2).sum()) for g, r in zip(got[:3], ref[:3]))
    den = sum(float((r ** 2).sum()) for r in ref[:3])
    return (num / den) ** 0.5


def ref_style_errors(got, ref):
    """utils/utils.cpp:60-168: rms(err)/rms(F_glst) per component."""
    import numpy as np
    out = {}
    for name, g, r in zip(("fx", "fy", "fz", "en"), got, ref):
        out[name] = float(np.sqrt(((r - g) ** 2).mean()) / np.sqrt((g ** 2).mean()))
    return out


def time_path(g, torch, d, natom, ct, steps, warmup, tune=False,
              schedule=None, profile=False):
    """Picks the schedule (the engine's autotuner, which also decides between
    eager steps and graph replay, or `schedule` pinned), then times `steps`
    steps after `warmup`: as tuned (the result), eagerly and as a replayed
    graph (both reported), eagerly with per-stage events (the stage
    breakdown; graph replay records no per-kernel events), and eagerly with
    the near field and the far-field stages one after another on one stream
    ("kernel_ms": each stage's events then bracket only its own kernels).
    `profile`: only the tuned and the stage runs (a pinned schedule)."""
    f = g.GLSTForce(natom, TOL, BOX, rcut=RCUT, ct=ct)
    keys = ("overlap_near", "lr_blocks", "rmt_shape", "pipeline", "graph")
    if schedule is not None:  # a pinned schedule (profiling runs)
        for k, v in zip(keys, schedule):
            f.set_option(k, v)
    elif tune:
        # 6 steps per candidate: with 3 the eager/graph decision followed
        # run-to-run noise of about 1 %
        f.autotune(*d, reps=6)
    tuned = {k: f.get_option(k) for k in keys}

    def run(graph, timing):
        f.set_option("graph", graph)
        f.set_timing(timing)
        for _ in range(warmup):
            f.zero_ef()
            f.calc_ener_force(*d)
        torch.cuda.synchronize()
        f.reset_timing()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            f.zero_ef()
            f.calc_ener_force(*d)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    def stage_times():
        out = {}
        for s in STAGES:
            ms, n = f.kernel_time(s)
            if n:
                out[s] = ms / n
        return out

    dt = run(tuned["graph"], False)
    got = [t.cpu().numpy() for t in f.get_ef()]
    modes = {tuned["graph"]: dt}
    if not profile:
        modes[1 - tuned["graph"]] = run(1 - tuned["graph"], False)
    run(0, True)
    stages = stage_times()
    kernel = stages
    if not profile:
        f.set_option("overlap_near", 0)
        f.set_option("pipeline", 0)
        run(0, True)
        kernel = stage_times()
    info = f.info()
    f.close()
    return {"dt": dt, "eager": modes.get(0), "graph": modes.get(1),
            "stages": stages, "kernel": kernel, "got": got, "info": info,
            "tuned": tuned}


def cpu_baseline(h, natom):
    """Oracle (C, OpenMP) on a bounded sample: assign + near field in full,
    the long-range part on every `stride`-th node, extrapolated."""
    import numpy as np
    from oracle import oracle as o
    threads = min(16, os.cpu_count() or 1)
    o.set_threads(threads)
    t0 = time.perf_counter()
    o.glst(*h, TOL, BOX, RCUT, far=False, near=True)
    t_near = time.perf_counter() - t0
    stride = 16
    t0 = time.perf_counter()
    o.glst(*h, TOL, BOX, RCUT, node_stride=stride, far=True, near=False)
    t_far = (time.perf_counter() - t0) * stride
    t_step = t_near + t_far
    return {
        "value": natom / t_step,
        "unit": "atom-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle FP64 restatement of the reference path, 1 step: "
                   f"assign+near field in full ({t_near:.2f} s), long range on "
                   f"every {stride}th node ({t_far / stride:.2f} s) x{stride}; "
                   f"est. {t_step:.1f} s/step"),
    }


def serve_bench(g, torch, np, n_requests=2000, seed=5, max_batch=8):
    """Served latency of the continuous-batching server (mixed path).

    Requests are systems of the batch size mix: waters log-uniform in
    [100, 10000] at 0.0334 waters/A^3, rcut 9 A, tol 1e-4.  30 % of the
    requests are repeat MD steps of 16 resident systems (positions perturbed
    by N(0, 0.01 A) per request); the rest draw from 32 other systems with
    the box jittered per request (their grid setups come from the server's
    cache).  The server computes at most `max_batch` requests per batch
    (8: with no limit one batch takes every queued request and the second
    slot idles; at most 8 keeps both slots busy and bounds a batch's time).
    Arrivals are Poisson at 75 % of the sustained capacity: the throughput
    of a closed loop that keeps 8 requests in the server, so that batches
    are as large as a loaded server forms them (a back-to-back burst forms
    larger batches; its rate is reported beside).  Latency =
    completion time - scheduled arrival time (queueing and setup
    included)."""
    rng = np.random.default_rng(seed)
    pool = []
    for k in range(48):
        nw = int(round(np.exp(rng.uniform(np.log(100), np.log(10000)))))
        L = (nw / 0.0334) ** (1 / 3)
        h = g.make_water_box(nw, L, seed=1000 + k)
        d = [torch.as_tensor(a, dtype=torch.float64, device="cuda") for a in h]
        o = [torch.empty(h[0].size, dtype=torch.float64, device="cuda")
             for _ in range(4)]
        pool.append((L, d, o))
    resident, fresh = pool[:16], pool[16:]
    jitter = [0]

    def make_request():
        if rng.uniform() < 0.3:
            L, d, o = resident[rng.integers(len(resident))]
            pos = [torch.clamp(x + 0.01 * torch.randn_like(x), min=0.0)
                   for x in d[:3]]
            return L, pos + [d[3]], o
        L, d, o = fresh[rng.integers(len(fresh))]
        jitter[0] += 1  # a new shape: its setup is built at submission
        return L * (1.0 + 1e-9 * jitter[0]), d, o

    srv = g.GLSTServer(TOL, 64 << 30, max_batch=max_batch, ct="float")
    for L, d, o in resident:  # resident systems are set up once
        srv.submit(L, 9.0, *d, *o)
    while srv.pending():
        srv.step()
    srv.completed(1 << 16)
    torch.cuda.synchronize()
    # burst capacity: 256 requests of the mix, back to back (large batches)
    burst = [make_request() for _ in range(256)]
    t0 = time.perf_counter()
    for L, d, o in burst[:32]:
        srv.submit(L, 9.0, *d, *o)
    for L, d, o in burst[32:]:  # arriving while the first batches run
        srv.submit(L, 9.0, *d, *o)
        srv.poll()
    while srv.pending():
        srv.step()
    burst_cap = len(burst) / (time.perf_counter() - t0)
    srv.completed(1 << 16)
    # sustained capacity: a closed loop that keeps `depth` requests in the
    # server, so batches are as large as a loaded server forms them
    depth, n_loop = 8, 512
    loop = [make_request() for _ in range(n_loop)]
    nxt = 0
    t0 = time.perf_counter()
    while nxt < depth:
        srv.submit(loop[nxt][0], 9.0, *loop[nxt][1], *loop[nxt][2])
        nxt += 1
    answered = 0
    while answered < n_loop:
        answered += srv.step()
        while nxt < n_loop and srv.pending() < depth:
            L, d, o = loop[nxt]
            srv.submit(L, 9.0, *d, *o)
            nxt += 1
    cap = n_loop / (time.perf_counter() - t0)
    srv.completed(1 << 16)
    # Poisson arrivals at 75 % of the sustained capacity
    rate = 0.75 * cap
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    reqs = [make_request() for _ in range(n_requests)]
    st0 = srv.stats()
    lat, when = [], {}
    i = 0
    t0 = time.perf_counter()
    while len(lat) < n_requests:
        now = time.perf_counter() - t0
        while i < n_requests and arrivals[i] <= now:
            L, d, o = reqs[i]
            when[srv.submit(L, 9.0, *d, *o)] = arrivals[i]
            i += 1
        if srv.pending():
            # launch into free slots, retire finished batches; never waits,
            # so arrivals keep being admitted while batches run
            srv.poll()
            t = time.perf_counter() - t0
            for tk in srv.completed(1 << 16):
                lat.append(t - when.pop(tk))
        elif i < n_requests:
            time.sleep(max(0.0, arrivals[i] - (time.perf_counter() - t0)))
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    lat = np.array(lat) * 1e3
    return {
        "workload": "Poisson arrivals of water systems, 100-10000 waters "
                    "(log-uniform), rcut 9, tol 1e-4, mixed; 30 % repeat "
                    "steps of 16 resident systems; server max_batch "
                    + str(max_batch),
        "requests": n_requests,
        "rate_req_s": rate,
        "capacity_req_s": cap,
        "burst_capacity_req_s": burst_cap,
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "mean_batch": (st["completed"] - st0["completed"]) /
                      max(1, st["batches"] - st0["batches"]),
        "atom_steps_per_s": (st["atoms"] - st0["atoms"]) / wall,
        "new_setups": st["setups"] - st0["setups"],
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--fp64-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the served-latency measurement")
    ap.add_argument("--schedule", default=None,
                    help="pin the mixed path's schedule instead of "
                         "autotuning: overlap_near,lr_blocks,rmt_shape"
                         "[,pipeline[,graph]]")
    ap.add_argument("--profile", action="store_true",
                    help="mixed path only (for rocprofv3 runs), plus one "
                         "1 GiB elementwise copy used to calibrate counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs one system on one GPU; --gpus "
              f"{args.gpus} is not supported (node sharding across GPUs is "
              f"designed, not built)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import glst_amd as g

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    h = g.make_water_box(NWATER, BOX, seed=SEED)
    natom = h[0].size
    d = [torch.as_tensor(a, dtype=torch.float64, device="cuda") for a in h]
    log(f"bench: {natom} atoms, box {BOX}, rcut {RCUT}, tol {TOL}")

    pinned = ([int(v) for v in args.schedule.split(",")]
              if args.schedule else None)
    tm = time_path(g, torch, d, natom, "float", args.steps, args.warmup,
                   tune=True, schedule=pinned, profile=args.profile)
    # the step as run for the result: the autotuned schedule, eager or
    # replayed as a HIP graph as the tuner chose (per-stage times come from
    # eager steps with events)
    dt_mixed, st_mixed, got_mixed = tm["dt"], tm["stages"], tm["got"]
    info, tuned = tm["info"], tm["tuned"]
    log(f"mixed: {dt_mixed * 1e3:.3f} ms/step  stages {st_mixed}")
    if args.profile:
        # calibration: reads 1 GiB and writes 1 GiB
        x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        torch.mul(x, 1.0, out=y)
        torch.cuda.synchronize()
        print(json.dumps({"profile": True, "value": natom / dt_mixed,
                          "ms_per_step": dt_mixed * 1e3,
                          "schedule": tuned,
                          "stages_ms": st_mixed,
                          "calibration_bytes": {"read": 1 << 30,
                                                "write": 1 << 30}}),
              flush=True)
        return
    # the FP64 path gets the same autotuning
    t64 = time_path(g, torch, d, natom, "double", args.fp64_steps, 1,
                    tune=True)
    dt_fp64, st_fp64, got_fp64 = t64["dt"], t64["stages"], t64["got"]
    tuned64 = dict(t64["tuned"], eager_ms_per_step=t64["eager"] * 1e3,
                   graph_ms_per_step=t64["graph"] * 1e3)
    log(f"fp64 : {dt_fp64 * 1e3:.3f} ms/step  stages {st_fp64}")

    ref = [t.cpu().numpy() for t in g.coulomb_direct(*d)]
    err_mixed = rel_force_err(got_mixed, ref)
    err_fp64 = rel_force_err(got_fp64, ref)
    log(f"relative force error vs direct FP64 sum: mixed {err_mixed:.3e} "
        f"fp64 {err_fp64:.3e}")

    # Roofline of the dominant kernel, the back-projection (calc_lr_ef:
    # lr_ag_kernel<NR> over atom groups, one launch per alpha group; avg_ms
    # is the sum of the group launches of one step, HIP events on the
    # handle's stream around them, from the steps run one stage after
    # another).
    # Algorithmic work: 20 flops (sin and cos counted as one each) per
    # (atom, node) pair over the padded node table, the count of the
    # reference's per-pair formulation.
    K = info.num_nodes
    Kp = (K + 255) // 256 * 256
    pairs = natom * Kp
    lr_ms = tm["kernel"].get("lr", float("nan"))
    achieved = 20.0 * pairs / (lr_ms * 1e-3) / 1e12
    traffic = None
    tf = os.path.join(REPO, "profiles", "lr_ag_288k_traffic.json")
    if os.path.exists(tf):
        with open(tf) as fh:
            traffic = json.load(fh).get("bytes_per_step")
    roof = {
        "kernel": "lr_ag_kernel<NR> (calc_lr_ef back-projection over atom "
                  "groups, all alpha-group launches of one step)",
        "bound": "valu",
        "achieved": achieved,
        "peak": PEAK_FP32_VALU_TFLOPS,
        "unit": "TFLOP/s",
        "frac": achieved / PEAK_FP32_VALU_TFLOPS,
        "traffic": traffic,
        "avg_ms": lr_ms,
        "pair_rate_per_s": pairs / (lr_ms * 1e-3),
    }

    served = None
    if not args.no_serve:
        served = serve_bench(g, torch, np)
        log(f"served: {served}")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(h, natom)
        log(f"cpu baseline: {cpu}")

    line = {
        "metric": "atom-steps/s @288k atoms mixed vs FP64; served p99 ms; "
                  "rel force err <=1e-5",
        "value": natom / dt_mixed,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt_mixed * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded rigid-water box)",
        "config": {
            "workload": "288,000-charge water box (96,000 SPC/E-geometry "
                        "waters), L=142 A, rcut=10 A, tol=1e-4, mixed path",
            "natom": natom, "box": BOX, "rcut": RCUT, "tol": TOL,
            "ncell": list(info.ncell), "num_nodes": K,
            "parallelism": "single GPU",
        },
        "rel_force_err": err_mixed,
        "rel_force_err_fp64": err_fp64,
        "ref_style_err": ref_style_errors(got_mixed, ref),
        "fp64_path": {"value": natom / dt_fp64, "ms_per_step": dt_fp64 * 1e3,
                      "steps": args.fp64_steps, "schedule": tuned64},
        "speedup_vs_fp64": dt_fp64 / dt_mixed,
        "schedule": dict(tuned, eager_ms_per_step=tm["eager"] * 1e3,
                         graph_ms_per_step=tm["graph"] * 1e3),
        "stages_ms": st_mixed,
        "kernel_ms": tm["kernel"],
        "stages_ms_fp64": st_fp64,
        "roofline": roof,
        "cpu_baseline": cpu,
        "served": served,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
