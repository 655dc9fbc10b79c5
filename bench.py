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
reference path, one full step on the host cores), the served p50/p99 latency
of the continuous-batching server under Poisson arrivals, the throughput
of a 512-system batch, and the 1,000-water and 1.0M-charge configurations on
both paths; every served answer, every batched system and every
configuration is checked against the direct FP64 sum of its own atoms, and
the run exits with status 1 when any error is above the metric's 1e-5 bar.
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
# Served stream and batched ensemble: water boxes of 100-10,000 waters at
# rcut 9 A.  tol 2e-5 keeps every such box at <= 4e-6 relative force error
# against the direct sum (oracle sweep over 300-10,000 waters); tol 1e-4
# gives 1-2 % there, because the radial-order test of the reference accepts
# an order at an accidental zero of its error for these cell edges.
SERVE_RCUT, SERVE_TOL = 9.0, 2e-5
ERR_BAR = 1e-5  # the metric's relative force error bar
PEAK_FP32_VALU_TFLOPS = 157.3   # MI355X published FP32 vector peak
PEAK_FP64_VALU_TFLOPS = 78.6    # MI355X published FP64 vector peak
PEAK_HBM_GBS = 8000.0
STAGES = ["bin", "sort", "sf", "rmt", "lr", "lr_reduce", "sr"]
# Schedules (overlap_near, lr_blocks, rmt_shape, pipeline, graph) the
# engine's autotuner chose for this workload on MI355X, pinned here: the
# mixed path's autotuner is refused by the engine (glst_autotune), and with
# the schedule pinned the result does not depend on a tuning pass.  For the
# mixed path eager steps and graph replay are within about 2 % and trade
# places from box to box, so both are timed and the faster is timed again
# for the result (time_path's pick_graph); both are reported.
MIXED_SCHEDULE = (2, 8192, 4, 1, 0)
FP64_SCHEDULE = (1, 16384, 0, 1, 0)


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


def error_report(got, ref):
    """Per-atom errors against the direct sum: the reference's report
    (utils/utils.cpp:60-168: absolute min/max/average/RMS of each
    component's error, each also normalised by the RMS of the GLST
    component) and the distribution of the per-atom relative force error
    |F_i - F_i^ref| / |F_i^ref|."""
    import numpy as np
    out = {}
    for name, gc, rc in zip(("fx", "fy", "fz", "en"), got, ref):
        e = np.abs(rc - gc)
        rms = float(np.sqrt((gc ** 2).mean()))
        st = {"abs_min": float(e.min()), "abs_max": float(e.max()),
              "avg": float(e.mean()), "rms": float(np.sqrt((e ** 2).mean()))}
        st.update({k + "_norm": v / rms for k, v in list(st.items())})
        out[name] = st
    dF = np.sqrt(sum((a - b) ** 2 for a, b in zip(got[:3], ref[:3])))
    Fr = np.sqrt(sum(b ** 2 for b in ref[:3]))
    rel = dF / Fr
    out["per_atom_rel_force_err"] = {
        "p50": float(np.percentile(rel, 50)),
        "p99": float(np.percentile(rel, 99)),
        "max": float(rel.max()),
        "argmax_force_over_rms": float(Fr[rel.argmax()] /
                                       np.sqrt((Fr ** 2).mean())),
    }
    out["max_abs_force_err_over_rms_force"] = float(
        dF.max() / np.sqrt((Fr ** 2).mean()))
    return out


def time_path(g, torch, d, natom, ct, steps, warmup, schedule,
              profile=False, pick_graph=False):
    """Sets `schedule` (overlap_near, lr_blocks, rmt_shape[, pipeline[,
    graph]]), then times `steps` steps after `warmup`: as scheduled (the
    result), eagerly and as a replayed graph (both reported), eagerly with
    per-stage events (the stage breakdown; graph replay records no
    per-kernel events), and eagerly with the near field and the far-field
    stages one after another on one stream ("kernel_ms": each stage's events
    then bracket only its own kernels).  `profile`: only the scheduled and
    the stage runs.  `pick_graph`: after both modes are timed, the faster
    one is timed once more and that run is the result (the two are within
    about 2 % and trade places from box to box)."""
    f = g.GLSTForce(natom, TOL, BOX, rcut=RCUT, ct=ct)
    keys = ("overlap_near", "lr_blocks", "rmt_shape", "pipeline", "graph")
    for k, v in zip(keys, schedule):
        f.set_option(k, v)
    sched = {k: f.get_option(k) for k in keys}

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

    dt = run(sched["graph"], False)
    got = [t.cpu().numpy() for t in f.get_ef()]
    modes = {sched["graph"]: dt}
    if not profile:
        modes[1 - sched["graph"]] = run(1 - sched["graph"], False)
        if pick_graph:
            sched["graph"] = min(modes, key=modes.get)
            dt = run(sched["graph"], False)
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
            "sched": sched}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(h, natom):
    """One full FP64 step of the oracle's C restatement of the reference
    path (OpenMP over the host cores) on the same 288k-charge box, timed per
    stage with the reference driver's labels."""
    from oracle import oracle as o
    threads = min(16, os.cpu_count() or 1)
    o.set_threads(threads)
    t0 = time.perf_counter()
    o.glst(*h, TOL, BOX, RCUT)
    t_step = time.perf_counter() - t0
    stages = o.last_stage_times()
    return {
        "value": natom / t_step,
        "unit": "atom-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"one full step of the oracle's FP64 C restatement of the "
                   f"reference path on the same 288,000-charge box (tol "
                   f"{TOL}): {t_step:.1f} s; {threads} OpenMP threads of "
                   f"{os.cpu_count()} host CPUs ({cpu_model()})"),
        "seconds_per_step": t_step,
        "stages_s": stages,
        "host_cpus": os.cpu_count(),
        "cpu_model": cpu_model(),
    }


def _water_mix(g, np, rng, n, seed0):
    """n water boxes of round(exp(U(ln 100, ln 10000))) waters at 0.0334
    waters/A^3: (L, host arrays)."""
    out = []
    for k in range(n):
        nw = int(round(np.exp(rng.uniform(np.log(100), np.log(10000)))))
        L = (nw / 0.0334) ** (1 / 3)
        out.append((L, g.make_water_box(nw, L, seed=seed0 + k)))
    return out


def _max_err_vs_direct(g, torch, cases):
    """Largest relative force error of (inputs, outputs) pairs against the
    direct FP64 all-pairs sum of the same inputs (HIP kernel), and the count
    checked."""
    worst = 0.0
    for ins, outs in cases:
        ref = g.coulomb_direct(*ins)
        num = sum(float(((a - b) ** 2).sum()) for a, b in zip(outs[:3], ref[:3]))
        den = sum(float((b ** 2).sum()) for b in ref[:3])
        worst = max(worst, (num / den) ** 0.5)
    return worst, len(cases)


def serve_bench(g, torch, np, n_requests=2000, seed=5, max_batch=8,
                fused=0, graph=None):
    """Served latency of the continuous-batching server (mixed path).

    Requests are systems of the batch size mix: waters log-uniform in
    [100, 10000] at 0.0334 waters/A^3, rcut 9 A, tol 2e-5.  30 % of the
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
    included).  Every request of the Poisson stream has its own output
    arrays; afterwards each answer is compared with the direct FP64 sum of
    its own inputs.  `fused`: the batches' launch mode (glst_batch option
    "fused"); `graph`: the slots' graph mode (None: the server's default)."""
    rng = np.random.default_rng(seed)
    pool = []
    for L, h in _water_mix(g, np, rng, 48, 1000):
        d = [torch.as_tensor(a, dtype=torch.float64, device="cuda") for a in h]
        pool.append((L, d))
    resident, fresh = pool[:16], pool[16:]
    jitter = [0]

    def outputs(n):
        return [torch.empty(n, dtype=torch.float64, device="cuda")
                for _ in range(4)]

    def make_request():
        if rng.uniform() < 0.3:
            L, d = resident[rng.integers(len(resident))]
            pos = [torch.clamp(x + 0.01 * torch.randn_like(x), min=0.0)
                   for x in d[:3]]
            return L, pos + [d[3]], outputs(d[0].numel())
        L, d = fresh[rng.integers(len(fresh))]
        jitter[0] += 1  # a new shape: its setup is built at submission
        return L * (1.0 + 1e-9 * jitter[0]), d, outputs(d[0].numel())

    srv = g.GLSTServer(SERVE_TOL, 64 << 30, max_batch=max_batch, ct="float")
    if fused:
        srv.set_option("fused", fused)
    if graph is not None:
        srv.set_option("graph", graph)
    for L, d in resident:  # resident systems are set up once
        srv.submit(L, SERVE_RCUT, *d, *outputs(d[0].numel()))
    while srv.pending():
        srv.step()
    srv.completed(1 << 16)
    torch.cuda.synchronize()
    # burst capacity: 256 requests of the mix, back to back (large batches)
    burst = [make_request() for _ in range(256)]
    t0 = time.perf_counter()
    for L, d, o in burst[:32]:
        srv.submit(L, SERVE_RCUT, *d, *o)
    for L, d, o in burst[32:]:  # arriving while the first batches run
        srv.submit(L, SERVE_RCUT, *d, *o)
        srv.poll()
    while srv.pending():
        srv.step()
    burst_cap = len(burst) / (time.perf_counter() - t0)
    srv.completed(1 << 16)
    del burst
    # sustained capacity: a closed loop that keeps `depth` requests in the
    # server, so batches are as large as a loaded server forms them
    depth, n_loop = 8, 512
    loop = [make_request() for _ in range(n_loop)]
    nxt = 0
    t0 = time.perf_counter()
    while nxt < depth:
        srv.submit(loop[nxt][0], SERVE_RCUT, *loop[nxt][1], *loop[nxt][2])
        nxt += 1
    answered = 0
    while answered < n_loop:
        answered += srv.step()
        while nxt < n_loop and srv.pending() < depth:
            L, d, o = loop[nxt]
            srv.submit(L, SERVE_RCUT, *d, *o)
            nxt += 1
    cap = n_loop / (time.perf_counter() - t0)
    srv.completed(1 << 16)
    del loop
    # Poisson arrivals at 75 % of the sustained capacity
    rate = 0.75 * cap
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    reqs = [make_request() for _ in range(n_requests)]
    torch.cuda.synchronize()
    st0 = srv.stats()
    lat, when = [], {}
    i = 0
    t0 = time.perf_counter()
    while len(lat) < n_requests:
        now = time.perf_counter() - t0
        while i < n_requests and arrivals[i] <= now:
            L, d, o = reqs[i]
            when[srv.submit(L, SERVE_RCUT, *d, *o)] = arrivals[i]
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
    # every served answer against the direct sum of its own inputs
    t_chk = time.perf_counter()
    worst, nchk = _max_err_vs_direct(
        g, torch, [(d, o) for _, d, o in reqs])
    t_chk = time.perf_counter() - t_chk
    lat = np.array(lat) * 1e3
    return {
        "workload": "Poisson arrivals of water systems, 100-10000 waters "
                    f"(log-uniform), rcut {SERVE_RCUT}, tol {SERVE_TOL}, "
                    "mixed; 30 % repeat steps of 16 resident systems; server "
                    "max_batch " + str(max_batch) + ", launch mode " +
                    str(fused),
        "tol": SERVE_TOL,
        "rcut": SERVE_RCUT,
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
        "max_rel_force_err": worst,
        "requests_checked": nchk,
        "check_seconds": t_chk,
        "meets_error_bar": worst <= ERR_BAR,
    }


def batch_bench(g, torch, np, nsys=512, steps=3, seed=4):
    """The batched ensemble: `nsys` water boxes of the size mix (waters
    log-uniform in [100, 10000], 0.0334 waters/A^3), rcut 9 A, tol 2e-5,
    packed into one batch on the mixed path; inputs resident on the device.
    Throughput over `steps` steps after one warm-up step; every system's
    forces are then compared with the direct FP64 sum of its own atoms."""
    rng = np.random.default_rng(seed)
    systems = _water_mix(g, np, rng, nsys, 100)
    natoms = [h[0].size for _, h in systems]
    cols = [torch.as_tensor(np.concatenate([h[i] for _, h in systems]),
                            device="cuda") for i in range(4)]
    t0 = time.perf_counter()
    b = g.GLSTBatch(natoms, [L for L, _ in systems], SERVE_RCUT, SERVE_TOL,
                    ct="float")
    t_setup = time.perf_counter() - t0
    b.zero_ef()
    b.calc_ener_force(*cols)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.zero_ef()
        b.calc_ener_force(*cols)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    got = b.get_ef()
    cases = []
    for s in range(nsys):
        o, n = b.offset(s), natoms[s]
        cases.append(([c[o:o + n] for c in cols], [t[o:o + n] for t in got]))
    worst, nchk = _max_err_vs_direct(g, torch, cases)
    b.close()
    return {
        "workload": f"{nsys} water boxes, 100-10000 waters (log-uniform, "
                    f"seed {seed}), rcut {SERVE_RCUT}, tol {SERVE_TOL}, "
                    "mixed, one batch",
        "systems": nsys,
        "atoms": int(sum(natoms)),
        "ms_per_step": dt * 1e3,
        "atom_steps_per_s": sum(natoms) / dt,
        "setup_s": t_setup,
        "max_rel_force_err": worst,
        "systems_checked": nchk,
        "meets_error_bar": worst <= ERR_BAR,
    }


def config_errors(g, torch, steps=5):
    """The two other single-system configurations, each on both paths with
    the default schedule: 1,000 waters (L 31.04 A, rcut 9, tol 1e-5) and the
    ionic solution of 333,000 waters and 898 Na+/Cl- pairs (1,000,796
    charges, L 215 A, rcut 12, tol 1e-5).  Time per step and relative force
    error against the direct FP64 sum of the same charges."""
    cases = [("1,000 waters, L 31.04 A, rcut 9 A, tol 1e-5", 1000, 0, 31.04,
              9.0, 1e-5),
             ("333,000 waters + 898 Na+/Cl- pairs, L 215 A, rcut 12 A, "
              "tol 1e-5", 333000, 898, 215.0, 12.0, 1e-5)]
    out = []
    for name, nw, nion, L, rc, tol in cases:
        h = g.make_water_box(nw, L, seed=1, nion=nion)
        d = [torch.as_tensor(a, dtype=torch.float64, device="cuda") for a in h]
        ref = g.coulomb_direct(*d)
        row = {"workload": name, "natom": int(h[0].size)}
        for ct in ("float", "double"):
            f = g.GLSTForce(h[0].size, tol, L, rcut=rc, ct=ct)
            f.zero_ef()
            f.calc_ener_force(*d)
            got = f.get_ef()
            num = sum(float(((a - b) ** 2).sum()) for a, b in zip(got[:3], ref[:3]))
            den = sum(float((b ** 2).sum()) for b in ref[:3])
            n = steps if ct == "float" else 2
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(n):
                f.zero_ef()
                f.calc_ener_force(*d)
            torch.cuda.synchronize()
            dt = (time.perf_counter() - t0) / n
            key = "mixed" if ct == "float" else "fp64"
            row[key] = {"ms_per_step": dt * 1e3,
                        "atom_steps_per_s": h[0].size / dt,
                        "rel_force_err": (num / den) ** 0.5,
                        "num_nodes": f.info().num_nodes}
            f.close()
        row["speedup_vs_fp64"] = (row["fp64"]["ms_per_step"] /
                                  row["mixed"]["ms_per_step"])
        out.append(row)
        del d, ref
        torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--fp64-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the served-latency and batch measurements")
    ap.add_argument("--schedule", default=None,
                    help="the mixed path's schedule in place of "
                         "MIXED_SCHEDULE: overlap_near,lr_blocks,rmt_shape"
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
              if args.schedule else MIXED_SCHEDULE)
    tm = time_path(g, torch, d, natom, "float", args.steps, args.warmup,
                   pinned, profile=args.profile,
                   pick_graph=args.schedule is None)
    # the step as run for the result: the pinned schedule, eager or
    # replayed as a HIP graph as it says (per-stage times come from eager
    # steps with events)
    dt_mixed, st_mixed, got_mixed = tm["dt"], tm["stages"], tm["got"]
    info, sched = tm["info"], tm["sched"]
    log(f"mixed: {dt_mixed * 1e3:.3f} ms/step  stages {st_mixed}")
    if args.profile:
        # calibration: reads 1 GiB and writes 1 GiB
        x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        torch.mul(x, 1.0, out=y)
        torch.cuda.synchronize()
        print(json.dumps({"profile": True, "value": natom / dt_mixed,
                          "ms_per_step": dt_mixed * 1e3,
                          "schedule": sched,
                          "stages_ms": st_mixed,
                          "calibration_bytes": {"read": 1 << 30,
                                                "write": 1 << 30}}),
              flush=True)
        return
    # the FP64 path at the schedule its autotuner chose
    t64 = time_path(g, torch, d, natom, "double", args.fp64_steps, 1,
                    FP64_SCHEDULE)
    dt_fp64, st_fp64, got_fp64 = t64["dt"], t64["stages"], t64["got"]
    sched64 = dict(t64["sched"], eager_ms_per_step=t64["eager"] * 1e3,
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
    # (atom, node) pair, the count of the reference's per-pair formulation,
    # over the K nodes of the cubature (the table's padding to a multiple of
    # 256 nodes is never visited).
    K = info.num_nodes
    pairs = natom * K
    lr_ms = tm["kernel"].get("lr", float("nan"))
    achieved = 20.0 * pairs / (lr_ms * 1e-3) / 1e12
    # HBM traffic is not measured in this run: it is read from the committed
    # summary of two rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE) over
    # the same kernel on this workload
    traffic, traffic_src = None, None
    tf = os.path.join(REPO, "profiles", "lr_ag_288k_traffic.json")
    if os.path.exists(tf):
        with open(tf) as fh:
            traffic = json.load(fh).get("bytes_per_step")
        traffic_src = ("profiles/lr_ag_288k_traffic.json: rocprofv3 --pmc "
                       "FETCH_SIZE and WRITE_SIZE of this kernel on this "
                       "workload, committed; not measured in this run")
    roof = {
        "kernel": "lr_ag_kernel<NR> (calc_lr_ef back-projection over atom "
                  "groups, all alpha-group launches of one step)",
        "bound": "valu",
        "achieved": achieved,
        "peak": PEAK_FP32_VALU_TFLOPS,
        "unit": "TFLOP/s",
        "frac": achieved / PEAK_FP32_VALU_TFLOPS,
        "traffic": traffic,
        "traffic_source": traffic_src,
        "avg_ms": lr_ms,
        "pairs": pairs,
        "pair_rate_per_s": pairs / (lr_ms * 1e-3),
    }

    # the FP64 path's back-projection (lr_f64_kernel, one launch per step)
    # at the same 20 flops per pair against the FP64 vector peak
    lr64_ms = t64["kernel"].get("lr", float("nan"))
    ach64 = 20.0 * pairs / (lr64_ms * 1e-3) / 1e12
    roof64 = {
        "kernel": "lr_f64_kernel (calc_lr_ef back-projection, FP64)",
        "bound": "valu",
        "achieved": ach64,
        "peak": PEAK_FP64_VALU_TFLOPS,
        "unit": "TFLOP/s",
        "frac": ach64 / PEAK_FP64_VALU_TFLOPS,
        "avg_ms": lr64_ms,
        "pairs": pairs,
    }

    configs = config_errors(g, torch)
    log(f"configs: {configs}")

    served = batched = None
    if not args.no_serve:
        served = serve_bench(g, torch, np)
        log(f"served: {served}")
        batched = batch_bench(g, torch, np)
        log(f"batched: {batched}")

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
        "per_atom_err": error_report(got_mixed, ref),
        "fp64_path": {"value": natom / dt_fp64, "ms_per_step": dt_fp64 * 1e3,
                      "steps": args.fp64_steps, "schedule": sched64,
                      "kernel_ms": t64["kernel"],
                      "roofline": roof64},
        "speedup_vs_fp64": dt_fp64 / dt_mixed,
        "schedule": dict(sched, eager_ms_per_step=tm["eager"] * 1e3,
                         graph_ms_per_step=tm["graph"] * 1e3),
        "stages_ms": st_mixed,
        "kernel_ms": tm["kernel"],
        "stages_ms_fp64": st_fp64,
        "roofline": roof,
        "cpu_baseline": cpu,
        "served": served,
        "batched": batched,
        "other_configs": configs,
    }
    print(json.dumps(line), flush=True)
    # a figure quoted above the metric's error bar is not a result
    over = [name for name, err in (("mixed", err_mixed), ("fp64", err_fp64))
            if err > ERR_BAR]
    over += [name for name, blk in (("served", served), ("batched", batched))
             if blk is not None and not blk["meets_error_bar"]]
    over += [f"{c['natom']} charges {k}" for c in configs
             for k in ("mixed", "fp64") if c[k]["rel_force_err"] > ERR_BAR]
    if over:
        log(f"bench: relative force error above {ERR_BAR} in {over}")
        sys.exit(1)


if __name__ == "__main__":
    main()
