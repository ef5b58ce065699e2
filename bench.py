"""Benchmark: synaptic events/s and realtime factor of the DTB single-block
step on one GPU.

Workload (one GPU's share of the whole-brain model): 2,000,000 LIF neurons
in 200 voxels (lognormal sizes, 80/20 E/I each), in-degree 500 (70% within
the voxel, 30% from a sparse lognormal voxel-to-voxel matrix at 5% density),
4 receptor channels, fp32 state and weights, dt = 0.1 ms, OU background
(mean 5.0, sigma 0.1, tau 4 ms, about 10 Hz), one parameter record per
population (400).  Synthetic, seeded.

A step is one launch of the step kernel (delta accumulation, neuron update
and spike binning) over the whole block with the network resident in HBM;
the timed steps are replayed from HIP graphs of 100 steps, as a run does.
Prints one JSON line.

The line also carries "ensemble": the data-assimilation ensemble (64
instances of 20k-200k neurons, log-uniform, in-degree 200, per-instance g_ui
uniform in 0.5-1.5 x the 10k-neuron block's, OU mean 3.0) stepped by one
launch per step for all instances ("batched") against the same kernels run
one instance at a time ("plain"), with every instance's counters and state
checked identical between the two; and "serving": a stream of 100 requests
(Poisson arrivals at 20/s, 10k-1M neurons, out-degree 100-1000, 100-2000 ms
of model time) served with continuous batching, with p50/p99 completion
latency, against its first 12 requests run alone one after the other; the
same 100 requests are also served as a backlog (all waiting at the start)
for the server's capacity.
"""
import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WL = importlib.import_module("digital_twin_brain-open_amd.workloads")
DT_MS = WL.DT_MS
OU_TAU = WL.OU_TAU
HBM_PEAK_GBS = 8000.0
GRAPH_CHUNK = 100
# The warm-up ends with the autotuner, in the firing regime that is timed:
# the schedules (2 queue-count layouts x 3 pass sizes, and gather) take turns
# in 4-step windows until each has run about TRIAL timed steps, the fastest
# total is kept; then GRAPH_CHUNK steps capture the graph.
AUTOTUNE_TRIAL = 32
AUTOTUNE_STEPS = 7 * (2 + (AUTOTUNE_TRIAL + 3) // 4 * 6)   # Engine.autotune_steps(AUTOTUNE_TRIAL)
# rocprofv3 FETCH_SIZE + WRITE_SIZE per launch, scaled by a device-copy
# calibration, of profiled runs: each summary names its window (workload,
# neurons, propagation layout, timed steps, warm-up steps, whether the
# autotuner ran) and is used only for a run of exactly that window.
PROFILES = ("k_step_voxel2M_in500.md", "k_step_voxel2M_in500_steps20_warmup5.md",
            "k_gather_k_step_uniform500k_in5000.md")

# Algorithmic HBM bytes of one step, as the step is defined on per-neuron
# arrays: per neuron 165 B (parameters 76 B: 7 scalars and g, V_u, decay of
# 4 channels; state 56 B read + write: V, refractory counter, i_ou, J x 4;
# delta accumulators 32 B read + reset; fired flag 1 B); per fired row 16 B
# (row offsets, fired-list entry); per delivered synapse 12 B (column index,
# weight pair).
ALGO_NEURON_BYTES, ALGO_ROW_BYTES, ALGO_SYN_BYTES = 165, 16, 12

# Bytes this layout has to move per step: per neuron 60 B (parameter class
# 2 B read; J 16 B, V, refractory counter, i_ou 3 x 4 B read + write; flag
# 1 B read + write; the shared parameter records and the delta accumulators
# stay on chip); per fired row 16 B (row offsets); per delivered synapse
# 36 B (column index + weight pair read, its 12-byte queue entry written and
# read).
LAYOUT_NEURON_BYTES = 2 + 2 * 16 + 2 * 12 + 2 * 1
LAYOUT_ROW_BYTES = 16
LAYOUT_SYN_BYTES = 12 + 2 * 12
# With gather propagation (k_gather, then k_step): per neuron the delta
# (16 B) is written by k_gather and read by k_step; per fired row 16 B (row
# offset, fired-list entry written and read) and its two 2-byte offsets for
# each group of 2048 neurons; per delivered synapse 12 B (column index +
# weight pair, no queue).
GATHER_NEURON_BYTES = LAYOUT_NEURON_BYTES + 2 * 16
GATHER_ROW_BYTES, GATHER_ROW_GROUP_BYTES = 16, 4
GATHER_SYN_BYTES = 12


def profiled_traffic(workload, neurons, layout, steps, warmup, autotuned):
    """(HBM bytes per launch measured with rocprofv3 counters over exactly
    this window (workload, size, layout, timed and warm-up steps, whether the
    autotuner ran), the summary that holds it), or (None, None)."""
    want = "window: workload=%s neurons=%d layout=%s steps=%d warmup=%d autotuned=%d" % (
        workload, neurons, layout, steps, warmup, int(autotuned))
    for name in PROFILES:
        path = os.path.join(ROOT, "profiles", name)
        if not os.path.exists(path):
            continue
        with open(path) as f:
            lines = [x.strip() for x in f]
        if want not in lines:
            continue
        for line in lines:
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1]), "profiles/" + name
    return None, None


def ensemble_bench(dtb, steps, warmup, chunk=64):
    """Batched (all instances in one launch per step, graph replay) vs plain
    (each instance alone on its own engine, same kernels, graph replay)."""
    t0 = time.perf_counter()
    nets, sizes = WL.ensemble_networks()
    log("ensemble: generated %d instances, %d neurons, %d synapses in %.1f s"
        % (len(nets), int(sizes.sum()), sum(x.nsyn for x in nets), time.perf_counter() - t0))

    def setup(eng, i, net):
        b = dtb.Block(net, delta_t=DT_MS, seed=WL.ENS_SIM_SEED + i, engine=eng)
        b.update_ou_background_stimuli(b.subblk_id, OU_TAU, WL.ENS_MU, WL.ENS_SIGMA)
        eng.configure(b.inst, 1, True)
        return b

    trial = 32
    tuned_steps = dtb.Engine.autotune_steps(trial)

    def tune(eng):
        # warm-up into the firing regime (the first ~300 steps are the
        # transient, whose costs rank the schedules differently), autotune
        # there, capture the graph
        if warmup >= tuned_steps + 2 * chunk:
            eng.run(warmup - tuned_steps - chunk, graph_chunk=0)
            eng.autotune(trial)
            eng.run(chunk, graph_chunk=chunk)
        else:
            eng.run(warmup, graph_chunk=chunk)
        eng.sync()
        k, _, layout = eng.schedule()
        return {"pass_synapses": 64 * k, "count_layout": layout, "autotuned": warmup >= tuned_steps + 2 * chunk}

    eng = dtb.Engine(0)
    blks = [setup(eng, i, net) for i, net in enumerate(nets)]
    schedule = tune(eng)
    c0 = [eng.counters(b.inst) for b in blks]
    t0 = time.perf_counter()
    eng.run(steps, graph_chunk=chunk)
    eng.sync()
    tb = time.perf_counter() - t0
    eng.check()
    c1 = [eng.counters(b.inst) for b in blks]
    vb = [b.state()["V"].view(np.uint32).copy() for b in blks]
    eng.close()
    ev_b = sum(c1[i][1] - c0[i][1] for i in range(len(blks)))

    tp, ev_p, same = 0.0, 0, True
    for i, net in enumerate(nets):
        e = dtb.Engine(0)
        b = setup(e, i, net)
        tune(e)
        s0 = e.counters(b.inst)
        t0 = time.perf_counter()
        e.run(steps, graph_chunk=chunk)
        e.sync()
        tp += time.perf_counter() - t0
        s1 = e.counters(b.inst)
        ev_p += s1[1] - s0[1]
        same = same and s1 == c1[i] and np.array_equal(b.state()["V"].view(np.uint32), vb[i])
        e.close()
    spikes = sum(c1[i][2] - c0[i][2] for i in range(len(blks)))
    return {"workload": WL.ensemble_description(sizes),
            "steps": steps, "warmup": warmup, "graph_chunk": chunk,
            "batched_schedule": schedule, "autotune": "every engine (batched, and each plain instance) "
                                                      "autotuned over %d warm-up steps" % tuned_steps,
            "batched_events_per_s": ev_b / tb, "plain_events_per_s": ev_p / tp, "speedup": tp / tb,
            "batched_ms_per_step": tb * 1e3 / steps, "plain_ms_per_step": tp * 1e3 / steps,
            "mean_rate_hz": spikes / (int(sizes.sum()) * steps * DT_MS * 1e-3),
            "identical_counters_and_state": bool(same)}


SERVE_REQUESTS, SERVE_RATE, SERVE_SEED, SERVE_PLAIN = 100, 20.0, 31, 12


def serving_bench(dtb, nreq, rate, plain, backlog=True):
    """Serving stream (Poisson arrivals at `rate`/s, networks of 10k-1M
    neurons at out-degree 100-1000 generated on the device, 100-2000 ms of
    model time each), continuous batching under the HBM budget; then the
    first `plain` requests again, one at a time, for the plain comparison."""
    serving = importlib.import_module("digital_twin_brain-open_amd.serving")
    reqs = serving.poisson_requests(nreq, rate, SERVE_SEED)
    srv = serving.Server(device=0)
    try:
        stats = srv.serve(reqs, realtime=True, progress=10.0)
    finally:
        srv.close()
    # The same requests as a backlog (all waiting at the start): the server's
    # capacity, without the stream's ramp-up and its last requests running
    # with few others resident.
    back, bst = None, None
    if backlog:
        back = serving.poisson_requests(nreq, rate, SERVE_SEED)
        srv = serving.Server(device=0)
        try:
            bst = srv.serve(back, realtime=False)
        finally:
            srv.close()
    t_plain, ev_plain, ns_plain, same = 0.0, 0, 0, True
    for r in reqs[:plain]:
        t0 = time.perf_counter()
        ev, sp, vh = serving.run_alone(r)
        t_plain += time.perf_counter() - t0
        ev_plain += ev
        ns_plain += r.n * r.steps
        same = same and (ev, sp, vh) == (r.events, r.spikes, r.v_hash)
    sizes = np.array([r.n for r in reqs])
    stats.update({"workload": "%d requests, Poisson %.0f/s, 10k-1M neurons log-uniform (mean %d), out-degree "
                              "100-1000, 100-2000 ms of model time, OU mu 3.0" % (nreq, rate, int(sizes.mean())),
                  "plain_requests": plain, "plain_events_per_s": ev_plain / t_plain,
                  "plain_neuron_steps_per_s": ns_plain / t_plain,
                  "speedup_vs_plain": stats["events_per_s"] / (ev_plain / t_plain),
                  "plain_same_results": bool(same)})
    if bst is not None:
        stats.update({"backlog_events_per_s": bst["events_per_s"], "backlog_wall_s": bst["wall_s"],
                      "backlog_mean_resident": bst["mean_resident"],
                      "speedup_vs_plain_backlog": bst["events_per_s"] / (ev_plain / t_plain),
                      "backlog_same_results": all((a.events, a.spikes, a.v_hash) == (b.events, b.spikes, b.v_hash)
                                                  for a, b in zip(reqs, back))})
    return stats


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_workload(name, neurons):
    """(network, OU mean, OU sigma, simulation seed, description) of a named workload."""
    if name == "block":
        net = WL.block_network(neurons)
        return net, WL.BLOCK_MU, WL.BLOCK_SIGMA, WL.BLOCK_SIM_SEED, WL.block_description(net)
    if name == "stress":
        net = WL.stress_network(neurons)
        return net, WL.STRESS_MU, WL.STRESS_SIGMA, WL.STRESS_SIM_SEED, WL.stress_description(net)
    raise ValueError(name)


def cpu_baseline(net, blk, steps, mu, sigma, seed):
    """The CPU oracle (a C port of the reference step, OpenMP) on this host,
    started from the GPU's state after warm-up so that it runs the same
    dynamics; returns events/s over `steps` steps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    orc = importlib.import_module("oracle")
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    on = orc.OracleNet.from_network(net, dt=DT_MS, seed=seed)
    on.set_ou(on.pop_ids, mu, sigma, OU_TAU)
    sim = orc.Sim(on, 32, nthreads=threads)
    st = blk.state()
    sim.load_state(st["V"], st["t"], st["f"], st["i_ou"], st["J"], st["I_syn"], st["step"])
    deg = np.diff(net.rowptr)
    t0 = time.perf_counter()
    out = sim.run(steps, raster=True)
    el = time.perf_counter() - t0
    events = int((out["raster"].astype(np.int64) * deg[None, :]).sum())
    return {"value": events / el, "unit": "synaptic events/s", "cores": threads, "kind": "port",
            "sample": "%d steps of the same %d-neuron network from the GPU state after warm-up "
                      "(%d events, %.1f s, s/step %.4f, RTF %.1f)" % (steps, net.n, events, el, el / steps,
                                                                     el / (steps * DT_MS * 1e-3))}


def copy_calibration(mib):
    """Device-to-device copies of `mib` MiB (HIP's copy kernel) for scaling the
    FETCH_SIZE / WRITE_SIZE counters of a profiled run to bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(n)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(n))
    hip.hipDeviceSynchronize()
    for _ in range(5):
        hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3)   # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes_per_copy": n, "copies": 5}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run device copies of this size (counter calibration)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="timed steps (default 10000; stress 2000)")
    ap.add_argument("--warmup", type=int, default=-1,
                    help="untimed steps (default 1000: the block passes its initial volley and settles after a "
                         "few hundred; the autotuner runs at the end)")
    ap.add_argument("--cpu-steps", type=int, default=1000, help="CPU oracle sample length (0 = skip)")
    ap.add_argument("--workload", choices=("block", "stress", "serving"), default="block",
                    help="block: the 2M-neuron block (headline); stress: 500k neurons at in-degree 5000; "
                         "serving: a request stream with continuous batching")
    ap.add_argument("--requests", type=int, default=SERVE_REQUESTS,
                    help="serving requests (the serving workload, and the block line's serving leg; 0 = skip "
                         "the leg)")
    ap.add_argument("--backlog", type=int, choices=(0, 1), default=1,
                    help="serving: also serve the same requests as a backlog, for the server's capacity")
    ap.add_argument("--neurons", type=int, default=0, help="override the workload's neuron count")
    ap.add_argument("--weights", choices=("float32", "float16", "uint8"), default="float32",
                    help="weight pairs kept on the device (uint8: codes round(255 w), as the reference's uint8 "
                         "blocks; 2 bytes per synapse instead of 8)")
    ap.add_argument("--ensemble-steps", type=int, default=5000,
                    help="ensemble comparison length, the ensemble configuration's 5000 steps (0 = skip)")
    ap.add_argument("--pull", type=int, choices=(0, 1), default=1,
                    help="pull propagation for synchronous volleys (in-edges grouped by target tile)")
    ap.add_argument("--defer-min", type=int, default=-1,
                    help="pull: emitted rows from which a tile defers them to the pull (default: the engine's)")
    ap.add_argument("--split", type=int, choices=(0, 1), default=0,
                    help="split binning: the fired rows' passes binned by a second launch per step")
    ap.add_argument("--autotune", type=int, choices=(0, 1), default=1,
                    help="0: keep the default schedule (128-synapse passes, dense counts) instead of autotuning, "
                         "so that separate profiled runs time the same kernels")
    args = ap.parse_args()
    if args.copy_calibration_mib > 0:
        copy_calibration(args.copy_calibration_mib)
        return

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this build steps one GPU's block; --gpus %d is not supported "
              "(multi-GPU sharding is design-only)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    dtb = importlib.import_module("digital_twin_brain-open_amd")
    if dtb.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)

    if args.workload == "serving":
        t0 = time.perf_counter()
        st = serving_bench(dtb, args.requests, SERVE_RATE, SERVE_PLAIN, bool(args.backlog))
        res = {"metric": "synaptic events/s (serving)", "value": st["events_per_s"], "unit": "synaptic events/s",
               "n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
               "data": "synthetic (seeded fan-out networks generated on the device)",
               "config": {"workload": st["workload"], "parallelism": "single GPU"}, "serving": st}
        log("serving %.1f s" % (time.perf_counter() - t0))
        print(json.dumps(res), flush=True)
        return
    if args.neurons == 0:
        args.neurons = WL.BLOCK_NEURONS if args.workload == "block" else WL.STRESS_NEURONS
    if args.steps <= 0:
        args.steps = 10000 if args.workload == "block" else 2000
    if args.warmup < 0:
        args.warmup = 1000
    t0 = time.perf_counter()
    net, mu, sigma, sim_seed, desc = make_workload(args.workload, args.neurons)
    if args.weights != "float32":
        net = net.with_weights(args.weights)
        desc += ", %s weights" % args.weights
    log("generated %d neurons / %d synapses in %.1f s" % (net.n, net.nsyn, time.perf_counter() - t0))
    t0 = time.perf_counter()
    eng = dtb.Engine(0)
    eng.split(bool(args.split))
    eng.pull(bool(args.pull), None if args.defer_min < 0 else args.defer_min)
    blk = dtb.Block(net, delta_t=DT_MS, seed=sim_seed, engine=eng)
    blk.update_ou_background_stimuli(blk.subblk_id, OU_TAU, mu, sigma)
    eng.configure(blk.inst, 1, True)
    log("uploaded in %.1f s (pull propagation %s)" % (time.perf_counter() - t0, eng.pull_info(blk.inst)))

    # warm-up: steps into the firing regime, the autotuner's trial steps,
    # then steps that also capture the graph the timed region replays
    k_best, table = None, {}
    if args.autotune and args.warmup >= AUTOTUNE_STEPS + 2 * GRAPH_CHUNK:
        eng.run(args.warmup - AUTOTUNE_STEPS - GRAPH_CHUNK, graph_chunk=0)
        k_best, _, table = eng.autotune(AUTOTUNE_TRIAL)
        eng.run(GRAPH_CHUNK, graph_chunk=GRAPH_CHUNK)
    else:
        eng.run(args.warmup, graph_chunk=GRAPH_CHUNK)
    eng.sync()
    pk, _, layout_used = eng.schedule()
    schedule = {"pass_synapses": 64 * pk, "count_layout": layout_used, "autotuned": bool(k_best),
                "trial_us_per_step": {"%s/%d" % (lay, 64 * kk): round(us, 2) for (lay, kk), us in table.items()}}
    step0, ev0, sp0 = eng.counters(blk.inst)
    prop0 = eng.propagation_stats()

    # Timed region: K steps (replayed from HIP graphs of GRAPH_CHUNK steps;
    # steps past the last whole chunk run eagerly), with HIP events on the
    # engine stream before the first step and after every step (eager) or
    # chunk (replayed): the step kernels and the gaps between them.
    t0 = time.perf_counter()
    ms_span, ms_steps = eng.run_span_steps(args.steps, graph_chunk=GRAPH_CHUNK)
    eng.sync()
    wall = time.perf_counter() - t0
    eng.check()
    step1, ev1, sp1 = eng.counters(blk.inst)
    prop1 = eng.propagation_stats()
    assert step1 - step0 == args.steps
    events = ev1 - ev0
    spikes = sp1 - sp0
    # The step kernel alone (events at its dispatch boundaries), after the
    # timed region: the figure rocprof's kernel trace reports.
    iso_steps = min(args.steps, 2000)
    us_iso = eng.time_kernels(iso_steps) * 1e3 / iso_steps
    value = events / wall
    rtf = wall / (args.steps * DT_MS * 1e-3)
    rate_hz = spikes / (net.n * args.steps * DT_MS * 1e-3)

    ev_step, sp_step = events / args.steps, spikes / args.steps
    us = ms_span * 1e3 / args.steps
    # a delivered synapse reads its column index and its weight pair as stored
    wpair = net.w.dtype.itemsize * 2
    syn_bytes = ALGO_SYN_BYTES - 8 + wpair
    algo_bytes = ALGO_NEURON_BYTES * net.n + ALGO_ROW_BYTES * sp_step + syn_bytes * ev_step
    gather = layout_used == "gather"
    if gather:
        ngroups = (net.n + 2047) // 2048
        layout_bytes = (GATHER_NEURON_BYTES * net.n + (GATHER_ROW_BYTES + GATHER_ROW_GROUP_BYTES * ngroups) * sp_step
                        + (GATHER_SYN_BYTES - 8 + wpair) * ev_step)
    else:
        layout_bytes = (LAYOUT_NEURON_BYTES * net.n + LAYOUT_ROW_BYTES * sp_step
                        + (LAYOUT_SYN_BYTES - 8 + wpair) * ev_step)
    kname = "k_gather+k_step" if gather else ("k_step+k_bin" if args.split else "k_step")
    achieved = algo_bytes / (us * 1e-6) / 1e9
    traffic, traffic_src = (profiled_traffic(args.workload, args.neurons, layout_used, args.steps, args.warmup,
                                             bool(k_best)) if args.weights == "float32" else (None, None))
    eager = args.steps < GRAPH_CHUNK

    res = {
        "metric": "synaptic events/s (%s)" % ("2M-neuron block" if args.workload == "block" else
                                              "propagation stress"),
        "value": value,
        "unit": "synaptic events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded %s network, random weights k*2^-24%s)" % (
            "voxel" if args.workload == "block" else "uniform",
            "" if args.weights == "float32" else ", stored as %s" % args.weights),
        "config": {"workload": desc, "neurons": net.n, "synapses": net.nsyn, "parallelism": "single GPU"},
        "realtime_factor": rtf,
        "mean_rate_hz": rate_hz,
        "events_per_step": ev_step,
        "kernel_us_per_step": {kname: us},
        "kernel_us_isolated": {kname: us_iso},
        "step_ms_max": float(ms_steps.max()),
        "step_ms": [round(float(x), 4) for x in ms_steps] if args.steps <= 50 else None,
        "steps_over_1ms": int((ms_steps > 1.0).sum()),
        "step_time_resolution": "one step (eager launches)" if eager else
                                "one graph of %d steps (its mean)" % GRAPH_CHUNK,
        "propagation": {"pull_tiles": prop1["pulled"] - prop0["pulled"],
                        "deferred_tiles": prop1["deferred"] - prop0["deferred"],
                        "synapses_past_queue_capacity": prop1["spilled"] - prop0["spilled"]},
        "schedule": schedule,
        "kernel_algo_bytes_per_step": {kname: algo_bytes},
        "roofline": {"kernel": kname, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algo_bytes_per_launch": algo_bytes,
                     "layout_bytes_per_launch": layout_bytes,
                     "layout_achieved": layout_bytes / (us * 1e-6) / 1e9,
                     "layout_frac": layout_bytes / (us * 1e-6) / 1e9 / HBM_PEAK_GBS,
                     "traffic_achieved": None if traffic is None else traffic / (us * 1e-6) / 1e9,
                     "traffic_source": traffic_src},
    }
    if args.cpu_steps > 0:
        t0 = time.perf_counter()
        res["cpu_baseline"] = cpu_baseline(net, blk, args.cpu_steps, mu, sigma, sim_seed)
        log("cpu baseline %.1f s" % (time.perf_counter() - t0))
    else:
        res["cpu_baseline"] = None
    # The reference's own CPU model (models/block_new.py, torch) beside this
    # port on the 10k-neuron block, timed where the reference is available
    # (it does not travel with this repository); recorded in a fixture.
    ref_timing = os.path.join(ROOT, "tests", "golden", "block_new_timing.json")
    if res["cpu_baseline"] is not None and os.path.exists(ref_timing):
        with open(ref_timing) as f:
            t = json.load(f)
        res["cpu_baseline"]["reference_cpu_model_10k"] = {
            "workload": t["workload"], "threads": t["threads"], "host": t["host"],
            "block_new_dense_s_per_step": t["block_new_dense"]["s_per_step"],
            "block_new_sparse_s_per_step": t["block_new_sparse"]["s_per_step"],
            "port_s_per_step": t["port"]["s_per_step"],
            "block_new_dense_events_per_s": t["block_new_dense"]["synaptic_events_per_s"],
            "port_events_per_s": t["port"]["synaptic_events_per_s"]}
    eng.close()
    del blk, net
    if args.ensemble_steps > 0 and args.workload == "block":
        t0 = time.perf_counter()
        res["ensemble"] = ensemble_bench(dtb, args.ensemble_steps, min(args.warmup, 800))
        log("ensemble %.1f s" % (time.perf_counter() - t0))
    if args.requests > 0 and args.workload == "block":
        t0 = time.perf_counter()
        res["serving"] = serving_bench(dtb, args.requests, SERVE_RATE, SERVE_PLAIN, bool(args.backlog))
        log("serving %.1f s" % (time.perf_counter() - t0))
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
