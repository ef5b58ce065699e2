"""Times the reference's CPU model (models/block_new.py, class `block`) on the
10k-neuron block and this project's C port of the step (oracle/, OpenMP) on
the same network, in a container that holds the reference (CPU only).
Writes tests/golden/block_new_timing.json; the numbers are quoted as the CPU
baseline of the 10k-neuron block.

Workload: 10,000 LIF neurons (80% E, 20% I), exactly 100 in-edges each
(gen_uniform seed 1), dt 0.1 ms, OU mean 1.5 (the literal 0.5 leaves the
block silent), sigma 0.1, tau 4 ms.  block_new.py holds the weights as a
dense [4, N, N] float32 tensor (1.6 GB) or as a torch.sparse_coo tensor (its
two variants, block_new.py:79-84).  50 warm-up steps, then 1000 timed steps,
median of 3 runs; torch and the port use the same thread count.

    python tests/golden/time_block_new.py [--threads 8]
"""
import argparse
import importlib
import json
import os
import platform
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_MODELS = "/root/reference/models"

N, K, NET_SEED, SEED = 10_000, 100, 1, 20240
DT, MU, SIGMA, TAU = 0.1, 1.5, 0.1, 4.0
WARMUP, STEPS, REPEATS = 50, 1000, 3


def cpu_name():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    args = ap.parse_args()
    import torch

    sys.path.insert(0, ROOT)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    dtb = importlib.import_module("digital_twin_brain-open_amd")
    orc = importlib.import_module("oracle")
    sys.path.insert(0, REF_MODELS)
    block_new = importlib.import_module("block_new")
    torch.set_num_threads(args.threads)

    net = dtb.gen_uniform(N, K, seed=NET_SEED)
    pre = np.repeat(np.arange(N), np.diff(net.rowptr))
    kind = net.kind_per_synapse().astype(np.int64)
    deg = np.diff(net.rowptr)
    idx = np.stack([np.concatenate([2 * kind, 2 * kind + 1]), np.concatenate([net.col, net.col]),
                    np.concatenate([pre, pre])])
    val = np.concatenate([net.w[:, 0], net.w[:, 1]])
    sparse_w = torch.sparse_coo_tensor(torch.from_numpy(idx), torch.from_numpy(val), (4, N, N)).coalesce()
    dense_w = sparse_w.to_dense()
    props = torch.from_numpy(net.props.copy())
    result = {"workload": "10000 LIF neurons (80/20 E/I), in-degree 100, dt 0.1 ms, OU mean 1.5, sigma 0.1, "
                          "tau 4 ms; %d warm-up + %d timed steps, median of %d runs" % (WARMUP, STEPS, REPEATS),
              "host": cpu_name(), "threads": args.threads, "torch": torch.__version__}

    def time_model(w):
        times, events = [], []
        for _ in range(REPEATS):
            m = block_new.block(props, w, delta_t=DT, i_mean=MU, i_sigma=SIGMA, tau_i=TAU)
            for _ in range(WARMUP):
                m.run(debug_mode=False)
            ev = 0
            t0 = time.perf_counter()
            for _ in range(STEPS):
                m.run(debug_mode=False)
                ev += int(deg[m.active.numpy()].sum())
            times.append(time.perf_counter() - t0)
            events.append(ev)
        k = int(np.argsort(times)[len(times) // 2])
        t = times[k]
        return {"s_per_step": t / STEPS, "realtime_factor": t / (STEPS * DT * 1e-3),
                "synaptic_events_per_s": events[k] / t, "events": events[k], "runs_s": times}

    result["block_new_dense"] = time_model(dense_w)
    print("dense", result["block_new_dense"], flush=True)
    del dense_w
    result["block_new_sparse"] = time_model(sparse_w)
    print("sparse", result["block_new_sparse"], flush=True)

    times, events = [], []
    for _ in range(REPEATS):
        on = orc.OracleNet.from_network(net, dt=DT, seed=SEED)
        on.set_ou(on.pop_ids, MU, SIGMA, TAU)
        sim = orc.Sim(on, 32, nthreads=args.threads)
        sim.run(WARMUP)
        t0 = time.perf_counter()
        out = sim.run(STEPS, raster=True)
        times.append(time.perf_counter() - t0)
        events.append(int((out["raster"].astype(np.int64) * deg[None, :]).sum()))
    k = int(np.argsort(times)[len(times) // 2])
    result["port"] = {"s_per_step": times[k] / STEPS, "realtime_factor": times[k] / (STEPS * DT * 1e-3),
                      "synaptic_events_per_s": events[k] / times[k], "events": events[k], "runs_s": times}
    print("port", result["port"], flush=True)
    with open(os.path.join(HERE, "block_new_timing.json"), "w") as f:
        json.dump(result, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
