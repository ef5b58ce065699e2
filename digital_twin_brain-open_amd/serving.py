"""Serving loop: simulation requests admitted by HBM budget and stepped
together on one GPU (continuous batching).

A request is a seeded network (generated in device memory, Engine.add_fanout)
to be simulated for a number of steps.  Requests wait in arrival order; a
request is admitted as soon as it has arrived and its device footprint fits
the HBM budget.  All admitted requests are stepped by the same launches (one
launch per step covers every resident instance); each request leaves at
exactly its own step count, its result is read back and its memory is
returned, and waiting requests are admitted between chunks of steps.  A
request's trajectory does not depend on what else is resident: every
instance has its own step counter and random streams, and a change of the
resident set first folds every pending input exactly (Engine settle), so a
result is the same alone and in the mix.
"""
import collections
import ctypes
import sys
import time

import numpy as np

from ._lib import OUT_OF_MEMORY, DtbError
from .engine import DEFAULT_G_UI, Block, Engine, device_memory_cached, device_memory_reserve

# Device bytes per neuron of an instance: parameter class, V, i_ou,
# refractory counter, flag, J, pending delta, spill accumulators (2 x 32),
# blocked flag, row offset and per-wave counters; per synapse: column index
# and weight pair; the two propagation queues hold 12-byte entries, per tile
# min(S_t, QUEUE_FLOOR + S_t / 2^QUEUE_SHIFT) of its S_t synapses (the
# engine's default capacity).
BYTES_PER_NEURON = 2 + 4 + 4 + 4 + 1 + 16 + 16 + 64 + 1 + 8 + 1
BYTES_PER_SYNAPSE = 4 + 8
QUEUE_ENTRY_BYTES, QUEUE_FLOOR, QUEUE_SHIFT = 12, 4096, 3
GATHER_GROUP, FIRED_LISTS = 2048, 64
# Gather records: each synapse's column and weight pair side by side.
GATHER_RECORD_BYTES = 12
# Pull propagation: the in-edges (source | target | channel key and weight
# pair per synapse), per tile its first in-edge, per wave of 64 neurons two
# bitmap words per parity.
PULL_BYTES_PER_SYNAPSE = 4 + 8


class Request:
    __slots__ = ("rid", "arrival", "n", "k", "steps", "net_seed", "seed", "mu", "sigma", "tau", "g_ui",
                 "admitted", "finished", "events", "spikes", "v_hash", "v_mean", "v", "shed")

    def __init__(self, rid, arrival, n, k, steps, net_seed, seed, mu, sigma=0.1, tau=4.0, g_ui=DEFAULT_G_UI):
        self.rid, self.arrival, self.n, self.k, self.steps = rid, arrival, int(n), int(k), int(steps)
        self.net_seed, self.seed, self.mu, self.sigma, self.tau = int(net_seed), int(seed), mu, sigma, tau
        self.g_ui = tuple(g_ui)
        self.admitted = self.finished = None
        self.events = self.spikes = None
        self.v_hash = self.v_mean = None
        self.v = None
        self.shed = False   # refused by admission control (waited past the server's limit)

    def device_bytes(self, gather=False, pull=False):
        """Device footprint; gather: also the gather buffers (2-byte offsets
        per row and group of 2048 neurons, 64 fired lists per parity, a
        12-byte record per synapse); pull
        (without gather): also the in-edges of the pull propagation."""
        syn = self.n * self.k
        tiles = (self.n + 255) // 256
        queue = min(syn, tiles * QUEUE_FLOOR + (syn >> QUEUE_SHIFT))   # bound on the sum over tiles
        b = self.n * BYTES_PER_NEURON + syn * BYTES_PER_SYNAPSE + 2 * QUEUE_ENTRY_BYTES * queue
        if gather:
            groups = (self.n + GATHER_GROUP - 1) // GATHER_GROUP
            fcap = (tiles + FIRED_LISTS - 1) // FIRED_LISTS * 256
            b += self.n * (groups + 1) * 2 + 2 * FIRED_LISTS * (fcap + 32) * 4 + syn * GATHER_RECORD_BYTES
        elif pull:
            b += syn * PULL_BYTES_PER_SYNAPSE + (tiles + 1) * 8 + 4 * ((self.n + 63) // 64) * 8 + 16
        return b

    def latency(self):
        return self.finished - self.arrival


def poisson_requests(count, rate, seed, n_range=(10_000, 1_000_000), k_range=(100, 1000),
                     duration_ms=(100.0, 2000.0), dt=0.1, mu=3.0):
    """Seeded request stream: Poisson arrivals at `rate` per second, sizes
    log-uniform in n_range, out-degree uniform in k_range, duration uniform
    in duration_ms of model time."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, count))
    n = np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]), count)).astype(np.int64)
    k = rng.integers(k_range[0], k_range[1] + 1, count)
    steps = np.round(rng.uniform(duration_ms[0], duration_ms[1], count) / dt).astype(np.int64)
    return [Request(i, float(t[i]), n[i], k[i], steps[i], net_seed=int(seed) * 1_000_003 + i,
                    seed=int(seed) * 7_919 + i, mu=mu) for i in range(count)]


def device_free_bytes(device=0):
    """Free device memory reported by the HIP runtime."""
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    hip.hipSetDevice(device)
    if hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) != 0:
        raise RuntimeError("hipMemGetInfo failed")
    return free.value


class Server:
    """Continuous batching of simulation requests on one engine."""

    def __init__(self, device=0, hbm_budget=None, chunk=100, dt=0.1, reserve=True, layout="auto", pull=True,
                 keep_v=False, pass_k=1):
        """layout: the engine's propagation layout (Engine.schedule); with
        "gather" every request propagates by gather and its footprint counts
        the gather buffers.  pass_k: synapses per lane in a binning pass;
        the fan-out networks of requests (uniformly random targets) make
        about one queue segment per synapse, where passes of one synapse per
        lane are fastest (100 requests served as a backlog: 1.21e10 events/s
        against 1.13e10 with two per lane and 0.99e10 with four).  Results
        never depend on either.  pull: requests get the in-edges of the pull
        propagation (their initial synchronous volley then costs one pass
        over the synapses instead of overflowing queues; 12 more bytes per
        synapse).  keep_v: each retired request keeps its final V
        (Request.v)."""
        self.engine = Engine(device)
        self.engine.schedule(pass_k=pass_k, layout=layout)
        self.engine.pull(pull)
        self.gather = layout == "gather"
        self.pull = bool(pull)
        self.keep_v = bool(keep_v)
        # the default budget counts the device memory the library already
        # holds unused (a previous server's reservation, removed instances)
        # as available, not only what the HIP runtime reports free
        self.budget = int(hbm_budget if hbm_budget is not None else
                          0.85 * (device_free_bytes(device) + device_memory_cached(device)))
        self.chunk = int(chunk)
        self.dt = dt
        if reserve:
            # one device allocation for the whole budget: admitting and
            # retiring requests then carves and returns ranges of it
            device_memory_reserve(self.budget, device)

    def close(self):
        self.engine.close()

    def _admit(self, r):
        eng = self.engine
        inst = eng.add_fanout(r.n, r.k, r.net_seed, g_ui=r.g_ui, dt=self.dt, seed=r.seed)
        b = Block(None, delta_t=self.dt, engine=eng, inst=inst)
        b.update_ou_background_stimuli(b.subblk_id, r.tau, r.mu, r.sigma)
        eng.configure(inst, 1, True)
        return b

    def _retire(self, r, b):
        eng = self.engine
        step, events, spikes = eng.counters(b.inst)
        assert step == r.steps
        V = eng.read_state(b.inst, r.n)["V"]
        r.events, r.spikes = events, spikes
        r.v_hash = int(np.bitwise_xor.reduce(V.view(np.uint32)))
        r.v_mean = float(V.astype(np.float64).mean())
        if self.keep_v:
            r.v = V
        eng.remove(b.inst)

    def serve(self, requests, realtime=True, max_resident=None, progress=None, shed_after=None):
        """Runs the stream to completion; arrival times are honoured on the
        wall clock when realtime (otherwise every request has arrived at
        start).  Returns summary statistics; per-request results and times
        are filled into the Request objects.  progress: seconds between
        one-line reports on stderr (None: silent).  shed_after: admission
        control under overload: a request still waiting this many seconds
        after its arrival is refused (Request.shed) instead of queued on,
        which bounds the latency of the served ones (None: every request is
        served)."""
        for r in requests:
            if r.device_bytes(self.gather, self.pull) > self.budget:
                raise ValueError("request %d needs %d bytes, more than the HBM budget"
                                 % (r.rid, r.device_bytes(self.gather, self.pull)))
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        active = {}    # inst -> (request, block, remaining steps)
        used = 0
        t0 = time.monotonic()
        offset = 0.0 if realtime else min(r.arrival for r in requests)
        steps_run = 0
        resident_steps = 0
        t_admit = t_run = t_retire = 0.0
        done_count = 0
        t_report = time.monotonic()
        wait_retire = False   # the last admission found no device range: wait for memory to free up

        def admit_arrived():
            nonlocal used, t_admit, wait_retire
            now = time.monotonic() - t0 + (0.0 if realtime else 1e30)
            if shed_after is not None:
                # the oldest waiting requests first (arrival order)
                clock = time.monotonic() - t0 + offset
                while pending and clock - pending[0].arrival > shed_after:
                    pending.popleft().shed = True
            while (pending and not wait_retire and pending[0].arrival - offset <= now
                   and used + pending[0].device_bytes(self.gather, self.pull) <= self.budget
                   and (max_resident is None or len(active) + len(joining) < max_resident)):
                r = pending.popleft()
                ta = time.monotonic()
                try:
                    b = self._admit(r)
                except DtbError as e:
                    # the budget fits but no contiguous device range does
                    # (fragmentation): wait for retirements, unless nothing
                    # is resident that could free memory
                    t_admit += time.monotonic() - ta
                    if e.status != OUT_OF_MEMORY or not (active or joining):
                        raise
                    pending.appendleft(r)
                    wait_retire = True
                    break
                t_admit += time.monotonic() - ta
                r.admitted = time.monotonic() - t0 + offset
                joining[b.inst] = [r, b, r.steps]
                used += r.device_bytes(self.gather, self.pull)

        joining = {}   # admitted while a chunk runs: they step from the next chunk on
        while pending or active or joining:
            if progress is not None and time.monotonic() - t_report >= progress:
                t_report = time.monotonic()
                print("serving: %.0f s, %d done, %d resident, %d waiting" % (t_report - t0, done_count, len(active),
                                                                            len(pending)), file=sys.stderr, flush=True)
            active.update(joining)
            joining = {}
            if not active:
                admit_arrived()
                active.update(joining)
                joining = {}
                if not active:
                    if not pending:
                        continue   # the rest was shed
                    wait = pending[0].arrival - offset - (time.monotonic() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 0.05))
                    continue
            s = min(self.chunk, min(v[2] for v in active.values()))
            ta = time.monotonic()
            self.engine.run(s, graph_chunk=0)
            # Admissions are prepared while the chunk runs (host work, and
            # device work queued behind it on the engine stream).
            t_run += time.monotonic() - ta
            admit_arrived()
            ta = time.monotonic()
            self.engine.sync()
            t_run += time.monotonic() - ta
            steps_run += s
            resident_steps += s * len(active)
            done = []
            for inst, v in active.items():
                v[2] -= s
                if v[2] == 0:
                    done.append(inst)
            for inst in done:
                r, b, _ = active.pop(inst)
                ta = time.monotonic()
                self._retire(r, b)
                t_retire += time.monotonic() - ta
                r.finished = time.monotonic() - t0 + offset
                used -= r.device_bytes(self.gather, self.pull)
                done_count += 1
                wait_retire = False
        wall = time.monotonic() - t0
        served = [r for r in requests if not r.shed]
        lat = np.array([r.latency() for r in served]) if served else np.zeros(1)
        events = sum(r.events for r in served)
        spikes = sum(r.spikes for r in served)
        nsteps = max(sum(r.n * r.steps for r in served), 1)
        return {"requests": len(requests), "shed": len(requests) - len(served), "wall_s": wall, "events": int(events),
                "events_per_s": events / wall, "neuron_steps_per_s": nsteps / wall,
                "mean_rate_hz": spikes / (nsteps * self.dt * 1e-3),
                "latency_p50_ms": float(np.percentile(lat, 50) * 1e3),
                "latency_p99_ms": float(np.percentile(lat, 99) * 1e3),
                "mean_resident": resident_steps / max(steps_run, 1), "engine_steps": steps_run,
                "admit_s": t_admit, "step_s": t_run, "retire_s": t_retire}


def run_alone(request, device=0, dt=0.1):
    """The request on an engine of its own: (events, spikes, V hash)."""
    srv = Server(device=device, hbm_budget=1 << 62, chunk=max(1, request.steps), dt=dt, reserve=False)
    try:
        r = Request(request.rid, 0.0, request.n, request.k, request.steps, request.net_seed, request.seed,
                    request.mu, request.sigma, request.tau, request.g_ui)
        srv.serve([r], realtime=False)
        return r.events, r.spikes, r.v_hash
    finally:
        srv.close()
