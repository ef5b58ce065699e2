"""Full-size parity: the benchmark configurations at their own shape against
the fp32 CPU oracle, from step 0, so the initial synchronous volley (every
neuron starts at V0 = (V_th + V_reset) / 2 and many cross threshold
together) is inside the compared window.

- the 2M-neuron block (1e9 synapses), 1000 steps replayed from HIP graphs:
  per-population counts of every step, a sampled raster and V trace, the
  event and spike counters and the final state, bit for bit;
- the propagation stress case (500k neurons, in-degree 5000, 2.5e9
  synapses), 200 steps, with queues (and the pull) and with gather;
- the 64-instance ensemble with per-instance g_ui after warm-up, autotuning
  and 1000 graph-replayed steps: every instance's counters and final state
  against its own oracle run;
- a serving stream with sizes from 10k to 1M neurons and out-degrees up to
  1000: every retired request's counters and final V against the oracle run
  of the same seeded network.

The oracle runs in chunks of steps so that only the sampled columns of its
per-step outputs are kept.
"""
import importlib
import time

import numpy as np
import pytest

from test_gpu_parity import DT, assert_state_equal, bits

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wl(dtb):
    return importlib.import_module("digital_twin_brain-open_amd.workloads")


@pytest.fixture(scope="module")
def serving(dtb):
    return importlib.import_module("digital_twin_brain-open_amd.serving")


def log(*a):
    print(*a, flush=True)


def oracle_sim(orc, net, seed, mu, sigma, tau=4.0):
    on = orc.OracleNet.from_network(net, dt=DT, seed=seed)
    on.set_ou(on.pop_ids, mu, sigma, tau)
    return orc.Sim(on, 32)


def oracle_chunks(sim, net, steps, sample=None, chunk=50):
    """Runs `steps` steps; returns per-population counts [steps][npop], the
    fired count of every step, the sampled raster and V trace ([steps][m],
    or None), the synaptic events (sum of the fired rows' out-degrees) and
    the spikes."""
    deg = np.diff(net.rowptr)
    freq, fired, rs, vs = [], [], [], []
    events = spikes = 0
    for k0 in range(0, steps, chunk):
        m = min(chunk, steps - k0)
        out = sim.run(m, raster=True, vtrace=sample is not None, pop_counts=True)
        r = out["raster"]
        freq.append(out["freq"])
        fired.append(r.sum(axis=1, dtype=np.int64))
        idx = np.flatnonzero(r)
        events += int(deg[idx % net.n].sum(dtype=np.int64))
        spikes += int(idx.shape[0])
        del idx
        if sample is not None:
            rs.append(r[:, sample].copy())
            vs.append(out["V"][:, sample].copy())
        del out, r
    return {"freq": np.concatenate(freq), "fired": np.concatenate(fired),
            "raster": np.concatenate(rs) if sample is not None else None,
            "V": np.concatenate(vs) if sample is not None else None, "events": events, "spikes": spikes}


def test_block_2m_1000_steps_from_rest(dtb, orc, gpu, wl):
    steps = 1000
    t0 = time.time()
    net = wl.block_network()
    sample = np.arange(0, net.n, 101, dtype=np.uint32)
    log("block: %d neurons, %d synapses generated in %.1f s" % (net.n, net.nsyn, time.time() - t0))
    eng = dtb.Engine(0)
    try:
        blk = dtb.Block(net, delta_t=DT, seed=wl.BLOCK_SIM_SEED, engine=eng)
        blk.update_ou_background_stimuli(blk.subblk_id, wl.OU_TAU, wl.BLOCK_MU, wl.BLOCK_SIGMA)
        blk.set_samples(sample)
        t0 = time.time()
        out = blk.run(steps, freqs=True, output_sample_spike=True, output_sample_vmemb=True, graph=True)
        st = blk.state()
        cnt = blk.counters()
        stats = eng.propagation_stats()
        pull = eng.pull_info(blk.inst)
        eng.check()
        log("gpu %.1f s, propagation %s, pull %s" % (time.time() - t0, stats, pull))
    finally:
        eng.close()
    t0 = time.time()
    sim = oracle_sim(orc, net, wl.BLOCK_SIM_SEED, wl.BLOCK_MU, wl.BLOCK_SIGMA, wl.OU_TAU)
    ref = oracle_chunks(sim, net, steps, sample)
    log("oracle %.1f s; fired per step: max %d at step %d, mean %.0f" % (
        time.time() - t0, ref["fired"].max(), int(ref["fired"].argmax()), ref["fired"].mean()))
    assert ref["fired"].max() > net.n // 4, "the block shows no initial volley"
    assert np.array_equal(out["freq"], ref["freq"]), "per-population counts differ"
    assert np.array_equal(out["spike"], ref["raster"]), "sampled raster differs"
    assert np.array_equal(bits(out["vmemb"]), bits(ref["V"])), "sampled V trace differs"
    assert cnt == (steps, ref["events"], ref["spikes"])
    assert_state_equal(st, sim)


@pytest.fixture(scope="module")
def stress_case(dtb, orc, wl):
    """The stress network and its oracle run of 200 steps (shared by the
    queue and gather tests)."""
    steps = 200
    t0 = time.time()
    net = wl.stress_network()
    log("stress: %d neurons, %d synapses generated in %.1f s" % (net.n, net.nsyn, time.time() - t0))
    sample = np.arange(0, net.n, 53, dtype=np.uint32)
    t0 = time.time()
    sim = oracle_sim(orc, net, wl.STRESS_SIM_SEED, wl.STRESS_MU, wl.STRESS_SIGMA, wl.OU_TAU)
    ref = oracle_chunks(sim, net, steps, sample)
    log("oracle %.1f s; fired per step: max %d at step %d, mean %.0f" % (
        time.time() - t0, ref["fired"].max(), int(ref["fired"].argmax()), ref["fired"].mean()))
    state = {"V": sim.st.V.copy(), "t": sim.st.t.copy(), "f": sim.st.f.copy(), "i_ou": sim.st.i_ou.copy(),
             "J": sim.st.J.copy(), "I_syn": sim.st.I_syn.copy()}
    return net, sample, steps, ref, state


@pytest.mark.parametrize("layout", ["auto", "gather"])
def test_stress_500k_in5000_from_rest(dtb, gpu, wl, stress_case, layout):
    net, sample, steps, ref, state = stress_case
    assert ref["fired"].max() > net.n // 2, "the stress case shows no synchronous volley"
    eng = dtb.Engine(0)
    try:
        eng.schedule(layout=layout)
        blk = dtb.Block(net, delta_t=DT, seed=wl.STRESS_SIM_SEED, engine=eng)
        blk.update_ou_background_stimuli(blk.subblk_id, wl.OU_TAU, wl.STRESS_MU, wl.STRESS_SIGMA)
        blk.set_samples(sample)
        t0 = time.time()
        out = blk.run(steps, freqs=True, output_sample_spike=True, output_sample_vmemb=True, graph=True)
        st = blk.state()
        cnt = blk.counters()
        stats = eng.propagation_stats()
        used = eng.schedule()[2]
        eng.check()
        log("%s (in use %s): gpu %.1f s, propagation %s" % (layout, used, time.time() - t0, stats))
    finally:
        eng.close()
    assert used == "gather" if layout == "gather" else used in ("dense", "spread")
    assert np.array_equal(out["freq"], ref["freq"]), "per-population counts differ"
    assert np.array_equal(out["spike"], ref["raster"]), "sampled raster differs"
    assert np.array_equal(bits(out["vmemb"]), bits(ref["V"])), "sampled V trace differs"
    assert cnt == (steps, ref["events"], ref["spikes"])
    for k in ("V", "i_ou", "J", "I_syn"):
        assert np.array_equal(bits(st[k]), bits(state[k])), k
    assert np.array_equal(st["t"], state["t"]) and np.array_equal(st["f"], state["f"])


def test_ensemble_64_instances_match_oracle(dtb, orc, gpu, wl):
    """All 64 instances in one engine: 100 eager steps, the autotuner (32
    trial steps per schedule), 1000 steps replayed from 64-step graphs; each
    instance against its own oracle run of the same total."""
    t0 = time.time()
    nets, sizes = wl.ensemble_networks()
    log("ensemble: %d neurons in %d instances generated in %.1f s" % (int(sizes.sum()), len(nets),
                                                                       time.time() - t0))
    eng = dtb.Engine(0)
    try:
        blks = []
        for i, net in enumerate(nets):
            b = dtb.Block(net, delta_t=DT, seed=wl.ENS_SIM_SEED + i, engine=eng)
            b.update_ou_background_stimuli(b.subblk_id, wl.OU_TAU, wl.ENS_MU, wl.ENS_SIGMA)
            eng.configure(b.inst, 1, True)
            blks.append(b)
        t0 = time.time()
        eng.run(100, graph_chunk=0)
        k, layout, _ = eng.autotune(32)
        eng.run(1000, graph_chunk=64)
        counters = [b.counters() for b in blks]
        states = [b.state() for b in blks]
        eng.check()
        log("gpu %.1f s, schedule %s/%d, propagation %s" % (time.time() - t0, layout, 64 * k,
                                                             eng.propagation_stats()))
    finally:
        eng.close()
    total = 100 + dtb.Engine.autotune_steps(32) + 1000
    t0 = time.time()
    for i, net in enumerate(nets):
        sim = oracle_sim(orc, net, wl.ENS_SIM_SEED + i, wl.ENS_MU, wl.ENS_SIGMA, wl.OU_TAU)
        ref = oracle_chunks(sim, net, total, chunk=250)
        assert counters[i] == (total, ref["events"], ref["spikes"]), "instance %d counters" % i
        assert_state_equal(states[i], sim)
        assert ref["spikes"] > 0, i
    log("oracle %.1f s for %d instances x %d steps" % (time.time() - t0, len(nets), total))


def test_serving_stream_matches_oracle(dtb, orc, gpu, serving):
    """Requests of 10k-1M neurons at out-degree 100-1000 admitted under a
    budget that keeps about three resident; every retired request's
    counters and final V equal the oracle run of its network (gen_fanout
    with the request's seed) for the request's steps."""
    shapes = [(1_000_000, 1000, 1000), (300_000, 700, 1500), (120_000, 400, 2000), (45_000, 900, 3000),
              (10_000, 100, 5000), (600_000, 150, 1200), (25_000, 550, 2500), (200_000, 1000, 1000)]
    reqs = [serving.Request(i, 0.001 * i, n, k, s, net_seed=5000 + i, seed=900 + i, mu=3.0)
            for i, (n, k, s) in enumerate(shapes)]
    budget = sum(sorted(r.device_bytes(pull=True) for r in reqs)[-3:])
    srv = serving.Server(device=0, hbm_budget=budget, chunk=100, keep_v=True)
    try:
        t0 = time.time()
        stats = srv.serve(reqs, realtime=False)
        log("served %d requests in %.1f s, mean resident %.2f" % (len(reqs), time.time() - t0,
                                                                   stats["mean_resident"]))
    finally:
        srv.close()
    assert stats["mean_resident"] > 1.0
    for r in reqs:
        t0 = time.time()
        net = dtb.gen_fanout(r.n, r.k, r.net_seed, g_ui=r.g_ui)
        sim = oracle_sim(orc, net, r.seed, r.mu, r.sigma, r.tau)
        ref = oracle_chunks(sim, net, r.steps, chunk=250)
        log("request %d (%d x %d, %d steps): oracle %.1f s, %d spikes, max fired per step %d" % (
            r.rid, r.n, r.k, r.steps, time.time() - t0, ref["spikes"], ref["fired"].max()))
        assert (r.events, r.spikes) == (ref["events"], ref["spikes"]), r.rid
        assert np.array_equal(bits(r.v), bits(sim.st.V)), r.rid
        del net, sim
