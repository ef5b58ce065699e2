"""Pull propagation (in-edges grouped by target tile) against the CPU oracle.

A tile whose queue overflowed takes every emitted row through its in-edges,
and a tile with many emitted rows leaves them to its targets' pull.  Both are
forced here (deferral from one emitted row, queues of capacity 0) and must
stay bit-exact: rasters, sampled V and I_syn (the materialised pending
input), per-population counts and the state, also across a settle with
deferred rows pending, a checkpoint and a second run.
"""
import numpy as np
import pytest

from test_gpu_parity import DT, assert_state_equal, bits, run_oracle

pytestmark = pytest.mark.gpu


def oracle_two_runs(orc, net, seed, steps, mu):
    on = orc.OracleNet.from_network(net, dt=DT, seed=seed)
    on.set_ou(on.pop_ids, mu, 0.1, 4.0)
    sim = orc.Sim(on, 32)
    ra = sim.run(steps, raster=True, vtrace=True, pop_counts=True)
    on.props[on.props[:, 3] == 1, 8] = -51.0
    rb = sim.run(steps, raster=True, vtrace=True, pop_counts=True)
    return ra, rb, sim


@pytest.mark.parametrize("defer_min,floor,shift", [(1, 4096, 3), (3, 4096, 3), (0, 0, 31), (1, 0, 31),
                                                   (2, 16, 31)])
def test_pull_forced_bit_exact(dtb, orc, gpu, defer_min, floor, shift):
    n, steps, seed, mu = 10000, 300, 31, 1.5
    net = dtb.gen_uniform(n, 100, seed=1)
    ids = np.arange(0, n, 7, dtype=np.uint32)
    print("case defer_min", defer_min, "floor", floor, "shift", shift, flush=True)
    eng = dtb.Engine(0)
    try:
        eng.pull(True, defer_min)
        eng.queue_capacity(floor, shift)
        blk = dtb.Block(net, delta_t=DT, seed=seed, engine=eng)
        assert eng.pull_info(blk.inst)["built"]
        blk.update_ou_background_stimuli(blk.subblk_id, 4.0, mu, 0.1)
        blk.set_samples(ids)
        a = blk.run(steps, freqs=True, output_sample_spike=True, output_sample_vmemb=True,
                    output_sample_isynaptic=True)
        assert eng.pull_info(blk.inst)["in_use"]
        blk.assign_property_by_subblk([1], [5], [-51.0])   # settles deferred rows pending
        b = blk.run(steps, freqs=True, output_sample_spike=True, output_sample_vmemb=True,
                    output_sample_isynaptic=True)
        st = blk.state()
        stats = eng.propagation_stats()
        eng.check()
    finally:
        eng.close()
    ra, rb, sim = oracle_two_runs(orc, net, seed, steps, mu)
    print("stats", stats, "spikes", int(ra["raster"].sum() + rb["raster"].sum()))
    assert stats["pulled"] > 0
    if defer_min:
        assert stats["deferred"] > 0
    if floor == 0 and defer_min != 1:   # (defer_min 1: every tile that emits defers, nothing is queued)
        assert stats["spilled"] > 0
    for out, ref in ((a, ra), (b, rb)):
        assert np.array_equal(out["spike"], ref["raster"][:, ids]), "raster differs"
        assert np.array_equal(bits(out["vmemb"]), bits(ref["V"][:, ids])), "V trace differs"
        assert np.array_equal(out["freq"], ref["freq"])
    assert_state_equal(st, sim)
    assert np.array_equal(bits(b["isynaptic"][-1]), bits(st["I_syn"][ids]))


def test_pull_checkpoint_with_deferred_rows_pending(dtb, orc, gpu):
    """state() with deferred rows pending folds them (settle); load_state and
    set_step resume bit-exactly, and a restore on a block whose own deferred
    rows are pending drops those."""
    n, seed, mu = 8000, 5, 1.5
    net = dtb.gen_uniform(n, 120, seed=3)

    def make():
        eng = dtb.Engine(0)
        eng.pull(True, 1)
        blk = dtb.Block(net, delta_t=DT, seed=seed, engine=eng)
        blk.update_ou_background_stimuli(blk.subblk_id, 4.0, mu, 0.1)
        return eng, blk

    e1, b1 = make()
    e2, b2 = make()
    try:
        b1.run(200)
        snap = b1.state()
        r1 = b1.run(100)
        s1 = b1.state()
        b2.run(77)
        b2.load_state(snap)
        r2 = b2.run(100)
        s2 = b2.state()
        assert e1.propagation_stats()["deferred"] > 0
    finally:
        e1.close()
        e2.close()
    assert snap["step"] == 200 and s2["step"] == 300
    assert np.array_equal(r1["freq"], r2["freq"])
    for k in ("V", "J", "i_ou", "I_syn"):
        assert np.array_equal(bits(s1[k]), bits(s2[k])), k
    on = orc.OracleNet.from_network(net, dt=DT, seed=seed)
    on.set_ou(on.pop_ids, mu, 0.1, 4.0)
    sim = orc.Sim(on, 32)
    sim.run(300)
    assert_state_equal(s1, sim)


def test_pull_volley_from_rest_matches_oracle(dtb, orc, gpu):
    """A block started at V0 = (V_th + V_reset) / 2 under strong drive fires
    almost all at once; with the default settings the dense tiles defer and
    the step after the volley pulls.  Bit-exact for 120 steps."""
    n, k, steps, seed, mu = 50_000, 500, 120, 9, 8.0
    net = dtb.gen_uniform(n, k, seed=12, g_ui=np.asarray(dtb.DEFAULT_G_UI, np.float32) * np.float32(0.05))
    sample = np.arange(0, n, 13, dtype=np.uint32)
    eng = dtb.Engine(0)
    try:
        blk = dtb.Block(net, delta_t=DT, seed=seed, engine=eng)
        blk.update_ou_background_stimuli(blk.subblk_id, 4.0, mu, 0.5)
        blk.set_samples(sample)
        out = blk.run(steps, freqs=True, output_sample_spike=True, output_sample_vmemb=True)
        st = blk.state()
        stats = eng.propagation_stats()
        eng.check()
    finally:
        eng.close()
    ref, sim = run_oracle(orc, net, steps, seed, mu, sigma=0.5, vtrace=True)
    per_step = ref["raster"].sum(axis=1)
    print("max fired in one step", int(per_step.max()), "of", n, "stats", stats)
    assert per_step.max() > n // 2, "no volley"
    assert stats["deferred"] > 0 and stats["pulled"] > 0
    assert np.array_equal(out["freq"], ref["freq"])
    assert np.array_equal(out["spike"], ref["raster"][:, sample])
    assert np.array_equal(bits(out["vmemb"]), bits(ref["V"][:, sample]))
    assert_state_equal(st, sim)


def test_pull_in_edges_built_in_tile_windows(dtb, orc, gpu):
    """An instance of more than 16,384 tiles (4.3M neurons): its in-edges are
    built in two windows of target tiles.  Deferral from one emitted row and
    queues of capacity 0 send every synapse through the pull; bit-exact for
    40 steps through the volley from rest."""
    n, k, steps, seed, mu = 4_300_000, 8, 40, 9, 8.0
    net = dtb.gen_uniform(n, k, seed=13)
    sample = np.arange(0, n, 4099, dtype=np.uint32)
    print("case n", n, "k", k, flush=True)
    eng = dtb.Engine(0)
    try:
        eng.pull(True, 1)
        eng.queue_capacity(0, 31)
        blk = dtb.Block(net, delta_t=DT, seed=seed, engine=eng)
        assert eng.pull_info(blk.inst)["built"]
        blk.update_ou_background_stimuli(blk.subblk_id, 4.0, mu, 0.5)
        blk.set_samples(sample)
        out = blk.run(steps, freqs=True, output_sample_spike=True, output_sample_vmemb=True)
        st = blk.state()
        stats = eng.propagation_stats()
        eng.check()
    finally:
        eng.close()
    ref, sim = run_oracle(orc, net, steps, seed, mu, sigma=0.5, vtrace=True)
    print("fired", int(ref["raster"].sum()), "stats", stats)
    assert ref["raster"].sum() > n // 2
    assert stats["deferred"] > 0 and stats["pulled"] > 0
    assert np.array_equal(out["freq"], ref["freq"])
    assert np.array_equal(out["spike"], ref["raster"][:, sample])
    assert np.array_equal(bits(out["vmemb"]), bits(ref["V"][:, sample]))
    assert_state_equal(st, sim)


def test_pull_batched_instances_match_oracle(dtb, orc, gpu):
    """Instances of different sizes and g_ui in one engine, every tile
    deferring; each instance against its own oracle run."""
    specs = [(3000, 80, 1.0), (7000, 150, 0.7), (12000, 60, 1.3)]
    steps, mu = 250, 1.5
    eng = dtb.Engine(0)
    nets, blks = [], []
    try:
        eng.pull(True, 1)
        for i, (n, k, gs) in enumerate(specs):
            g = (np.asarray(dtb.DEFAULT_G_UI, np.float32) * np.float32(gs)).astype(np.float32)
            net = dtb.gen_uniform(n, k, seed=40 + i, g_ui=g)
            b = dtb.Block(net, delta_t=DT, seed=70 + i, engine=eng)
            b.update_ou_background_stimuli(b.subblk_id, 4.0, mu, 0.1)
            eng.configure(b.inst, 1, True)
            nets.append(net)
            blks.append(b)
        eng.run(steps, graph_chunk=50)
        states = [b.state() for b in blks]
        stats = eng.propagation_stats()
        eng.check()
    finally:
        eng.close()
    assert stats["deferred"] > 0
    for i, net in enumerate(nets):
        ref, sim = run_oracle(orc, net, steps, 70 + i, mu)
        assert_state_equal(states[i], sim)


def test_pull_off_keeps_spill_path(dtb, gpu):
    """With the pull off nothing builds in-edges and overflowing queues go
    through the spill accumulators; the results equal the pull's."""
    net = dtb.gen_uniform(6000, 150, seed=9)
    res = []
    for on in (False, True):
        eng = dtb.Engine(0)
        try:
            eng.pull(on, 1)
            eng.queue_capacity(0, 31)
            blk = dtb.Block(net, delta_t=DT, seed=17, engine=eng)
            info = eng.pull_info(blk.inst)
            assert info["built"] == on and info["allowed"]
            blk.update_ou_background_stimuli(blk.subblk_id, 4.0, 1.5, 0.1)
            out = blk.run(200)
            res.append((out["freq"], blk.state(), eng.propagation_stats()))
        finally:
            eng.close()
    assert res[0][2]["pulled"] == 0 and res[0][2]["spilled"] > 0
    assert res[1][2]["pulled"] > 0
    assert np.array_equal(res[0][0], res[1][0])
    for k in ("V", "J", "I_syn"):
        assert np.array_equal(bits(res[0][1][k]), bits(res[1][1][k])), k
