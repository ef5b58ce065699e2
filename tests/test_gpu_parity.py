"""GPU parity: the HIP step (through the C ABI) against the CPU oracle.

Bit-exact for rasters, V, J, i_ou, refractory counters and per-population
spike counts on the same seeded inputs; V within 1e-5 relative of the fp64
oracle (rtol = 1e-5 written below).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DT = 0.1


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint8)


def run_gpu(dtb, net, steps, seed, mu, sigma=0.1, tau=4.0, t_steps=1, use_ou=True, sample=None, graph=False,
            engine=None):
    blk = dtb.Block(net, delta_t=DT, seed=seed, engine=engine)
    if use_ou:
        blk.update_ou_background_stimuli(blk.subblk_id, tau, mu, sigma)
    if sample is not None:
        blk.set_samples(sample)
    out = blk.run(steps, freqs=True, output_sample_spike=sample is not None,
                  output_sample_vmemb=sample is not None, output_sample_isynaptic=sample is not None,
                  output_sample_iou=sample is not None and use_ou, iou=use_ou, t_steps=t_steps, graph=graph)
    st = blk.state()
    cnt = blk.counters()
    blk.engine.check()
    if engine is None:
        blk.engine.close()
    return out, st, cnt


def run_oracle(orc, net, steps, seed, mu, sigma=0.1, tau=4.0, t_steps=1, use_ou=True, precision=32,
               raster=True, vtrace=False, force_raster=None, margins=False):
    on = orc.OracleNet.from_network(net, dt=DT, t_steps=t_steps, seed=seed)
    if use_ou:
        on.set_ou(on.pop_ids, mu, sigma, tau)
    sim = orc.Sim(on, precision)
    out = sim.run(steps, use_ou=use_ou, raster=raster, vtrace=vtrace, pop_counts=(precision == 32),
                  force_raster=force_raster, margins=margins)
    return out, sim


def assert_state_equal(st, sim):
    assert np.array_equal(bits(st["V"]), bits(sim.st.V)), "V differs"
    assert np.array_equal(st["t"], sim.st.t), "refractory counters differ"
    assert np.array_equal(st["f"], sim.st.f), "spike flags differ"
    assert np.array_equal(bits(st["i_ou"]), bits(sim.st.i_ou)), "i_ou differs"
    assert np.array_equal(bits(st["J"]), bits(sim.st.J)), "J differs"
    assert np.array_equal(bits(st["I_syn"]), bits(sim.st.I_syn)), "I_syn differs"


def test_ou_normals_bit_identical(dtb, orc, gpu):
    for seed, n0, sub, cnt in [(0, 0, 0, 1 << 16), (1234567, 1000, 999, 1 << 18), (2**63 + 5, 2**31, 2**40, 4096)]:
        g = dtb.ou_normals(seed, n0, sub, cnt)
        c = orc.normals(seed, n0, sub, cnt)
        assert np.array_equal(bits(g), bits(c)), (seed, n0, sub)


@pytest.mark.parametrize("mu,label", [(0.5, "config0_literal"), (1.5, "config0_active")])
def test_config0_1000_steps_bit_exact(dtb, orc, gpu, mu, label):
    """10k LIF, 80/20 E/I, in-degree 100, dt 0.1 ms, 1000 steps; full raster
    and V trace of every neuron via the sample path."""
    n, steps, seed = 10000, 1000, 20240
    net = dtb.gen_uniform(n, 100, seed=1)
    ids = np.arange(n, dtype=np.uint32)
    out, st, cnt = run_gpu(dtb, net, steps, seed, mu, sample=ids)
    ref, sim = run_oracle(orc, net, steps, seed, mu, vtrace=True)
    print(label, "spikes", int(ref["raster"].sum()))
    assert np.array_equal(out["spike"], ref["raster"]), "raster differs"
    assert np.array_equal(bits(out["vmemb"]), bits(ref["V"])), "V trace differs"
    assert np.array_equal(out["freq"], ref["freq"])
    assert_state_equal(st, sim)
    # events counter = sum of out-degrees of fired rows
    deg = np.diff(net.rowptr)
    assert cnt[0] == steps
    assert cnt[1] == int((ref["raster"].astype(np.int64) * deg[None, :]).sum())
    assert cnt[2] == int(ref["raster"].sum())
    if mu > 1.0:
        assert ref["raster"].sum() > 1000


def test_config1_cut_voxel_bit_exact(dtb, orc, gpu):
    """200k-neuron cut of the whole-brain block (200 voxels, in-degree 500,
    70/30 intra/inter voxel), 1000 steps."""
    n, steps, seed = 200_000, 1000, 99
    net = dtb.gen_voxel(n, 200, 500, seed=5)
    sample = np.arange(0, n, 97, dtype=np.uint32)
    out, st, cnt = run_gpu(dtb, net, steps, seed, mu=5.0, sample=sample)
    ref, sim = run_oracle(orc, net, steps, seed, mu=5.0)
    assert np.array_equal(out["freq"], ref["freq"]), "per-population counts differ"
    assert np.array_equal(out["spike"], ref["raster"][:, sample])
    assert_state_equal(st, sim)
    assert ref["raster"].sum() > 10000


def test_stress_cut_high_degree(dtb, orc, gpu):
    """Propagation stress cut: in-degree 2000 on 20k neurons, strong drive."""
    n, steps, seed = 20_000, 300, 5
    net = dtb.gen_uniform(n, 2000, seed=8)
    out, st, _ = run_gpu(dtb, net, steps, seed, mu=8.0, sigma=0.5)
    ref, sim = run_oracle(orc, net, steps, seed, mu=8.0, sigma=0.5)
    assert np.array_equal(out["freq"], ref["freq"])
    assert_state_equal(st, sim)
    assert ref["raster"].sum() > 1000


def test_substeps_and_graph(dtb, orc, gpu):
    n, steps, seed = 5000, 300, 3
    net = dtb.gen_uniform(n, 100, seed=2)
    out, st, _ = run_gpu(dtb, net, steps, seed, mu=1.5, t_steps=3, graph=True)
    ref, sim = run_oracle(orc, net, steps, seed, mu=1.5, t_steps=3)
    assert np.array_equal(out["freq"], ref["freq"])
    assert_state_equal(st, sim)


def test_noise_spike_mode(dtb, orc, gpu):
    n, steps, seed = 5000, 300, 4
    net = dtb.gen_uniform(n, 100, seed=4)
    net.props[:, 0] = 0.002   # noise_rate column
    ids = np.arange(n, dtype=np.uint32)
    out, st, _ = run_gpu(dtb, net, steps, seed, mu=None, use_ou=False, sample=ids)
    ref, sim = run_oracle(orc, net, steps, seed, mu=None, use_ou=False)
    assert np.array_equal(out["spike"], ref["raster"])
    assert_state_equal(st, sim)
    assert ref["raster"].sum() > 100


def test_graph_equals_eager(dtb, gpu):
    net = dtb.gen_uniform(6000, 150, seed=9)
    res = []
    for graph in (False, True):
        out, st, _ = run_gpu(dtb, net, 257, 17, mu=1.5, graph=graph)
        res.append((out["freq"], st))
    freq, st = res[1]
    assert np.array_equal(freq, res[0][0])
    assert np.array_equal(bits(st["V"]), bits(res[0][1]["V"]))
    assert np.array_equal(bits(st["J"]), bits(res[0][1]["J"]))


def test_graph_chunks_both_parities(dtb, gpu):
    """Replayed graphs of even and odd chunk lengths, entered at either launch
    parity and mixed with eager steps, step exactly like eager stepping."""
    net = dtb.gen_uniform(6000, 150, seed=9)
    res = []
    for plan in ([(257, 0)], [(1, 0), (256, 64)], [(3, 0), (254, 7)], [(100, 5), (157, 2)]):
        eng = dtb.Engine(0)
        blk = dtb.Block(net, delta_t=DT, seed=17, engine=eng)
        blk.update_ou_background_stimuli(blk.subblk_id, 4.0, 1.5, 0.1)
        eng.configure(blk.inst, 1, True)
        for steps, chunk in plan:
            eng.run(steps, graph_chunk=chunk)
        eng.check()
        res.append((blk.state(), blk.counters()))
        eng.close()
    assert res[0][1][0] == 257 and res[0][1][2] > 100
    for st, cnt in res[1:]:
        assert cnt == res[0][1]
        for k in ("V", "J", "i_ou", "I_syn"):
            assert np.array_equal(bits(st[k]), bits(res[0][0][k])), k
        assert np.array_equal(st["t"], res[0][0]["t"])


def test_checkpoint_resume(dtb, gpu):
    """state() + load_state() (with the step index) resumes bit-exactly on a
    block whose own state is different."""
    n, seed = 5000, 21
    net = dtb.gen_uniform(n, 100, seed=6)
    blk = dtb.Block(net, delta_t=DT, seed=seed)
    blk.update_ou_background_stimuli(blk.subblk_id, 4.0, 1.5, 0.1)
    blk.run(150)
    snap = blk.state()
    a = blk.run(100)
    st_a = blk.state()
    blk2 = dtb.Block(net, delta_t=DT, seed=seed)
    blk2.update_ou_background_stimuli(blk2.subblk_id, 4.0, 1.5, 0.1)
    blk2.run(37)
    blk2.load_state(snap)
    b = blk2.run(100)
    st_b = blk2.state()
    assert snap["step"] == 150 and st_b["step"] == 250
    assert np.array_equal(a["freq"], b["freq"])
    assert np.array_equal(bits(st_a["V"]), bits(st_b["V"]))
    assert np.array_equal(bits(st_a["J"]), bits(st_b["J"]))
    blk.engine.close()
    blk2.engine.close()


def test_split_run_equals_single_run(dtb, gpu):
    """Pending synaptic input survives everything between runs: a run split by
    no-op property updates, OU re-settings and sample changes equals one
    uninterrupted run, bit for bit."""
    net = dtb.gen_uniform(6000, 120, seed=13)

    def fresh():
        b = dtb.Block(net, delta_t=DT, seed=5)
        b.update_ou_background_stimuli(b.subblk_id, 4.0, 1.5, 0.1)
        return b

    a = fresh()
    ra = a.run(240)
    sa = a.state()
    b = fresh()
    parts = [b.run(60)]
    b.mul_property_by_subblk([0, 1], [7, 8], [1.0, 1.0])
    parts.append(b.run(60, graph=True))
    b.update_ou_background_stimuli(b.subblk_id, 4.0, 1.5, 0.1)
    b.set_samples(np.arange(0, 6000, 7))
    parts.append(b.run(60, output_sample_spike=True, output_sample_isynaptic=True))
    _ = b.state()
    parts.append(b.run(60))
    sb = b.state()
    assert np.array_equal(ra["freq"], np.concatenate([p["freq"] for p in parts]))
    for k in ("V", "J", "i_ou", "I_syn"):
        assert np.array_equal(bits(sa[k]), bits(sb[k])), k
    a.engine.close()
    b.engine.close()


def test_property_updates_match_oracle(dtb, orc, gpu):
    n, steps, seed = 5000, 200, 8
    net = dtb.gen_uniform(n, 100, seed=12)
    blk = dtb.Block(net, delta_t=DT, seed=seed)
    blk.update_ou_background_stimuli(blk.subblk_id, 4.0, 1.5, 0.1)
    # per-population multiply of g_AMPA on population 0, assign V_th on 1,
    # per-neuron i_ext on a few neurons
    blk.mul_property_by_subblk([0], [7], [1.3])
    blk.assign_property_by_subblk([1], [5], [-51.0])
    blk.update_property([5, 17, 4999], [0, 0, 0], [0.3, 0.4, 0.5])
    out = blk.run(steps)
    st = blk.state()
    blk.engine.close()

    # The oracle initialises its state from the original properties (as the
    # block did at load time), then runs with the updated ones: property
    # updates change parameters, not the state (property.cpp:57-466).
    on = orc.OracleNet(net.props.copy(), net.rowptr, net.col, net.w, net.kind_per_synapse(), dt=DT, seed=seed)
    on.set_ou(on.pop_ids, 1.5, 0.1, 4.0)
    sim = orc.Sim(on, 32)
    props = on.props
    p0 = props[:, 3] == 0
    props[p0, 10] *= np.float32(1.3)
    props[props[:, 3] == 1, 8] = -51.0
    props[[5, 17, 4999], 2] = [0.3, 0.4, 0.5]
    ref = sim.run(steps, pop_counts=True)
    assert np.array_equal(out["freq"], ref["freq"])
    assert_state_equal(st, sim)


def test_parameter_classes_and_per_neuron_records(dtb, orc, gpu):
    """Parameters shared by populations (a few records) and parameters that
    differ on every neuron (more records than the class index holds, kept per
    neuron) give the oracle's results; switching from one layout to the other
    between runs, with spikes still pending, changes nothing either."""
    n, seed = 70000, 5
    net = dtb.gen_uniform(n, 100, seed=14)
    rng = np.random.default_rng(3)
    vth = (-50.0 + rng.uniform(-0.5, 0.5, n)).astype(np.float32)
    gamp = np.where(net.props[:, 3] == 0, 0.011, 0.009).astype(np.float32)

    blk = dtb.Block(net, delta_t=DT, seed=seed)
    blk.update_ou_background_stimuli(blk.subblk_id, 4.0, 1.5, 0.1)
    blk.assign_property_by_subblk([0, 1], [7, 7], [0.011, 0.009])   # two records
    a = blk.run(100)
    blk.update_property(np.arange(n, dtype=np.uint32), np.full(n, 5, np.uint32), vth)   # 70000 records
    b = blk.run(100)
    st = blk.state()
    blk.engine.close()

    on = orc.OracleNet(net.props.copy(), net.rowptr, net.col, net.w, net.kind_per_synapse(), dt=DT, seed=seed)
    on.set_ou(on.pop_ids, 1.5, 0.1, 4.0)
    sim = orc.Sim(on, 32)
    on.props[:, 10] = gamp
    ra = sim.run(100, pop_counts=True)
    on.props[:, 8] = vth
    rb = sim.run(100, pop_counts=True)
    assert np.array_equal(a["freq"], ra["freq"])
    assert np.array_equal(b["freq"], rb["freq"])
    assert_state_equal(st, sim)
    assert rb["freq"].sum() > 100


def test_fp64_oracle_v_within_1e5(dtb, orc, gpu):
    """V of the GPU fp32 path within 1e-5 relative of the fp64 oracle driven
    by the same spike raster; every spike decision the fp64 model disagrees
    with is a near tie at threshold."""
    n, steps, seed = 10000, 1000, 31
    net = dtb.gen_uniform(n, 100, seed=1)
    ids = np.arange(n, dtype=np.uint32)
    out, st, _ = run_gpu(dtb, net, steps, seed, mu=1.5, sample=ids)
    ref64, sim64 = run_oracle(orc, net, steps, seed, mu=1.5, precision=64, vtrace=True,
                              force_raster=out["spike"], margins=True)
    rtol = 1e-5
    v32 = out["vmemb"].astype(np.float64)
    rel = np.abs(v32 - ref64["V"]) / np.abs(ref64["V"])
    assert rel.max() < rtol, rel.max()
    assert ref64["margin"].max() < 1e-3


def test_batched_equals_alone(dtb, gpu):
    nets = [dtb.gen_uniform(n, k, seed=s) for n, k, s in [(3000, 100, 1), (7777, 200, 2), (12345, 50, 3)]]
    alone = []
    for i, net in enumerate(nets):
        out, st, _ = run_gpu(dtb, net, 200, 100 + i, mu=1.5)
        alone.append((out, st))
    eng = dtb.Engine(0)
    blks = [dtb.Block(net, delta_t=DT, seed=100 + i, engine=eng) for i, net in enumerate(nets)]
    for b in blks:
        b.update_ou_background_stimuli(b.subblk_id, 4.0, 1.5, 0.1)
        eng.configure(b.inst, 1, True)
    eng.run(200, graph_chunk=50)
    eng.sync()
    for b, (out, st) in zip(blks, alone):
        st2 = b.state()
        assert np.array_equal(bits(st2["V"]), bits(st["V"]))
        assert np.array_equal(bits(st2["J"]), bits(st["J"]))
        assert b.counters()[0] == 200
    eng.close()


def test_population_means_match_oracle(dtb, orc, gpu):
    """vmean / imean / receptor imeans (stat_result.cpp:266-391) bit-exact
    against the oracle, with blocked neurons, 12 populations, two runs (the
    second replayed from graphs) so every run's last-step tail is covered."""
    n, seed = 6000, 12
    net = dtb.gen_voxel(n, 6, 80, seed=7)
    net.props[::9, 1] = 1.0
    blk = dtb.Block(net, delta_t=DT, seed=seed)
    blk.update_ou_background_stimuli(blk.subblk_id, 4.0, 3.0, 0.2)
    kw = dict(freqs=True, vmean=True, imean=True, receptor_imean=True)
    a = blk.run(150, **kw)
    b = blk.run(150, graph=True, **kw)
    st = blk.state()
    blk.engine.close()

    on = orc.OracleNet.from_network(net, dt=DT, seed=seed)
    on.set_ou(on.pop_ids, 3.0, 0.2, 4.0)
    sim = orc.Sim(on, 32)
    ra = sim.run(150, pop_counts=True, stats=True)
    rb = sim.run(150, pop_counts=True, stats=True)
    for got, ref in ((a, ra), (b, rb)):
        assert np.array_equal(got["freq"], ref["freq"])
        assert np.array_equal(bits(got["vmean"]), bits(ref["stat_means"][:, :, 0])), "vmean"
        assert np.array_equal(bits(got["imean"]), bits(ref["stat_means"][:, :, 1])), "imean"
        for k in range(4):
            assert np.array_equal(bits(got["receptor_imean"][k]), bits(ref["stat_means"][:, :, 2 + k])), k
    assert_state_equal(st, sim)
    assert rb["freq"].sum() > 1000


def test_blockwrapper_mirror(dtb, orc, gpu, tmp_path):
    """BlockWrapper (dist_blockwrapper.py signatures) over reference block
    files: unsorted OU population ids, relative mul_property_by_subblk,
    samples with repeats in caller order, per-step tuples in the reference
    order."""
    net = dtb.gen_uniform(3000, 100, seed=3)
    path = dtb.save_npz(net, tmp_path, "single")
    bw = dtb.BlockWrapper(None, path, DT, seed=5)
    assert bw.total_neurons == 3000 and list(bw.subblk_id) == [0, 1]
    bw.update_ou_background_stimuli(np.array([1, 0]), np.array([4.0, 4.0], np.float32),
                                     np.array([1.6, 1.5], np.float32), np.array([0.1, 0.1], np.float32))
    bw.mul_property_by_subblk(np.array([[0, 10]]), np.array([1.2], np.float32))
    bw.mul_property_by_subblk(np.array([[0, 10]]), np.array([1.5], np.float32))
    samples = np.array([2999, 5, 5, 100])
    bw.set_samples(samples)
    outs = list(bw.run(200, freqs=True, vmean=True, output_sample_spike=True, output_sample_vmemb=True, iou=True))
    bw.close()
    assert len(outs) == 200 and len(outs[0]) == 4
    freq, vmean, spike, vm = (np.stack([o[k] for o in outs]) for k in range(4))

    on = orc.OracleNet.from_network(net, dt=DT, seed=5)
    on.set_ou(on.pop_ids, np.array([1.5, 1.6], np.float32), 0.1, 4.0)
    sim = orc.Sim(on, 32)
    p0 = on.props[:, 3] == 0
    g = on.props[p0, 10]
    step1 = (g * np.float32(1.2)).astype(np.float32)
    on.props[p0, 10] = (step1 * (np.float32(1.5) / np.float32(1.2))).astype(np.float32)
    ref = sim.run(200, raster=True, vtrace=True, pop_counts=True, stats=True)
    assert np.array_equal(freq, ref["freq"].astype(np.int32))
    assert np.array_equal(bits(vmean), bits(ref["stat_means"][:, :, 0]))
    assert np.array_equal(spike, ref["raster"][:, samples])
    assert np.array_equal(bits(vm), bits(ref["V"][:, samples]))
    assert ref["freq"].sum() > 100


def test_device_generated_network_equals_host(dtb, orc, gpu):
    """A fan-out network generated in device memory steps exactly like the
    same network built on the host (gen_fanout) and uploaded, and both match
    the oracle."""
    n, k, nseed, seed, steps = 20000, 300, 77, 9, 300
    host = dtb.gen_fanout(n, k, nseed)
    eng = dtb.Engine(0)
    bd = dtb.Block(None, delta_t=DT, engine=eng, inst=eng.add_fanout(n, k, nseed, dt=DT, seed=seed))
    bh = dtb.Block(host, delta_t=DT, seed=seed, engine=eng)
    for b in (bd, bh):
        b.update_ou_background_stimuli(b.subblk_id, 4.0, 2.0, 0.1)
        eng.configure(b.inst, 1, True)
    # queue capacities: the device counts the synapses per target tile
    expect = np.bincount(host.col.astype(np.int64) >> 8, minlength=(n + 255) // 256)
    assert np.array_equal(eng.tile_synapses(bd.inst), expect)
    assert np.array_equal(eng.tile_synapses(bh.inst), expect)
    eng.run(steps)
    eng.check()
    sd, sh = bd.state(), bh.state()
    assert bd.counters() == bh.counters()
    for key in ("V", "J", "i_ou", "I_syn"):
        assert np.array_equal(bits(sd[key]), bits(sh[key])), key
    eng.close()
    ref, sim = run_oracle(orc, host, steps, seed, mu=2.0)
    assert_state_equal(sd, sim)
    assert ref["raster"].sum() > 1000


def _edge_network(dtb, case):
    """Hand-built networks at the edges of the layout (see the cases below);
    weights k * 2^-24, channel bit = presynaptic neuron inhibitory."""
    rng = np.random.default_rng(77)
    if case == "single":
        n = 1
        pre = post = np.zeros(0, dtype=np.int64)
    elif case == "ragged":
        # two tiles, the second partial; empty rows, a one-synapse row, a
        # 700-synapse row into one tile (a segment across passes), a row onto
        # every neuron, repeated targets and a self-connection
        n = 300
        pre, post = [], []
        pre += [5] * 700
        post += list(rng.integers(0, 256, 700))
        pre += [17] * n
        post += list(range(n))
        pre += [40]
        post += [299]
        pre += [41] * 6
        post += [3, 3, 3, 280, 280, 41]
        for r in range(60, 300, 3):
            k = int(rng.integers(1, 90))
            pre += [r] * k
            post += list(rng.integers(0, n, k))
        pre, post = np.asarray(pre), np.asarray(post)
    elif case == "hot_tile":
        # every synapse targets the first tile: its queue takes them all
        n = 2048
        pre = np.repeat(np.arange(n), 64)
        post = rng.integers(0, 256, pre.size)
    elif case == "ragged_groups":
        # three gather groups of 2048 neurons, the last partial: a row of
        # 65535 synapses (the longest the 16-bit group offsets index), a row
        # onto every neuron, rows only into the last group or only into the
        # middle one, one-synapse rows, ragged rows and many empty rows
        n = 5000
        pre, post = [], []
        pre += [7] * 65535
        post += list(rng.integers(0, n, 65535))
        pre += [13] * n
        post += list(range(n))
        pre += [11] * 40
        post += list(rng.integers(4096, n, 40))
        pre += [12]
        post += [2100]
        pre += [14] * 3
        post += [2047, 2048, 4095]
        for r in range(100, 3000, 7):
            k = int(rng.integers(1, 300))
            pre += [r] * k
            post += list(rng.integers(0, n, k))
        pre, post = np.asarray(pre), np.asarray(post)
    elif case == "long_row":
        # one row longer than the group offsets can index: gather is not
        # allowed and the queues carry it
        n = 3000
        pre = np.concatenate([np.full(70000, 5), np.repeat(np.arange(100, 3000, 3), 20)])
        post = rng.integers(0, n, pre.size)
    else:
        raise ValueError(case)
    props = dtb.gen_uniform(max(n, 2), 1, seed=3).props[:n]
    if n == 1:
        props = props.copy()
        props[:, 3] = 0
    kind = (pre >= int(n * 0.8)).astype(np.uint8)
    w = rng.integers(0, 1 << 24, (pre.size, 2)).astype(np.float32) * np.float32(2.0 ** -24)
    order = np.lexsort((post, pre))   # rows ordered by target, as the block files are
    return dtb.NetworkArrays.from_edges(props, pre[order], post[order], w[order], kind[order])


@pytest.mark.parametrize("layout", ["auto", "gather"])
@pytest.mark.parametrize("case,mu", [("single", 3.0), ("ragged", 2.0), ("hot_tile", 2.0), ("ragged_groups", 2.0),
                                     ("long_row", 2.0)])
def test_edge_shapes_bit_exact(dtb, orc, gpu, case, mu, layout):
    """A single neuron without synapses, ragged rows over a partial tile, all
    synapses into one tile, ragged rows over partial gather groups (up to the
    longest row the group offsets index), a row too long for them; by queue
    and by gather: raster, V trace and final state bit-exact."""
    net = _edge_network(dtb, case)
    steps, seed = 400, 99
    ids = np.arange(net.n, dtype=np.uint32)
    eng = dtb.Engine(0)
    try:
        eng.schedule(layout=layout)
        out, st, cnt = run_gpu(dtb, net, steps, seed, mu, sample=ids, engine=eng)
        prop = eng.propagation(0)
        assert prop["allowed"] == (net.nsyn > 0 and case != "long_row")
        assert prop["in_use"] == (layout == "gather" and prop["allowed"])
    finally:
        eng.close()
    ref, sim = run_oracle(orc, net, steps, seed, mu, vtrace=True)
    print(case, "n", net.n, "synapses", net.nsyn, "spikes", int(ref["raster"].sum()))
    assert ref["raster"].sum() > 0
    assert np.array_equal(out["spike"], ref["raster"]), "raster differs"
    assert np.array_equal(bits(out["vmemb"]), bits(ref["V"])), "V trace differs"
    assert_state_equal(st, sim)
    deg = np.diff(net.rowptr)
    assert cnt[1] == int((ref["raster"].astype(np.int64) * deg[None, :]).sum())
    assert cnt[2] == int(ref["raster"].sum())


@pytest.mark.parametrize("pull", [False, True])
@pytest.mark.parametrize("floor,shift", [(0, 31), (16, 31), (64, 4)])
def test_queue_capacity_spill_bit_exact(dtb, orc, gpu, floor, shift, pull):
    """Queues smaller than their worst case: the synapses past a tile's
    capacity take the spill accumulators (every synapse with capacity 0; a
    segment split between queue and spill with 16 entries), or with the pull
    on the overflowing tile takes every emitted row through its in-edges.
    Raster, sampled V and I_syn, and the state stay bit-exact, also across a
    parameter change with spikes pending (settle) and a second run."""
    n, steps, seed = 10000, 300, 31
    net = dtb.gen_uniform(n, 100, seed=1)
    ids = np.arange(0, n, 7, dtype=np.uint32)
    eng = dtb.Engine(0)
    eng.pull(pull, 0)
    eng.queue_capacity(floor, shift)
    try:
        blk = dtb.Block(net, delta_t=DT, seed=seed, engine=eng)
        blk.update_ou_background_stimuli(blk.subblk_id, 4.0, 1.5, 0.1)
        blk.set_samples(ids)
        a = blk.run(steps, freqs=True, output_sample_spike=True, output_sample_vmemb=True,
                    output_sample_isynaptic=True)
        blk.assign_property_by_subblk([1], [5], [-51.0])   # settles the pending delta
        b = blk.run(steps, freqs=True, output_sample_spike=True, output_sample_vmemb=True,
                    output_sample_isynaptic=True)
        st = blk.state()
        spilled = eng.spilled()
        eng.check()
    finally:
        eng.close()
    on = orc.OracleNet.from_network(net, dt=DT, seed=seed)
    on.set_ou(on.pop_ids, 1.5, 0.1, 4.0)
    sim = orc.Sim(on, 32)
    ra = sim.run(steps, raster=True, vtrace=True, pop_counts=True)
    on.props[on.props[:, 3] == 1, 8] = -51.0
    rb = sim.run(steps, raster=True, vtrace=True, pop_counts=True)
    print("floor", floor, "shift", shift, "spilled", spilled, "spikes", int(ra["raster"].sum() + rb["raster"].sum()))
    assert spilled > 0
    for out, ref in ((a, ra), (b, rb)):
        assert np.array_equal(out["spike"], ref["raster"][:, ids]), "raster differs"
        assert np.array_equal(bits(out["vmemb"]), bits(ref["V"][:, ids])), "V trace differs"
        assert np.array_equal(out["freq"], ref["freq"])
    assert_state_equal(st, sim)
    # the sampled I_syn of the last step (pending delta materialised with the
    # spill) is the state's I_syn
    assert np.array_equal(bits(b["isynaptic"][-1]), bits(st["I_syn"][ids]))


def test_gamma_resampling_matches_oracle(dtb, orc, gpu):
    """Per-population gamma resampling of properties (Snn.Updategamma): the
    population's neurons hold the sampler's draws (one stream per call), and
    the run with those per-neuron properties is bit-exact against the oracle
    given the same property matrix."""
    n, steps, seed = 6000, 300, 13
    net = dtb.gen_uniform(n, 100, seed=21)
    blk = dtb.Block(net, delta_t=DT, seed=seed)
    try:
        blk.update_ou_background_stimuli(blk.subblk_id, 4.0, 1.5, 0.1)
        pops = blk.engine.info(blk.inst)
        b0, b1 = int(pops["pop_begin"][0]), int(pops["pop_begin"][1])
        # g_AMPA (prop id 7) of population 0 ~ Gamma(4, 400): mean 0.01;
        # i_ext (prop id 0) of population 1 ~ Gamma(0.5, 1)
        blk.gamma_property_by_subblk([0, 1], [7, 0], [4.0, 0.5], [400.0, 1.0])
        props = blk.props()
        out = blk.run(steps)
        st = blk.state()
    finally:
        blk.engine.close()
    assert np.array_equal(props[b0:b1, 10], dtb.gen_gamma(seed, 0, b0, b1 - b0, 4.0, 400.0))
    assert np.array_equal(props[b1:, 2], dtb.gen_gamma(seed, 1, b1, n - b1, 0.5, 1.0))
    untouched = np.ones(dtb.NCOL, bool)
    untouched[[2, 10]] = False
    assert np.array_equal(props[:, untouched], net.props[:, untouched])
    on = orc.OracleNet(net.props.copy(), net.rowptr, net.col, net.w, net.kind_per_synapse(), dt=DT, seed=seed)
    on.set_ou(on.pop_ids, 1.5, 0.1, 4.0)
    sim = orc.Sim(on, 32)
    on.props[:] = props
    ref = sim.run(steps, pop_counts=True)
    assert np.array_equal(out["freq"], ref["freq"])
    assert_state_equal(st, sim)


def test_population_rates_over_10s_within_1pct_of_fp64(dtb, orc, gpu):
    """10 s of model time (100,000 steps, about 17 Hz): per-population firing
    rates of the GPU path within 1 % (rtol = 0.01) of the free-running fp64
    oracle with the same noise stream."""
    n, steps, seed, mu = 2000, 100_000, 5, 2.0
    net = dtb.gen_uniform(n, 100, seed=1)
    blk = dtb.Block(net, delta_t=DT, seed=seed)
    try:
        blk.update_ou_background_stimuli(blk.subblk_id, 4.0, mu, 0.1)
        out = blk.run(steps, freqs=True, graph=True)
        begin = blk.engine.info(blk.inst)["pop_begin"].astype(np.int64)
    finally:
        blk.engine.close()
    ref, _ = run_oracle(orc, net, steps, seed, mu, precision=64)
    sizes = np.diff(begin)
    gpu_rate = out["freq"].astype(np.int64).sum(axis=0) / (sizes * steps * DT * 1e-3)
    r = ref["raster"]
    ref_rate = np.array([r[:, a:b].sum() for a, b in zip(begin[:-1], begin[1:])]) / (sizes * steps * DT * 1e-3)
    print("rates Hz gpu", gpu_rate, "fp64", ref_rate)
    assert (ref_rate > 5).all()
    np.testing.assert_allclose(gpu_rate, ref_rate, rtol=0.01)


@pytest.mark.parametrize("pass_k", [1, 2, 4])
@pytest.mark.parametrize("layout", ["dense", "spread", "gather"])
def test_every_schedule_bit_exact(dtb, orc, gpu, pass_k, layout):
    """Every step-kernel schedule (binning pass size x propagation layout)
    gives the oracle's raster and state, also when the schedule changes
    between runs with spikes pending."""
    n, steps, seed = 10000, 200, 20240
    net = dtb.gen_uniform(n, 100, seed=1)
    ids = np.arange(0, n, 5, dtype=np.uint32)
    eng = dtb.Engine(0)
    try:
        eng.schedule(pass_k, layout)
        blk = dtb.Block(net, delta_t=DT, seed=seed, engine=eng)
        blk.update_ou_background_stimuli(blk.subblk_id, 4.0, 1.5, 0.1)
        blk.set_samples(ids)
        a = blk.run(steps, output_sample_spike=True)
        assert eng.schedule()[0] == pass_k and eng.schedule()[2] == layout
        eng.schedule(2 if pass_k != 2 else 1, "spread" if layout == "dense" else "dense")
        b = blk.run(steps, output_sample_spike=True, graph=True)
        st = blk.state()
    finally:
        eng.close()
    ref, sim = run_oracle(orc, net, 2 * steps, seed, 1.5)
    assert np.array_equal(np.concatenate([a["spike"], b["spike"]]), ref["raster"][:, ids])
    assert np.array_equal(np.concatenate([a["freq"], b["freq"]]), ref["freq"])
    assert_state_equal(st, sim)


def test_autotune_changes_nothing_but_the_schedule(dtb, gpu):
    """The autotuner's trial steps are ordinary steps: two instances stepped
    through autotuning end bit-identical to the same number of steps run with
    the default schedule."""
    nets = [dtb.gen_uniform(6000, 150, seed=4), dtb.gen_voxel(20000, 10, 200, seed=6)]
    states = []
    for tune in (False, True):
        eng = dtb.Engine(0)
        try:
            blks = [dtb.Block(net, delta_t=DT, seed=60 + i, engine=eng) for i, net in enumerate(nets)]
            for b in blks:
                b.update_ou_background_stimuli(b.subblk_id, 4.0, 2.0, 0.1)
                eng.configure(b.inst, 1, True)
            if tune:
                k, layout, table = eng.autotune(8)
                print("autotune", k, layout, table)
                assert k in (1, 2, 4) and layout in ("dense", "spread", "gather")
                assert ("gather", 2) in table and len(table) == 7
                assert eng.schedule()[:2] == (k, layout)
                eng.run(100, graph_chunk=0)
            else:
                eng.run(dtb.Engine.autotune_steps(8) + 100, graph_chunk=0)
            eng.sync()
            states.append([(b.state(), b.counters()) for b in blks])
        finally:
            eng.close()
    for (s0, c0), (s1, c1) in zip(*states):
        assert c0 == c1
        for key in ("V", "J", "i_ou", "t", "f", "I_syn"):
            assert np.array_equal(np.asarray(s0[key]).view(np.uint8), np.asarray(s1[key]).view(np.uint8)), key


@pytest.mark.parametrize("n,k", [(3000, 8000), (2_200_000, 40)])
def test_device_generator_tile_counts(dtb, gpu, n, k):
    """The generator's per-tile synapse counts on both of its paths: one
    global atomic per synapse when the tile counts do not fit beside a long
    row's sort buffer (out-degree 8000: 64 KB of keys), and counts kept in
    LDS per workgroup (2.2M neurons: 8594 tile counts beside 64 keys)."""
    host = dtb.gen_fanout(n, k, 5)
    expect = np.bincount(host.col.astype(np.int64) >> 8, minlength=(n + 255) // 256)
    eng = dtb.Engine(0)
    try:
        inst = eng.add_fanout(n, k, 5, dt=DT, seed=1)
        assert np.array_equal(eng.tile_synapses(inst), expect)
    finally:
        eng.close()


def test_gather_stress_cut_bit_exact(dtb, orc, gpu):
    """Gather propagation on the propagation stress cut (in-degree 2000 on
    20k neurons, strong drive): rates and state equal the oracle's."""
    n, steps, seed = 20_000, 300, 5
    net = dtb.gen_uniform(n, 2000, seed=8)
    eng = dtb.Engine(0)
    try:
        eng.schedule(layout="gather")
        out, st, cnt = run_gpu(dtb, net, steps, seed, mu=8.0, sigma=0.5, engine=eng, graph=True)
        prop = eng.propagation(0)
        assert prop["allowed"] and prop["in_use"] and prop["group_coverage"] > 0.9
        assert eng.schedule()[2] == "gather"
    finally:
        eng.close()
    ref, sim = run_oracle(orc, net, steps, seed, mu=8.0, sigma=0.5)
    assert np.array_equal(out["freq"], ref["freq"])
    assert_state_equal(st, sim)
    assert ref["raster"].sum() > 1000
    deg = np.diff(net.rowptr)
    assert cnt[1] == int((ref["raster"].astype(np.int64) * deg[None, :]).sum())


def test_gather_mixed_engine_switching_and_checkpoint(dtb, orc, gpu):
    """One engine holding an instance that gathers and one that cannot (its
    group offsets would outweigh its synapses), switched between gather and
    queues with spikes pending, sampled, read back and checkpointed: every
    result equals the oracle (gathering instance) and the instance run alone
    with queues (the other), bit for bit."""
    na, nb, seed_a, seed_b = 8000, 200_000, 31, 32
    net_a = dtb.gen_uniform(na, 150, seed=3)
    net_b = dtb.gen_uniform(nb, 2, seed=4)
    ids = np.arange(0, na, 3, dtype=np.uint32)
    eng = dtb.Engine(0)
    try:
        eng.schedule(layout="gather")
        a = dtb.Block(net_a, delta_t=DT, seed=seed_a, engine=eng)
        b = dtb.Block(net_b, delta_t=DT, seed=seed_b, engine=eng)
        a.update_ou_background_stimuli(a.subblk_id, 4.0, 1.5, 0.1)
        b.update_ou_background_stimuli(b.subblk_id, 4.0, 9.0, 0.1)
        a.set_samples(ids)
        pa, pb = eng.propagation(a.inst), eng.propagation(b.inst)
        assert pa["allowed"] and pa["in_use"] and not pb["allowed"] and not pb["in_use"]
        parts = [a.run(120, output_sample_spike=True, output_sample_isynaptic=True)]
        mid = a.state()                        # read back with spikes pending (gather)
        eng.schedule(layout="dense")
        assert not eng.propagation(a.inst)["in_use"]
        parts.append(a.run(120, output_sample_spike=True, graph=True))
        eng.schedule(layout="gather")
        parts.append(a.run(120, output_sample_spike=True, graph=True))
        snap = a.state()
        tail = a.run(60, output_sample_spike=True)
        st_a, st_b = a.state(), b.state()
        cnt_b = b.counters()
    finally:
        eng.close()
    steps = 360
    ref, sim = run_oracle(orc, net_a, steps, seed_a, 1.5)
    raster = np.concatenate([p["spike"] for p in parts])
    assert np.array_equal(raster, ref["raster"][:, ids])
    assert np.array_equal(np.concatenate([p["freq"] for p in parts]), ref["freq"])
    assert_state_equal(snap, sim)
    ref_mid, sim_mid = run_oracle(orc, net_a, 120, seed_a, 1.5)
    assert_state_equal(mid, sim_mid)
    # the sampled I_syn of the first part equals that of a run with queues
    eng3 = dtb.Engine(0)
    try:
        a3 = dtb.Block(net_a, delta_t=DT, seed=seed_a, engine=eng3)
        a3.update_ou_background_stimuli(a3.subblk_id, 4.0, 1.5, 0.1)
        a3.set_samples(ids)
        q0 = a3.run(120, output_sample_spike=True, output_sample_isynaptic=True)
    finally:
        eng3.close()
    assert np.array_equal(bits(parts[0]["isynaptic"]), bits(q0["isynaptic"]))
    assert np.abs(q0["isynaptic"]).max() > 0
    # checkpoint taken under gather, resumed in a fresh gathering engine
    eng2 = dtb.Engine(0)
    try:
        eng2.schedule(layout="gather")
        a2 = dtb.Block(net_a, delta_t=DT, seed=seed_a, engine=eng2)
        a2.update_ou_background_stimuli(a2.subblk_id, 4.0, 1.5, 0.1)
        a2.set_samples(ids)
        a2.run(17)
        a2.load_state(snap)
        tail2 = a2.run(60, output_sample_spike=True)
        st_a2 = a2.state()
    finally:
        eng2.close()
    assert np.array_equal(tail["spike"], tail2["spike"])
    for k in ("V", "J", "i_ou", "I_syn"):
        assert np.array_equal(bits(st_a[k]), bits(st_a2[k])), k
    # the instance that cannot gather equals itself run alone with queues
    out_b, st_b_alone, cnt_b_alone = run_gpu(dtb, net_b, steps + 60, seed_b, 9.0)
    assert cnt_b == cnt_b_alone
    for k in ("V", "J", "i_ou", "I_syn"):
        assert np.array_equal(bits(st_b[k]), bits(st_b_alone[k])), k
    assert cnt_b[2] > 0


def test_gather_device_generated_stress_shape(dtb, gpu):
    """A device-generated fan-out network at the stress shape's degree (in
    5000 out of 60k neurons): gather and queue propagation give bit-identical
    state, counters and spike counts over 200 strongly driven steps."""
    n, k, nseed, seed, steps = 60_000, 5000, 11, 3, 200
    res = []
    for layout in ("dense", "gather"):
        eng = dtb.Engine(0)
        try:
            eng.schedule(layout=layout)
            blk = dtb.Block(None, delta_t=DT, engine=eng, inst=eng.add_fanout(n, k, nseed, dt=DT, seed=seed))
            blk.update_ou_background_stimuli(blk.subblk_id, 4.0, 6.0, 0.5)
            eng.configure(blk.inst, 1, True)
            out = blk.run(steps, freqs=True, graph=True)
            assert eng.propagation(blk.inst)["in_use"] == (layout == "gather")
            res.append((out["freq"], blk.state(), blk.counters()))
        finally:
            eng.close()
    (f0, s0, c0), (f1, s1, c1) = res
    assert c0 == c1 and c0[2] > 10_000
    assert np.array_equal(f0, f1)
    for key in ("V", "J", "i_ou", "I_syn"):
        assert np.array_equal(bits(s0[key]), bits(s1[key])), key


def test_run_stream_equals_run(dtb, orc, gpu):
    """Block.run_stream (dtb_run_stream, chunks of 37 steps) yields, step by
    step, what one run() returns, and both match the oracle; closing the
    stream early stops the run after the chunk in progress."""
    n, steps, seed = 6000, 300, 23
    net = dtb.gen_uniform(n, 120, seed=14)
    ids = np.arange(0, n, 5, dtype=np.uint32)

    def block():
        b = dtb.Block(net, delta_t=DT, seed=seed)
        b.update_ou_background_stimuli(b.subblk_id, 4.0, 1.5, 0.1)
        b.set_samples(ids)
        return b

    kw = dict(freqs=True, output_sample_spike=True, output_sample_vmemb=True, output_sample_isynaptic=True,
              vmean=True, imean=True, receptor_imean=True)
    b1 = block()
    whole = b1.run(steps, **kw)
    s1 = b1.state()
    b1.engine.close()
    b2 = block()
    rows = list(b2.run_stream(steps, chunk=37, **kw))
    s2 = b2.state()
    assert [k for k, _ in rows] == list(range(steps))
    for key in ("freq", "spike", "vmemb", "isynaptic", "vmean", "imean"):
        got = np.stack([r[key] for _, r in rows])
        assert np.array_equal(bits(got), bits(whole[key])), key
    rec = np.stack([r["receptor_imean"] for _, r in rows], axis=1)
    assert np.array_equal(bits(rec), bits(whole["receptor_imean"]))
    for k in ("V", "J", "I_syn", "i_ou"):
        assert np.array_equal(bits(s1[k]), bits(s2[k])), k
    # early stop: 2 steps consumed from chunks of 37 -> at most 2 chunks + 1 run
    gen = b2.run_stream(1000, chunk=37)
    next(gen)
    next(gen)
    gen.close()
    assert s2["step"] + 37 <= b2.counters()[0] <= s2["step"] + 4 * 37
    b2.engine.close()
    ref, sim = run_oracle(orc, net, steps, seed, 1.5, vtrace=True)
    assert np.array_equal(whole["spike"], ref["raster"][:, ids])
    assert np.array_equal(bits(whole["vmemb"]), bits(ref["V"][:, ids]))
