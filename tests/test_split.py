"""Split binning (the fired rows' binning passes listed by the step kernel
and binned by a second launch over the whole GPU) against the CPU oracle:
bit-exact for every pass size, with the pass list large enough and with a
list so small that most passes fall back to their own tile; switched on and
off between runs with spikes pending; in a batched engine, with the pull and
with queue overflow."""
import numpy as np
import pytest

from test_gpu_parity import DT, assert_state_equal, bits, run_oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("capacity", [0, 3])
@pytest.mark.parametrize("pass_k", [1, 2, 4])
def test_split_binning_bit_exact(dtb, orc, gpu, pass_k, capacity):
    n, steps, seed, mu = 10000, 200, 20240, 1.5
    net = dtb.gen_uniform(n, 100, seed=1)
    ids = np.arange(0, n, 5, dtype=np.uint32)
    print("case pass_k", pass_k, "capacity", capacity, flush=True)
    eng = dtb.Engine(0)
    try:
        eng.schedule(pass_k, "dense")
        eng.split(True, capacity)
        blk = dtb.Block(net, delta_t=DT, seed=seed, engine=eng)
        blk.update_ou_background_stimuli(blk.subblk_id, 4.0, mu, 0.1)
        blk.set_samples(ids)
        a = blk.run(steps, output_sample_spike=True, output_sample_vmemb=True, output_sample_isynaptic=True)
        assert eng.split()
        eng.split(False)   # off with spikes pending, then on again in graphs
        b = blk.run(steps // 2, output_sample_spike=True, output_sample_vmemb=True, output_sample_isynaptic=True)
        eng.split(True, capacity)
        c = blk.run(steps // 2, output_sample_spike=True, output_sample_vmemb=True, output_sample_isynaptic=True,
                    graph=True)
        st = blk.state()
        eng.check()
    finally:
        eng.close()
    ref, sim = run_oracle(orc, net, 2 * steps, seed, mu, vtrace=True)
    assert ref["raster"].sum() > 1000
    got = {k: np.concatenate([a[k], b[k], c[k]]) for k in ("spike", "vmemb", "freq")}
    assert np.array_equal(got["spike"], ref["raster"][:, ids])
    assert np.array_equal(bits(got["vmemb"]), bits(ref["V"][:, ids]))
    assert np.array_equal(got["freq"], ref["freq"])
    assert_state_equal(st, sim)
    assert np.array_equal(bits(c["isynaptic"][-1]), bits(st["I_syn"][ids]))


@pytest.mark.parametrize("pull", [False, True])
def test_split_binning_batched_volley_and_overflow(dtb, orc, gpu, pull):
    """Three instances of different sizes in one engine from rest (their
    volleys overflow small queues), split binning with the spill path or the
    pull; each instance against its own oracle run."""
    nets = [dtb.gen_uniform(9000, 120, seed=2), dtb.gen_voxel(30000, 12, 150, seed=3),
            dtb.gen_uniform(3000, 60, seed=4)]
    steps, mus = 300, (1.5, 2.5, 2.0)
    eng = dtb.Engine(0)
    try:
        eng.pull(pull, 16)
        eng.queue_capacity(64, 4)
        eng.split(True)
        blks = []
        for i, net in enumerate(nets):
            b = dtb.Block(net, delta_t=DT, seed=70 + i, engine=eng)
            b.update_ou_background_stimuli(b.subblk_id, 4.0, mus[i], 0.1)
            eng.configure(b.inst, 1, True)
            blks.append(b)
        eng.run(steps, graph_chunk=50)
        res = [(b.state(), b.counters()) for b in blks]
        stats = eng.propagation_stats()
        eng.check()
    finally:
        eng.close()
    print("stats", stats)
    if pull:
        assert stats["pulled"] > 0    # volleys deferred to the pull, or queues overflowed into it
    else:
        assert stats["spilled"] > 0   # queues overflowed into the spill accumulators
    for i, net in enumerate(nets):
        ref, sim = run_oracle(orc, net, steps, 70 + i, mus[i])
        deg = np.diff(net.rowptr)
        events = int((ref["raster"].astype(np.int64) * deg[None, :]).sum())
        assert res[i][1] == (steps, events, int(ref["raster"].sum())), i
        assert_state_equal(res[i][0], sim)
