"""Compact weight pairs on the device (uint8 codes and float16, decoded per
synapse as presynaptic_voltage.cpp:153-176 does) against the CPU oracle,
which decodes them to fp32 once: bit-exact under queue propagation, gather
and the pull, from reference block files loaded with weights="native"."""
import numpy as np
import pytest

from test_gpu_parity import DT, assert_state_equal, bits, run_oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", ["queue", "gather", "pull"])
@pytest.mark.parametrize("dtype", ["uint8", "half"])
def test_compact_weights_bit_exact(dtb, orc, gpu, tmp_path, dtype, mode):
    n, k, steps, seed, mu = 12000, 150, 300, 19, 1.5
    net = dtb.gen_uniform(n, k, seed=6)
    path = dtb.save_npz(net, tmp_path, dtype)
    native, _ = dtb.load_npz(path, weights="native")
    assert native.w.dtype == {"uint8": np.uint8, "half": np.float16}[dtype]
    ids = np.arange(0, n, 11, dtype=np.uint32)
    print("case", dtype, mode, flush=True)
    eng = dtb.Engine(0)
    try:
        if mode == "gather":
            eng.schedule(layout="gather")
        if mode == "pull":
            eng.pull(True, 1)
            eng.queue_capacity(16, 31)
        blk = dtb.Block(native, delta_t=DT, seed=seed, engine=eng)
        blk.update_ou_background_stimuli(blk.subblk_id, 4.0, mu, 0.1)
        blk.set_samples(ids)
        out = blk.run(steps, freqs=True, output_sample_spike=True, output_sample_vmemb=True,
                      output_sample_isynaptic=True)
        st = blk.state()
        used = eng.schedule()[2]
        stats = eng.propagation_stats()
        info = eng.pull_info(blk.inst)
        eng.check()
    finally:
        eng.close()
    # queues: the automatic count layout (dense, or spread when rows reach
    # most tiles, as here: 150 synapses over 47 tiles)
    assert used == "gather" if mode == "gather" else used in ("dense", "spread")
    if mode == "pull":
        assert stats["pulled"] > 0 and stats["deferred"] > 0
        assert info["bytes"] < native.nsyn * (4 + 8)   # in-edges keep the compact weights
    ref, sim = run_oracle(orc, native, steps, seed, mu, vtrace=True)
    assert ref["raster"].sum() > 1000
    assert np.array_equal(out["freq"], ref["freq"])
    assert np.array_equal(out["spike"], ref["raster"][:, ids])
    assert np.array_equal(bits(out["vmemb"]), bits(ref["V"][:, ids]))
    assert_state_equal(st, sim)
    assert np.array_equal(bits(out["isynaptic"][-1]), bits(st["I_syn"][ids]))


def test_mixed_weight_types_in_one_engine(dtb, orc, gpu, tmp_path):
    """An fp32, a half and a uint8 block stepped by the same launches (the
    kernels then decode each instance's pairs by its own type): each against
    its own oracle run, by queues and by gather."""
    steps, mu = 250, 1.5
    base = [dtb.gen_uniform(7000, 120, seed=11), dtb.gen_uniform(5000, 90, seed=12), dtb.gen_uniform(9000, 150, seed=13)]
    nets = [base[0]]
    for net, dtype in ((base[1], "half"), (base[2], "uint8")):
        path = dtb.save_npz(net, tmp_path / dtype, dtype)
        nets.append(dtb.load_npz(path, weights="native")[0])
    assert [n.w.dtype for n in nets] == [np.float32, np.float16, np.uint8]
    for layout in ("auto", "gather"):
        eng = dtb.Engine(0)
        try:
            eng.schedule(layout=layout)
            blks = []
            for i, net in enumerate(nets):
                b = dtb.Block(net, delta_t=DT, seed=40 + i, engine=eng)
                b.update_ou_background_stimuli(b.subblk_id, 4.0, mu, 0.1)
                eng.configure(b.inst, 1, True)
                blks.append(b)
            eng.run(steps, graph_chunk=50)
            res = [(b.state(), b.counters()) for b in blks]
            eng.check()
        finally:
            eng.close()
        for i, net in enumerate(nets):
            ref, sim = run_oracle(orc, net, steps, 40 + i, mu)
            assert ref["raster"].sum() > 500, (layout, i)
            assert res[i][1][2] == int(ref["raster"].sum()), (layout, i)
            assert_state_equal(res[i][0], sim)
