"""BlockWrapper population ids across block files (dist_blockwrapper.py:79-117
rebase, :153-177 merge).  The rebase cases are CPU-only; the merged outputs
are checked on the GPU against the oracle's per-part results."""
import numpy as np
import pytest

from test_gpu_parity import DT, bits


def parts(dtb, blocks, **kw):
    """Caller population id of every (block, sub-block) part for blocks given
    as lists of sub-block ids (one neuron per id)."""
    sub = np.concatenate([np.asarray(b, dtype=np.float32) for b in blocks])
    begin = np.concatenate([[0], np.cumsum([len(b) for b in blocks])])
    args = dict(allow_rebase=True, force_rebase=False, overlap=2)
    args.update(kw)
    part_of, part_id, part_block = dtb.BlockWrapper._parts(sub, begin, **args)
    return part_id.tolist(), part_of.tolist(), part_block.tolist()


@pytest.mark.parametrize("blocks,kw,expect", [
    # a second block starting at 0 again lands above the ids so far
    ([[0, 1], [0, 1]], {}, [0, 1, 2, 3]),
    # rounded up to a multiple of overlap: max 2 -> base 4
    ([[0, 1, 2], [0, 1]], {}, [0, 1, 2, 4, 5]),
    # rebasing off: repeated ids stay (and are merged)
    ([[0, 1], [0, 1]], {"allow_rebase": False}, [0, 1, 0, 1]),
    # the ids so far end at the start id itself: no rebase unless forced
    ([[0], [0, 1]], {}, [0, 0, 1]),
    ([[0], [0, 1]], {"force_rebase": True}, [0, 2, 3]),
    # overlap 1: the cortical start id is 2
    ([[2, 3], [2, 3]], {"overlap": 1}, [2, 3, 6, 7]),
    # a block not starting at the start id keeps its ids, and the base of an
    # earlier rebase carries over to later blocks
    ([[0, 1], [5, 6]], {}, [0, 1, 5, 6]),
    ([[0, 1], [0, 1], [7]], {}, [0, 1, 2, 3, 9]),
])
def test_rebase_cases(dtb, blocks, kw, expect):
    part_id, part_of, part_block = parts(dtb, blocks, **kw)
    assert part_id == expect
    assert part_of == list(range(len(expect)))
    assert part_block == [b for b, ids in enumerate(blocks) for _ in ids]


@pytest.mark.gpu
def test_merged_populations_match_oracle(dtb, orc, gpu, tmp_path):
    """Two block files with the same sub-block ids {0, 1} loaded without
    rebasing: the caller sees populations 0 and 1, each the union of one
    part per block.  Spike counts are the parts' sums, V means the
    neuron-weighted means of the parts' means, and an OU or multiplier
    update by population reaches both parts."""
    net = dtb.gen_uniform(4000, 100, seed=12)
    begin = [0, 1500, 4000]
    sub = np.zeros(4000, np.float32)
    sub[800:1500] = 1   # block 0: 0..799 -> 0, 800..1499 -> 1
    sub[2700:] = 1      # block 1: 1500..2699 -> 0, 2700..3999 -> 1
    net.props[:, 3] = sub
    path = dtb.save_npz(net, tmp_path, "single", block_begin=begin)
    bw = dtb.BlockWrapper(None, path, DT, seed=4, allow_rebase=False)
    try:
        assert list(bw.subblk_id) == [0, 1]
        assert list(bw.neurons_per_subblk) == [800 + 1200, 700 + 1300]
        bw.update_ou_background_stimuli(np.array([0, 1]), np.array([4.0, 4.0], np.float32),
                                         np.array([1.5, 1.6], np.float32), np.array([0.1, 0.1], np.float32))
        bw.mul_property_by_subblk(np.array([[1, 10]]), np.array([1.3], np.float32))
        outs = list(bw.run(200, freqs=True, vmean=True, iou=True))
    finally:
        bw.close()
    freq = np.stack([o[0] for o in outs])
    vmean = np.stack([o[1] for o in outs])
    # the oracle sees the four parts as populations 0..3
    part = np.zeros(4000, np.float32)
    part[800:1500], part[1500:2700], part[2700:] = 1, 2, 3
    net.props[:, 3] = part
    on = orc.OracleNet.from_network(net, dt=DT, seed=4)
    on.set_ou(np.array([0, 1, 2, 3]), np.array([1.5, 1.6, 1.5, 1.6], np.float32), 0.1, 4.0)
    for q in (1, 3):
        sel = on.props[:, 3] == q
        on.props[sel, 10] = on.props[sel, 10] * np.float32(1.3)
    ref = orc.Sim(on, 32).run(200, raster=True, pop_counts=True, stats=True)
    pf = ref["freq"].astype(np.int32)
    assert np.array_equal(freq, np.stack([pf[:, 0] + pf[:, 2], pf[:, 1] + pf[:, 3]], axis=1))
    n = np.array([800, 700, 1200, 1300], np.float32)
    ratio = n / np.array([2000, 2000, 2000, 2000], np.float32)
    m = ref["stat_means"][:, :, 0] * ratio
    expect = np.stack([m[:, 0] + m[:, 2], m[:, 1] + m[:, 3]], axis=1).astype(np.float32)
    assert np.array_equal(bits(vmean), bits(expect))
    assert freq.sum() > 100
