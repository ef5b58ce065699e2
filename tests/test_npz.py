"""The reference block-file reader (dtb_npz_size / dtb_npz_read) on the CPU.

Block files follow python/make_block.py:608-614 (np.savez with property,
output_neuron_idx, input_block_idx (relative), input_neuron_idx,
input_channel_offset, weight) as read by src/configuration.cpp:1079-1667.
"""
import os

import numpy as np
import pytest


def _props(n, sub):
    p = np.zeros((n, 22), dtype=np.float32)
    p[:, 3] = sub
    p[:, 4] = 1.0
    return p


def _rows(net):
    """Per presynaptic row: sorted list of (post, w0, w1, kind)."""
    kind = net.kind_per_synapse()
    out = []
    for r in range(net.n):
        a, b = net.rowptr[r], net.rowptr[r + 1]
        out.append(sorted(zip(net.col[a:b].tolist(), net.w[a:b, 0].tolist(), net.w[a:b, 1].tolist(),
                              kind[a:b].tolist())))
    return out


def test_hand_built_two_blocks(dtb, tmp_path):
    """Two block files written by hand; the reader must produce exactly the
    CSR that the format defines (global id = block base + local id)."""
    d = tmp_path / "single"
    d.mkdir()
    # block 0: neurons 0..2, block 1: neurons 3..4 (global)
    # in-edges of block 0: 1->0 (E), 3->2 (I, from block 1 = rel +1), 4->0 (E)
    np.savez(d / "block_0", property=_props(3, 0),
             output_neuron_idx=np.array([0, 2, 0], np.uint32),
             input_block_idx=np.array([0, 1, 1], np.int16),
             input_neuron_idx=np.array([1, 0, 1], np.uint32),
             input_channel_offset=np.array([0, 2, 0], np.uint8),
             weight=np.array([[0.5, 0.25], [0.125, 1.0], [0.75, 0.0]], np.float32))
    # in-edges of block 1: 0->3 (E, rel -1), 3->4 (I, rel 0)
    np.savez(d / "block_1", property=_props(2, 1),
             output_neuron_idx=np.array([0, 1], np.uint32),
             input_block_idx=np.array([-1, 0], np.int16),
             input_neuron_idx=np.array([0, 0], np.uint32),
             input_channel_offset=np.array([0, 2], np.uint8),
             weight=np.array([[0.375, 0.0625], [0.5, 0.5]], np.float32))
    net, begin = dtb.load_npz(d)
    assert begin.tolist() == [0, 3, 5]
    assert net.n == 5 and net.nsyn == 5
    assert net.rowptr.tolist() == [0, 1, 2, 2, 4, 5]
    rows = _rows(net)
    assert rows[0] == [(3, 0.375, 0.0625, 0)]
    assert rows[1] == [(0, 0.5, 0.25, 0)]
    assert rows[2] == []
    assert rows[3] == [(2, 0.125, 1.0, 1), (4, 0.5, 0.5, 1)]
    assert rows[4] == [(0, 0.75, 0.0, 0)]
    assert np.array_equal(net.props[:3, 3], [0, 0, 0]) and np.array_equal(net.props[3:, 3], [1, 1])


def test_roundtrip_multi_block(dtb, tmp_path):
    net = dtb.gen_voxel(3000, 12, 40, seed=3)
    begin = np.array([0, 1100, 2300, 3000])
    out = dtb.save_npz(net, tmp_path, "single", block_begin=begin)
    assert sorted(os.listdir(out)) == ["block_0.npz", "block_1.npz", "block_2.npz"]
    back, b2 = dtb.load_npz(out)
    assert b2.tolist() == begin.tolist()
    assert np.array_equal(back.props, net.props)
    assert np.array_equal(back.rowptr, net.rowptr)
    assert _rows(back) == _rows(net)
    # every row comes back sorted by postsynaptic id
    for r in range(0, back.n, 97):
        c = back.col[back.rowptr[r]:back.rowptr[r + 1]]
        assert np.all(c[1:] >= c[:-1])
    # a single block file of a multi-block model refers to other blocks
    with pytest.raises(dtb.DtbError):
        dtb.load_npz(os.path.join(out, "block_1.npz"))


@pytest.mark.parametrize("dtype", ["double", "half", "uint8"])
def test_weight_dtypes(dtb, tmp_path, dtype):
    """float64 / float16 weights are cast, uint8 weights are code / 255
    (presynaptic_voltage.cpp:153-176)."""
    net = dtb.gen_uniform(500, 20, seed=2)
    rng = np.random.default_rng(0)
    codes = rng.integers(0, 256, size=net.w.shape)
    if dtype == "uint8":
        net.w[:] = (codes.astype(np.float32) / np.float32(255.0))
    elif dtype == "half":
        net.w[:] = codes.astype(np.float16).astype(np.float32) / np.float32(256.0)
    out = dtb.save_npz(net, tmp_path, dtype)
    back, _ = dtb.load_npz(out)
    if dtype == "double":
        expect = net.w.astype(np.float64).astype(np.float32)
    else:
        expect = net.w
    ref = dtb.NetworkArrays(net.props, net.rowptr, net.col, expect, net.kinds)
    assert _rows(back) == _rows(ref)


def test_deflated_archive(dtb, tmp_path):
    net = dtb.gen_uniform(800, 30, seed=5)
    d = tmp_path / "single"
    d.mkdir()
    kind = net.kind_per_synapse()
    pre = np.repeat(np.arange(net.n), np.diff(net.rowptr))
    np.savez_compressed(d / "block_0", property=net.props, output_neuron_idx=net.col,
                        input_block_idx=np.zeros(net.nsyn, np.int16), input_neuron_idx=pre.astype(np.uint32),
                        input_channel_offset=(kind * 2).astype(np.uint8), weight=net.w)
    back, _ = dtb.load_npz(d / "block_0.npz")
    assert _rows(back) == _rows(net)


def test_rejects_malformed(dtb, tmp_path):
    d = tmp_path / "single"
    d.mkdir()
    good = dict(property=_props(2, 0), output_neuron_idx=np.array([1], np.uint32),
                input_block_idx=np.array([0], np.int16), input_neuron_idx=np.array([0], np.uint32),
                input_channel_offset=np.array([0], np.uint8), weight=np.array([[0.5, 0.5]], np.float32))
    cases = {
        "missing": {k: v for k, v in good.items() if k != "weight"},
        "channel": dict(good, input_channel_offset=np.array([1], np.uint8)),
        "post_range": dict(good, output_neuron_idx=np.array([2], np.uint32)),
        "pre_range": dict(good, input_neuron_idx=np.array([5], np.uint32)),
        "prop_cols": dict(good, property=np.zeros((2, 21), np.float32)),
        "weight_shape": dict(good, weight=np.array([0.5, 0.5], np.float32)),
    }
    for name, arrays in cases.items():
        np.savez(d / ("bad_" + name), **arrays)
        with pytest.raises(dtb.DtbError):
            dtb.load_npz(d / ("bad_" + name + ".npz"))
    (d / "junk.npz").write_bytes(b"not a zip archive at all, just bytes")
    with pytest.raises(dtb.DtbError):
        dtb.load_npz(d / "junk.npz")
    with pytest.raises(dtb.DtbError):
        dtb.load_npz(tmp_path / "does_not_exist")


def test_uint8_weight_decode_is_correctly_rounded(dtb):
    """uint8 weight codes u decode to the fp32 value nearest to u / 255
    (presynaptic_voltage.cpp:171-176 divides in float), for all 256 codes:
    checked with exact rational arithmetic, independently of any float
    division."""
    from fractions import Fraction
    w = dtb.engine.decode_weights(np.arange(256, dtype=np.uint8))
    for u in range(256):
        x = Fraction(u, 255)
        got = Fraction(float(w[u]))
        below = np.nextafter(np.float32(w[u]), np.float32(-1))
        above = np.nextafter(np.float32(w[u]), np.float32(2))
        # no float32 is closer to u / 255 than the decoded value
        assert abs(got - x) <= abs(Fraction(float(below)) - x), u
        assert abs(got - x) <= abs(Fraction(float(above)) - x), u


@pytest.mark.parametrize("dtype", ["uint8", "half"])
def test_native_weight_types_round_trip(dtb, tmp_path, dtype):
    """A network stored as uint8 or float16 block files loads with its codes
    unchanged (weights="native"), and the fp32 load equals the decode of those
    codes: the reader's fp32 conversion and decode_weights (the oracle's
    decode) agree for every synapse."""
    np_dtype = {"uint8": np.uint8, "half": np.float16}[dtype]
    net = dtb.gen_uniform(3000, 40, seed=5).with_weights(np_dtype)
    assert net.w.dtype == np_dtype
    path = dtb.save_npz(net, tmp_path, dtype, block_begin=[0, 1200, 3000])
    native, _ = dtb.load_npz(path, weights="native")
    f32, _ = dtb.load_npz(path)
    assert native.w.dtype == np_dtype and f32.w.dtype == np.float32
    assert np.array_equal(native.rowptr, net.rowptr) and np.array_equal(native.col, net.col)
    assert np.array_equal(native.w.view(np.uint8), net.w.view(np.uint8))
    assert np.array_equal(f32.w.view(np.uint32), dtb.engine.decode_weights(native.w).view(np.uint32))
    assert np.array_equal(native.kinds, f32.kinds)


def test_native_load_rejects_mixed_weight_types(dtb, tmp_path):
    net = dtb.gen_uniform(1000, 20, seed=2)
    d = dtb.save_npz(net, tmp_path, "single", block_begin=[0, 500, 1000])
    # replace block 1's weights by uint8 codes: the blocks then disagree
    z = dict(np.load(os.path.join(d, "block_1.npz")))
    z["weight"] = np.clip(np.rint(z["weight"] * 255), 0, 255).astype(np.uint8)
    np.savez(os.path.join(d, "block_1"), **z)
    with pytest.raises(dtb.DtbError):
        dtb.load_npz(d, weights="native")
