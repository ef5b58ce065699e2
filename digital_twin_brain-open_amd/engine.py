"""Python host mirror of the reference block API over the C ABI (libdtb.so).

``Block`` keeps the method vocabulary of the reference client
(``python/dist_blockwrapper.py:31-766``: ``run``, ``update_property``,
``mul_property_by_subblk``, ``assign_property_by_subblk``, ``set_samples``,
``update_ou_background_stimuli``) for one block resident on one GPU, calling
the native engine in-process instead of the gRPC/MPI service.
``Engine`` exposes the multi-instance engine (batched stepping).
"""
import ctypes
import os
from ctypes import byref, c_float, c_int, c_int32, c_int64, c_uint64, c_void_p

import numpy as np

from . import _lib

NCOL = 22
# dist_blockwrapper.py:32-33: property column -> PropType (-1: not updatable).
COLUMN_TO_PROP = np.array([19, -1, 0, -1, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18],
                          dtype=np.int32)
DEFAULT_G_UI = (0.01, 0.0005, 0.1, 0.005)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(c_void_p)


# Weight pair types kept on the device (include/weights.hpp:14-21; the
# decode of presynaptic_voltage.cpp:153-176): numpy dtype -> dtb_weight_type.
WEIGHT_TYPES = {np.dtype(np.float32): 0, np.dtype(np.float16): 1, np.dtype(np.uint8): 2}


def decode_weights(w):
    """fp32 weight pairs of stored pairs: float16 converted exactly, uint8
    codes u as float32(u) / float32(255) (correctly rounded, as the
    reference's float block decodes them)."""
    w = np.asarray(w)
    if w.dtype == np.uint8:
        return w.astype(np.float32) / np.float32(255.0)
    return w.astype(np.float32)


class NetworkArrays:
    """A block in the reference layout (src/configuration.cpp:1079-1667):
    property matrix [n][22] float32, presynaptic-major CSR (int64 rowptr,
    uint32 postsynaptic ids), weight pair per synapse (float32, float16, or
    uint8 codes of weight u / 255; kept in that type on the device), channel
    bits packed 32 per uint32 word."""

    def __init__(self, props, rowptr, col, w, kinds):
        self.props = np.ascontiguousarray(props, dtype=np.float32)
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
        self.col = np.ascontiguousarray(col, dtype=np.uint32)
        w = np.asarray(w)
        if w.dtype not in WEIGHT_TYPES:
            w = w.astype(np.float32)
        self.w = np.ascontiguousarray(w).reshape(-1, 2)
        self.kinds = np.ascontiguousarray(kinds, dtype=np.uint32)
        assert self.props.ndim == 2 and self.props.shape[1] == NCOL
        assert self.rowptr.shape[0] == self.n + 1
        assert self.col.shape[0] == self.nsyn and self.w.shape[0] == self.nsyn
        assert self.kinds.shape[0] >= (self.nsyn + 31) // 32

    @property
    def n(self):
        return int(self.props.shape[0])

    @property
    def nsyn(self):
        return int(self.col.shape[0])

    @property
    def weight_type(self):
        return WEIGHT_TYPES[self.w.dtype]

    def weights_f32(self):
        """The weight pairs as the propagation uses them (fp32)."""
        return decode_weights(self.w)

    def with_weights(self, dtype):
        """The same network with its weight pairs stored as dtype: "uint8"
        (codes round(255 w), clipped to 0..255, as make_block.py writes uint8
        blocks), "float16" or "float32" (decoded)."""
        dtype = np.dtype(dtype)
        if dtype == self.w.dtype:
            w = self.w
        elif dtype == np.uint8:
            w = np.clip(np.rint(self.weights_f32().astype(np.float64) * 255.0), 0, 255).astype(np.uint8)
        elif dtype in (np.float16, np.float32):
            w = self.weights_f32().astype(dtype)
        else:
            raise ValueError("weights are stored as float32, float16 or uint8")
        out = NetworkArrays(self.props, self.rowptr, self.col, w, self.kinds)
        if hasattr(self, "voxel_sizes"):
            out.voxel_sizes = self.voxel_sizes
        return out

    def kind_per_synapse(self):
        """Channel bit of every synapse as uint8 (0: AMPA/NMDA, 1: GABA_A/GABA_B)."""
        bits = np.unpackbits(self.kinds.view(np.uint8), bitorder="little")
        return bits[: self.nsyn].astype(np.uint8)

    def populations(self):
        """(brain ids, begin offsets [npop+1]) from property column 3."""
        sub = self.props[:, 3].astype(np.int64)
        starts = np.flatnonzero(np.r_[True, sub[1:] != sub[:-1]])
        return sub[starts].astype(np.uint32), np.r_[starts, self.n].astype(np.int64)

    def struct(self):
        return _lib.Network(self.n, _ptr(self.props), self.nsyn, _ptr(self.rowptr), _ptr(self.col),
                            _ptr(self.w), _ptr(self.kinds), self.weight_type)

    @staticmethod
    def from_edges(props, pre, post, w, kind):
        """Builds the CSR from an edge list (presynaptic id, postsynaptic id,
        weight pair, channel bit); synapses stay in input order within a row."""
        props = np.asarray(props, dtype=np.float32)
        n = props.shape[0]
        pre = np.asarray(pre, dtype=np.int64)
        order = np.argsort(pre, kind="stable")
        counts = np.bincount(pre, minlength=n)
        rowptr = np.zeros(n + 1, dtype=np.int64)
        np.cumsum(counts, out=rowptr[1:])
        kind_s = np.asarray(kind, dtype=np.uint8)[order]
        nw = (len(pre) + 31) // 32
        padded = np.zeros(nw * 32, dtype=np.uint8)
        padded[: len(pre)] = kind_s
        kinds = np.packbits(padded, bitorder="little").view(np.uint32)
        w = np.asarray(w)
        if w.dtype not in WEIGHT_TYPES:
            w = w.astype(np.float32)
        return NetworkArrays(props, rowptr, np.asarray(post, dtype=np.uint32)[order], w.reshape(-1, 2)[order],
                             kinds)


def gen_uniform(n, k, seed, g_ui=DEFAULT_G_UI, sub_base=0):
    """n neurons (80% E, 20% I), exactly k uniformly drawn in-edges each."""
    lib = _lib.load()
    nsyn = int(n) * int(k)
    props = np.empty((n, NCOL), dtype=np.float32)
    rowptr = np.empty(n + 1, dtype=np.int64)
    col = np.empty(nsyn, dtype=np.uint32)
    w = np.empty((nsyn, 2), dtype=np.float32)
    kinds = np.empty((nsyn + 31) // 32, dtype=np.uint32)
    g = np.asarray(g_ui, dtype=np.float32)
    _lib.check(lib.dtb_gen_uniform(n, k, seed, _ptr(g), sub_base, _ptr(props), _ptr(rowptr), _ptr(col), _ptr(w),
                                   _ptr(kinds)))
    return NetworkArrays(props, rowptr, col, w, kinds)


def gen_fanout(n, k, seed, g_ui=DEFAULT_G_UI, sub_base=0):
    """n neurons (80% E, 20% I); every presynaptic neuron has exactly k
    synapses to uniformly drawn targets (mean in-degree k); the network that
    Engine.add_fanout generates on the device."""
    lib = _lib.load()
    nsyn = int(n) * int(k)
    props = np.empty((n, NCOL), dtype=np.float32)
    rowptr = np.empty(n + 1, dtype=np.int64)
    col = np.empty(nsyn, dtype=np.uint32)
    w = np.empty((nsyn, 2), dtype=np.float32)
    kinds = np.empty((nsyn + 31) // 32, dtype=np.uint32)
    g = np.asarray(g_ui, dtype=np.float32)
    _lib.check(lib.dtb_gen_fanout(n, k, seed, _ptr(g), sub_base, _ptr(props), _ptr(rowptr), _ptr(col), _ptr(w),
                                  _ptr(kinds)))
    return NetworkArrays(props, rowptr, col, w, kinds)


def load_npz(path, weights="float32"):
    """Reads the reference's block files through the native reader: one
    ``block_<i>.npz`` or a directory of ``block_0.npz .. block_<B-1>.npz``
    (python/make_block.py:608-614; src/configuration.cpp:1079-1667), blocks
    concatenated in block order.  weights: "float32" converts every weight
    type to fp32 as the reference's propagation does; "native" keeps uint8
    and float16 blocks in their stored codes (the device then reads 2 or 4
    bytes per weight pair instead of 8; src/configuration.cpp:1411-1435 picks
    the type by the block path).  Returns (NetworkArrays, block_begin)."""
    lib = _lib.load()
    p = os.fsencode(os.fspath(path))
    n, s, b = c_int64(0), c_int64(0), c_int32(0)
    _lib.check(lib.dtb_npz_size(p, byref(n), byref(s), byref(b)))
    wt = c_int32(0)
    if weights == "native":
        _lib.check(lib.dtb_npz_weight_type(p, byref(wt)))
    elif weights != "float32":
        raise ValueError("weights must be 'float32' or 'native'")
    wdtype = {0: np.float32, 1: np.float16, 2: np.uint8}[wt.value]
    props = np.empty((n.value, NCOL), dtype=np.float32)
    rowptr = np.empty(n.value + 1, dtype=np.int64)
    col = np.empty(s.value, dtype=np.uint32)
    w = np.empty((s.value, 2), dtype=wdtype)
    kinds = np.empty(max((s.value + 31) // 32, 1), dtype=np.uint32)
    begin = np.empty(b.value + 1, dtype=np.int64)
    _lib.check(lib.dtb_npz_read_typed(p, _ptr(props), _ptr(rowptr), _ptr(col), _ptr(w), wt.value, _ptr(kinds),
                                      _ptr(begin)))
    return NetworkArrays(props, rowptr, col, w, kinds), begin


def save_npz(net, directory, dtype="single", block_begin=None):
    """Writes ``net`` as the reference's block files
    ``<directory>/<dtype>/block_<i>.npz`` (python/make_block.py:580-614): per
    block its properties and in-edges sorted by (postsynaptic neuron,
    relative presynaptic block), input_block_idx relative to the block,
    input_channel_offset 0 (AMPA/NMDA) or 2 (GABA_A/GABA_B).  dtype "single"
    stores the weights as float32, "double" float64, "half" float16, "uint8"
    as round(255 w).  Returns the directory holding the files."""
    wtype = {"single": np.float32, "double": np.float64, "half": np.float16, "uint8": np.uint8}[dtype]
    begin = np.asarray([0, net.n] if block_begin is None else block_begin, dtype=np.int64)
    out = os.path.join(os.fspath(directory), dtype)
    os.makedirs(out, exist_ok=True)
    pre = np.repeat(np.arange(net.n, dtype=np.int64), np.diff(net.rowptr))
    post = net.col.astype(np.int64)
    pre_blk = np.searchsorted(begin, pre, side="right") - 1
    post_blk = np.searchsorted(begin, post, side="right") - 1
    kind = net.kind_per_synapse()
    for i in range(len(begin) - 1):
        sel = np.flatnonzero(post_blk == i)
        rel = (pre_blk[sel] - i).astype(np.int16)
        onid = (post[sel] - begin[i]).astype(np.uint32)
        order = np.lexsort((rel, onid))
        sel, rel, onid = sel[order], rel[order], onid[order]
        w = net.w[sel]
        if w.dtype == wtype:
            pass   # stored codes kept as they are
        elif wtype is np.uint8:
            w = np.clip(np.rint(decode_weights(w).astype(np.float64) * 255.0), 0, 255).astype(np.uint8)
        else:
            w = decode_weights(w).astype(wtype)
        np.savez(os.path.join(out, "block_%d" % i),
                 property=net.props[begin[i]:begin[i + 1]],
                 output_neuron_idx=onid,
                 input_block_idx=rel,
                 input_neuron_idx=(pre[sel] - begin[pre_blk[sel]]).astype(np.uint32),
                 input_channel_offset=(kind[sel] * 2).astype(np.uint8),
                 weight=w)
    return out


def gen_voxel(n, nvox, k, seed, intra_frac=0.7, density=0.05, g_ui=DEFAULT_G_UI):
    """nvox lognormal-sized E/I voxels, k in-edges per neuron (intra_frac
    within the voxel, the rest from a sparse lognormal voxel matrix)."""
    lib = _lib.load()
    nsyn = int(n) * int(k)
    props = np.empty((n, NCOL), dtype=np.float32)
    rowptr = np.empty(n + 1, dtype=np.int64)
    col = np.empty(nsyn, dtype=np.uint32)
    w = np.empty((nsyn, 2), dtype=np.float32)
    kinds = np.empty((nsyn + 31) // 32, dtype=np.uint32)
    sizes = np.empty(nvox, dtype=np.uint32)
    g = np.asarray(g_ui, dtype=np.float32)
    _lib.check(lib.dtb_gen_voxel(n, nvox, k, intra_frac, density, seed, _ptr(g), _ptr(props), _ptr(rowptr),
                                 _ptr(col), _ptr(w), _ptr(kinds), _ptr(sizes)))
    net = NetworkArrays(props, rowptr, col, w, kinds)
    net.voxel_sizes = sizes
    return net


def device_count():
    c = c_int(0)
    _lib.check(_lib.load().dtb_device_count(byref(c)))
    return c.value


def gen_gamma(seed, call, first, count, alpha, beta):
    """Gamma(alpha, rate beta) draws of the property resampler for neurons
    first .. first + count - 1 of stream `call` (host)."""
    out = np.empty(count, dtype=np.float32)
    _lib.check(_lib.load().dtb_gen_gamma(seed, call, first, count, alpha, beta, _ptr(out)))
    return out


def device_memory_cached(device=0):
    """Bytes of freed device memory the library keeps for reuse on `device`."""
    b = c_uint64(0)
    _lib.check(_lib.load().dtb_device_memory_cached(device, byref(b)))
    return b.value


def device_memory_reserve(nbytes, device=0):
    """Keeps at least nbytes of device memory for the library's large buffers
    (one allocation now; instances added later are carved from it)."""
    _lib.check(_lib.load().dtb_device_memory_reserve(device, int(nbytes)))


def device_memory_trim():
    """Returns every cached device block to the HIP runtime."""
    _lib.check(_lib.load().dtb_device_memory_trim())


def ou_normals(seed, neuron0, substep, count, device=0):
    out = np.empty(count, dtype=np.float32)
    _lib.check(_lib.load().dtb_ou_normals(device, seed, neuron0, substep, count, _ptr(out)))
    return out


class Engine:
    """One GPU, one HIP stream, any number of resident instances."""

    def __init__(self, device=0):
        self._lib = _lib.load()
        h = c_void_p()
        _lib.check(self._lib.dtb_engine_create(device, byref(h)))
        self._h = h
        self.device = device

    def close(self):
        if self._h:
            _lib.check(self._lib.dtb_engine_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def add(self, net, dt=0.1, seed=0):
        s = net.struct()
        iid = c_int(-1)
        _lib.check(self._lib.dtb_instance_add(self._h, byref(s), dt, seed, byref(iid)))
        return iid.value

    def add_fanout(self, n, k, net_seed, g_ui=DEFAULT_G_UI, sub_base=0, dt=0.1, seed=0):
        """Adds the fan-out network gen_fanout(n, k, net_seed, g_ui, sub_base),
        generated in device memory."""
        g = np.ascontiguousarray(g_ui, dtype=np.float32)
        iid = c_int(-1)
        _lib.check(self._lib.dtb_instance_add_fanout(self._h, n, k, net_seed, _ptr(g), sub_base, dt, seed,
                                                     byref(iid)))
        return iid.value

    def remove(self, inst):
        _lib.check(self._lib.dtb_instance_remove(self._h, inst))

    def configure(self, inst, t_steps=1, use_ou=True):
        _lib.check(self._lib.dtb_configure(self._h, inst, t_steps, int(bool(use_ou))))

    def run(self, steps, graph_chunk=0):
        _lib.check(self._lib.dtb_engine_run(self._h, steps, graph_chunk))

    def sync(self):
        _lib.check(self._lib.dtb_engine_sync(self._h))

    def check(self):
        _lib.check(self._lib.dtb_engine_check(self._h))

    def queue_capacity(self, floor, shift):
        """Per-tile queue capacity min(S_t, floor + (S_t >> shift)); synapses
        past it go through the exact spill accumulators."""
        _lib.check(self._lib.dtb_engine_queue_capacity(self._h, int(floor), int(shift)))

    def pull(self, on=None, defer_min=None):
        """Sets (when given) and returns (on, defer_min) of the pull
        propagation: instances get their synapses grouped by target tile, a
        tile whose queue overflowed takes every emitted row through them, and
        a tile with at least defer_min emitted rows in a step leaves them to
        the pull (0: never).  Every setting gives the same results."""
        o, d = c_int32(0), ctypes.c_uint32(0)
        _lib.check(self._lib.dtb_engine_get_pull(self._h, byref(o), byref(d)))
        if on is not None or defer_min is not None:
            o = c_int32(int(bool(on)) if on is not None else o.value)
            d = ctypes.c_uint32(int(defer_min) if defer_min is not None else d.value)
            _lib.check(self._lib.dtb_engine_pull(self._h, o, d))
        return bool(o.value), int(d.value)

    def split(self, on=None, capacity=0):
        """Sets (when given) and returns split binning: the fired rows'
        binning passes run in a second launch per step over the whole GPU
        instead of in the tiles whose neurons fired (capacity: pass-list
        entries, 0 = sized by the instances; passes past it are binned by
        their own tile).  Same results either way."""
        if on is not None:
            _lib.check(self._lib.dtb_engine_split(self._h, c_int32(int(bool(on))), ctypes.c_uint32(int(capacity))))
        o, c = c_int32(0), ctypes.c_uint32(0)
        _lib.check(self._lib.dtb_engine_get_split(self._h, byref(o), byref(c)))
        return bool(o.value)

    def pull_info(self, inst):
        """{"allowed", "built", "in_use", "bytes"} of an instance's pull
        propagation (in-edges)."""
        a, b, u, n = c_int32(0), c_int32(0), c_int32(0), c_uint64(0)
        _lib.check(self._lib.dtb_instance_pull_info(self._h, inst, byref(a), byref(b), byref(u), byref(n)))
        return {"allowed": bool(a.value), "built": bool(b.value), "in_use": bool(u.value), "bytes": int(n.value)}

    LAYOUTS = {"auto": 0, "dense": 1, "spread": 2, "gather": 3}

    def schedule(self, pass_k=None, layout=None):
        """Sets (when given) and returns the step kernel's schedule: synapses
        per lane in a binning pass (1, 2, 4) and the propagation layout
        ("dense" or "spread" queue counts, "gather" for every instance that
        allows it, "auto"); every schedule gives the same results.
        Returns (pass_k, requested layout, layout in use)."""
        names = {v: k for k, v in self.LAYOUTS.items()}
        k, l, u = c_int32(0), c_int32(0), c_int32(0)
        _lib.check(self._lib.dtb_engine_get_schedule(self._h, byref(k), byref(l), byref(u)))
        if pass_k is not None or layout is not None:
            k = c_int32(int(pass_k) if pass_k is not None else k.value)
            l = c_int32(self.LAYOUTS[layout] if layout is not None else l.value)
            _lib.check(self._lib.dtb_engine_schedule(self._h, k, l))
            _lib.check(self._lib.dtb_engine_get_schedule(self._h, byref(k), byref(l), byref(u)))
        return k.value, names[l.value], names.get(u.value, "dense")

    def autotune(self, trial_steps=50):
        """Times every schedule on about trial_steps real steps, the schedules
        taking turns in 4-step windows (autotune_steps(trial_steps) steps in
        all; they advance the simulation), and keeps the fastest; returns
        (pass_k, layout, mean us per step as {(layout, pass_k): us}); the
        gather schedule is a candidate (and in the table) only when some
        instance allows it."""
        k, l = c_int32(0), c_int32(0)
        us = np.zeros(7, dtype=np.float32)
        _lib.check(self._lib.dtb_engine_autotune(self._h, int(trial_steps), byref(k), byref(l), _ptr(us)))
        names = {v: n for n, v in self.LAYOUTS.items()}
        table = {(lay, kk): float(us[i * 3 + j]) for i, lay in enumerate(("dense", "spread"))
                 for j, kk in enumerate((1, 2, 4))}
        if us[6] > 0:
            table[("gather", 2)] = float(us[6])
        return k.value, names[l.value], table

    @staticmethod
    def autotune_steps(trial_steps):
        """Steps that autotune(trial_steps) runs."""
        rounds = (int(trial_steps) + 3) // 4
        return 7 * (2 + rounds * 6)

    def propagation(self, inst):
        """Gather propagation of an instance: {"allowed", "in_use",
        "group_coverage"} (rows ordered by target group and at most 65535
        long, offsets no larger than the synapses; in use under the current
        layout; mean fraction of the 2048-neuron groups a row reaches)."""
        a, u, c = c_int32(0), c_int32(0), c_float(0.0)
        _lib.check(self._lib.dtb_instance_propagation(self._h, inst, byref(a), byref(u), byref(c)))
        return {"allowed": bool(a.value), "in_use": bool(u.value), "group_coverage": float(c.value)}

    def tile_synapses(self, inst):
        """Synapses targeting each tile of 256 neurons of an instance."""
        out = np.empty((self.info(inst)["n"] + 255) // 256, dtype=np.uint32)
        _lib.check(self._lib.dtb_instance_tile_synapses(self._h, inst, _ptr(out)))
        return out

    def spilled(self):
        """Synapses delivered through the spill accumulators so far."""
        v = c_uint64(0)
        _lib.check(self._lib.dtb_engine_spilled(self._h, byref(v)))
        return v.value

    def propagation_stats(self):
        """{"spilled": synapses past a queue's capacity, "deferred": tile steps
        that left their rows to the pull, "pulled": tile steps that took input
        through their in-edges}, since the engine was created."""
        a, b, c = c_uint64(0), c_uint64(0), c_uint64(0)
        _lib.check(self._lib.dtb_engine_propagation_stats(self._h, byref(a), byref(b), byref(c)))
        return {"spilled": a.value, "deferred": b.value, "pulled": c.value}

    def run_span(self, steps, graph_chunk=0):
        """Runs `steps` steps like run(); returns the GPU milliseconds between
        events recorded on the engine stream before and after them."""
        ms = c_float(0.0)
        _lib.check(self._lib.dtb_engine_run_span(self._h, steps, graph_chunk, byref(ms)))
        return ms.value

    def run_span_steps(self, steps, graph_chunk=0):
        """run_span with an event after every step run eagerly or graph chunk
        replayed; returns (total GPU ms, per-step ms [steps]: each step's
        share of its unit)."""
        ms = c_float(0.0)
        per = np.zeros(max(int(steps), 0), dtype=np.float32)
        _lib.check(self._lib.dtb_engine_run_span_steps(self._h, steps, graph_chunk, byref(ms), _ptr(per)))
        return ms.value, per

    def time_kernels(self, steps):
        """Runs `steps` steps; returns the total ms of the step kernel
        measured with HIP events on the engine stream."""
        a = c_float(0)
        _lib.check(self._lib.dtb_engine_time_kernels(self._h, steps, byref(a)))
        return a.value

    def set_step(self, inst, step):
        _lib.check(self._lib.dtb_set_step(self._h, inst, step))

    def counters(self, inst):
        st, ev, sp = c_int64(0), c_uint64(0), c_uint64(0)
        _lib.check(self._lib.dtb_counters(self._h, inst, byref(st), byref(ev), byref(sp)))
        return st.value, ev.value, sp.value

    def info(self, inst):
        n, ns, npop = c_int64(0), c_int64(0), c_int32(0)
        _lib.check(self._lib.dtb_instance_info(self._h, inst, byref(n), byref(ns), byref(npop), None, None))
        ids = np.empty(npop.value, dtype=np.uint32)
        beg = np.empty(npop.value + 1, dtype=np.uint32)
        _lib.check(self._lib.dtb_instance_info(self._h, inst, None, None, None, _ptr(ids), _ptr(beg)))
        return {"n": n.value, "nsyn": ns.value, "pop_ids": ids, "pop_begin": beg}

    def set_ou(self, inst, brain_ids, mean, deviation, correlation_time):
        ids = np.ascontiguousarray(brain_ids, dtype=np.uint32)
        m = np.broadcast_to(np.asarray(mean, dtype=np.float32), ids.shape).copy()
        d = np.broadcast_to(np.asarray(deviation, dtype=np.float32), ids.shape).copy()
        t = np.broadcast_to(np.asarray(correlation_time, dtype=np.float32), ids.shape).copy()
        _lib.check(self._lib.dtb_set_ou(self._h, inst, _ptr(ids), _ptr(m), _ptr(d), _ptr(t), ids.size))

    def set_samples(self, inst, ids):
        ids = np.ascontiguousarray(ids, dtype=np.uint32)
        _lib.check(self._lib.dtb_set_samples(self._h, inst, _ptr(ids), ids.size))

    def update_props(self, inst, neuron_ids, prop_ids, vals):
        nid = np.ascontiguousarray(neuron_ids, dtype=np.uint32)
        pid = np.broadcast_to(np.asarray(prop_ids, dtype=np.uint32), nid.shape).copy()
        v = np.broadcast_to(np.asarray(vals, dtype=np.float32), nid.shape).copy()
        _lib.check(self._lib.dtb_update_props(self._h, inst, _ptr(nid), _ptr(pid), _ptr(v), nid.size))

    def update_prop_cols(self, inst, prop_ids, brain_ids, vals, assign):
        bid = np.ascontiguousarray(brain_ids, dtype=np.uint32)
        pid = np.broadcast_to(np.asarray(prop_ids, dtype=np.uint32), bid.shape).copy()
        v = np.broadcast_to(np.asarray(vals, dtype=np.float32), bid.shape).copy()
        _lib.check(self._lib.dtb_update_prop_cols(self._h, inst, _ptr(pid), _ptr(bid), _ptr(v), bid.size,
                                                  int(bool(assign))))

    def gamma_prop_cols(self, inst, prop_ids, brain_ids, concentration, rate):
        bid = np.ascontiguousarray(brain_ids, dtype=np.uint32)
        pid = np.broadcast_to(np.asarray(prop_ids, dtype=np.uint32), bid.shape).copy()
        a = np.broadcast_to(np.asarray(concentration, dtype=np.float32), bid.shape).copy()
        b = np.broadcast_to(np.asarray(rate, dtype=np.float32), bid.shape).copy()
        _lib.check(self._lib.dtb_gamma_prop_cols(self._h, inst, _ptr(pid), _ptr(bid), _ptr(a), _ptr(b), bid.size))

    def read_props(self, inst):
        """The instance's property matrix [n][22] with every update applied."""
        out = np.empty((self.info(inst)["n"], NCOL), dtype=np.float32)
        _lib.check(self._lib.dtb_read_props(self._h, inst, _ptr(out)))
        return out

    def run_block(self, inst, iters, t_steps=1, use_ou=True, freqs=True, sample_spike=False, sample_vmemb=False,
                  sample_isynaptic=False, sample_iou=False, graph=False, nsample=0, npop=0, vmean=False,
                  imean=False, receptor_imean=False, noalloc=False):
        """dtb_run; noalloc: dtb_run_noalloc (records in the library's static
        device array, no device allocation; eager steps)."""
        flags = (_lib.RUN_USE_OU if use_ou else 0) | (_lib.RUN_GRAPH if graph else 0)
        out = {}
        o = _lib.RunOutputs()
        if freqs:
            flags |= _lib.RUN_FREQ
            out["freq"] = np.zeros((iters, npop), dtype=np.uint32)
            o.freq = _ptr(out["freq"])
        if vmean:
            flags |= _lib.RUN_VMEAN
            out["vmean"] = np.zeros((iters, npop), dtype=np.float32)
            o.vmean = _ptr(out["vmean"])
        if imean:
            flags |= _lib.RUN_IMEAN
            out["imean"] = np.zeros((iters, npop), dtype=np.float32)
            o.imean = _ptr(out["imean"])
        if receptor_imean:
            flags |= _lib.RUN_RECEPTOR_IMEAN
            out["receptor_imean"] = np.zeros((4, iters, npop), dtype=np.float32)
            o.ampa_imean, o.nmda_imean, o.gabaa_imean, o.gabab_imean = (_ptr(out["receptor_imean"][k])
                                                                        for k in range(4))
        for key, flag, dt, field in (("spike", _lib.RUN_SAMPLE_SPIKE, np.uint8, "sample_spike"),
                                     ("vmemb", _lib.RUN_SAMPLE_VMEMB, np.float32, "sample_vmemb"),
                                     ("isynaptic", _lib.RUN_SAMPLE_ISYN, np.float32, "sample_isynaptic"),
                                     ("iou", _lib.RUN_SAMPLE_IOU, np.float32, "sample_iou")):
            if {"spike": sample_spike, "vmemb": sample_vmemb, "isynaptic": sample_isynaptic,
                    "iou": sample_iou}[key]:
                flags |= flag
                out[key] = np.zeros((iters, nsample), dtype=dt)
                setattr(o, field, _ptr(out[key]))
        fn = self._lib.dtb_run_noalloc if noalloc else self._lib.dtb_run
        _lib.check(fn(self._h, inst, iters, t_steps, flags, byref(o)))
        return out

    def run_block_stream(self, inst, iters, chunk=100, t_steps=1, use_ou=True, freqs=True, sample_spike=False,
                         sample_vmemb=False, sample_isynaptic=False, sample_iou=False, graph=False, nsample=0, npop=0,
                         vmean=False, imean=False, receptor_imean=False, queue_chunks=2):
        """run_block streamed: yields (step index within the run, dict of
        that step's rows) while the engine keeps stepping.  The native run
        (dtb_run_stream) goes on in a worker thread, chunk by chunk; at most
        queue_chunks chunks wait for the consumer before the run pauses.
        Closing the generator early stops the run after the current chunk.
        Results are those of run_block."""
        import queue
        import threading

        flags = (_lib.RUN_USE_OU if use_ou else 0) | (_lib.RUN_GRAPH if graph else 0)
        fields = []   # (key, flag, dtype, RunOutputs field, row width)
        if freqs:
            fields.append(("freq", _lib.RUN_FREQ, np.uint32, "freq", npop))
        if vmean:
            fields.append(("vmean", _lib.RUN_VMEAN, np.float32, "vmean", npop))
        if imean:
            fields.append(("imean", _lib.RUN_IMEAN, np.float32, "imean", npop))
        for key, flag, dt, field, on in (("spike", _lib.RUN_SAMPLE_SPIKE, np.uint8, "sample_spike", sample_spike),
                                         ("vmemb", _lib.RUN_SAMPLE_VMEMB, np.float32, "sample_vmemb", sample_vmemb),
                                         ("isynaptic", _lib.RUN_SAMPLE_ISYN, np.float32, "sample_isynaptic",
                                          sample_isynaptic),
                                         ("iou", _lib.RUN_SAMPLE_IOU, np.float32, "sample_iou", sample_iou)):
            if on:
                fields.append((key, flag, dt, field, nsample))
        if receptor_imean:
            flags |= _lib.RUN_RECEPTOR_IMEAN
        for f in fields:
            flags |= f[1]
        rec_fields = ("ampa_imean", "nmda_imean", "gabaa_imean", "gabab_imean")
        q = queue.Queue(maxsize=max(1, int(queue_chunks)))
        stop = threading.Event()
        _END = object()

        def on_chunk(user, first, rows, out):
            o = out.contents
            data = {}
            for key, _, dt, field, width in fields:
                ptr = getattr(o, field)
                data[key] = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(np.ctypeslib.as_ctypes_type(dt))),
                                                  shape=(rows, width)).copy() if width else np.zeros((rows, 0), dt)
            if receptor_imean:
                data["receptor_imean"] = np.stack([np.ctypeslib.as_array(
                    ctypes.cast(getattr(o, f), ctypes.POINTER(ctypes.c_float)), shape=(rows, npop)).copy()
                    for f in rec_fields]) if npop else np.zeros((4, rows, 0), np.float32)
            while not stop.is_set():
                try:
                    q.put((first, rows, data), timeout=0.1)
                    break
                except queue.Full:
                    continue
            return 1 if stop.is_set() else 0

        cb = _lib.RUN_CHUNK_FN(on_chunk)
        result = {}

        def work():
            try:
                result["status"] = self._lib.dtb_run_stream(self._h, inst, iters, t_steps, flags, chunk, cb, None)
                if result["status"] != 0:
                    result["error"] = _lib.load().dtb_last_error()
            finally:
                q.put(_END)

        th = threading.Thread(target=work, daemon=True)
        th.start()
        self._stream_thread = th
        try:
            while True:
                item = q.get()
                if item is _END:
                    break
                first, rows, data = item
                for k in range(rows):
                    row = {key: v[k] for key, v in data.items() if key != "receptor_imean"}
                    if "receptor_imean" in data:
                        row["receptor_imean"] = data["receptor_imean"][:, k]
                    yield first + k, row
        finally:
            stop.set()
            while th.is_alive():
                try:
                    q.get(timeout=0.1)
                except queue.Empty:
                    pass
            th.join()
        if result.get("status", 0) != 0:
            msg = result.get("error")
            raise _lib.DtbError(result["status"], msg.decode() if msg else "")

    def read_state(self, inst, n):
        st = {
            "V": np.empty(n, dtype=np.float32),
            "t": np.empty(n, dtype=np.int32),
            "f": np.empty(n, dtype=np.uint8),
            "emitted": np.empty(n, dtype=np.uint8),
            "i_ou": np.empty(n, dtype=np.float32),
            "J": np.empty((n, 4), dtype=np.float32),
            "I_syn": np.empty(n, dtype=np.float32),
        }
        _lib.check(self._lib.dtb_read_state(self._h, inst, _ptr(st["V"]), _ptr(st["t"]), _ptr(st["f"]),
                                            _ptr(st["emitted"]), _ptr(st["i_ou"]), _ptr(st["J"]),
                                            _ptr(st["I_syn"])))
        return st

    def write_state(self, inst, V=None, t=None, f=None, i_ou=None, J=None):
        conv = lambda a, dt: None if a is None else np.ascontiguousarray(a, dtype=dt)
        V, t, f, i_ou, J = (conv(V, np.float32), conv(t, np.int32), conv(f, np.uint8), conv(i_ou, np.float32),
                            conv(J, np.float32))
        _lib.check(self._lib.dtb_write_state(self._h, inst, _ptr(V), _ptr(t), _ptr(f), _ptr(i_ou), _ptr(J)))


class Block:
    """One block on one GPU with the reference client vocabulary."""

    def __init__(self, net, delta_t=0.1, seed=0, device=0, engine=None, inst=None):
        """net: NetworkArrays uploaded as a new instance; or net=None with
        `inst` an instance already on `engine` (e.g. Engine.add_fanout)."""
        self.engine = engine if engine is not None else Engine(device)
        self.inst = self.engine.add(net, delta_t, seed) if net is not None else inst
        self.delta_t = delta_t
        info = self.engine.info(self.inst)
        self._n = info["n"]
        self.pop_ids = info["pop_ids"]
        self.pop_begin = info["pop_begin"]
        self._nsample = 0

    # dist_blockwrapper.py:128-151
    @property
    def total_neurons(self):
        return self._n

    @property
    def subblk_id(self):
        return self.pop_ids

    @property
    def neurons_per_subblk(self):
        return np.diff(self.pop_begin.astype(np.int64))

    def freq_char(self, freq):
        """Per-population spike counts [steps][npop] in the reference's char
        form (stat_result.cpp:103-130): lroundf(255 c / (n - blocked)) as
        uint8, with n the population size and blocked its neurons excluded
        from statistics (property column 1; no property update changes it)."""
        if getattr(self, "_stat_neurons", None) is None:
            blocked = self.props()[:, 1] != 0
            begin = self.pop_begin.astype(np.int64)
            per = np.array([int(blocked[a:b].sum()) for a, b in zip(begin[:-1], begin[1:])], dtype=np.int64)
            self._stat_neurons = (self.neurons_per_subblk - per).astype(np.float32)
        x = np.float32(255.0) * np.asarray(freq).astype(np.float32)
        with np.errstate(divide="ignore", invalid="ignore"):
            q = x / self._stat_neurons   # float32 division, as the reference's
        q = np.where(np.isfinite(q), q, np.float32(0.0))
        return np.floor(q.astype(np.float64) + 0.5).astype(np.uint8)   # lroundf of a value >= 0

    def update_ou_background_stimuli(self, population_id, correlation_time, mean, deviation):
        """dist_blockwrapper.py:583-612 (argument order kept)."""
        self.engine.set_ou(self.inst, population_id, mean, deviation, correlation_time)

    def set_samples(self, sample_idx):
        """dist_blockwrapper.py:556-581."""
        sample_idx = np.asarray(sample_idx, dtype=np.uint32)
        self.engine.set_samples(self.inst, sample_idx)
        self._nsample = int(sample_idx.size)

    def update_property(self, neuron_idx, property_idx, property_weight):
        """Per-neuron set (Snn.Updateprop).  property_idx are PropType ids."""
        self.engine.update_props(self.inst, neuron_idx, property_idx, property_weight)

    def mul_property_by_subblk(self, population_id, property_idx, hyper_parameter):
        """dist_blockwrapper.py:502-515, multiplies per population."""
        self.engine.update_prop_cols(self.inst, property_idx, population_id, hyper_parameter, assign=False)

    def assign_property_by_subblk(self, population_id, property_idx, hyper_parameter):
        """dist_blockwrapper.py:517-527, assigns per population."""
        self.engine.update_prop_cols(self.inst, property_idx, population_id, hyper_parameter, assign=True)

    def gamma_property_by_subblk(self, population_id, property_idx, concentration, rate):
        """dist_blockwrapper.py:529-554: each neuron of the population draws
        the property from Gamma(concentration, rate)."""
        self.engine.gamma_prop_cols(self.inst, property_idx, population_id, concentration, rate)

    def props(self):
        """Property matrix [n][22] with every update applied."""
        return self.engine.read_props(self.inst)

    def run(self, iterations, freqs=True, output_sample_spike=False, output_sample_vmemb=False,
            output_sample_iou=False, output_sample_isynaptic=False, iou=True, t_steps=1, graph=False,
            vmean=False, imean=False, receptor_imean=False):
        """dist_blockwrapper.py:179-350; returns a dict of per-step arrays
        ("freq", "spike", "vmemb", "isynaptic", "iou", "vmean", "imean",
        "receptor_imean" [4][iters][npop])."""
        return self.engine.run_block(self.inst, iterations, t_steps=t_steps, use_ou=iou, freqs=freqs,
                                     sample_spike=output_sample_spike, sample_vmemb=output_sample_vmemb,
                                     sample_isynaptic=output_sample_isynaptic, sample_iou=output_sample_iou,
                                     graph=graph, nsample=self._nsample, npop=len(self.pop_ids), vmean=vmean,
                                     imean=imean, receptor_imean=receptor_imean)

    def run_stream(self, iterations, chunk=100, freqs=True, output_sample_spike=False, output_sample_vmemb=False,
                   output_sample_iou=False, output_sample_isynaptic=False, iou=True, t_steps=1, graph=False,
                   vmean=False, imean=False, receptor_imean=False):
        """run() streamed: yields (step, per-step dict) while the run goes
        on, chunk steps at a time (dist_simulator.cpp:3343-3455 sends one
        RunResponse per step); the same values as run()."""
        return self.engine.run_block_stream(self.inst, iterations, chunk=chunk, t_steps=t_steps, use_ou=iou,
                                            freqs=freqs, sample_spike=output_sample_spike,
                                            sample_vmemb=output_sample_vmemb, sample_isynaptic=output_sample_isynaptic,
                                            sample_iou=output_sample_iou, graph=graph, nsample=self._nsample,
                                            npop=len(self.pop_ids), vmean=vmean, imean=imean,
                                            receptor_imean=receptor_imean)

    def state(self):
        st = self.engine.read_state(self.inst, self._n)
        st["step"] = self.engine.counters(self.inst)[0]
        return st

    def load_state(self, st):
        """Restores a checkpoint taken with state() (state and step index)."""
        self.engine.write_state(self.inst, st.get("V"), st.get("t"), st.get("f"), st.get("i_ou"), st.get("J"))
        if "step" in st:
            self.engine.set_step(self.inst, int(st["step"]))

    def counters(self):
        return self.engine.counters(self.inst)
