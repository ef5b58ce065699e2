"""In-process ``BlockWrapper`` with the signatures of the reference client
(python/dist_blockwrapper.py:31-766), stepping the blocks on the local GPU
through libdtb.so instead of a gRPC/MPI simulator.

Arguments keep the reference's meaning:
- ``property_idx`` is an int64 [k, 2] array of (neuron id or population id,
  property-matrix column); columns map to PropType through
  ``property_idx_trans`` (dist_blockwrapper.py:32-33);
- ``mul_property_by_subblk`` multiplies relative to the value the block was
  loaded with unless ``not_divided`` (it divides by the last multiplier it
  sent, dist_blockwrapper.py:502-515);
- ``set_samples`` takes int64 neuron ids in any order with repeats; outputs
  come back in that order (dist_blockwrapper.py:556-581);
- ``run`` is a generator yielding one item per step: a single array, or a
  tuple in the order freqs, vmean, sample spike, sample V, sample I_syn,
  sample i_ou, imean, receptor imeans (AMPA, NMDA, GABA_A, GABA_B)
  (dist_blockwrapper.py:179-350).
Inputs may be numpy arrays or torch tensors; outputs are numpy arrays.
Several block files in one directory are concatenated into one block on the
one GPU (their connections become ordinary synapses).

Population ids follow the reference client (dist_blockwrapper.py:79-117):
each block reports its own sub-block ids; a block whose first id is the
cortical start (0 with overlap 2, else 2) after others is rebased above the
largest id so far (rounded up to a multiple of overlap) when allow_rebase and
(force_rebase or that id differs from the start).  Populations that end up
with the same id in several blocks are one population to the caller: their
spike counts are summed and their means weighted by neuron counts
(_merge_sbblk, :153-177), and updates by population reach every part.  On
the device every (block, sub-block) part stays its own population (property
column 3 holds the part's index).
"""
import numpy as np

from .engine import COLUMN_TO_PROP, Block, NetworkArrays, load_npz
from . import _lib


def _np(x, dtype=None):
    if hasattr(x, "detach"):   # torch tensor
        x = x.detach().cpu().numpy()
    a = np.asarray(x)
    return a if dtype is None else a.astype(dtype, copy=False)


class BlockWrapper:
    """dist_blockwrapper.py:31-766 on one local GPU."""

    property_idx_trans = COLUMN_TO_PROP
    buffersize = 1 << 20

    def __init__(self, address, path, delta_t, use_route=False, print_stat=False, force_rebase=False,
                 allow_rebase=True, overlap=2, seed=0, device=0):
        """``address`` is not used (no server: the block runs on ``device`` of
        this process); ``path`` is a block file, a directory of block files,
        or a NetworkArrays; ``seed`` keys the OU/noise streams."""
        self.print_stat = print_stat
        if isinstance(path, NetworkArrays):
            net, begin = path, np.array([0, path.n], dtype=np.int64)
        else:
            net, begin = load_npz(path)
        begin = np.asarray(begin, dtype=np.int64)
        part_of, part_id, part_block = self._parts(net.props[:, 3], begin, allow_rebase, force_rebase, overlap)
        props = net.props.copy()
        props[:, 3] = part_of.astype(np.float32)   # device populations: one per (block, sub-block) part
        net = NetworkArrays(props, net.rowptr, net.col, net.w, net.kinds)
        self._block = Block(net, delta_t=delta_t, seed=seed, device=device)
        self._neurons_per_block = np.diff(begin).astype(np.int64)
        self._block_id = np.arange(len(self._neurons_per_block), dtype=np.int64)
        self._neurons_thrush = begin.astype(np.int64)
        # parts (device populations, in order) -> caller population ids
        self._part_id = part_id
        self._part_block = part_block
        self._part_neurons = self._block.neurons_per_subblk.astype(np.int64)
        assert self._part_neurons.shape[0] == part_id.shape[0]
        self._subblk_id, self._part_row = np.unique(part_id, return_inverse=True)
        if self._part_row.shape[0] == self._subblk_id.shape[0]:
            self._subblk_idx = None
            self._neurons_per_subblk = self._part_neurons
            self.merge_popu_ratio = None
        else:
            self._subblk_idx = self._part_row
            self._neurons_per_subblk = np.bincount(self._part_row, weights=self._part_neurons,
                                                   minlength=self._subblk_id.shape[0]).astype(np.int64)
            self.merge_popu_ratio = (self._part_neurons.astype(np.float32)
                                     / self._neurons_per_subblk[self._part_row].astype(np.float32))
        self._reset_hyper_parameter()
        self._sample_order = None
        self._sample_num = 0
        self._iterations = 0

    @staticmethod
    def _parts(sub, begin, allow_rebase, force_rebase, overlap):
        """Per neuron its part index; per part its caller population id and
        block (dist_blockwrapper.py:79-98: each block's sub-block ids in
        ascending order, rebased block by block)."""
        sub = np.asarray(sub).astype(np.int64)
        part_of = np.empty(sub.shape[0], dtype=np.int64)
        part_id, part_block = [], []
        seen = []           # caller ids so far (subblk_info)
        base = 0
        start = 0 if overlap == 2 else 2
        for b in range(len(begin) - 1):
            lo, hi = int(begin[b]), int(begin[b + 1])
            local = np.unique(sub[lo:hi])
            for j, sid in enumerate(local):
                if j == 0 and sid == start and seen:
                    new_base = max(seen)
                    if allow_rebase and (force_rebase or new_base != start):
                        base = (new_base + 1 + overlap - 1) // overlap * overlap
                seen.append(int(sid) + base)
            # parts in neuron order (a sub-block is a contiguous range of its block)
            starts = lo + np.flatnonzero(np.r_[True, sub[lo + 1:hi] != sub[lo:hi - 1]]) if hi > lo else []
            for k, a in enumerate(starts):
                e = starts[k + 1] if k + 1 < len(starts) else hi
                part_of[a:e] = len(part_id)
                part_id.append(int(sub[a]) + base)
                part_block.append(b)
        return part_of, np.asarray(part_id, dtype=np.int64), np.asarray(part_block, dtype=np.int64)

    def _merge_sbblk(self, array, weight=None):
        """dist_blockwrapper.py:165-177: per-part columns [..., parts] summed
        into the caller's populations [..., populations], weighted first when
        given."""
        if self._subblk_idx is None:
            return array
        if weight is not None:
            array = array * weight
        out = np.zeros(array.shape[:-1] + (self._subblk_id.shape[0],), dtype=array.dtype)
        for k, r in enumerate(self._subblk_idx):
            out[..., r] += array[..., k]
        return out

    def _parts_of(self, population_id):
        """Device population ids (parts) of caller population ids, with the
        index of the caller id each part serves."""
        pop = np.asarray(population_id, dtype=np.int64).reshape(-1)
        rows = self._subblk_rows(pop)
        parts, which = [], []
        for i, r in enumerate(rows):
            for k in np.flatnonzero(self._part_row == r):
                parts.append(k)
                which.append(i)
        return np.asarray(parts, dtype=np.uint32), np.asarray(which, dtype=np.int64)

    def _reset_hyper_parameter(self):
        self._last_hyper_parameter = np.ones([self._subblk_id.shape[0], 20], dtype=np.float32)

    # dist_blockwrapper.py:127-151
    @property
    def total_neurons(self):
        return int(self._neurons_per_block.sum())

    @property
    def block_id(self):
        return self._block_id

    @property
    def subblk_id(self):
        return self._subblk_id

    @property
    def total_subblks(self):
        return int(self._subblk_id.shape[0])

    @property
    def neurons_per_subblk(self):
        return self._neurons_per_subblk

    @property
    def neurons_per_block(self):
        return self._neurons_per_block

    @property
    def block(self):
        """The underlying single-GPU Block."""
        return self._block

    def run(self, iterations, freqs=True, freq_char=False, vmean=False, output_sample_spike=False,
            output_sample_vmemb=False, output_sample_iou=False, output_sample_isynaptic=False,
            receptor_imean=False, imean=False, iou=False, t_steps=1, equal_sample=False):
        if (output_sample_spike or output_sample_vmemb or output_sample_iou or output_sample_isynaptic) \
                and self._sample_num == 0:
            raise _lib.DtbError(_lib.UNSET_SAMPLE_FAILURE, "samples are not set")
        out = self._block.run(iterations, freqs=freqs, vmean=vmean, output_sample_spike=output_sample_spike,
                              output_sample_vmemb=output_sample_vmemb, output_sample_iou=output_sample_iou,
                              output_sample_isynaptic=output_sample_isynaptic, imean=imean,
                              receptor_imean=receptor_imean, iou=iou, t_steps=t_steps)
        self._iterations += iterations
        seq = []
        ratio = self.merge_popu_ratio
        if freqs:
            f = self._merge_sbblk(out["freq"].astype(np.int32))
            if freq_char:
                # stat_result.cpp:103-130 sends lroundf(255 c / (n - blocked))
                # per population; the client scales it back by the
                # population size (dist_blockwrapper.py:285-292).
                n = self._part_neurons.astype(np.float32)
                ch = self._block.freq_char(out["freq"])
                f = self._merge_sbblk(ch.astype(np.float32) / np.float32(255.0) * n).astype(np.int32)
            seq.append(f)
        if vmean:
            seq.append(self._merge_sbblk(out["vmean"], ratio))
        for flag, key in ((output_sample_spike, "spike"), (output_sample_vmemb, "vmemb"),
                          (output_sample_isynaptic, "isynaptic"), (output_sample_iou, "iou")):
            if flag:
                a = out[key]
                if self._sample_order is not None:
                    a = a[:, self._sample_order]
                seq.append(a)
        if imean:
            seq.append(self._merge_sbblk(out["imean"], ratio))
        if receptor_imean:
            seq.extend(self._merge_sbblk(out["receptor_imean"][k], ratio) for k in range(4))
        for j in range(iterations):
            if len(seq) == 1:
                yield seq[0][j]
            else:
                yield tuple(a[j] for a in seq)

    def _prop_ids(self, cols):
        p = self.property_idx_trans[cols]
        if (p < 0).any():
            raise _lib.DtbError(_lib.INVALID_PARAMETER, "property column is not updatable")
        return p.astype(np.uint32)

    def update_property(self, property_idx, property_weight, bid=None):
        """dist_blockwrapper.py:369-433: (neuron id, column) pairs and values."""
        idx = _np(property_idx, np.int64).reshape(-1, 2)
        w = _np(property_weight, np.float32).reshape(-1)
        assert w.shape[0] == idx.shape[0]
        nid = idx[:, 0]
        if bid is not None:
            nid = nid + self._neurons_thrush[int(bid)]
        self._block.engine.update_props(self._block.inst, nid.astype(np.uint32), self._prop_ids(idx[:, 1]), w)
        self._reset_hyper_parameter()   # dist_blockwrapper.py:432

    def _subblk_rows(self, pop):
        rows = np.searchsorted(self._subblk_id, pop)
        if (rows >= len(self._subblk_id)).any() or (self._subblk_id[np.minimum(rows, len(self._subblk_id) - 1)]
                                                     != pop).any():
            raise _lib.DtbError(_lib.INVALID_PARAMETER, "unknown population id")
        return rows

    def mul_property_by_subblk(self, property_idx, property_hyper_parameter, not_divided=False):
        """dist_blockwrapper.py:502-515."""
        idx = _np(property_idx, np.int64).reshape(-1, 2)
        hp = _np(property_hyper_parameter, np.float32).reshape(-1).copy()
        prop = self._prop_ids(idx[:, 1])
        if not not_divided:
            rows = self._subblk_rows(idx[:, 0])
            hp = hp / self._last_hyper_parameter[rows, prop]
            self._last_hyper_parameter[rows, prop] *= hp
        parts, which = self._parts_of(idx[:, 0])
        self._block.engine.update_prop_cols(self._block.inst, prop[which], parts, hp[which], assign=False)

    def assign_property_by_subblk(self, property_idx, property_hyper_parameter):
        """dist_blockwrapper.py:517-527."""
        idx = _np(property_idx, np.int64).reshape(-1, 2)
        hp = _np(property_hyper_parameter, np.float32).reshape(-1)
        prop = self._prop_ids(idx[:, 1])
        parts, which = self._parts_of(idx[:, 0])
        self._block.engine.update_prop_cols(self._block.inst, prop[which], parts, hp[which], assign=True)

    def gamma_property_by_subblk(self, property_idx, gamma_concentration, gamma_rate):
        """dist_blockwrapper.py:529-554: every neuron of each listed
        population draws the property from Gamma(concentration, rate);
        concentration None means 1."""
        idx = _np(property_idx, np.int64).reshape(-1, 2)
        rate = _np(gamma_rate, np.float32).reshape(-1)
        conc = np.ones_like(rate) if gamma_concentration is None else _np(gamma_concentration, np.float32).reshape(-1)
        if not ((conc >= 0).all() and (rate >= 0).all()):
            raise ValueError("gamma hyperparameters must be non-negative")
        prop = self._prop_ids(idx[:, 1])
        rows = self._subblk_rows(idx[:, 0])
        self._last_hyper_parameter[rows, prop] = 1.0   # dist_blockwrapper.py:538-540
        parts, which = self._parts_of(idx[:, 0])
        self._block.engine.gamma_prop_cols(self._block.inst, prop[which], parts, conc[which], rate[which])

    def set_samples(self, sample_idx, bid=None):
        """dist_blockwrapper.py:556-581."""
        s = _np(sample_idx, np.int64).reshape(-1)
        if bid is not None:
            s = s + self._neurons_thrush[int(bid)]
        order, recover = np.unique(s, return_inverse=True)
        if order.shape[0] == s.shape[0] and (order == s).all():
            self._sample_order = None
        else:
            self._sample_order = recover
        self._sample_num = int(order.shape[0])
        self._block.set_samples(order.astype(np.uint32))

    def set_samples_by_specifying_popu_idx(self, population_id, sample_number=None):
        """dist_blockwrapper.py:713-747: the first sample_number neurons of
        each listed population."""
        pop = _np(population_id, np.int64).reshape(-1)
        assert np.unique(pop).shape[0] == pop.shape[0], "population_id should be unique!"
        parts, _ = self._parts_of(pop)
        if parts.shape[0] != pop.shape[0]:
            raise NotImplementedError("one population has been splited into more than one sub-populations in this "
                                      "network structure and this method is invalid in this case.")
        if sample_number is None:
            m = min(int(self._part_neurons[parts].min()) // 2, 50)
            sample_number = np.full(parts.shape, m, dtype=np.int64)
        sample_number = _np(sample_number, np.int64).reshape(-1)
        assert (sample_number < self._part_neurons[parts]).all(), "sample number invalid, too big to sample"
        base = np.concatenate([[0], np.cumsum(self._part_neurons)])
        ids = np.concatenate([np.arange(base[r], base[r] + k) for r, k in zip(parts, sample_number)])
        self.set_samples(ids)
        return int(sample_number.sum())

    def update_ou_background_stimuli(self, population_id, correlation_time, mean, deviation):
        """dist_blockwrapper.py:583-612."""
        pop = _np(population_id, np.int64).reshape(-1)
        order = np.argsort(pop, kind="stable")
        pop = pop[order]
        if np.unique(pop).shape[0] != pop.shape[0]:
            raise _lib.DtbError(_lib.INVALID_PARAMETER, "population ids must be unique")
        tau = _np(correlation_time, np.float32).reshape(-1)[order]
        mu = _np(mean, np.float32).reshape(-1)[order]
        sd = _np(deviation, np.float32).reshape(-1)[order]
        parts, which = self._parts_of(pop)
        self._block.update_ou_background_stimuli(parts, tau[which], mu[which], sd[which])

    def close(self):
        if self._block is not None:
            self._block.engine.close()
            self._block = None

    def shutdown(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, exc_type, exc_val, exc_tb):
        self.close()

    def get_version(self):
        """dist_blockwrapper.py:764-767 formatting of dtb_version()."""
        v = _lib.load().dtb_version()
        return f"Version: {v // 1000}.{(v % 1000) // 100}.{v % 100}"
