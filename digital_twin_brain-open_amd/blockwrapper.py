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
        self._block = Block(net, delta_t=delta_t, seed=seed, device=device)
        self._neurons_per_block = np.diff(begin).astype(np.int64)
        self._block_id = np.arange(len(self._neurons_per_block), dtype=np.int64)
        self._neurons_thrush = begin.astype(np.int64)
        self._subblk_id = self._block.subblk_id.astype(np.int64)
        self._neurons_per_subblk = self._block.neurons_per_subblk
        self._reset_hyper_parameter()
        self._sample_order = None
        self._sample_num = 0
        self._iterations = 0

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
        if freqs:
            f = out["freq"].astype(np.int32)
            if freq_char:
                # stat_result.cpp:103-130 sends lroundf(255 c / (n - blocked))
                # per population; the client scales it back by the
                # population size (dist_blockwrapper.py:285-289).
                n = self._neurons_per_subblk.astype(np.float32)
                ch = self._block.freq_char(out["freq"])
                f = (ch.astype(np.float32) / np.float32(255.0) * n).astype(np.int32)
            seq.append(f)
        if vmean:
            seq.append(out["vmean"])
        for flag, key in ((output_sample_spike, "spike"), (output_sample_vmemb, "vmemb"),
                          (output_sample_isynaptic, "isynaptic"), (output_sample_iou, "iou")):
            if flag:
                a = out[key]
                if self._sample_order is not None:
                    a = a[:, self._sample_order]
                seq.append(a)
        if imean:
            seq.append(out["imean"])
        if receptor_imean:
            seq.extend(out["receptor_imean"][k] for k in range(4))
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
        self._block.engine.update_prop_cols(self._block.inst, prop, idx[:, 0].astype(np.uint32), hp,
                                            assign=False)

    def assign_property_by_subblk(self, property_idx, property_hyper_parameter):
        """dist_blockwrapper.py:517-527."""
        idx = _np(property_idx, np.int64).reshape(-1, 2)
        hp = _np(property_hyper_parameter, np.float32).reshape(-1)
        self._block.engine.update_prop_cols(self._block.inst, self._prop_ids(idx[:, 1]),
                                            idx[:, 0].astype(np.uint32), hp, assign=True)

    def gamma_property_by_subblk(self, property_idx, gamma_concentration, gamma_rate):
        """dist_blockwrapper.py:529-554: every neuron of each listed
        population draws the property from Gamma(concentration, rate);
        concentration None means 1."""
        idx = _np(property_idx, np.int64).reshape(-1, 2)
        rate = _np(gamma_rate, np.float32).reshape(-1)
        conc = np.ones_like(rate) if gamma_concentration is None else _np(gamma_concentration, np.float32).reshape(-1)
        if not ((conc >= 0).all() and (rate >= 0).all()):
            raise ValueError("gamma hyperparameters must be non-negative")
        self._block.engine.gamma_prop_cols(self._block.inst, self._prop_ids(idx[:, 1]),
                                           idx[:, 0].astype(np.uint32), conc, rate)

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
        rows = self._subblk_rows(pop)
        if sample_number is None:
            m = min(int(self._neurons_per_subblk[rows].min()) // 2, 50)
            sample_number = np.full(rows.shape, m, dtype=np.int64)
        sample_number = _np(sample_number, np.int64).reshape(-1)
        assert (sample_number < self._neurons_per_subblk[rows]).all(), "sample number invalid, too big to sample"
        base = np.concatenate([[0], np.cumsum(self._neurons_per_subblk)])
        ids = np.concatenate([np.arange(base[r], base[r] + k) for r, k in zip(rows, sample_number)])
        self.set_samples(ids)
        return int(sample_number.sum())

    def update_ou_background_stimuli(self, population_id, correlation_time, mean, deviation):
        """dist_blockwrapper.py:583-612."""
        pop = _np(population_id, np.int64).reshape(-1)
        order = np.argsort(pop, kind="stable")
        pop = pop[order]
        if np.unique(pop).shape[0] != pop.shape[0]:
            raise _lib.DtbError(_lib.INVALID_PARAMETER, "population ids must be unique")
        self._subblk_rows(pop)
        tau = _np(correlation_time, np.float32).reshape(-1)[order]
        mu = _np(mean, np.float32).reshape(-1)[order]
        sd = _np(deviation, np.float32).reshape(-1)[order]
        self._block.update_ou_background_stimuli(pop.astype(np.uint32), tau, mu, sd)

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
