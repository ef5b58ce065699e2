"""dtb_run_noalloc (records in the library's static device array, no device
allocation) against dtb_run: two engines with the same network and seed, the
same calls in the same order, every output bit for bit; two calls in a row,
the second shorter and with fewer samples; the device-memory cache does not
grow during the calls; a run whose records exceed the array is refused
before anything runs."""
import ctypes

import numpy as np
import pytest

from test_gpu_parity import DT, bits

pytestmark = pytest.mark.gpu

FLAGS_ALL = ("freq", "spike", "vmemb", "isynaptic", "iou", "vmean", "imean", "receptor")


def outputs(lib, iters, npop, nsample):
    """Host arrays for every output and the dtb_run_outputs pointing at them."""
    a = {"freq": np.zeros((iters, npop), np.uint32),
         "spike": np.zeros((iters, nsample), np.uint8),
         "vmemb": np.zeros((iters, nsample), np.float32),
         "isynaptic": np.zeros((iters, nsample), np.float32),
         "iou": np.zeros((iters, nsample), np.float32),
         "vmean": np.zeros((iters, npop), np.float32),
         "imean": np.zeros((iters, npop), np.float32),
         "receptor": np.zeros((4, iters, npop), np.float32)}
    o = lib.RunOutputs()
    p = lambda x: x.ctypes.data_as(ctypes.c_void_p)
    o.freq, o.sample_spike, o.sample_vmemb = p(a["freq"]), p(a["spike"]), p(a["vmemb"])
    o.sample_isynaptic, o.sample_iou, o.vmean, o.imean = p(a["isynaptic"]), p(a["iou"]), p(a["vmean"]), p(a["imean"])
    o.ampa_imean, o.nmda_imean, o.gabaa_imean, o.gabab_imean = (p(a["receptor"][k]) for k in range(4))
    flags = (lib.RUN_USE_OU | lib.RUN_FREQ | lib.RUN_SAMPLE_SPIKE | lib.RUN_SAMPLE_VMEMB | lib.RUN_SAMPLE_ISYN
             | lib.RUN_SAMPLE_IOU | lib.RUN_VMEAN | lib.RUN_IMEAN | lib.RUN_RECEPTOR_IMEAN)
    return a, o, flags


def test_noalloc_run_equals_run(dtb, gpu):
    lib = dtb._lib
    L = lib.load()
    n, k, seed, mu = 6000, 120, 23, 1.5
    net = dtb.gen_uniform(n, k, seed=4)
    engines, blocks = [], []
    try:
        for _ in range(2):
            e = dtb.Engine(0)
            b = dtb.Block(net, delta_t=DT, seed=seed, engine=e)
            b.update_ou_background_stimuli(b.subblk_id, 4.0, mu, 0.1)
            engines.append(e)
            blocks.append(b)
        npop = len(blocks[0].subblk_id)
        calls = [(400, np.arange(0, n, 7, dtype=np.uint32)), (150, np.arange(3, n, 29, dtype=np.uint32))]
        for c, (iters, ids) in enumerate(calls):
            print("call", c, "iters", iters, "samples", ids.size, flush=True)
            res = []
            for which, (e, b) in enumerate(zip(engines, blocks)):
                b.set_samples(ids)
                a, o, flags = outputs(lib, iters, npop, ids.size)
                if which == 0:
                    lib.check(L.dtb_run(e.handle, b.inst, iters, 1, flags, ctypes.byref(o)))
                else:
                    # the tables are built once after the sample change (this
                    # may allocate); the no-allocation call follows
                    e.run(0)
                    cached = dtb.device_memory_cached(0)
                    lib.check(L.dtb_run_noalloc(e.handle, b.inst, iters, 1, flags, ctypes.byref(o)))
                    assert dtb.device_memory_cached(0) == cached, "the call took device memory"
                res.append(a)
            for key in FLAGS_ALL:
                assert np.array_equal(bits(res[0][key]) if res[0][key].dtype == np.float32 else res[0][key],
                                      bits(res[1][key]) if res[1][key].dtype == np.float32 else res[1][key]), \
                    (c, key)
            assert res[0]["freq"].sum() > 100, "the network did not fire"
            assert blocks[0].counters() == blocks[1].counters()
        s0, s1 = blocks[0].state(), blocks[1].state()
        for key in ("V", "J", "i_ou", "I_syn"):
            assert np.array_equal(bits(s0[key]), bits(s1[key])), key
    finally:
        for e in engines:
            e.close()


def test_noalloc_refuses_records_past_the_array(dtb, gpu):
    lib = dtb._lib
    L = lib.load()
    net = dtb.gen_uniform(3000, 50, seed=8)
    e = dtb.Engine(0)
    try:
        b = dtb.Block(net, delta_t=DT, seed=2, engine=e)
        b.update_ou_background_stimuli(b.subblk_id, 4.0, 1.5, 0.1)
        ids = np.arange(0, 3000, dtype=np.uint32)
        b.set_samples(ids)
        e.run(0)
        # 13 bytes per sample and step: 3000 samples x 2000 steps = 78 MB
        iters = 2000
        a, o, flags = outputs(lib, iters, len(b.subblk_id), ids.size)
        before = b.counters()
        st = L.dtb_run_noalloc(e.handle, b.inst, iters, 1, flags, ctypes.byref(o))
        assert st == lib.INVALID_PARAMETER
        assert b"static buffer" in L.dtb_last_error()
        assert b.counters() == before, "a refused call ran steps"
    finally:
        e.close()
