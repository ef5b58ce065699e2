"""Writes tests/golden/block_new_small.npz: inputs and per-step outputs of the
reference's CPU model (models/block_new.py, class `block`) on a tiny seeded
network.

It pins the composition of the step (stage order and what each stage reads),
which the closed-form checks of the oracle do not.  The model is run as
published; only torch.randn, its OU noise source (block_new.py:93-95), is
replaced by the OU normals of this project's Philox stream, so that the
oracle can be driven by the same noise.  dt = 0.125 ms keeps the model's
float32 clock exact (its refractory test compares float times,
block_new.py:111-113).

Run in a container that holds the reference (CPU only):
    python tests/golden/make_block_new_fixture.py
"""
import importlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_MODELS = "/root/reference/models"

N, K, STEPS, SEED, NET_SEED = 120, 12, 300, 77, 4
DT = 0.125
I_MEAN, I_SIGMA, TAU_I = 1.5, 0.1, 4.0


def main():
    import torch

    sys.path.insert(0, ROOT)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    dtb = importlib.import_module("digital_twin_brain-open_amd")
    orc = importlib.import_module("oracle")
    sys.path.insert(0, REF_MODELS)
    block_new = importlib.import_module("block_new")

    net = dtb.gen_uniform(N, K, seed=NET_SEED)
    props = net.props.copy()
    pre = np.repeat(np.arange(N), np.diff(net.rowptr))
    kind = net.kind_per_synapse()
    w = np.zeros((4, N, N), dtype=np.float64)
    for j in range(net.nsyn):
        c = 2 * int(kind[j])
        w[c, net.col[j], pre[j]] += net.w[j, 0]
        w[c + 1, net.col[j], pre[j]] += net.w[j, 1]

    torch.set_num_threads(1)
    step = {"k": 0}

    def philox_randn(*shape, **kw):
        z = orc.normals(SEED, 0, step["k"], N)
        step["k"] += 1
        return torch.from_numpy(z.astype(np.float32))

    model = block_new.block(torch.from_numpy(props), torch.from_numpy(w.astype(np.float32)), delta_t=DT,
                            i_mean=I_MEAN, i_sigma=I_SIGMA, tau_i=TAU_I)
    block_new.torch.randn = philox_randn
    raster = np.zeros((STEPS, N), dtype=np.uint8)
    vtrace = np.zeros((STEPS, N), dtype=np.float32)
    for s in range(STEPS):
        model.run(debug_mode=False)
        raster[s] = model.active.numpy().astype(np.uint8)
        vtrace[s] = model.V_i.numpy()
    np.savez_compressed(os.path.join(HERE, "block_new_small.npz"), props=props, rowptr=net.rowptr, col=net.col,
                        w=net.w, kind=kind, dt=np.float32(DT), seed=np.uint64(SEED), i_mean=np.float32(I_MEAN),
                        i_sigma=np.float32(I_SIGMA), tau_i=np.float32(TAU_I), raster=raster, V=vtrace,
                        J_final=model.J_ui.numpy().T.copy(), i_bg_final=model.i_background.numpy())
    print("wrote block_new_small.npz: %d spikes in %d steps" % (int(raster.sum()), STEPS))


if __name__ == "__main__":
    main()
