"""C-ABI boundary and host logic (CPU only: no kernel launches)."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dtb_abi.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dtb_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("dtb_engine_create", "dtb_instance_add", "dtb_set_ou", "dtb_run", "dtb_update_props",
                 "dtb_update_prop_cols", "dtb_set_samples", "dtb_read_state", "dtb_write_state"):
        assert must in names


def test_library_exports_every_declared_symbol(dtb):
    lib_path = dtb.LIB_PATH
    assert os.path.exists(lib_path)
    out = subprocess.run(["nm", "-D", "--defined-only", lib_path], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (dtb_[a-z0-9_]+)\b", out))
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing
    # the ctypes binding covers the same set
    assert set(dtb.EXPORTED) == set(declared_functions())


def test_library_loads_and_reports_version(dtb):
    lib = dtb.load()
    assert lib.dtb_version() == 1
    assert dtb.device_count() >= 0


def test_engine_without_device_fails_loudly(dtb):
    if dtb.device_count() > 0:
        pytest.skip("a device is visible")
    with pytest.raises(dtb.DtbError) as e:
        dtb.Engine(0)
    assert e.value.status == 102   # DTB_NO_DEVICE


def test_generator_rejects_bad_sizes(dtb):
    with pytest.raises(dtb.DtbError) as e:
        dtb.gen_uniform(1, 10, seed=0)
    assert e.value.status == 1


def test_gen_uniform_structure(dtb):
    n, k = 3000, 37
    net = dtb.gen_uniform(n, k, seed=11)
    assert net.nsyn == n * k
    pre = np.repeat(np.arange(n), np.diff(net.rowptr))
    post = net.col.astype(np.int64)
    assert np.all(np.bincount(post, minlength=n) == k)          # exact in-degree
    assert not np.any(pre == post)                               # no self-connections
    n_exc = n * 8 // 10
    assert np.array_equal(net.kind_per_synapse(), (pre >= n_exc).astype(np.uint8))
    q = net.w.astype(np.float64) * 2 ** 24
    assert np.all(q == np.floor(q)) and net.w.min() >= 0 and net.w.max() < 1
    # rows sorted by postsynaptic id
    for r in range(0, n, 97):
        seg = net.col[net.rowptr[r]:net.rowptr[r + 1]]
        assert np.all(np.diff(seg.astype(np.int64)) >= 0)
    ids, begin = net.populations()
    assert ids.tolist() == [0, 1] and begin.tolist() == [0, n_exc, n]
    p = net.props[0]
    assert p[4] == 1.0 and p[5] == 2.0 and p[8] == -50.0 and p[9] == -65.0
    assert np.allclose(net.props[:, 10:14], [0.01, 0.0005, 0.1, 0.005])


def test_gen_is_deterministic(dtb):
    a = dtb.gen_uniform(2000, 20, seed=5)
    b = dtb.gen_uniform(2000, 20, seed=5)
    c = dtb.gen_uniform(2000, 20, seed=6)
    assert np.array_equal(a.col, b.col) and np.array_equal(a.w, b.w) and np.array_equal(a.rowptr, b.rowptr)
    assert not np.array_equal(a.col, c.col)


def test_gen_voxel_structure(dtb):
    n, nv, k = 40000, 20, 50
    net = dtb.gen_voxel(n, nv, k, seed=3)
    sizes = net.voxel_sizes.astype(np.int64)
    assert sizes.sum() == n and sizes.min() >= 8
    vb = np.r_[0, np.cumsum(sizes)]
    vox = np.repeat(np.arange(nv), sizes)
    pre = np.repeat(np.arange(n), np.diff(net.rowptr))
    post = net.col.astype(np.int64)
    assert np.all(np.bincount(post, minlength=n) == k)
    same = vox[pre] == vox[post]
    k_intra = int(round(k * 0.7))
    assert np.all(np.bincount(post[same], minlength=n) == k_intra)
    exc_end = vb[:-1] + sizes * 8 // 10
    inh = np.arange(n) >= exc_end[vox]
    assert np.array_equal(net.kind_per_synapse(), inh[pre].astype(np.uint8))
    ids, begin = net.populations()
    assert len(ids) == 2 * nv and ids.tolist() == list(range(2 * nv))


def test_network_from_edges_roundtrip(dtb):
    src = dtb.gen_uniform(500, 9, seed=2)
    pre = np.repeat(np.arange(500), np.diff(src.rowptr))
    net = dtb.NetworkArrays.from_edges(src.props, pre, src.col, src.w, src.kind_per_synapse())
    assert np.array_equal(net.rowptr, src.rowptr)
    assert np.array_equal(net.col, src.col)
    assert np.array_equal(net.kinds[: len(src.kinds)], src.kinds)


def test_gen_fanout_structure(dtb):
    """Fan-out generator: k synapses per row, rows sorted by target, no
    self-connections, channel by presynaptic index (80/20), weights k 2^-24,
    deterministic."""
    n, k = 3000, 37
    a = dtb.gen_fanout(n, k, seed=5)
    b = dtb.gen_fanout(n, k, seed=5)
    c = dtb.gen_fanout(n, k, seed=6)
    assert np.array_equal(a.col, b.col) and np.array_equal(a.w, b.w)
    assert not np.array_equal(a.col, c.col)
    assert np.array_equal(a.rowptr, np.arange(n + 1) * k)
    rows = a.col.reshape(n, k)
    assert np.all(np.diff(rows.astype(np.int64), axis=1) >= 0)
    assert not np.any(rows == np.arange(n)[:, None])
    kind = a.kind_per_synapse().reshape(n, k)
    assert np.array_equal(kind[:, 0], (np.arange(n) >= n * 8 // 10).astype(np.uint8))
    assert np.all(kind == kind[:, :1])
    q = a.w * 2.0 ** 24
    assert np.array_equal(q, np.floor(q)) and a.w.min() >= 0 and a.w.max() < 1
    # targets roughly uniform: mean in-degree k
    indeg = np.bincount(a.col, minlength=n)
    assert abs(indeg.mean() - k) < 1e-9 and indeg.std() < 3 * np.sqrt(k)
    pops, begin = a.populations()
    assert list(pops) == [0, 1] and list(begin) == [0, n * 8 // 10, n]


@pytest.mark.parametrize("alpha", [0.3, 1.0, 2.5, 20.0])
def test_gamma_sampler_distribution(dtb, alpha):
    """The property resampler's Gamma(alpha, rate) draws: Kolmogorov-Smirnov
    against the gamma CDF, mean alpha / rate, and distinct streams per call
    and per neuron; alpha = 0 gives 0."""
    scipy_stats = pytest.importorskip("scipy.stats")
    rate = 2.0
    x = dtb.gen_gamma(7, 3, 0, 100000, alpha, rate).astype(np.float64)
    ks = scipy_stats.kstest(x, scipy_stats.gamma(alpha, scale=1.0 / rate).cdf)
    assert ks.pvalue > 1e-4, ks
    assert abs(x.mean() - alpha / rate) < 5 * np.sqrt(alpha) / rate / np.sqrt(x.size)
    y = dtb.gen_gamma(7, 4, 0, 1000, alpha, rate)
    assert not np.array_equal(y, x[:1000].astype(np.float32))
    z = dtb.gen_gamma(7, 3, 500, 500, alpha, rate)
    assert np.array_equal(z, x[500:1000].astype(np.float32))   # per-neuron streams
    assert (dtb.gen_gamma(7, 3, 0, 10, 0.0, rate) == 0).all()


def test_gamma_sampler_rejects_bad_parameters(dtb):
    for a, b in [(-1.0, 1.0), (1.0, 0.0), (float("nan"), 1.0)]:
        with pytest.raises(dtb.DtbError):
            dtb.gen_gamma(1, 0, 0, 4, a, b)
