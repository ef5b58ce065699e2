"""Pins the CPU oracle (CPU only).

Each check is independent of the oracle's own formulas: published or
third-party known-answer vectors (Philox), library functions (numpy log,
sin, cos), distribution statistics (KS), closed forms of the discrete-time
model, and brute force on tiny networks.  Reference semantics being pinned
are cited per test (paths relative to the reference repository).
"""
import json
import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


# ---- Philox4x32-10 -------------------------------------------------------

def test_philox_random123_published_vectors(orc):
    """Known-answer vectors published with Random123 (Salmon et al., SC'11)."""
    cases = [
        ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
        ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
        ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
         (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
    ]
    for ctr, key, want in cases:
        assert tuple(orc.philox(ctr, key)) == want


def test_philox_matches_rocrand_fixture(orc):
    """tests/golden/philox_kat.json, written by tests/golden/make_philox_vectors.sh
    from rocRAND's Philox4x32-10 engine (the generator behind the reference's
    hiprand calls, src/stage/random.cpp:7-19)."""
    with open(os.path.join(HERE, "golden", "philox_kat.json")) as f:
        cases = json.load(f)
    assert len(cases) >= 10
    for c in cases:
        assert orc.philox(c["ctr"], c["key"]) == c["out"]


# ---- normal transform ----------------------------------------------------

def test_log_against_numpy(orc):
    lib = orc.load()
    rng = np.random.default_rng(1)
    u = np.concatenate([rng.random(20000), np.exp(-rng.random(2000) * 16), [2.0 ** -24, 0.5, 1 - 2.0 ** -24]])
    u = u.astype(np.float32)
    got = np.array([lib.dtbo_log(float(x)) for x in u], dtype=np.float64)
    want = np.log(u.astype(np.float64))
    err = np.abs(got - want) / np.maximum(np.abs(want), 1e-3)
    assert err.max() < 4e-7, err.max()


def test_sin_cos_halfpi_against_numpy(orc):
    lib = orc.load()
    x = np.linspace(0.0, 0.5, 4097, dtype=np.float32)
    s = np.array([lib.dtbo_sin_halfpi(float(v)) for v in x])
    c = np.array([lib.dtbo_cos_halfpi(float(v)) for v in x])
    xs = x.astype(np.float64) * np.pi / 2
    assert np.abs(s - np.sin(xs)).max() < 2e-7
    assert np.abs(c - np.cos(xs)).max() < 2e-7


def test_normal_distribution(orc):
    from scipy import stats

    z = orc.normals(seed=12345, neuron0=0, substep=7, count=1 << 20).astype(np.float64)
    n = z.size
    assert abs(z.mean()) < 5 / math.sqrt(n)
    assert abs(z.var() - 1.0) < 5 * math.sqrt(2.0 / n)
    assert stats.kstest(z, "norm").pvalue > 1e-3
    tail = np.mean(np.abs(z) > 3.0)
    assert abs(tail - 0.0026998) < 5 * math.sqrt(0.0027 / n)
    # Box-Muller with u1 >= 2^-24: |z| <= sqrt(-2 ln 2^-24)
    assert np.abs(z).max() <= math.sqrt(-2 * math.log(2.0 ** -24)) + 1e-5


def test_normals_independent_across_neurons_and_steps(orc):
    a = orc.normals(99, 0, 10, 1 << 16).astype(np.float64)
    b = orc.normals(99, 0, 11, 1 << 16).astype(np.float64)
    c = orc.normals(100, 0, 10, 1 << 16).astype(np.float64)
    bound = 5 / math.sqrt(a.size)
    assert abs(np.corrcoef(a, b)[0, 1]) < bound
    assert abs(np.corrcoef(a, c)[0, 1]) < bound
    assert abs(np.corrcoef(a[:-1], a[1:])[0, 1]) < bound


# ---- model closed forms ----------------------------------------------------

def _props(n, **over):
    """Property matrix with make_block.py:414-465 defaults (T_ref 2 ms)."""
    p = np.zeros((n, 22), dtype=np.float32)
    p[:, 4], p[:, 5], p[:, 6], p[:, 7], p[:, 8], p[:, 9] = 1.0, 2.0, 0.03, -75.0, -50.0, -65.0
    p[:, 10:14] = (0.01, 0.0005, 0.1, 0.005)
    p[:, 14:18] = (0.0, 0.0, -70.0, -100.0)
    p[:, 18:22] = (2.0, 40.0, 10.0, 50.0)
    for k, v in over.items():
        p[:, int(k[1:])] = v
    return p


def _empty_net(orc, props, dt=0.1, seed=0):
    n = props.shape[0]
    return orc.OracleNet(props, np.zeros(n + 1, np.int64), np.zeros(0, np.uint32), np.zeros((0, 2), np.float32),
                         np.zeros(0, np.uint8), dt=dt, seed=seed)


def test_lif_zero_input_relaxation(orc):
    """membrane_voltage.cpp:159-165 with I = 0: V_k = V_L + (V_0 - V_L)(1 - dt g_L / C)^k."""
    net = _empty_net(orc, _props(4))
    for prec, tol in ((64, 1e-12), (32, 2e-6)):
        sim = orc.Sim(net, prec)
        out = sim.run(500, vtrace=True)
        k = np.arange(1, 501, dtype=np.float64)
        dt, g, vl, v0 = float(np.float32(0.1)), float(np.float32(0.03)), -75.0, -57.5
        want = vl + (v0 - vl) * (1 - dt * g) ** k
        rel = np.abs(out["V"][:, 0].astype(np.float64) - want) / np.abs(want)
        assert rel.max() < tol, (prec, rel.max())


def test_lif_constant_input_spike_train_and_refractory(orc):
    """Euler map with constant input: V_k = V_inf + (V_r - V_inf)(1-a)^k from the
    reset; first spike when V >= V_th; V clamps to V_th at the spike, then R-1
    held steps at V_reset (membrane_voltage.cpp:140-181), R = (int)(T_ref/dt)."""
    iext = 1.2
    net = _empty_net(orc, _props(1, c2=iext))
    sim = orc.Sim(net, 64)
    out = sim.run(400, raster=True, vtrace=True)
    spikes = np.flatnonzero(out["raster"][:, 0])
    dt, g = float(np.float32(0.1)), float(np.float32(0.03))
    a = dt * g
    vinf = -75.0 + float(np.float32(iext)) / g
    R = int(np.float32(2.0) / np.float32(0.1))
    # first spike from V0 = -57.5
    k0 = math.ceil(math.log((-50.0 - vinf) / (-57.5 - vinf)) / math.log(1 - a))
    # after a reset at V_r: R-1 held steps, then k1 integration steps
    k1 = math.ceil(math.log((-50.0 - vinf) / (-65.0 - vinf)) / math.log(1 - a))
    assert spikes[0] == k0 - 1
    assert np.all(np.diff(spikes) == (R - 1) + k1), np.diff(spikes)[:5]
    V = out["V"][:, 0]
    s = spikes[0]
    assert V[s] == -50.0
    assert np.all(V[s + 1:s + R] == -65.0)
    assert V[s + R] > -65.0


def test_ou_deterministic_relaxation_and_statistics(orc):
    """extern_stimuli.cpp:58-65: i <- i + (1-e^{-dt/tau})(mu-i) + sqrt(1-e^{-2dt/tau}) sigma xi.
    sigma = 0: i_k = mu + (i_0 - mu) e^{-k dt/tau}.  sigma > 0: stationary
    mean mu, variance sigma^2, lag-1 autocorrelation e^{-dt/tau}."""
    n = 4096
    props = _props(n, c8=1e9)   # never fires
    net = _empty_net(orc, props, seed=5)
    net.set_ou([0], 0.7, 0.0, 4.0)
    sim = orc.Sim(net, 64)
    sim.st.i_ou[:] = 0.0
    sim.run(50)
    want = float(np.float32(0.7)) * (1 - math.exp(-50 * float(np.float32(0.1)) / 4.0))
    assert abs(sim.st.i_ou[0] - want) < 1e-12

    net2 = _empty_net(orc, props, seed=6)
    net2.set_ou([0], 0.5, 0.2, 4.0)
    sim2 = orc.Sim(net2, 64)
    sim2.run(2000)   # 50 correlation times: stationary
    xs = []
    for _ in range(20):
        sim2.run(40)  # 1 tau apart, roughly independent samples
        xs.append(sim2.st.i_ou.copy())
    x = np.concatenate(xs)
    assert abs(x.mean() - 0.5) < 5 * 0.2 / math.sqrt(x.size / 2)
    assert abs(x.std() - 0.2) < 0.01
    # lag-1 autocorrelation across one step
    a0 = sim2.st.i_ou.copy()
    sim2.run(1)
    a1 = sim2.st.i_ou.copy()
    rho = np.corrcoef(a0 - 0.5, a1 - 0.5)[0, 1]
    assert abs(rho - math.exp(-float(np.float32(0.1)) / 4.0)) < 0.01


def test_synaptic_trace_closed_form(orc):
    """A single presynaptic spike at step s: J_u(k) = w_u d_u^(k-s) for k >= s
    (presynaptic_voltage.cpp:62-73 with the delta of step s), and
    I_syn = sum_u g_u (V_u - V) J_u (synaptic_current.cpp:39-43)."""
    props = _props(2)
    props[0, 2] = 200.0         # neuron 0 fires at step 0
    props[1, 8] = 1e9           # neuron 1 never fires
    w = np.array([[0.75, 0.25]], dtype=np.float32)
    net = orc.OracleNet(props, np.array([0, 1, 1], np.int64), np.array([1], np.uint32), w, np.array([0], np.uint8))
    sim = orc.Sim(net, 64)
    out = sim.run(1, raster=True)
    assert out["raster"][0].tolist() == [1, 0]
    props[0, 2] = 0.0
    js = []
    for _ in range(30):
        js.append(sim.st.J[1].copy())
        sim.run(1)
    js = np.array(js)
    dt = float(np.float32(0.1))
    k = np.arange(30)
    for c, tau in enumerate((2.0, 40.0)):
        want = float(w[0, c]) * np.exp(-dt / tau) ** k
        assert np.allclose(js[:, c], want, rtol=1e-12, atol=0)
    assert np.all(js[:, 2:] == 0)


def _random_net(orc, n, nsyn, seed):
    rng = np.random.default_rng(seed)
    pre = rng.integers(0, n, nsyn)
    post = rng.integers(0, n, nsyn)
    order = np.argsort(pre, kind="stable")
    rowptr = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(pre, minlength=n), out=rowptr[1:])
    w = (rng.integers(0, 1 << 24, (nsyn, 2)) * 2.0 ** -24).astype(np.float32)
    kind = rng.integers(0, 2, nsyn).astype(np.uint8)
    return orc.OracleNet(_props(n), rowptr, post[order].astype(np.uint32), w[order], kind[order]), rng


def test_scatter_equals_dense_matvec(orc):
    """CSR scatter (presynaptic_voltage.cpp:137-188) == dense W @ f
    (models/block_new.py:79-84), exactly, including repeated synapses."""
    net, rng = _random_net(orc, 60, 900, 3)
    for _ in range(5):
        fired = (rng.random(60) < 0.3).astype(np.uint8)
        q = orc.scatter_q40(net, fired)
        d = orc.dense_delta(net, fired)
        assert np.array_equal(q.astype(np.float64) * 2.0 ** -40, d)


def test_scatter_channel_split_invariant(orc):
    """Sum of delta over targets = sum of the fired rows' weight pairs, split
    by channel bit (0: AMPA/NMDA, 1: GABA_A/GABA_B)."""
    net, rng = _random_net(orc, 300, 20000, 4)
    fired = (rng.random(300) < 0.2).astype(np.uint8)
    q = orc.scatter_q40(net, fired).astype(np.float64) * 2.0 ** -40
    src = np.repeat(np.arange(300), np.diff(net.rowptr))
    m = fired[src].astype(bool)
    for kind, cols in ((0, (0, 1)), (1, (2, 3))):
        sel = m & (net.kind == kind)
        assert q[:, cols[0]].sum() == pytest.approx(net.w[sel, 0].astype(np.float64).sum(), rel=0, abs=1e-9)
        assert q[:, cols[1]].sum() == pytest.approx(net.w[sel, 1].astype(np.float64).sum(), rel=0, abs=1e-9)


def test_fp32_canonical_within_1e5_of_fp64(dtb, orc):
    """The fp32 canonical order stays within 1e-5 relative of fp64 on V when
    both follow the same raster; fp64 disagrees only at near ties."""
    net = dtb.gen_uniform(3000, 100, seed=1)
    a = orc.OracleNet.from_network(net, seed=31)
    a.set_ou(a.pop_ids, 1.5, 0.1, 4.0)
    s32 = orc.Sim(a, 32)
    o32 = s32.run(400, raster=True, vtrace=True)
    assert o32["raster"].sum() > 100
    b = orc.OracleNet.from_network(net, seed=31)
    b.set_ou(b.pop_ids, 1.5, 0.1, 4.0)
    s64 = orc.Sim(b, 64)
    o64 = s64.run(400, vtrace=True, force_raster=o32["raster"], margins=True)
    rel = np.abs(o32["V"].astype(np.float64) - o64["V"]) / np.abs(o64["V"])
    assert rel.max() < 1e-5
    assert o64["margin"].max() < 1e-3


def test_population_counts_exclude_blocked(dtb, orc):
    """stat_result.cpp:161-206: per-population spike counts skip neurons whose
    blocked_in_stat column is set."""
    net = dtb.gen_uniform(2000, 50, seed=2)
    net.props[::7, 1] = 1.0
    a = orc.OracleNet.from_network(net, seed=3)
    a.set_ou(a.pop_ids, 1.5, 0.1, 4.0)
    s = orc.Sim(a, 32)
    out = s.run(300, raster=True, pop_counts=True)
    blocked = net.props[:, 1] != 0
    for q in range(a.npop):
        lo, hi = a.pop_begin[q], a.pop_begin[q + 1]
        assert np.array_equal(out["freq"][:, q], out["raster"][:, lo:hi][:, ~blocked[lo:hi]].sum(1))


def test_noise_spike_mode_rate(orc):
    """spike.cpp:55-62: without OU, f_rec = f | (U < noise_rate)."""
    props = _props(4000, c0=0.01)
    net = _empty_net(orc, props, seed=9)
    sim = orc.Sim(net, 32)
    out = sim.run(500, use_ou=False, raster=True)
    rate = out["raster"].mean()
    assert abs(rate - 0.01) < 5 * math.sqrt(0.01 / out["raster"].size)


def test_substeps_equal_finer_dt_without_synapses(orc):
    """t_steps substeps of dt/t_steps equal t_steps plain steps of dt/t_steps for
    the membrane and OU when there is no recurrent input (block.cpp:1272-1278,
    dist_simulator.cpp:2086-2112)."""
    props = _props(500)
    a = _empty_net(orc, props, dt=0.2, seed=4)
    a.t_steps = 2
    a.set_ou([0], 1.5, 0.1, 4.0)
    s2 = orc.Sim(a, 64)
    s2.run(100)
    b = _empty_net(orc, props, dt=0.1, seed=4)
    b.set_ou([0], 1.5, 0.1, 4.0)
    # refractory steps: (int)(2/0.2)*2 = 20 = (int)(2/0.1)
    s1 = orc.Sim(b, 64)
    s1.run(200)
    assert np.allclose(s2.st.V, s1.st.V, rtol=1e-6, atol=1e-6)
    assert np.allclose(s2.st.i_ou, s1.st.i_ou, rtol=1e-6, atol=1e-6)


def test_oracle_matches_reference_cpu_model_fixture(orc):
    """tests/golden/block_new_small.npz holds the reference CPU model
    (models/block_new.py:122-158) run on a 120-neuron network with the same
    Philox OU normals (tests/golden/make_block_new_fixture.py).  It pins the
    step composition: identical raster; V, J, i_ou within float32 rounding."""
    z = np.load(os.path.join(HERE, "golden", "block_new_small.npz"))
    steps = z["raster"].shape[0]
    for prec, vtol in ((64, 5e-6), (32, 5e-6)):
        on = orc.OracleNet(z["props"], z["rowptr"], z["col"], z["w"], z["kind"], dt=float(z["dt"]),
                           seed=int(z["seed"]))
        on.set_ou(on.pop_ids, float(z["i_mean"]), float(z["i_sigma"]), float(z["tau_i"]))
        sim = orc.Sim(on, prec)
        out = sim.run(steps, raster=True, vtrace=True)
        assert z["raster"].sum() > 100
        assert np.array_equal(out["raster"], z["raster"])
        rel = np.abs(out["V"].astype(np.float64) - z["V"]) / np.abs(z["V"].astype(np.float64))
        assert rel.max() < vtol
        assert np.abs(sim.st.J - z["J_final"]).max() <= 1e-5 * np.abs(z["J_final"]).max()
        assert np.abs(sim.st.i_ou - z["i_bg_final"]).max() < 1e-5


def test_population_means_of_v_isyn_and_receptor_currents(dtb, orc):
    """stat_result.cpp:266-391: per-population means over non-blocked neurons
    of V, I_syn and the receptor currents g_u (V_u - V) J_u
    (synaptic_current.cpp:39-43) at the end of every step, checked against
    float64 means of the state after each single step."""
    net = dtb.gen_uniform(1500, 60, seed=4)
    net.props[::5, 1] = 1.0
    on = orc.OracleNet.from_network(net, seed=9)
    on.set_ou(on.pop_ids, 2.0, 0.2, 4.0)
    sim = orc.Sim(on, 32)
    blocked = net.props[:, 1] != 0
    g = net.props[:, 10:14].astype(np.float64)
    vu = net.props[:, 14:18].astype(np.float64)
    spikes = 0
    for step in range(120):
        out = sim.run(1, raster=True, stats=True)
        spikes += int(out["raster"].sum())
        m = out["stat_means"][0]
        V = sim.st.V.astype(np.float64)
        J = sim.st.J.astype(np.float64)
        rc = g * (vu - V[:, None]) * J
        for q in range(on.npop):
            sel = np.zeros(net.n, bool)
            sel[on.pop_begin[q]:on.pop_begin[q + 1]] = True
            sel &= ~blocked
            expect = [V[sel].mean(), sim.st.I_syn[sel].astype(np.float64).mean()] + list(rc[sel].mean(0))
            for k in range(6):
                assert abs(m[q, k] - expect[k]) <= 1e-5 * abs(expect[k]) + 1e-6, (step, q, k, m[q, k], expect[k])
        # the four receptor sums add up to the I_syn sum within one unit per neuron
        s = out["stat_sums"][0]
        assert np.all(np.abs(s[:, 2:].sum(1) - s[:, 1]) <= 4 * np.diff(on.pop_begin))
    assert spikes > 50
