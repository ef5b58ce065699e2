"""Serving: continuous batching of simulation requests on one GPU."""
import importlib

import numpy as np
import pytest


@pytest.fixture(scope="module")
def serving(dtb):
    return importlib.import_module("digital_twin_brain-open_amd.serving")


def test_request_stream_is_seeded(dtb, serving):
    a = serving.poisson_requests(50, 20.0, seed=3)
    b = serving.poisson_requests(50, 20.0, seed=3)
    assert [(r.arrival, r.n, r.k, r.steps, r.net_seed) for r in a] == [(r.arrival, r.n, r.k, r.steps, r.net_seed)
                                                                      for r in b]
    arr = np.array([r.arrival for r in a])
    assert np.all(np.diff(arr) > 0)
    assert all(10_000 <= r.n <= 1_000_000 and 100 <= r.k <= 1000 and 1000 <= r.steps <= 20000 for r in a)
    assert all(r.device_bytes() > r.n * r.k * 12 for r in a)


@pytest.mark.gpu
@pytest.mark.parametrize("layout", ["auto", "gather"])
def test_results_same_alone_and_in_the_mix(dtb, gpu, serving, layout):
    """Requests admitted while others are resident, under a budget that
    forces waiting, retire with exactly the counters and final V they have
    when run alone (alone: queue propagation; in the mix: queues or
    gather)."""
    reqs = [serving.Request(i, 0.01 * i, n, k, steps, net_seed=100 + i, seed=7 + i, mu=3.0)
            for i, (n, k, steps) in enumerate([(30000, 200, 700), (12000, 150, 300), (50000, 300, 450),
                                               (20000, 120, 900), (9000, 500, 250), (40000, 250, 600)])]
    gather = layout == "gather"
    budget = sum(sorted(r.device_bytes(gather, pull=True) for r in reqs)[-3:])   # at most about three resident at once
    srv = serving.Server(device=0, hbm_budget=budget, chunk=64, layout=layout)
    try:
        stats = srv.serve(reqs, realtime=False)
    finally:
        srv.close()
    assert stats["requests"] == len(reqs) and stats["mean_resident"] > 1.0
    # admissions happened while others were resident
    assert any(r.admitted > reqs[0].admitted for r in reqs[1:])
    for r in reqs:
        assert r.events > 0
        assert (r.events, r.spikes, r.v_hash) == serving.run_alone(r), r.rid


@pytest.mark.gpu
def test_device_memory_is_reused_and_zeroed(dtb, gpu, serving):
    """Memory of a removed instance is cached and handed to the next one; the
    new instance starts from the same zeroed state as on fresh memory."""
    dtb.device_memory_trim()
    r = serving.Request(0, 0.0, 20000, 300, 200, net_seed=9, seed=4, mu=3.0)
    first = serving.run_alone(r)
    held = dtb.device_memory_cached(0)
    assert held > r.n * r.k * 12
    again = serving.run_alone(r)   # same sizes: served from the cache
    assert again == first and first[0] > 0
    dtb.device_memory_trim()
    assert dtb.device_memory_cached(0) < held


@pytest.mark.gpu
def test_second_server_budget_counts_the_cache(dtb, gpu, serving):
    """A server's reservation stays with the library after it closes; the
    next server's default budget counts it as available instead of seeing
    only what the HIP runtime reports free."""
    a = serving.Server(device=0)
    budget_a = a.budget
    a.close()
    assert dtb.device_memory_cached(0) >= budget_a // 2
    b = serving.Server(device=0)
    try:
        assert b.budget >= 0.95 * budget_a, (b.budget, budget_a)
    finally:
        b.close()
    dtb.device_memory_trim()
