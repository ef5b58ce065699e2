"""Host logic of the serving loop (no GPU): admission by HBM budget, every
request leaving at exactly its own step count, and admission control that
sheds requests waiting past the limit.  The engine is a stand-in that counts
steps per instance."""
import importlib

import numpy as np
import pytest


class FakeEngine:
    """Counts steps per resident instance; nothing runs on a device."""

    def __init__(self, device=0):
        self.steps = {}
        self.n = {}
        self.next_id = 0
        self.max_resident = 0
        self.staged = set()

    def schedule(self, pass_k=None, layout=None):
        return None

    def pull(self, on=None, defer_min=None):
        return None

    def add_fanout(self, n, k, net_seed, g_ui=None, sub_base=0, dt=0.1, seed=0):
        i = self.next_id
        self.next_id += 1
        self.steps[i] = 0
        self.n[i] = n
        self.staged.add(i)
        return i

    def info(self, inst):
        return {"n": self.n[inst], "pop_ids": np.array([0, 1], np.uint32),
                "pop_begin": np.array([0, self.n[inst] // 2, self.n[inst]], np.uint32)}

    def set_ou(self, *a):
        return None

    def configure(self, *a):
        return None

    def run(self, steps, graph_chunk=0):
        self.staged.clear()
        for i in self.steps:
            self.steps[i] += steps
        self.max_resident = max(self.max_resident, len(self.steps))

    def sync(self):
        return None

    def counters(self, inst):
        return self.steps[inst], 7 * self.steps[inst], self.steps[inst]

    def read_state(self, inst, n):
        return {"V": np.full(n, -60.0, np.float32)}

    def remove(self, inst):
        del self.steps[inst]

    def close(self):
        return None


@pytest.fixture
def serving(monkeypatch):
    mod = importlib.import_module("digital_twin_brain-open_amd.serving")
    monkeypatch.setattr(mod, "Engine", FakeEngine)
    return mod


def requests(serving, count, steps=(300, 50, 120, 700, 90, 410)):
    return [serving.Request(i, 0.0, 20_000 + 1000 * i, 100, steps[i % len(steps)], net_seed=i, seed=i, mu=3.0)
            for i in range(count)]


def test_every_request_runs_exactly_its_steps(serving):
    reqs = requests(serving, 12)
    budget = sum(sorted(r.device_bytes(pull=True) for r in reqs)[-3:])
    srv = serving.Server(hbm_budget=budget, chunk=100, reserve=False)
    stats = srv.serve(reqs, realtime=False)
    assert stats["shed"] == 0 and stats["requests"] == 12
    for r in reqs:
        assert r.events == 7 * r.steps and r.spikes == r.steps   # retired at exactly its step count
        assert r.finished >= r.admitted
    # the budget (the three largest footprints) holds a few at a time
    smallest = sorted(r.device_bytes(pull=True) for r in reqs)
    assert 1 < srv.engine.max_resident and sum(smallest[:srv.engine.max_resident]) <= budget


def test_requests_past_the_wait_limit_are_shed(serving, monkeypatch):
    clock = {"t": 0.0}
    monkeypatch.setattr(serving.time, "monotonic", lambda: clock["t"])
    monkeypatch.setattr(serving.time, "sleep", lambda s: clock.__setitem__("t", clock["t"] + s))
    reqs = requests(serving, 10)
    for i, r in enumerate(reqs):
        r.arrival = 0.01 * i
    budget = max(r.device_bytes(pull=True) for r in reqs)   # one resident at a time

    srv = serving.Server(hbm_budget=budget, chunk=100, reserve=False)
    run0 = srv.engine.run

    def slow_run(steps, graph_chunk=0):   # every chunk takes 1 s of the fake clock
        clock["t"] += 1.0
        run0(steps, graph_chunk)

    srv.engine.run = slow_run
    stats = srv.serve(reqs, realtime=True, shed_after=2.5)
    served = [r for r in reqs if not r.shed]
    assert stats["shed"] == len(reqs) - len(served) > 0
    assert reqs[0] in served
    for r in served:
        assert r.admitted - r.arrival <= 2.5 + 1.0   # admitted before its limit (checked between chunks)
        assert r.events == 7 * r.steps
    for r in reqs:
        if r.shed:
            assert r.events is None and r.finished is None
    assert stats["events"] == sum(7 * r.steps for r in served)
