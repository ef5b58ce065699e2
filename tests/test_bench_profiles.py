"""The bench line's measured traffic comes only from a profile of the same
window: every committed summary under profiles/ names its window and its
traffic, and the lookup returns nothing for any other window."""
import importlib.util
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def load_bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


WINDOW = re.compile(r"^window: workload=(\w+) neurons=(\d+) layout=(\w+) steps=(\d+) warmup=(\d+) autotuned=([01])$")


def test_every_profile_names_its_window_and_traffic():
    bench = load_bench()
    for name in bench.PROFILES:
        path = os.path.join(ROOT, "profiles", name)
        assert os.path.exists(path), name
        lines = [x.strip() for x in open(path)]
        windows = [m for m in map(WINDOW.match, lines) if m]
        assert len(windows) == 1, name
        traffic = [x for x in lines if x.startswith("traffic_bytes_per_launch:")]
        assert len(traffic) == 1, name
        assert float(traffic[0].split(":", 1)[1]) > 0
        assert any(x.startswith("rocprofv3 --kernel-trace") for x in lines), name
        assert any(x.startswith("rocprofv3 --pmc FETCH_SIZE") for x in lines), name
        assert any(x.startswith("rocprofv3 --pmc WRITE_SIZE") for x in lines), name


def test_traffic_only_for_the_profiled_window():
    bench = load_bench()
    for name in bench.PROFILES:
        lines = [x.strip() for x in open(os.path.join(ROOT, "profiles", name))]
        m = next(m for m in map(WINDOW.match, lines) if m)
        wl, n, layout, steps, warmup, tuned = m.groups()
        value, src = bench.profiled_traffic(wl, int(n), layout, int(steps), int(warmup), tuned == "1")
        assert src == "profiles/" + name and value > 0
        # any other window: nothing
        assert bench.profiled_traffic(wl, int(n), layout, int(steps) + 1, int(warmup), tuned == "1") == (None, None)
        assert bench.profiled_traffic(wl, int(n), layout, int(steps), int(warmup) + 1, tuned == "1") == (None, None)
        assert bench.profiled_traffic(wl, int(n), layout, int(steps), int(warmup), tuned != "1") == (None, None)
        assert bench.profiled_traffic(wl, int(n) + 1, layout, int(steps), int(warmup), tuned == "1") == (None, None)
