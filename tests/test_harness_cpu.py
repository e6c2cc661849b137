"""CPU tests of the timing helper shared with the reference
(benchmark_functions, src/_utils.py:6-32)."""
from warpvbd_amd import benchmark_functions


def test_benchmark_functions_samples_and_skip():
    calls = []

    def f(x, y=0):
        calls.append((x, y))
        return x + y

    def g(x, y=0):
        return x * y

    t = benchmark_functions([f, g], 2, num_samples=3, y=5)
    assert set(t) == {"f", "g"} and len(t["f"]) == 3 and len(t["g"]) == 3
    assert all(v >= 0.0 for v in t["f"])
    assert calls == [(2, 5)] * 3
    calls.clear()
    t = benchmark_functions([f], 1, num_samples=2, skipOne=True)
    assert len(t["f"]) == 2 and len(calls) == 3
