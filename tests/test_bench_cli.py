"""bench.py on the CPU: its command line measures one GPU and refuses more, and
its precision modes are warmed and timed interleaved."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("n", [2, 8])
def test_bench_refuses_several_gpus(n):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", str(n)],
                       capture_output=True, text=True, timeout=60, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""  # no JSON line for a configuration that was not measured


def test_best_interleaved_warms_and_alternates():
    sys.path.insert(0, ROOT)
    import bench

    calls = []
    times = iter([9.0, 9.0, 8.0, 8.0, 3.0, 5.0, 4.0, 2.0])

    def run(name):
        calls.append(name)
        return {"total_s": next(times), "i": len(calls)}

    best = bench.best_interleaved(("fp64", "mixed"), run, warm=2, timed=2)
    assert calls == ["fp64", "mixed"] * 4
    # warm-up runs (the first four calls) never count, however fast
    assert best["fp64"] == {"total_s": 3.0, "i": 5}
    assert best["mixed"] == {"total_s": 2.0, "i": 8}


def test_best_interleaved_ignores_fast_warmup():
    sys.path.insert(0, ROOT)
    import bench

    times = iter([0.1, 0.1, 7.0, 6.0])
    best = bench.best_interleaved(("a", "b"), lambda name: {"total_s": next(times)},
                                  warm=1, timed=1)
    assert best == {"a": {"total_s": 7.0}, "b": {"total_s": 6.0}}
