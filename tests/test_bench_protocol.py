"""The benchmark's trial protocol (bench.run_trials), restating the reference's
benchmarks/harness/protocol.cpp and statistics.cpp, on a synthetic body."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_warmup_settling_and_nearest_rank_statistics(monkeypatch):
    # A body whose k-th call "lasts" durations[k]: the clock advances by it.
    durations = [9.0, 9.0, 1.0] + [1.0] * 40
    calls = []
    now = [0.0]

    def fake_clock():
        return now[0]

    def body():
        now[0] += durations[len(calls)]
        calls.append(1)

    monkeypatch.setattr(bench.time, "perf_counter", fake_clock)
    st = bench.run_trials(body, warmup=2, settling_limit=24, settled_within=0.05, trials=11,
                          minimum_trial=0.0)
    # 2 warm-ups, one sizing call, then settling until two agreements in a row
    # (three calls of 1.0), then 11 trials.
    assert st["warmup"] == 2 and st["settling_trials"] == 3 and st["trials"] == 11
    assert len(calls) == 2 + 1 + 3 + 11
    assert st["median_s"] == 1.0 and st["q1_s"] == 1.0 and st["q3_s"] == 1.0
    assert st["drift"] == 0.0 and st["calls_per_trial"] == 1


def test_quartiles_and_drift_follow_the_reference():
    # statistics.cpp: quantile = sorted[floor(f n)] (clamped), drift = median of
    # the later half over the earlier half's, minus one.
    seq = iter([1, 1, 1, 1, 5, 3, 4, 1, 2, 6, 7, 8, 9, 10, 11])
    now = [0.0]

    def body():
        now[0] += next(seq)

    real = bench.time.perf_counter
    bench.time.perf_counter = lambda: now[0]
    try:
        st = bench.run_trials(body, warmup=1, settling_limit=2, trials=11, minimum_trial=0.0)
    finally:
        bench.time.perf_counter = real
    trials = [5, 3, 4, 1, 2, 6, 7, 8, 9, 10, 11][:11]
    ordered = sorted(trials)
    assert st["median_s"] == ordered[5] and st["q1_s"] == ordered[2] and st["q3_s"] == ordered[8]
    early, late = sorted(trials[:5]), sorted(trials[-5:])
    assert abs(st["drift"] - (late[2] / early[2] - 1.0)) < 1e-12
