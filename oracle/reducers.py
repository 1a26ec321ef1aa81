"""Oracle of the downstream stream reducers (mission duration, safe haven).

TEST INFRASTRUCTURE: a numpy restatement of the reference's streaming
reducers, fed with whole [time][y][x] signal stacks. Only the tests (and the
benchmark's CPU-baseline leg) use it, as the checker; the product reduces
inside the lighting walk (``lsh_durations_core.h``) and never imports this.

* ``longest_candidate_durations`` -- reduce_longest_candidate_duration_stream
  (src/lunarscout/_numba_horizon/mission_duration.py:178-256), with the
  candidate-start intervals given as sample-index ranges [first_k, end_k)
  (the samples whose time lies in the interval; mission_duration.py:225-227
  tests start <= time < stop on sorted samples);
* ``safe_haven`` -- reduce_safe_haven_patch_stream (safe_haven.py:79-243).

Both follow the reference's per-time state machines step by step (same
float64 accumulation order for durations, same integer run lengths), so they
pin the fused reductions bit for bit given the same signal stacks. The
reference's own known-answer cases pin these functions
(tests/test_durations.py).
"""

from __future__ import annotations

import numpy as np


def longest_candidate_durations(conditions, sample_hours, interval_first, interval_end, *,
                                output_days: bool = False) -> np.ndarray:
    """float32 [K, ...]: per interval, the longest right-censored candidate duration.

    conditions: bool [T, ...]; sample_hours: float64 [T] (hours sample t stands for).
    """
    cond = np.asarray(conditions, dtype=bool)
    hours = np.asarray(sample_hours, dtype=np.float64)
    T = cond.shape[0]
    K = len(interval_first)
    shape = cond.shape[1:]
    longest = np.zeros((K,) + shape, dtype=np.float64)
    for k in range(K):
        active = np.zeros(shape, dtype=bool)
        current = np.zeros(shape, dtype=np.float64)
        best = longest[k]
        for t in range(T):
            c = cond[t]
            ended = active & ~c                       # mission_duration.py:232-238
            best[ended] = np.maximum(best[ended], current[ended])
            current[ended] = 0.0
            active[ended] = False
            if interval_first[k] <= t < interval_end[k]:  # accepts a start (:239-240)
                active |= c
            accruing = active & c                     # :241-242
            current[accruing] += hours[t]
        np.maximum(best, current, out=best)           # right-censored at the stop (:251)
    if output_days:
        longest /= 24.0
    return longest.astype(np.float32)


def safe_haven(fractions, earth_margins, month_of, band_count: int, *, sunlight_threshold: float,
               earth_threshold_deg: float, time_step_hours: float) -> np.ndarray:
    """float32 [B, ...]: per month band, the longest low-Sun run (in hours) that
    overlaps an Earth outage in that month; NaN where the band is ill-posed."""
    fr = np.asarray(fractions, dtype=np.float32)
    em = np.asarray(earth_margins, dtype=np.float32)
    T = fr.shape[0]
    shape = fr.shape[1:]
    run_length = np.zeros(shape, dtype=np.int32)
    touching = np.zeros((band_count,) + shape, dtype=bool)
    best = np.zeros((band_count,) + shape, dtype=np.int32)
    had_outage = np.zeros((band_count,) + shape, dtype=bool)
    was_above = np.zeros((band_count,) + shape, dtype=bool)
    for t in range(T):
        low = fr[t] < np.float32(sunlight_threshold)           # safe_haven.py:184-185
        below = em[t] < np.float32(earth_threshold_deg)
        run_length = np.where(low, run_length + 1, 0)          # :190
        m = int(month_of[t])
        if m >= 0:                                             # :194-198
            was_above[m] |= ~below
            had_outage[m] |= below
            touching[m] |= below & low
        for b in range(band_count):                            # :203-208
            active = touching[b] & low
            best[b] = np.maximum(best[b], np.where(active, run_length, 0))
            touching[b] &= low
    censored = run_length > 0                                  # :211-220
    for b in range(band_count):
        best[b] = np.maximum(best[b], np.where(censored & touching[b], run_length, 0))
    out = best.astype(np.float32) * np.float32(time_step_hours)  # :232-236
    out[~had_outage | ~was_above] = np.nan
    return out
