"""Autotuner choice rule (host logic): a candidate clearly faster than the rest wins; times
within the noise band of the fastest are ties, resolved by the fixed preference order, so
that the same matrix gets the same launch shape run after run."""
from inching_amd import autotune as T


def test_clearly_fastest_wins():
    times = {1232: 2.00, 232: 1.99, 16: 1.80}
    assert T._pick(times, T.BSR_PREFERENCE) == 16


def test_ties_resolved_by_preference():
    # 1.5% apart: a tie, the preferred 16-bit XCD shape wins whichever was measured faster
    assert T._pick({1232: 2.00, 16: 1.97}, T.BSR_PREFERENCE) == 1232
    assert T._pick({1232: 1.97, 16: 2.00}, T.BSR_PREFERENCE) == 1232
    assert T._pick({6104: 0.8267, 6001: 0.8379, 4001: 0.880}, T.CLUSTER_PREFERENCE) == 6001


def test_candidates_outside_the_order_still_compete():
    assert T._pick({999: 1.0, 1232: 1.5}, T.BSR_PREFERENCE) == 999


def test_preference_orders_cover_every_candidate():
    bsr = set(T.BSR_LANES) | {200 + g for g in T.BSR_LANES} | {1200 + g for g in T.BSR_LANES}
    assert set(T.BSR_PREFERENCE) == bsr
    cl = {f + c for f in (4000, 6000) for c in T.CLUSTER_CONFIGS}
    assert set(T.CLUSTER_PREFERENCE) == cl
