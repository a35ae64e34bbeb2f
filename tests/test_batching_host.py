"""Host logic of the uneven-batch aligner: bucketing and march schedules."""
from matcha_amd.batching import march_levels, plan_buckets


def test_plan_buckets_groups_by_box_and_L_in_first_seen_order():
    specs = [(64, 16), (96, 24), (64, 16), (128, 32), (96, 24), (64, 24), (64, 16)]
    b = plan_buckets(specs, batch=2)
    assert list(b) == [(64, 16), (96, 24), (128, 32), (64, 24)]
    assert b[(64, 16)] == [[0, 2], [6]]
    assert b[(96, 24)] == [[1, 4]]
    assert b[(128, 32)] == [[3]] and b[(64, 24)] == [[5]]
    # every particle exactly once
    flat = sorted(i for chunks in b.values() for c in chunks for i in c)
    assert flat == list(range(len(specs)))


def test_march_levels_truncate_below_the_end():
    assert march_levels(16) == [8]
    assert march_levels(24) == [8, 16]
    assert march_levels(32) == [8, 16, 24]
    assert march_levels(8) == [8]
    assert march_levels(12, Li=(16, 8)) == [8]
