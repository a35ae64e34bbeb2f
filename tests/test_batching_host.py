"""Host logic of the uneven-batch aligner: bucketing and march schedules."""
from matcha_amd.batching import grid_degree, march_levels, plan_buckets, plan_groups


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


def test_plan_groups_mix_march_ends_of_one_box():
    specs = [(64, 16), (96, 32), (64, 32), (64, 24), (96, 16), (64, 16), (64, 8), (64, 6)]
    g = plan_groups(specs, batch=3, Li=(8, 16, 24))
    # ends >= min(Li) share the grid bandlimit 8; an end below it has its own grid
    assert list(g) == [(64, 8), (96, 8), (64, 6)]
    assert g[(64, 8)] == [[0, 2, 3], [5, 6]]
    assert g[(96, 8)] == [[1, 4]]
    assert g[(64, 6)] == [[7]]


def test_grid_degree_is_the_smallest_level_of_the_schedule():
    assert grid_degree(32) == 8
    assert grid_degree(8) == 8
    assert grid_degree(6) == 6
    assert grid_degree(20, Li=(12, 16)) == 12
