"""Host logic of the differentiable rollout's segment plan (no GPU): which checkpoint segments
are recorded in the forward pass, with or without their grids."""
from nuclearmpm_amd import autodiff


def tape(k, with_grid):
    # a segment of k steps: 10 bytes per inner state and sort, 4 per step of grids
    return 10 * (k - 1) + (4 * k if with_grid else 0) + 1


def test_segment_sizes():
    assert autodiff.segment_sizes(256, 32) == [32] * 8
    assert autodiff.segment_sizes(36, 7) == [7, 7, 7, 7, 7, 1]
    assert autodiff.segment_sizes(5, 8) == [5]


def test_fixed_modes():
    sizes = autodiff.segment_sizes(36, 7)
    assert autodiff.plan_records(sizes, "none", tape) == [None] * 6
    assert autodiff.plan_records(sizes, "grid", tape) == [1] * 6
    assert autodiff.plan_records(sizes, "nogrid", tape) == [0] * 6


def test_auto_fills_the_budget_without_grids_first():
    sizes = [8, 8, 8, 8]  # 71 bytes without grids, 103 with
    assert autodiff.plan_records(sizes, "auto", tape, budget=0) == [None] * 4
    assert autodiff.plan_records(sizes, "auto", tape, budget=70) == [None] * 4
    assert autodiff.plan_records(sizes, "auto", tape, budget=150) == [0, 0, None, None]
    # all four without grids (284), then grids (32 each) while they fit
    assert autodiff.plan_records(sizes, "auto", tape, budget=284) == [0, 0, 0, 0]
    assert autodiff.plan_records(sizes, "auto", tape, budget=350) == [1, 1, 0, 0]
    assert autodiff.plan_records(sizes, "auto", tape, budget=412) == [1, 1, 1, 1]


def test_auto_takes_a_short_last_segment_that_fits():
    sizes = autodiff.segment_sizes(25, 8)  # 8, 8, 8, 1: 71, 71, 71 and 1 bytes without grids
    assert autodiff.plan_records(sizes, "auto", tape, budget=72) == [0, None, None, 0]
    assert autodiff.plan_records(sizes, "auto", tape, budget=76) == [0, None, None, 1]
