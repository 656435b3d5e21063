"""Autotuner search logic (CPU): coordinate descent finds the minimum of a
separable cost, never ends worse than its start, and evaluates each point
once."""
from branson_amd import autotune


def test_coordinate_descent_finds_separable_minimum():
    cand = {"a": [1, 2, 3, 4], "b": [10, 20, 30], "c": [0, 1]}
    target = {"a": 3, "b": 10, "c": 1}
    calls = []

    def cost(p):
        calls.append(p)
        return sum((p[k] - target[k]) ** 2 for k in p)

    best, c, log = autotune.coordinate_descent(cost, cand, start={"a": 1, "b": 30, "c": 0})
    assert best == target and c == 0
    keys = [tuple(sorted(p.items())) for p in calls]
    assert len(keys) == len(set(keys))  # no point measured twice
    assert log[0][0] == {"a": 1, "b": 30, "c": 0}


def test_coordinate_descent_never_worse_than_start():
    cand = {"x": [0, 1, 2], "y": [0, 1, 2]}
    # a non-separable cost with a trap: the start is already a local minimum
    table = {(0, 0): 5.0, (1, 0): 6.0, (2, 0): 7.0, (0, 1): 6.0, (0, 2): 9.0,
             (1, 1): 1.0, (2, 2): 0.5, (1, 2): 8.0, (2, 1): 8.0}
    best, c, _ = autotune.coordinate_descent(lambda p: table[(p["x"], p["y"])], cand,
                                             start={"x": 0, "y": 0})
    assert c <= table[(0, 0)]


def test_defaults_are_candidates():
    for k, v in autotune.DEFAULTS.items():
        assert v in autotune.CANDIDATES[k]
