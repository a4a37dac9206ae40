"""Host logic of the batched path (CPU): how StreamedBatches splits plans."""

import itertools

import numpy as np


def _check(costs, n, groups):
    m = len(costs)
    assert len(groups) == min(n, m)
    assert all(groups), "every group holds a plan"
    assert [i for g in groups for i in g] == list(range(m)), "contiguous, in order, each plan once"


def test_split_contiguous_cases():
    from DoseCUDA.batch import split_contiguous as split
    assert split([1, 1, 1, 1], 2) == [[0, 1], [2, 3]]
    assert split([1, 1, 1, 1], 4) == [[0], [1], [2], [3]]
    assert split([5, 1, 1, 1], 2) == [[0], [1, 2, 3]]
    assert split([1, 1, 1, 5], 2) == [[0, 1, 2], [3]]
    assert split([3], 4) == [[0]]
    assert split([0, 0, 0], 2) == [[0, 1], [2]]      # no cost information: by count
    assert split([2, 2], 1) == [[0, 1]]


def test_split_contiguous_is_near_balanced():
    """On random costs every split is a partition into min(n, m) contiguous
    groups, and a group other than the last with two or more plans has not yet
    reached its share (j + 1) / n of the total before its last plan (a group
    closes at the first plan whose running sum reaches its share)."""
    from DoseCUDA.batch import split_contiguous as split
    rng = np.random.default_rng(0)
    for m, n in itertools.product(range(1, 9), range(1, 6)):
        for _ in range(20):
            costs = rng.uniform(0.1, 10.0, size=m).tolist()
            groups = split(costs, n)
            _check(costs, n, groups)
            total = sum(costs)
            run = 0.0
            for j, g in enumerate(groups[:-1]):
                if len(g) >= 2:
                    assert run + sum(costs[i] for i in g[:-1]) < total * (j + 1) / len(groups)
                run += sum(costs[i] for i in g)
