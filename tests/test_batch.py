"""The training-batch sampler (Batch, src/ml/batch.hpp:18-55, as used by
src/nnp/nnptefr2.cpp:676-678 and 1309-1315)."""
import numpy as np
import pytest

from ann_amd.batch import BatchSampler


def test_epochs_partition_the_set_when_size_divides_capacity():
    s = BatchSampler(4, 12, seed=3)
    for _ in range(5):
        seen = []
        for _ in range(3):
            b = s.next()
            assert np.all(np.diff(b) > 0)          # sorted, no repeats
            seen.extend(b.tolist())
        assert sorted(seen) == list(range(12))
        assert s.count == 0


def test_wraparound_batch_then_reshuffle():
    """With 5 of 7, the second draw wraps: indices 5, 6 of the current
    permutation, then 0, 1, 2 of the same permutation; then it reshuffles."""
    s = BatchSampler(5, 7, seed=1)
    perm = s.data.copy()
    b1 = s.next()
    assert sorted(b1.tolist()) == sorted(perm[:5].tolist()) and s.count == 5
    b2 = s.next()
    assert sorted(b2.tolist()) == sorted(perm[[5, 6, 0, 1, 2]].tolist())
    assert s.count == 0 and sorted(s.data.tolist()) == list(range(7))


def test_reference_argument_checks():
    with pytest.raises(ValueError, match="smaller than capacity"):
        BatchSampler(8, 4)
    with pytest.raises(ValueError, match="negative"):
        BatchSampler(-1, 4)
    s = BatchSampler(0, 3)
    assert len(s.next()) == 0
