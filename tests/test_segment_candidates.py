"""Candidate segment counts of the storage-order autotuner (host logic): one
per target segment length, rounded as dmpm_sort_segments rounds them (equal
runs of layers, whole checkpoint segments), without duplicates."""
from dmpm.autotune import segment_candidates


def _segments(L, K, S):
    """The runs dmpm_sort_segments makes of L layers (its documented rule)."""
    P = -(-L // S)
    P = -(-P // K) * K
    return [(t, min(t + P, L)) for t in range(0, L, P)]


def test_whole_tape():
    # 301 layers: one order, 13 runs of 24 layers (the last 13), 31 of 10
    assert segment_candidates(301, 1) == [1, 13, 31]
    assert [len(_segments(301, 1, S)) for S in (1, 13, 31)] == [1, 13, 31]


def test_checkpointed_runs_are_whole_checkpoint_segments():
    # 501 layers, K = 25: 25 and 10 layers both round to 25-layer runs
    assert segment_candidates(501, 25) == [1, 21]
    for S in (1, 21):
        runs = _segments(501, 25, S)
        assert all(a % 25 == 0 for a, _ in runs) and runs[-1][1] == 501
    assert segment_candidates(101, 10) == [1, 5, 11]


def test_short_episodes():
    assert segment_candidates(13, 3) == [1, 2]
    assert segment_candidates(2, 1) == [1]
    assert segment_candidates(1, 1) == [1]
