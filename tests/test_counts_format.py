"""Counts formatting of StateVectorSimulator.Run (StateVectorSimulator.cs:92-98):
binary strings padded to the qubit count, qubit 0 rightmost, ordinal order."""
from quantummatter_amd.simulator import counts_of


def test_padding_order_and_counts():
    got = counts_of([3, 1, 3, 0, 5, 3, 1], 4)
    assert list(got.items()) == [("0000", 1), ("0001", 2), ("0011", 3), ("0101", 1)]


def test_ordinal_order_matches_numeric_order_for_equal_lengths():
    outs = [9, 2, 15, 0, 8, 7]
    keys = list(counts_of(outs, 4))
    assert keys == sorted(keys) == [format(o, "04b") for o in sorted(set(outs))]


def test_wide_outcomes_keep_every_bit():
    assert list(counts_of([1 << 29], 30)) == ["1" + "0" * 29]
