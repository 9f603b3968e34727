"""check_data (utils.h:85-336): the oracle's restatement and the host
checker against violation counts worked out by hand for planted
corruptions, and against each other on random corruptions that include
out-of-range ids."""
import sevc
from _util import oracle_check, planted_corruptions, random_corruptions, ref2k_options


def _ref2k():
    return sevc.load_files(ref2k_options())


def test_valid_mesh_has_no_violations():
    m = _ref2k()
    assert oracle_check(m) == 0
    assert sevc.check_mesh(m)[0] == 0


def test_planted_corruptions_match_hand_counts():
    m = _ref2k()
    for name, c, expected in planted_corruptions(m):
        assert oracle_check(c) == expected, name
        assert sevc.check_mesh(c)[0] == expected, name


def test_random_corruptions_oracle_equals_host():
    m = _ref2k()
    counts = []
    for c in random_corruptions(m, 200, seed=7):
        a, b = oracle_check(c), sevc.check_mesh(c)[0]
        assert a == b
        counts.append(a)
    assert sum(1 for a in counts if a > 0) > 150      # most single edits are caught
