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


def test_openmp_oracle_build_gives_the_same_bits():
    """The OpenMP build of the oracle (a timing baseline) runs the physics
    loops on threads; its states equal the serial build's bit for bit."""
    from _util import Oracle, diff_fields
    import sevc as _s
    o = ref2k_options(dt=2e-4, adaptive_dt=1, dt_safety=0.5, flip_threshold=1e-4, do_nucleation=1,
                      nucleation_mode=_s.NUCL_RATE, nucleation_rate=1e-3, nucleation_scaling=0.9, seed=4,
                      extra_capacity=0.05)
    m = _ref2k()
    a, b = Oracle(o, m), Oracle(o, m, omp=True)
    for n in (10, 60):
        ra, sa = a.step(n)
        rb, sb = b.step(n)
        assert ra == 0 and rb == 0
        assert diff_fields(a.get_mesh(), b.get_mesh()) == {}
        assert (sa.n_flips, sa.n_removals, sa.n_nucleations) == (sb.n_flips, sb.n_removals, sb.n_nucleations)
    assert sa.n_flips > 0
