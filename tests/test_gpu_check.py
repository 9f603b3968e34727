"""check_data on the GPU: the device checker against the oracle's
restatement (utils.h:85-336) on planted and random corruptions, the check
of a live state, and the in-step checks of debug_lvl >= 1 (main.cu:185-196,
237-248, 269-280, 297-308), which must leave every result unchanged."""
import pytest

import sevc
from _util import Oracle, diff_fields, oracle_check, planted_corruptions, random_corruptions, ref2k_options
from test_gpu_parity import _generated, _run_pair

pytestmark = pytest.mark.gpu


def test_device_checker_matches_oracle_on_corruptions():
    m = sevc.load_files(ref2k_options())
    print("  valid mesh", flush=True)
    assert sevc.check_mesh_device(m) == 0
    for name, c, expected in planted_corruptions(m):
        print("  planted:", name, flush=True)
        assert sevc.check_mesh_device(c) == expected, name
    print("  random corruptions (ids in and out of range)", flush=True)
    for k, c in enumerate(random_corruptions(m, 200, seed=7)):
        assert sevc.check_mesh_device(c) == oracle_check(c), k


def test_live_state_check_after_events():
    o = ref2k_options(dt=2e-4, leave_percent=0.5)
    mesh = sevc.load_files(o)
    st = sevc.State(o, mesh)
    assert st.check() == [0]
    s = st.step(300)
    assert s.n_flips > 100 and s.n_removals > 100, (s.n_flips, s.n_removals)
    assert st.check() == [0]
    assert oracle_check(st.get_mesh()) == 0
    st.close()


def test_debug_checks_leave_results_unchanged():
    """debug_lvl = 1 runs the four checks in every step; with a valid
    state they change nothing, so the state still equals the oracle's
    bit for bit (plain and graph replay)."""
    o = ref2k_options(dt=2e-4, leave_percent=0.5, debug_lvl=1)
    mesh = sevc.load_files(o)
    tot = _run_pair(o, mesh, [50, 150])
    assert tot["removals"] >= 300 and tot["flips"] >= 300, tot
    tot = _run_pair(o, mesh, [40, 40], graph=True)
    assert tot["flips"] >= 10 and tot["removals"] >= 10, tot


def test_debug_checks_with_nucleation_in_a_batch():
    """A debugged member next to a plain one: both match their oracles."""
    kw = dict(dt=1e-4, adaptive_dt=1, dt_safety=0.5, flip_threshold=2e-3, do_nucleation=1,
              nucleation_mode=sevc.NUCL_RATE, nucleation_rate=0.02, nucleation_scaling=0.9, seed=11,
              extra_capacity=0.1)
    o1, m1 = _generated(3000, 5, debug_lvl=1, **kw)
    o2, m2 = _generated(2000, 6, **kw)
    st = sevc.State([o1, o2], [m1, m2])
    orcs = [Oracle(o1, m1), Oracle(o2, m2)]
    for n in (20, 80):
        print(f"  chunk of {n} steps", flush=True)
        st.step(n)
        for k, orc in enumerate(orcs):
            rc, s2 = orc.step(n)
            assert rc == 0
            s1 = st.member_stats(k)
            assert s1.status == 0 and s1.n == s2.n and s1.n_nucleations == s2.n_nucleations
            assert diff_fields(st.get_mesh(k), orc.get_mesh()) == {}, k
    assert st.check() == [0, 0]
    st.close()
