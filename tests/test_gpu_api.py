"""Input and argument validation of the C ABI on a device: the codes the
library returns where the reference calls exit() or would read invalid
pointers (utils.h:17-44, check_data at load).  A state that rejected a
call keeps working."""
import pytest

import sevc
from sevc._abi import HostMesh
from _util import planted_corruptions, ref2k_options

pytestmark = pytest.mark.gpu


def _code(fn):
    with pytest.raises(sevc.SevcError) as e:
        fn()
    return e.value.code


def test_create_rejects_invalid_meshes_and_options():
    o = ref2k_options()
    m = sevc.load_files(o)
    print("  empty mesh", flush=True)
    assert _code(lambda: sevc.State(o, HostMesh(0, 0, 0, 32))) == sevc.E_INVALID_ARG
    assert _code(lambda: sevc.State([], [])) == sevc.E_INVALID_ARG
    for name, c, _ in planted_corruptions(m):
        print("  corrupted:", name, flush=True)
        assert _code(lambda: sevc.State(o, c)) == sevc.E_TOPOLOGY, name
    print("  list stride above 128", flush=True)
    o129 = ref2k_options(max_vrt_per_grn=129)
    assert _code(lambda: sevc.State(o129, sevc.load_files(o129))) == sevc.E_INVALID_ARG
    print("  dt = 0", flush=True)
    o0 = ref2k_options(dt=0.0)
    assert _code(lambda: sevc.State(o0, m)) == sevc.E_INVALID_ARG
    print("  a corrupted member rejects the whole batch", flush=True)
    bad = planted_corruptions(m)[0][1]
    assert _code(lambda: sevc.State([o, o], [m, bad])) == sevc.E_TOPOLOGY


def test_calls_reject_bad_arguments_and_state_keeps_working():
    o = ref2k_options(dt=2e-4)
    m = sevc.load_files(o)
    st = sevc.State(o, m)
    assert _code(lambda: st.step(-1)) == sevc.E_INVALID_ARG
    assert _code(lambda: st.get_mesh(1)) == sevc.E_INVALID_ARG
    assert _code(lambda: st.member_stats(3)) == sevc.E_INVALID_ARG
    assert _code(lambda: st.positions(2)) == sevc.E_INVALID_ARG
    assert _code(lambda: sevc.check_mesh_device(HostMesh(0, 0, 0, 32))) == sevc.E_INVALID_ARG
    s = st.step(150)
    assert s.status == 0 and s.steps_done == 150 and s.n_flips > 0
    assert st.check() == [0]
    st.close()


def test_sampled_kernel_timing_counts_launches():
    """Kernel timing of the hot kernels on one step in `every`: the launch
    counts follow the sample; bad arguments are rejected."""
    o = ref2k_options(dt=2e-4)
    st = sevc.State(o, sevc.load_files(o))
    st.set_velocity_config(0, 64)
    assert _code(lambda: st.set_timing_sample(0)) == sevc.E_INVALID_ARG
    assert _code(lambda: st.set_search_round(-1)) == sevc.E_INVALID_ARG
    st.enable_timing(2)
    st.set_timing_sample(4)
    st.step(10)                         # steps 0, 4 and 8 are timed
    kt = st.kernel_times()
    assert kt["vel_grain"][1] == 3 and kt["velocity"][1] == 3 and kt["text"][1] == 3
    assert kt["vel_grain"][0] > 0 and "flip_par" not in kt
    st.enable_timing(2)                 # counts restart; the sample is kept
    st.step(5)
    assert st.kernel_times()["text"][1] == 2
    st.set_timing_sample(1)
    st.enable_timing(0)
    st.close()


def test_member_mesh_restore_replays_the_same_steps():
    """sevc_set_member_mesh restores a checkpoint taken with sevc_get_mesh:
    stepping again from it gives the same state as the first time (no
    nucleation, so the step counter does not enter), in a batch whose other
    member is left alone."""
    from _util import diff_fields
    o = ref2k_options(dt=2e-4)
    m = sevc.load_files(o)
    st = sevc.State([o, o], [m, m])
    st.step(50)
    ck = st.get_mesh(0)
    st.step(50)
    a = st.get_mesh(0)
    st.set_mesh(0, ck)
    assert diff_fields(st.get_mesh(0), ck, ["pos", "v_bnd", "v_grn", "b_vtx", "g_vtx", "g_vlen", "g_enabled"]) == {}
    st.step(50)
    assert diff_fields(st.get_mesh(0), a) == {}
    bad = ck.copy()
    bad["b_enabled"][10] = 0
    assert _code(lambda: st.set_mesh(0, bad)) == sevc.E_TOPOLOGY
    st.close()


def test_device_checkpoint_restore_repeats_steps():
    """sevc_checkpoint / sevc_restore keep and put back the whole device state
    (meshes, scratch, loop state, with the step counter the nucleation stream
    is keyed by): the steps after a restore repeat the first run bit for
    bit, nucleations included, in a batch of two members."""
    from _util import diff_fields
    o = ref2k_options(dt=2e-4, adaptive_dt=1, dt_safety=0.5, flip_threshold=1e-4, do_nucleation=1,
                      nucleation_mode=sevc.NUCL_RATE, nucleation_rate=2e-2, seed=4, extra_capacity=0.05)
    m = sevc.load_files(o)
    o2 = ref2k_options(dt=1e-4)
    st = sevc.State([o, o2], [m, sevc.load_files(o2)])
    st.step(20)
    st.checkpoint(3)
    s_a = st.step(40)
    a = [st.get_mesh(k) for k in range(2)]
    st.restore(3)
    s_b = st.step(40)
    for k in range(2):
        assert diff_fields(st.get_mesh(k), a[k]) == {}, k
    assert s_a.n == s_b.n and s_a.n_nucleations == s_b.n_nucleations > 0 and s_a.t == s_b.t
    assert _code(lambda: st.restore(9)) == sevc.E_INVALID_ARG
    st.close()


@pytest.mark.parametrize("graph", [False, True])
def test_call_ends_when_every_member_has_stopped(graph):
    """A call asked for far more steps than the stop rule allows (the
    2,000-grain run to 50 % removed) stops launching once every member has
    finished: it returns in well under the time 100,000 launched steps take
    (about 10 s), with the oracle's step count and end state."""
    import time
    from _util import Oracle, diff_fields
    o = ref2k_options(dt=2e-4, adaptive_dt=1, dt_safety=0.5, flip_threshold=1e-4, do_nucleation=1,
                      nucleation_mode=sevc.NUCL_RATE, nucleation_rate=1e-3, nucleation_scaling=0.9, seed=3,
                      extra_capacity=0.05, leave_percent=0.5)
    mesh = sevc.load_files(o)
    st, orc = sevc.State(o, mesh), Oracle(o, mesh)
    t0 = time.perf_counter()
    s1 = st.step(100000, graph=graph)
    el = time.perf_counter() - t0
    rc, s2 = orc.step(100000)
    assert rc == 0 and s1.status == 0 and s1.done
    assert s1.n == s2.n and s1.n_grains == s2.n_grains <= 1000
    assert diff_fields(st.get_mesh(), orc.get_mesh()) == {}
    assert el < 3.0, el
    st.close()
