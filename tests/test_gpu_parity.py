"""GPU parity: the HIP time step against the serial FP64 oracle.

Every test runs the same seeded input through the C ABI on the GPU and
through the oracle, then compares the whole state.  Integer topology and
every FP64 field must match bit for bit: both sides evaluate the same
operations in the same order with no fused multiply-add, and the device's
division, square root and elementary functions are checked against the
host's bits first."""
import numpy as np
import pytest

import sevc
from _util import Oracle, diff_fields, oracle_lib, ref2k_options

pytestmark = pytest.mark.gpu

STAT_KEYS = ("n", "n_grains", "n_vertices", "n_flips", "n_removals", "n_two_grain", "n_nucleations",
             "n_vertex_steps", "steps_done", "done")


def _run_pair(opts, mesh, chunks, graph=False, setup=None):
    """Step GPU and oracle in the given chunk sizes; compare after each.
    `setup`, if given, is called with the GPU state before the first step."""
    st, orc = sevc.State(opts, mesh), Oracle(opts, mesh)
    if setup is not None:
        setup(st)
    tot = {"flips": 0, "removals": 0, "nucl": 0, "two": 0}
    for n in chunks:
        print(f"  chunk of {n} steps", flush=True)
        s1 = st.step(n, graph=graph)
        rc, s2 = orc.step(n)
        assert rc == 0 and s1.status == 0, (rc, s1.status)
        a, b = st.get_mesh(), orc.get_mesh()
        bad = diff_fields(a, b)
        assert bad == {}, f"state differs after chunk: {bad}"
        for k in STAT_KEYS:
            assert getattr(s1, k) == getattr(s2, k), (k, getattr(s1, k), getattr(s2, k))
        assert s1.t == s2.t and s1.dt_last == s2.dt_last
        assert s1.nucleated_area_pct == s2.nucleated_area_pct
        tot["flips"] += s1.n_flips
        tot["removals"] += s1.n_removals
        tot["nucl"] += s1.n_nucleations
        tot["two"] += s1.n_two_grain
    st.close()
    return tot


def test_selftest_math_bits():
    rng = np.random.default_rng(0)
    n = 20000
    x = rng.uniform(-4, 4, n)
    y = rng.uniform(-4, 4, n)
    L = oracle_lib()
    cases = {0: lambda a, b: L.orc_atan2(a, b), 1: lambda a, b: L.orc_sin(a), 2: lambda a, b: L.orc_cos(a),
             3: lambda a, b: L.orc_exp(a), 4: lambda a, b: a / b, 6: lambda a, b: L.orc_mag(a, b),
             7: lambda a, b: L.orc_wrap_dist(a, b, 1.0)}
    for op, f in cases.items():
        dev = sevc.selftest_math(op, x, y)
        host = np.array([f(a, b) for a, b in zip(x, y)])
        assert np.array_equal(dev.view(np.uint64), host.view(np.uint64)), f"op {op} differs"
    # correctly rounded sqrt over many binades, and division by tiny/huge numbers
    xs = np.abs(x) * 10.0 ** rng.integers(-300, 300, n).astype(np.float64)
    assert np.array_equal(sevc.selftest_math(5, xs, y).view(np.uint64), np.sqrt(xs).view(np.uint64))
    ys = y * 10.0 ** rng.integers(-150, 150, n).astype(np.float64)
    with np.errstate(over="ignore", under="ignore"):
            assert np.array_equal(sevc.selftest_math(4, xs, ys).view(np.uint64), (xs / ys).view(np.uint64))
    # the square root of mag's argument 1 + (q/p)^2 in [1, 2], shortened for
    # that range, equals the correctly rounded sqrt bit for bit
    r = rng.uniform(0, 1, 200000)
    s12 = np.concatenate([1.0 + r * r, rng.uniform(1, 2, 200000),
                          [1.0, 2.0, np.nextafter(1.0, 2.0), np.nextafter(2.0, 1.0), 1.5, 1.25]])
    assert np.array_equal(sevc.selftest_math(9, s12, np.zeros_like(s12)).view(np.uint64),
                          np.sqrt(s12).view(np.uint64))
    # mag on axis-aligned, equal-component, zero, tiny and huge vectors
    special = np.array([0.0, -0.0, 1e-300, -1e-300, 3e-17, 0.5, -0.5, 1.0, 1e300, 5e-324])
    ax, ay = np.meshgrid(special, special)
    ax, ay = ax.ravel(), ay.ravel()
    dev = sevc.selftest_math(6, ax, ay)
    host = np.array([L.orc_mag(a, b) for a, b in zip(ax, ay)])
    assert np.array_equal(dev.view(np.uint64), host.view(np.uint64))
    # sin and cos from one reduction equal sevc_sin and sevc_cos, over the
    # atan2 range (the edge tangents) and wider, with signed zeros and the
    # octant edges
    a = np.concatenate([rng.uniform(-np.pi, np.pi, n), rng.uniform(-50, 50, n),
                        np.arange(-16, 17) * (np.pi / 4), [0.0, -0.0, np.pi, -np.pi, 1e-300, -1e-300]])
    z = np.zeros_like(a)
    assert np.array_equal(sevc.selftest_math(12, a, z).view(np.uint64), sevc.selftest_math(1, a, z).view(np.uint64))
    assert np.array_equal(sevc.selftest_math(13, a, z).view(np.uint64), sevc.selftest_math(2, a, z).view(np.uint64))
    # Philox word 0 on the device equals the host restatement
    ctr = rng.integers(0, 2 ** 32, n).astype(np.float64)
    dev = sevc.selftest_math(8, ctr, ctr[::-1].copy())
    import ctypes as C
    host = []
    for a, b in zip(ctr, ctr[::-1]):
        c = (C.c_uint32 * 4)(int(a), int(b), 0, 0)
        k = (C.c_uint32 * 2)(1, 2)
        o = (C.c_uint32 * 4)()
        L.orc_philox(c, k, o)
        host.append(float(o[0]))
    assert np.array_equal(dev, np.array(host))


def test_wave_scan_lane_layout():
    """The DPP wave scan used by the free-slot search and the workgroup
    scans: inclusive prefix sums and totals over each wave of 64 lanes equal
    numpy's, for counts up to 256 per lane, zeros, and a partial last
    workgroup (lanes past the input hold 0); and the DPP wave minimum of
    64-bit patterns that the extinction-time minimum uses."""
    rng = np.random.default_rng(3)
    for x in (rng.integers(0, 257, 64 * 300), np.zeros(640), rng.integers(0, 2, 64 * 40),
              rng.integers(0, 1000, 64 * 7 + 19)):
        x = x.astype(np.float64)
        pad = np.concatenate([x, np.zeros(-x.size % 64)]).reshape(-1, 64)
        scan = sevc.selftest_math(10, x, np.zeros_like(x))
        assert np.array_equal(scan, np.cumsum(pad, axis=1).ravel()[:x.size])
        tot = sevc.selftest_math(11, x, np.zeros_like(x))
        assert np.array_equal(tot, np.repeat(pad.sum(axis=1), 64)[:x.size])
    # wave minimum of 64-bit patterns (non-negative doubles and +inf, the
    # extinction-time minimum), with a partial last wave
    for x in (rng.uniform(0, 5, 64 * 50), np.full(128, np.inf), np.concatenate([rng.uniform(0, 1, 64 * 3 + 5)])):
        x = x.copy()
        x[rng.random(x.size) < 0.3] = np.inf
        bits = x.view(np.uint64)
        pad = np.concatenate([bits, np.full(-x.size % 64, np.iinfo(np.uint64).max, dtype=np.uint64)]).reshape(-1, 64)
        got = sevc.selftest_math(14, x, np.zeros_like(x)).view(np.uint64)
        assert np.array_equal(got, np.repeat(pad.min(axis=1), 64)[:x.size])


def test_step0_velocities_bitwise():
    o = ref2k_options(events=0)
    mesh = sevc.load_files(o)
    _run_pair(o, mesh, [1])


def test_run_until_half_removed():
    """Config-1 style run of the 2,000-grain fixture to 50 % removal."""
    o = ref2k_options(dt=2e-4, leave_percent=0.5)
    mesh = sevc.load_files(o)
    tot = _run_pair(o, mesh, [50, 150, 400])
    assert tot["removals"] >= 900 and tot["flips"] >= 1000, tot


def test_mc_nucleation_recrystallisation():
    """Reference Metropolis nucleation (nucleation2.cu) at the example.ini
    scale: domain 40, SE in [3, 6], gate at 5/6 of the grains."""
    o = ref2k_options(domain_bound=40.0, dt=1e-3, dtau=1e-3, do_nucleation=1, leave_percent=0.05,
                      SE_eps=1.0, MC_k=1.0)
    o.SE_min, o.SE_max = 3.0, 6.0
    mesh = sevc.load_files(o)
    tot = _run_pair(o, mesh, [400, 400, 400])
    assert tot["nucl"] >= 20, tot


def _generated(n, gen_seed, **kw):
    pos, rid = sevc.voronoi(n, gen_seed)
    rng = np.random.default_rng(gen_seed)
    se = np.exp(np.log(0.3) - 0.125 + 0.5 * rng.standard_normal(n))
    o = sevc.default_options()
    o.GB_eps = 0.0
    o.n_grains = n
    for k, v in kw.items():
        setattr(o, k, v)
    return o, sevc.build_mesh(o, pos, rid, rng.random(n), se, n)


@pytest.mark.parametrize("path,tile", [(0, 64), (1, 64), (2, 0)])
def test_vertex_on_the_domain_edge(path, tile):
    """Vertices whose coordinate dom_mod moves (a coordinate stored as D
    itself, which an addition of D to a tiny negative value rounds to) take
    the reference's two energy_diff passes (calculus.h:19-73) in
    k_velocity_fallback, listed by pass 2 or by the one-pass tiles.  Every
    path equals the oracle."""
    o = ref2k_options(dt=1e-4)
    mesh = sevc.load_files(o)
    D = o.domain_bound
    pos = mesh["pos"]
    pos[2 * 5] = D                        # x of vertex 5
    pos[2 * 40 + 1] = D                   # y of vertex 40
    pos[2 * 77] = D
    pos[2 * 77 + 1] = D
    st, orc = sevc.State(o, mesh), Oracle(o, mesh)
    st.set_velocity_config(path, tile)
    for chunk in (1, 1, 5):
        s1 = st.step(chunk)
        rc, s2 = orc.step(chunk)
        assert rc == 0 and s1.status == 0
        assert diff_fields(st.get_mesh(), orc.get_mesh()) == {}
    st.close()


def test_rate_nucleation_adaptive_threshold():
    """The benchmark's model options (rate nucleation, adaptive dt, flip
    threshold) on a 3,000-grain tessellation."""
    o, mesh = _generated(3000, 5, dt=1e-4, adaptive_dt=1, dt_safety=0.5, flip_threshold=2e-3,
                         do_nucleation=1, nucleation_mode=sevc.NUCL_RATE, nucleation_rate=0.02,
                         nucleation_scaling=0.9, seed=11, extra_capacity=0.1)
    tot = _run_pair(o, mesh, [20, 80, 100])
    assert tot["nucl"] >= 10 and tot["flips"] >= 50, tot


@pytest.mark.parametrize("path,tile", [(0, 16), (0, 64), (1, 32), (1, 64), (2, 0)])
def test_velocity_paths_bitwise(path, tile):
    """Both FD velocity paths -- two passes (grain tiles, then vertices) and
    one pass over vertex tiles -- at every tile size give the oracle's bits
    through flips, removals and nucleations: on a generated tessellation
    (vertex ids follow space, so every one-pass tile fits its staging) with
    rate nucleation and adaptive dt, and on the reference's 2,000-grain
    fixture with anisotropic energies (vertex ids in file order: one-pass
    tiles overflow and their vertices take the two-pass fallback)."""
    cases = [_generated(3000, 5, dt=1e-4, adaptive_dt=1, dt_safety=0.5, flip_threshold=2e-3, do_nucleation=1,
                        nucleation_mode=sevc.NUCL_RATE, nucleation_rate=0.02, nucleation_scaling=0.9, seed=11,
                        extra_capacity=0.1)]
    o2 = ref2k_options(dt=2e-4, GB_eps=0.3)
    cases.append((o2, sevc.load_files(o2)))
    for ci, (o, mesh) in enumerate(cases):
        st, orc = sevc.State(o, mesh), Oracle(o, mesh)
        assert st.velocity_config() == (0, 64)       # a new state's path
        st.set_velocity_config(path, tile)
        assert st.velocity_config() == (path, tile)
        tot = 0
        for chunk in (1, 20, 60):
            print(f"  case {ci} path {path} tile {tile}: chunk of {chunk}", flush=True)
            s1 = st.step(chunk)
            rc, s2 = orc.step(chunk)
            assert rc == 0 and s1.status == 0
            assert diff_fields(st.get_mesh(), orc.get_mesh()) == {}
            for k in STAT_KEYS:
                assert getattr(s1, k) == getattr(s2, k), k
            tot += s1.n_flips + s1.n_removals + s1.n_nucleations
        assert tot > 0
        st.close()


def test_autotune_keeps_results():
    """The autotuner times every path and tile that fits the state and keeps
    the fastest; the state is left as it was, and the run then equals the
    oracle bit for bit."""
    for o, mesh in (_generated(3000, 6, dt=1e-4), (ref2k_options(dt=2e-4), None)):
        mesh = mesh or sevc.load_files(o)
        st, orc = sevc.State(o, mesh), Oracle(o, mesh)
        before = st.get_mesh()
        tile = st.autotune()
        path, t2 = st.velocity_config()
        assert t2 == tile and path in (0, 1)
        assert diff_fields(st.get_mesh(), before) == {}
        s1 = st.step(30)
        rc, _ = orc.step(30)
        assert rc == 0 and s1.status == 0
        assert diff_fields(st.get_mesh(), orc.get_mesh()) == {}
        st.close()


@pytest.mark.parametrize("limit,lds", [(0, 0), (64, 0), (0, 16384), (64, 16384), (0, 200)])
def test_poll_general_rounds_match_lds_rounds(limit, lds):
    """The three ways of running the vote rounds and the conflict check --
    mapped neighbours in registers (lists up to `limit`), flags in LDS
    (lists up to `lds`) and the general rounds over global memory -- give
    the same flips: the same inputs stay bit-identical to the oracle with
    each way switched on or off, as the default run does."""
    o, mesh = _generated(3000, 21, dt=1e-5, flip_threshold=0.3 * np.sqrt(1.0 / 3000))
    st, orc = sevc.State(o, mesh), Oracle(o, mesh)
    st.set_poll_limit(limit, lds)
    flips = 0
    for chunk in (5, 20):
        print(f"  poll limits {limit}, {lds}: chunk of {chunk}", flush=True)
        s1 = st.step(chunk)
        rc, s2 = orc.step(chunk)
        assert rc == 0 and s1.status == 0
        assert diff_fields(st.get_mesh(), orc.get_mesh()) == {}
        flips += s1.n_flips
    assert flips > 20
    st.close()


def test_graph_replay():
    o = ref2k_options(dt=2e-4)
    mesh = sevc.load_files(o)
    tot = _run_pair(o, mesh, [30, 30, 60], graph=True)
    assert tot["flips"] > 0


def test_diagnostics_fields():
    o = ref2k_options(dt=1e-4)
    mesh = sevc.load_files(o)
    st, orc = sevc.State(o, mesh), Oracle(o, mesh)
    st.step(7)
    orc.step(7)
    st.request_diagnostics()
    orc.request_diagnostics()
    st.step(1)
    orc.step(1)
    a, b = st.get_mesh(), orc.get_mesh()
    assert diff_fields(a, b, ["v_energy", "v_vel_parts", "g_dAdt", "pos", "g_area"]) == {}


def test_analytic_velocity_model():
    o = ref2k_options(dt=1e-5, velocity_model=sevc.VEL_ANALYTIC, events=0)
    mesh = sevc.load_files(o)
    _run_pair(o, mesh, [10])


def test_batch_members_match_separate_oracles():
    o1 = ref2k_options(dt=2e-4)
    m1 = sevc.load_files(o1)
    o2, m2 = _generated(1500, 3, dt=5e-5, do_nucleation=1, nucleation_mode=sevc.NUCL_RATE,
                        nucleation_rate=0.02, seed=5, extra_capacity=0.1)
    o3, m3 = _generated(800, 4, dt=1e-4, adaptive_dt=1, flip_threshold=3e-3)
    st = sevc.State([o1, o2, o3], [m1, m2, m3])
    orcs = [Oracle(o1, m1), Oracle(o2, m2), Oracle(o3, m3)]
    for chunk in (40, 80):
        st.step(chunk)
        for k, orc in enumerate(orcs):
            rc, s2 = orc.step(chunk)
            assert rc == 0
            s1 = st.member_stats(k)
            assert diff_fields(st.get_mesh(k), orc.get_mesh()) == {}, k
            for key in STAT_KEYS:
                assert getattr(s1, key) == getattr(s2, key), (k, key)


def test_list_strides_reference_and_largest():
    """Grain-list strides of 32 (default), 100 (the reference's
    MAX_VRT_PER_GRN) and 128 (the largest supported) in one batch, with
    flips, removals and rate nucleation editing the lists: each member
    equals its own oracle run at its own stride."""
    ms, os_ = [], []
    for S, gen_seed in ((32, 21), (100, 22), (128, 23)):
        o, m = _generated(1200, gen_seed, dt=1e-4, adaptive_dt=1, dt_safety=0.5, flip_threshold=2e-3,
                          do_nucleation=1, nucleation_mode=sevc.NUCL_RATE, nucleation_rate=0.02,
                          nucleation_scaling=0.9, seed=gen_seed, extra_capacity=0.1, max_vrt_per_grn=S)
        os_.append(o)
        ms.append(m)
    assert [m.S for m in ms] == [32, 100, 128]
    st = sevc.State(os_, ms)
    orcs = [Oracle(o, m) for o, m in zip(os_, ms)]
    tot = 0
    for chunk in (30, 90):
        print(f"  chunk of {chunk} steps", flush=True)
        st.step(chunk)
        for k, orc in enumerate(orcs):
            rc, s2 = orc.step(chunk)
            assert rc == 0
            s1 = st.member_stats(k)
            assert s1.status == 0
            assert diff_fields(st.get_mesh(k), orc.get_mesh()) == {}, k
            for key in STAT_KEYS:
                assert getattr(s1, key) == getattr(s2, key), (k, key)
            tot += s1.n_flips + s1.n_removals + s1.n_nucleations
    assert tot >= 100, tot
    st.close()


def _half_removal_2k(seed, dt=2e-4):
    """2,000-grain fixture with the model options of the small-scale runs:
    gamma = mu = 1, SE = U[0,1) * 0.5, rate nucleation 1e-3 * SE per grain
    and step with Philox seed `seed`, adaptive dt (upper bound dt), flip
    threshold 1e-4, stop at 50 % of the grains."""
    return ref2k_options(dt=dt, adaptive_dt=1, dt_safety=0.5, flip_threshold=1e-4, do_nucleation=1,
                         nucleation_mode=sevc.NUCL_RATE, nucleation_rate=1e-3, nucleation_scaling=0.9,
                         seed=seed, extra_capacity=0.05, leave_percent=0.5)


def test_seed_ensemble_curves_match_oracle():
    """32 Philox seeds of the 2,000-grain workload in one batched state: every
    member's state, grain count and mean grain area equal its own oracle
    run after each chunk (so the ensemble curves agree exactly)."""
    seeds = list(range(32))
    opts = [_half_removal_2k(s) for s in seeds]
    meshes = [sevc.load_files(o) for o in opts]
    st = sevc.State(opts, meshes)
    orcs = [Oracle(o, m) for o, m in zip(opts, meshes)]
    curves = {"gpu": [], "oracle": []}
    events = 0
    for chunk in (40, 40, 40):
        print(f"  chunk of {chunk} steps x {len(seeds)} members", flush=True)
        st.step(chunk)
        for k, orc in enumerate(orcs):
            rc, s2 = orc.step(chunk)
            assert rc == 0
            s1 = st.member_stats(k)
            a, b = st.get_mesh(k), orc.get_mesh()
            assert diff_fields(a, b) == {}, seeds[k]
            for key in STAT_KEYS:
                assert getattr(s1, key) == getattr(s2, key), (seeds[k], key)
            events += s1.n_nucleations + s1.n_flips + s1.n_removals
            for name, h, s in (("gpu", a, s1), ("oracle", b, s2)):
                en = h["g_enabled"] == 1
                curves[name].append((s.n_grains, float(np.abs(h["g_area"][en]).mean())))
    assert curves["gpu"] == curves["oracle"]
    assert events > 0
    st.close()


def test_autotuned_tiles_keep_results():
    """sevc_autotune changes only the tiling of the velocity's grain pass:
    the state it leaves and every later step equal the oracle's."""
    o, mesh = _generated(2500, 8, dt=1e-4, adaptive_dt=1, flip_threshold=2e-3, do_nucleation=1,
                         nucleation_mode=sevc.NUCL_RATE, nucleation_rate=0.01, seed=3, extra_capacity=0.1)
    st, orc = sevc.State(o, mesh), Oracle(o, mesh)
    before = st.get_mesh()
    tile = st.autotune()
    assert 16 <= tile <= 64
    assert diff_fields(st.get_mesh(), before) == {}
    for chunk in (30, 60):
        st.step(chunk)
        rc, _ = orc.step(chunk)
        assert rc == 0
        assert diff_fields(st.get_mesh(), orc.get_mesh()) == {}
    st.close()


def test_anisotropic_boundary_energy():
    """GB_eps != 0 (geometry.h:548-564): per-edge gamma from the grains'
    orientations in the velocity, with T1/T2 events."""
    o = ref2k_options(dt=2e-4, GB_eps=0.3)
    mesh = sevc.load_files(o)
    tot = _run_pair(o, mesh, [40, 80])
    assert tot["flips"] > 0, tot


@pytest.mark.parametrize("n,gen_seed", [(12, 3), (16, 11), (20, 5), (30, 13), (40, 7)])
def test_tiny_meshes(n, gen_seed):
    """Periodic tessellations of a dozen to a few dozen grains with a large
    dt: T1 and T2 events crowd and must still follow the oracle bit for bit
    until the stop rule at half the grains (below that these meshes
    degenerate, which the oracle reports as an inconsistent topology)."""
    o, mesh = _generated(n, gen_seed, dt=1e-2, leave_percent=0.5)
    st, orc = sevc.State(o, mesh), Oracle(o, mesh)
    for chunk in (5, 20, 100):
        print(f"  {n} grains: chunk of {chunk} steps", flush=True)
        s1 = st.step(chunk)
        rc, s2 = orc.step(chunk)
        assert rc == 0 and s1.status == 0
        assert diff_fields(st.get_mesh(), orc.get_mesh()) == {}
        for k in STAT_KEYS:
            assert getattr(s1, k) == getattr(s2, k), k
    st.close()


def test_member_stopping_early_leaves_others_running():
    """Per-member stop rules in a batch: a member that reaches its grain rule
    stops while the other keeps stepping; both equal their own oracle run."""
    o1 = ref2k_options(dt=2e-4, leave_percent=0.9)
    m1 = sevc.load_files(o1)
    o2, m2 = _generated(1500, 3, dt=1e-4)
    st = sevc.State([o1, o2], [m1, m2])
    orcs = [Oracle(o1, m1), Oracle(o2, m2)]
    for chunk in (100, 200):
        st.step(chunk)
        for k, orc in enumerate(orcs):
            rc, s2 = orc.step(chunk)
            assert rc == 0
            s1 = st.member_stats(k)
            assert diff_fields(st.get_mesh(k), orc.get_mesh()) == {}, k
            for key in STAT_KEYS:
                assert getattr(s1, key) == getattr(s2, key), (k, key)
    assert st.member_stats(0).done and not st.member_stats(1).done
    st.close()


def test_grain_list_overflow_is_reported():
    """A grain list capacity equal to the initial largest grain: the first
    grain to grow past it stops the run with SEVC_E_GRAIN_FULL, in the same
    chunk of steps as the oracle."""
    o = ref2k_options(dt=2e-4)
    o.max_vrt_per_grn = int(sevc.load_files(o)["g_vlen"].max())
    mesh = sevc.load_files(o)
    st, orc = sevc.State(o, mesh), Oracle(o, mesh)
    gpu_fail = orc_fail = None
    for chunk in range(40):
        if gpu_fail is None:
            try:
                st.step(10)
            except sevc.SevcError as e:
                gpu_fail = (chunk, e.code)
        if orc_fail is None:
            rc, _ = orc.step(10)
            if rc != 0:
                orc_fail = (chunk, rc)
        if gpu_fail and orc_fail:
            break
    assert gpu_fail is not None and gpu_fail == orc_fail, (gpu_fail, orc_fail)
    assert gpu_fail[1] == sevc.E_GRAIN_FULL
    st.close()


def test_member_error_stops_only_that_member():
    """In a batch, a member whose grain list overflows (capacity equal to its
    initial largest grain) stops with SEVC_E_GRAIN_FULL in the same chunk of
    steps as its own oracle, and the call that meets it names the member;
    the other members keep stepping and stay bit-identical to their own
    oracle runs, before and after."""
    o0 = ref2k_options(dt=2e-4)
    o0.max_vrt_per_grn = int(sevc.load_files(o0)["g_vlen"].max())
    m0 = sevc.load_files(o0)
    o1, m1 = _generated(1500, 3, dt=1e-4)
    o2 = ref2k_options(dt=1e-4)
    m2 = sevc.load_files(o2)
    st = sevc.State([o1, o0, o2], [m1, m0, m2])
    orcs = [Oracle(o1, m1), Oracle(o0, m0), Oracle(o2, m2)]
    gpu_fail = orc_fail = None
    for chunk in range(40):
        print(f"  chunk {chunk} of 10 steps", flush=True)
        try:
            st.step(10)
        except sevc.SevcError as e:
            assert gpu_fail is None
            gpu_fail = (chunk, e.code)
            assert "member 1" in str(e)
        for k, orc in enumerate(orcs):
            if k == 1 and orc_fail is not None:
                continue
            rc, _ = orc.step(10)
            if k == 1 and rc != 0:
                orc_fail = (chunk, rc)
                continue
            assert rc == 0, k
            if k != 1:
                assert diff_fields(st.get_mesh(k), orc.get_mesh()) == {}, (chunk, k)
                assert st.member_stats(k).status == 0
        if gpu_fail and chunk >= gpu_fail[0] + 3:
            break
    assert gpu_fail is not None and gpu_fail == orc_fail, (gpu_fail, orc_fail)
    assert gpu_fail[1] == sevc.E_GRAIN_FULL
    s = st.member_stats(1)
    assert s.status == sevc.E_GRAIN_FULL and s.done
    st.close()


@pytest.mark.parametrize("n,gen_seed", [(12, 3), (30, 13)])
def test_degenerate_mesh_error_matches_oracle(n, gen_seed):
    """Run a tiny tessellation down to a handful of grains, where a flip
    meets an edge whose ends no longer share two grains
    (topological.h:174-291 has no case for it): both sides match bit for bit
    until then and report SEVC_E_STATE in the same chunk of steps.  (The
    oracle, built with AddressSanitizer, stays in bounds on these runs.)"""
    o, mesh = _generated(n, gen_seed, dt=2e-2, leave_percent=0.1)
    st, orc = sevc.State(o, mesh), Oracle(o, mesh)
    gpu_fail = orc_fail = None
    for chunk in range(30):
        print(f"  chunk {chunk} of 10 steps", flush=True)
        try:
            st.step(10)
        except sevc.SevcError as e:
            gpu_fail = (chunk, e.code)
        rc, _ = orc.step(10)
        if rc != 0:
            orc_fail = (chunk, rc)
        if gpu_fail or orc_fail:
            break
        assert diff_fields(st.get_mesh(), orc.get_mesh()) == {}, chunk
    assert gpu_fail == orc_fail and gpu_fail[1] == sevc.E_STATE, (gpu_fail, orc_fail)
    st.close()


@pytest.mark.parametrize("n,factor", [(8000, 1.0), (5000, 1.5)])
def test_long_candidate_lists(n, factor):
    """A flip threshold of the order of the mean edge length makes every
    short shrinking edge a candidate: about 2,000 listed candidates per step
    for (8000, 1.0), the size where polling switches from the in-LDS lists
    (at most 2,048) to the general rounds, and 3,000-4,000 for (5000, 1.5).
    Flips, removals and the whole state match the oracle bit for bit."""
    o, mesh = _generated(n, 21, dt=1e-5, flip_threshold=factor * 0.6 * np.sqrt(1.0 / n))
    st, orc = sevc.State(o, mesh), Oracle(o, mesh)
    for c in (1, 1, 3, 5):
        print(f"  chunk of {c} steps", flush=True)
        s1 = st.step(c)
        rc, s2 = orc.step(c)
        assert rc == 0 and s1.status == 0, (rc, s1.status)
        a, b = st.get_mesh(), orc.get_mesh()
        assert diff_fields(a, b) == {}
        for k in STAT_KEYS:
            assert getattr(s1, k) == getattr(s2, k), (k, getattr(s1, k), getattr(s2, k))
        if c == 1 and factor > 1.0:
            # the candidate rule on the state after one step, as an estimate of
            # the list length polling sees: well above 2,048
            bt, bl, en = b["b_t_ext"], b["b_arclength"], b["b_enabled"]
            nc = np.sum((((bt >= 0) & (bt <= o.dt)) | ((bt >= 0) & (bl > 0) & (bl < o.flip_threshold))) & (en == 1))
            assert nc > 2048, nc
    st.close()


@pytest.mark.parametrize("rate", [0.05, 0.4])
def test_nucleation_bursts(rate):
    """Hundreds to thousands of rate-mode nucleations per step on 20,000
    grains: about 270 kept events per step at rate 0.05 (more free slots
    wanted than the in-LDS slot list holds), about 2,400 candidates per step
    at rate 0.4 (beyond the in-LDS candidate table), where the free grain
    slots run out in the second step and the remaining events are skipped.
    State and event counts match the oracle bit for bit."""
    o, mesh = _generated(20000, 31, dt=1e-6, adaptive_dt=1, dt_safety=0.5, flip_threshold=1e-4, do_nucleation=1,
                         nucleation_mode=sevc.NUCL_RATE, nucleation_rate=rate, nucleation_scaling=0.9, seed=5,
                         extra_capacity=0.15)
    tot = _run_pair(o, mesh, [1, 1, 2])
    assert tot["nucl"] >= (1000 if rate < 0.1 else 3000), tot


@pytest.mark.parametrize("blocks", [1, 3])
def test_free_slot_search_in_rounds(blocks):
    """The nucleation free-slot search read in rounds of `blocks` table
    blocks of each kind (instead of one round of up to 4096, 12288 and 8192
    blocks), so that it walks several rounds to find its ranks: rate mode
    with tens of events per step, more than the in-LDS slot list holds, and
    the Metropolis mode, all bit-equal to the oracle."""
    def rounds(st):
        st.set_search_round(blocks)
    o, mesh = _generated(3000, 5, dt=1e-4, adaptive_dt=1, dt_safety=0.5, flip_threshold=2e-3,
                         do_nucleation=1, nucleation_mode=sevc.NUCL_RATE, nucleation_rate=0.02,
                         nucleation_scaling=0.9, seed=11, extra_capacity=0.1)
    tot = _run_pair(o, mesh, [1, 1, 20, 40], setup=rounds)
    assert tot["nucl"] >= 10 and tot["removals"] >= 5, tot
    o, mesh = _generated(20000, 31, dt=1e-6, adaptive_dt=1, dt_safety=0.5, flip_threshold=1e-4, do_nucleation=1,
                         nucleation_mode=sevc.NUCL_RATE, nucleation_rate=0.05, nucleation_scaling=0.9, seed=5,
                         extra_capacity=0.15)
    tot = _run_pair(o, mesh, [1, 1], setup=rounds)
    assert tot["nucl"] >= 400, tot
    o = ref2k_options(domain_bound=40.0, dt=1e-3, dtau=1e-3, do_nucleation=1, leave_percent=0.05,
                      SE_eps=1.0, MC_k=1.0)
    o.SE_min, o.SE_max = 3.0, 6.0
    tot = _run_pair(o, sevc.load_files(o), [600, 600], setup=rounds)
    assert tot["nucl"] >= 10, tot


def test_boundary_energy_scale_with_flips():
    """GB_scaling != 1 with GB_eps = 0: a flipped edge's energy is left
    without the scale until the next step recomputes it (topological.h:286),
    and nucleation reads those energies.  Checked after every step so the
    exported energies of each step's flipped edges are compared too."""
    o, mesh = _generated(3000, 17, dt=1e-4, adaptive_dt=1, dt_safety=0.5, flip_threshold=2e-3, GB_scaling=1.7,
                         do_nucleation=1, nucleation_mode=sevc.NUCL_RATE, nucleation_rate=0.02,
                         nucleation_scaling=0.9, seed=3, extra_capacity=0.1)
    tot = _run_pair(o, mesh, [1] * 6 + [40, 60])
    assert tot["flips"] >= 50 and tot["nucl"] >= 10, tot


def test_batch_scaled_member_beside_anisotropic_member():
    """A batch of a GB_scaling = 1.7 member (GB_eps = 0) and an anisotropic
    member (GB_eps = 0.3): every step then recomputes all boundary energies
    at its start, which replaces the rescaling of the flipped edges alone;
    each member stays equal to its own oracle after every chunk."""
    o1, m1 = _generated(3000, 17, dt=1e-4, adaptive_dt=1, dt_safety=0.5, flip_threshold=2e-3, GB_scaling=1.7,
                        do_nucleation=1, nucleation_mode=sevc.NUCL_RATE, nucleation_rate=0.02,
                        nucleation_scaling=0.9, seed=3, extra_capacity=0.1)
    o2 = ref2k_options(dt=2e-4, GB_eps=0.3)
    m2 = sevc.load_files(o2)
    st = sevc.State([o1, o2], [m1, m2])
    orcs = [Oracle(o1, m1), Oracle(o2, m2)]
    flips = 0
    for chunk in (1, 1, 1, 20, 40):
        print(f"  chunk of {chunk}", flush=True)
        s1 = st.step(chunk)
        assert s1.status == 0
        for k, orc in enumerate(orcs):
            rc, s2 = orc.step(chunk)
            assert rc == 0
            assert diff_fields(st.get_mesh(k), orc.get_mesh()) == {}, k
            flips += s2.n_flips
    assert flips >= 20
    st.close()


def test_two_triangle_collapse():
    """Flip threshold twice the mean edge length: flips leave pairs of
    adjacent 3-sided grains, which the two-triangle stage collapses (about
    20 in 60 steps here; T2 labels are then taken again after the stage).
    Compared after every chunk."""
    n = 8000
    o, mesh = _generated(n, 21, dt=1e-5, flip_threshold=2.0 * 0.6 * np.sqrt(1.0 / n))
    tot = _run_pair(o, mesh, [1] * 10 + [10] * 5)
    assert tot["two"] >= 5, tot


def test_batch_mc_member_after_rate_member():
    """Metropolis nucleation in a member whose slots start past another
    member's (free-slot search and events at non-zero offsets), next to a
    rate-mode member; each matches its own oracle."""
    o1, m1 = _generated(1500, 3, dt=5e-5, do_nucleation=1, nucleation_mode=sevc.NUCL_RATE,
                        nucleation_rate=0.02, seed=5, extra_capacity=0.1)
    o2 = ref2k_options(domain_bound=40.0, dt=1e-3, dtau=1e-3, do_nucleation=1, leave_percent=0.05,
                       SE_eps=1.0, MC_k=1.0)
    o2.SE_min, o2.SE_max = 3.0, 6.0
    m2 = sevc.load_files(o2)
    st = sevc.State([o1, o2], [m1, m2])
    orcs = [Oracle(o1, m1), Oracle(o2, m2)]
    nucl = 0
    for chunk in (400, 400, 400):
        print(f"  chunk of {chunk} steps", flush=True)
        st.step(chunk)
        for k, orc in enumerate(orcs):
            rc, s2 = orc.step(chunk)
            assert rc == 0
            s1 = st.member_stats(k)
            assert diff_fields(st.get_mesh(k), orc.get_mesh()) == {}, k
            for key in STAT_KEYS:
                assert getattr(s1, key) == getattr(s2, key), (k, key)
            if k == 1:
                nucl += s1.n_nucleations
    st.close()
    assert nucl >= 20, nucl


def test_full_size_two_steps():
    """The benchmark workload itself (periodic Voronoi of 10^6 seeds,
    lognormal SE, rate nucleation, adaptive dt, flip threshold 1e-4): two
    steps bit for bit against the oracle, with flips, removals and
    nucleations in both.  In these first steps the free slots lie only at the
    ends of the arrays, so the free-slot search runs over several rounds of
    blocks (3M edge slots)."""
    import os
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    o, mesh = bench.make_mesh(sevc, 1000000, 1)
    tot = _run_pair(o, mesh, [1, 1])
    assert tot["flips"] > 50 and tot["removals"] > 10 and tot["nucl"] > 20, tot
