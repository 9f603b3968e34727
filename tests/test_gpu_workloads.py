"""The workloads at their real shapes on the GPU, against the oracle.

* Small-scale runs: 32 Philox seeds of the 2,000-grain workload (rate
  nucleation, adaptive dt, flip threshold) in one batched state, each run to
  its stop at 50 % of the grains (about 1,000 steps): after every chunk of
  100 steps each member's whole state equals its own oracle run bit for bit
  (sha256 of every field), so its grain-count, mean-area and area-histogram
  curves are the oracle's exactly.
* Batched ensemble: the benchmark's 256 members (1,000 to 50,000 grains,
  3.2M grains in all) stepped together; after 3 steps every member equals
  its own oracle run.
* Adjoint: a 10,000-grain window of 500 steps with a checkpoint every 50;
  forward positions and the reverse pass (dL/dx_0, dL/dSE_g, dL/dgamma,
  dL/dmu) equal the oracle's bit for bit.

The oracle runs of the ensembles are spread over worker processes
(_oracle_pool)."""
import os
import sys
import time

import numpy as np
import pytest

import sevc
from _oracle_pool import STAT_KEYS, half_removal_options, half_removal_runs, ensemble_runs, record
from _util import Oracle, REPO, diff_fields

pytestmark = pytest.mark.gpu

sys.path.insert(0, REPO)


def test_32_seeds_of_2k_grains_to_half_the_grains():
    seeds = list(range(32))
    chunk, max_chunks = 100, 30
    t0 = time.time()
    ref = half_removal_runs(seeds, chunk, max_chunks)
    print(f"  oracle runs: {time.time() - t0:.1f} s, chunks per member {[len(r) for r in ref]}", flush=True)
    opts = [half_removal_options(s) for s in seeds]
    st = sevc.State(opts, [sevc.load_files(o) for o in opts])
    done = [False] * len(seeds)
    for c in range(max_chunks):
        if all(done):
            break
        st.step(chunk)
        for k in range(len(seeds)):
            if done[k]:
                continue
            s = st.member_stats(k)
            got = record(s, st.get_mesh(k))
            want = ref[k][c]
            assert s.status == 0 and want["rc"] == 0, (seeds[k], c)
            assert got["stats"] == want["stats"], (seeds[k], c, got["stats"], want["stats"])
            assert got["curves"] == want["curves"], (seeds[k], c)
            assert got["digest"] == want["digest"], (seeds[k], c)
            if s.done:
                done[k] = True
                assert c == len(ref[k]) - 1
                assert s.n_grains <= 1000 < ref[k][0]["curves"][0] + 1000
        print(f"  chunk {c}: {sum(done)} of {len(seeds)} members at the stop", flush=True)
    assert all(done), "every member reaches the 50 % stop"
    st.close()


def test_ensemble_256_members_three_steps():
    import bench
    t0 = time.time()
    ref = ensemble_runs(256, 3)
    print(f"  oracle runs: {time.time() - t0:.1f} s", flush=True)
    opts, meshes = bench.make_ensemble(sevc, 256)
    st = sevc.State(opts, meshes)
    st.step(3)
    events = 0
    for k in range(256):
        s = st.member_stats(k)
        got = record(s, st.get_mesh(k))
        assert s.status == 0 and ref[k]["rc"] == 0, k
        assert got["stats"] == ref[k]["stats"], (k, got["stats"], ref[k]["stats"])
        assert got["digest"] == ref[k]["digest"], k
        events += s.n_flips + s.n_removals + s.n_nucleations
    assert events > 0
    st.close()


def test_adjoint_window_10k_grains_500_steps():
    n, steps, ckpt = 10000, 500, 50
    pos, rid = sevc.voronoi(n, 3)
    rng = np.random.default_rng(3)
    o = bench_window_options(n)
    mesh = sevc.build_mesh(o, pos, rid, rng.random(n), 0.3 * rng.random(n), n)
    st, orc = sevc.State(o, mesh), Oracle(o, mesh)
    st.window_forward(steps, ckpt)
    assert orc.window_forward(steps) == 0
    assert diff_fields(st.get_mesh(), orc.get_mesh(), ["pos", "vel", "g_area"]) == {}
    w = np.random.default_rng(5).standard_normal(2 * mesh.nv) * np.repeat(mesh["v_enabled"], 2)
    (dp, ds, dgb, dmu), = st.adjoint(w)
    op, os_, ogb, omu = orc.adjoint(w)
    same = lambda a, b: np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))
    assert same(dp, op) and same(ds, os_) and dgb == ogb and dmu == omu
    st.close()
    orc.close()


def bench_window_options(n):
    import bench
    return bench.window_options(sevc, n, 1e-6)


def test_full_size_anisotropic_two_steps():
    """The 1M-grain workload with anisotropic boundary energy (GB_eps = 0.3,
    geometry.h:548-564: gamma from the two grains' orientations, recomputed
    every step): two steps bit for bit against the oracle."""
    import bench
    from test_gpu_parity import _run_pair
    o, mesh = bench.make_mesh(sevc, 1000000, 1)
    o.GB_eps = 0.3
    tot = _run_pair(o, mesh, [1, 1])
    assert tot["flips"] > 50 and tot["removals"] > 5, tot


def test_full_size_mc_nucleation():
    """The reference's Monte Carlo nucleation (nucleation2.cu:239-443) at
    1M grains: one Metropolis site per step from the first step, edges O(1)
    long in a domain of side 1000, SE in [3, 6], GB_eps = 0.3; three steps
    bit for bit against the oracle, with nucleations."""
    import bench
    from test_gpu_parity import _run_pair
    o, mesh = bench.make_mc_mesh(sevc, 1000000, 1)
    tot = _run_pair(o, mesh, [1, 2])
    assert tot["nucl"] >= 1 and tot["flips"] > 50, tot
