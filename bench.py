"""Benchmark of the MI355X vertex-model time step (FP64).

Workload (one GPU): periodic Voronoi tessellation of 1,000,000 random seeds
(about 2M vertices, 3M boundaries), boundary energy 1, mobility 1, per-grain
stored energy lognormal (mean 0.3, sigma 0.5), nucleation rate 1e-4 per
grain per step, adaptive dt from the minimum extinction time, flip
threshold 1e-4 edge length.  A step is one full time step of the model
(velocity, extinction times, votes, T1/T2, nucleation, translation) over all
grains; the value is enabled vertex-steps per second over the timed steps,
with the state resident in HBM.

Prints one JSON line.  Run `python bench.py --help` for options."""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "se-grain-growth-gpu_amd"))

METRIC = "FP64 vertex-steps/s at 1M grains & 256-run batch vs plain; adjoint/fwd ratio"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per enabled vertex and step (FP64 values, int32 ids,
# each datum once): velocity 98 B, translation 48 B.
B_VELOCITY = 98.0
B_TRANSLATE = 48.0
# HBM bytes per enabled vertex of the two velocity passes at the 1M-grain
# workload, from rocprofv3 FETCH_SIZE/WRITE_SIZE scaled by a copy of known
# size (profiles/step_1M_hbm_traffic.md): 515.0 MB for 2.0M vertices.
TRAFFIC_VELOCITY_B_PER_VERTEX = 258.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_options(sevc, n_grains):
    o = sevc.default_options()
    o.dt = 1e-6                 # upper bound of the adaptive step
    o.adaptive_dt = 1
    o.dt_safety = 0.5
    o.flip_threshold = 1e-4
    o.GB_eps = 0.0
    o.GB_scaling = 1.0
    o.mobility = 1.0
    o.do_nucleation = 1
    o.nucleation_mode = sevc.NUCL_RATE
    o.nucleation_rate = 1e-4         # p_g = 1e-4 * SE_g per grain and step
    o.nucleation_scaling = 0.9
    o.seed = 1
    o.max_vrt_per_grn = 32
    o.extra_capacity = 0.02
    o.n_grains = n_grains
    o.leave_percent = -1.0
    return o


def make_mesh(sevc, n_grains, seed):
    pos, rid = sevc.voronoi(n_grains, seed)
    rng = np.random.default_rng(seed)
    sigma = 0.5
    mu = np.log(0.3) - 0.5 * sigma * sigma
    se = np.exp(mu + sigma * rng.standard_normal(n_grains))
    ori = rng.random(n_grains)
    o = make_options(sevc, n_grains)
    return o, sevc.build_mesh(o, pos, rid, ori, se, n_grains)


def make_ensemble(sevc, members, seed0=1000, g_lo=1000, g_hi=50000):
    """Ensemble workload: member i has G_i = round(exp(ln g_lo + U_i ln(g_hi/g_lo)))
    grains (log-uniform), Voronoi seed seed0+i, SE = s_i * U[0,1) with
    s_i ~ U[0,1], and the model options of the single-domain workload."""
    opts, meshes = [], []
    for i in range(members):
        rng = np.random.default_rng(seed0 + i)
        g = int(round(np.exp(np.log(g_lo) + rng.random() * np.log(g_hi / g_lo))))
        s_i = rng.random()
        pos, rid = sevc.voronoi(g, seed0 + i)
        o = make_options(sevc, g)
        o.seed = seed0 + i
        meshes.append(sevc.build_mesh(o, pos, rid, rng.random(g), s_i * rng.random(g), g))
        opts.append(o)
    return opts, meshes


def bench_ensemble(sevc, members, steps, b2b_steps, warmup=5):
    """All members stepped together in one state (one launch sequence per
    step for the whole ensemble) against the same members stepped one state
    at a time, back to back."""
    t0 = time.time()
    opts, meshes = make_ensemble(sevc, members)
    grains = sum(o.n_grains for o in opts)
    log(f"ensemble: {members} members, {grains} grains ({time.time() - t0:.1f} s)")
    st = sevc.State(opts, meshes)
    tile = st.autotune()
    st.step(warmup)
    st.step(0)
    t1 = time.perf_counter()
    s = st.step(steps)
    t2 = time.perf_counter()
    per = [st.member_stats(k) for k in range(members)]     # a batched call reports member 0 only
    bad = [k for k, p in enumerate(per) if p.status != 0]
    if bad:
        raise SystemExit(f"ensemble members {bad[:8]} failed (status {per[bad[0]].status})")
    batched = sum(p.n_vertex_steps for p in per) / (t2 - t1)
    # the same batch replayed as a captured HIP graph of one step
    gsteps = max(1, steps // 5)
    st.step(1, graph=True)
    g1 = time.perf_counter()
    sg = st.step(gsteps, graph=True)
    g2 = time.perf_counter()
    perg = [st.member_stats(k) for k in range(members)]
    if any(p.status != 0 for p in perg):
        raise SystemExit("ensemble graph replay: a member failed")
    graph_rate = sum(p.n_vertex_steps for p in perg) / (g2 - g1)
    st.close()
    vs, el = 0, 0.0
    for o, m in zip(opts, meshes):
        one = sevc.State(o, m)
        one.step(warmup)
        one.step(0)
        a = time.perf_counter()
        r = one.step(b2b_steps)
        el += time.perf_counter() - a
        vs += r.n_vertex_steps
        one.close()
    b2b = vs / el
    log(f"ensemble: batched {batched:.3e} vertex-steps/s over {steps} steps (graph replay {graph_rate:.3e}); "
        f"back-to-back {b2b:.3e} over {b2b_steps} steps per member")
    return {"value": batched, "unit": "vertex-steps/s", "members": members, "grains": grains, "steps": steps,
            "tile_grains": tile,
            "ms_per_step": (t2 - t1) / max(1, s.steps_done) * 1e3,
            "graph_replay": {"value": graph_rate, "steps": gsteps,
                             "ms_per_step": (g2 - g1) / max(1, sg.steps_done) * 1e3},
            "back_to_back": {"value": b2b, "steps_per_member": b2b_steps, "seconds": el},
            "speedup": batched / b2b,
            "workload": f"{members} members, G log-uniform in [1000, 50000] (Voronoi seed 1000+i), "
                        "SE = s_i*U[0,1) with s_i ~ U[0,1]; model options of the single-domain workload"}


def window_options(sevc, n_grains, dt):
    o = sevc.default_options()
    o.GB_eps = 0.0
    o.GB_scaling = 1.0
    o.mobility = 1.0
    o.dt = dt
    o.n_grains = n_grains
    o.velocity_model = sevc.VEL_ANALYTIC
    o.events = 0
    o.do_nucleation = 0
    o.leave_percent = -1.0
    return o


def bench_adjoint(sevc, grains, seed, steps, ckpt, dt, n_fd=20, reps=3):
    """Adjoint over a topology-fixed window: a smooth 20-bin histogram of
    grain areas after `steps` steps, against the histogram the same window
    gives with (1.1 gamma, 0.9 mu).  Times the window forward (with its
    checkpoints) and the reverse pass (with the segment recomputation), and
    checks the reverse pass against differences of the forward."""
    from sevc.calib import AreaHistogram, grain_areas
    pos, rid = sevc.voronoi(grains, seed)
    rng = np.random.default_rng(seed)
    o = window_options(sevc, grains, dt)
    mesh = sevc.build_mesh(o, pos, rid, rng.random(grains), 0.3 * rng.random(grains), grains)
    am = float(np.abs(grain_areas(mesh)).mean())
    hist = AreaHistogram(20, 0.0, 3.0 * am, 0.1 * am)
    st = sevc.State(o, mesh)
    st.set_params(0, gb_scaling=1.1, mobility=0.9)
    st.window_forward(steps, ckpt)
    target = hist(mesh, st.positions(0))
    x0 = mesh["pos"].copy()

    def forward(pos=x0, se=None, gb=1.0, mu=1.0):
        st.set_params(0, pos=pos, g_SE=mesh["g_SE"] if se is None else se, gb_scaling=gb, mobility=mu)
        st.window_forward(steps, ckpt)
        return hist.loss_and_grad(mesh, st.positions(0), target)

    L, gN = forward()
    t_f, t_a = [], []
    for _ in range(reps):
        st.set_params(0, pos=x0, g_SE=mesh["g_SE"], gb_scaling=1.0, mobility=1.0)
        a = time.perf_counter()
        st.window_forward(steps, ckpt)
        b = time.perf_counter()
        (dp, dse, dgb, dmu), = st.adjoint(gN)
        c = time.perf_counter()
        t_f.append(b - a)
        t_a.append(c - b)
    # accuracy: a seeded linear functional w . x_N, whose reverse pass is
    # compared with Richardson central differences of the window forward on
    # n_fd - 2 seeded position coordinates (h = 1e-6), on GB_scaling and
    # mobility (h = 1e-3), and along one seeded direction in SE (h = 3e-3;
    # a single SE_g moves x_N by less than the rounding of a 20k-term sum).
    # The rounding of w . x_N over a 500-step window is ~2e-13, so a
    # difference quotient carries ~3e-13 / h of noise: parameter steps are
    # taken large enough to keep that below 1e-6 of the derivative, which
    # the O(h^4) Richardson truncation allows.
    pick = np.random.default_rng(seed + 1)
    w = pick.standard_normal(2 * mesh.nv) * np.repeat(mesh["v_enabled"], 2)
    u = pick.standard_normal(mesh.ng) * mesh["g_enabled"]
    st.set_params(0, pos=x0, g_SE=mesh["g_SE"], gb_scaling=1.0, mobility=1.0)
    st.window_forward(steps, ckpt)
    xN = st.positions(0).reshape(-1).copy()
    (lp, ls, lgb, lmu), = st.adjoint(w)
    D = mesh.D

    def lin(p, s, gbv, muv):
        st.set_params(0, pos=p, g_SE=s, gb_scaling=gbv, mobility=muv)
        st.window_forward(steps, ckpt)
        d = st.positions(0).reshape(-1) - xN
        return float(np.dot(w, d - D * np.round(d / D)))

    checks = [("pos", int(c), 1e-6) for c in pick.choice(np.nonzero(np.repeat(mesh["v_enabled"], 2))[0],
                                                          n_fd - 2, replace=False)]
    checks += [("gb", 0, 1e-3), ("mu", 0, 1e-3), ("SE-dir", 0, 3e-3)]
    worst, detail = 0.0, []
    for kind, i, h in checks:
        def f(e):
            p, s, gbv, muv = x0.copy(), mesh["g_SE"].copy(), 1.0, 1.0
            if kind == "pos":
                p[i] += e
            elif kind == "SE-dir":
                s = s + e * u
            elif kind == "gb":
                gbv += e
            else:
                muv += e
            return lin(p, s, gbv, muv)
        c1 = (f(h) - f(-h)) / (2 * h)
        c2 = (f(h / 2) - f(-h / 2)) / h
        fd = (4 * c2 - c1) / 3
        ad = {"pos": lambda: lp[i], "SE-dir": lambda: float(np.dot(u, ls)), "gb": lambda: lgb,
              "mu": lambda: lmu}[kind]()
        rel = abs(ad - fd) / max(abs(fd), 1e-300)
        worst = max(worst, rel)
        detail.append([kind, i, ad, fd, rel])
    coords = checks
    st.close()
    tf, ta = min(t_f), min(t_a)
    log(f"adjoint: {grains} grains, {steps} steps (checkpoint every {ckpt}): forward {tf * 1e3:.1f} ms, "
        f"reverse {ta * 1e3:.1f} ms, ratio {ta / tf:.2f}; worst FD rel. error {worst:.2e}")
    return {"ratio": ta / tf, "forward_ms": tf * 1e3, "adjoint_ms": ta * 1e3, "grains": grains, "steps": steps,
            "ckpt_every": ckpt, "dt": dt, "loss": L, "fd_max_rel_err": worst, "fd_coords": len(coords),
            "fd_checks": detail,
            "workload": f"periodic Voronoi of {grains} seeds (seed {seed}), SE = 0.3*U[0,1), 20-bin smooth "
                        "area histogram vs the window's histogram at (1.1 gamma, 0.9 mu)"}


def bench_calibration(sevc, members, iters, steps, ckpt, dt, seed0=2000):
    """Batched calibration: members with G ~ U[5000, 20000] grains and
    lognormal SE; each member's target is its own 20-bin area histogram
    after the window at (gamma, mu) = (1.1, 0.9).  Starting from (1, 1),
    Adam on (log gamma, log mu) with the summed loss; every iteration is
    one batched window forward, the device histogram loss (its gradient
    seeds the reverse pass on the device) and one batched reverse pass."""
    from sevc.calib import grain_areas
    opts, meshes, ranges = [], [], []
    t0 = time.time()
    for i in range(members):
        rng = np.random.default_rng(seed0 + i)
        g = int(rng.integers(5000, 20001))
        pos, rid = sevc.voronoi(g, seed0 + i)
        o = window_options(sevc, g, dt)
        se = np.exp(np.log(0.3) - 0.125 + 0.5 * rng.standard_normal(g))
        m = sevc.build_mesh(o, pos, rid, rng.random(g), se, g)
        am = float(np.abs(grain_areas(m)).mean())
        opts.append(o)
        meshes.append(m)
        ranges.append((0.0, 3.0 * am, 0.1 * am))
    grains = sum(o.n_grains for o in opts)
    log(f"calibration: {members} members, {grains} grains ({time.time() - t0:.1f} s)")
    st = sevc.State(opts, meshes)

    def set_all(gb, mu):
        for k, m in enumerate(meshes):
            st.set_params(k, pos=m["pos"], gb_scaling=gb, mobility=mu)

    set_all(1.1, 0.9)
    st.window_forward(steps, ckpt)
    targets = np.zeros((members, 20))
    targets, _ = st.area_histogram(20, ranges, targets)
    theta = np.zeros(2)                        # log gamma, log mu
    m1, m2 = np.zeros(2), np.zeros(2)
    lr, b1, b2 = 0.02, 0.9, 0.999
    trace = []
    t1 = time.perf_counter()
    t_dev = 0.0
    for it in range(iters):
        gb, mu = float(np.exp(theta[0])), float(np.exp(theta[1]))
        set_all(gb, mu)
        a = time.perf_counter()
        st.window_forward(steps, ckpt)
        _, losses = st.area_histogram(20, ranges, targets, seed=True)    # dL/dx stays on the device
        L = float(losses.sum())
        res = st.adjoint(None)
        t_dev += time.perf_counter() - a
        g = np.array([sum(r[2] for r in res) * gb, sum(r[3] for r in res) * mu])
        trace.append((L, gb, mu))
        m1 = b1 * m1 + (1 - b1) * g
        m2 = b2 * m2 + (1 - b2) * g * g
        theta -= lr * (m1 / (1 - b1 ** (it + 1))) / (np.sqrt(m2 / (1 - b2 ** (it + 1))) + 1e-12)
    el = time.perf_counter() - t1
    st.close()
    log(f"calibration: loss {trace[0][0]:.3e} -> {trace[-1][0]:.3e}, (gamma, mu) -> "
        f"({trace[-1][1]:.4f}, {trace[-1][2]:.4f}); {el / iters:.3f} s per iteration")
    return {"members": members, "grains": grains, "iterations": iters, "steps": steps, "ckpt_every": ckpt,
            "dt": dt, "s_per_iteration": el / iters, "device_s_per_iteration": t_dev / iters,
            "loss_first": trace[0][0], "loss_last": trace[-1][0],
            "gamma_mu_last": [trace[-1][1], trace[-1][2]], "gamma_mu_true": [1.1, 0.9],
            "gamma_times_mu_last": trace[-1][1] * trace[-1][2], "gamma_times_mu_true": 1.1 * 0.9,
            "workload": f"{members} members, G ~ U[5000, 20000], lognormal SE (mean 0.3, sigma 0.5), "
                        "20-bin area histograms; Adam on (log gamma, log mu) from (1, 1)"}


def cpu_baseline(opts, mesh, steps, omp=False):
    """The oracle timed on the host: serial (the checker itself), or its
    OpenMP build (per-entity physics loops on all threads OpenMP is given,
    topology serial; the same results), a multi-core baseline only."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from _util import Oracle  # the oracle is the CPU comparator
    orc = Oracle(opts, mesh, omp=omp)
    t0 = time.perf_counter()
    rc, st = orc.step(steps)
    dt = time.perf_counter() - t0
    orc.close()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)) if omp else 1
    how = f"OpenMP build, {threads} threads" if omp else "1 thread"
    return {"value": st.n_vertex_steps / dt, "unit": "vertex-steps/s", "cores": threads, "kind": "port",
            "sample": f"{steps} steps of the same 1M-grain workload on the FP64 oracle "
                      f"(gcc -O2 -ffp-contract=off, {how}), {dt:.1f} s"}


def copy_calibration(mib, reps=10):
    """Plain device-to-device copy of a known size: the measured HBM ceiling,
    and the reference point for scaling FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = mib * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    print(json.dumps({"copy_MiB": mib, "bytes_moved": 2 * n * 8, "avg_us": dt * 1e6,
                      "GBs": 2 * n * 8 / dt / 1e9}), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--grains", type=int, default=1_000_000)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-steps", type=int, default=6, help="oracle steps for cpu_baseline (0: skip)")
    ap.add_argument("--graph", type=int, default=0, help="replay the step as a HIP graph (no per-kernel timing)")
    ap.add_argument("--graph-steps", type=int, default=200, help="extra steps timed as HIP graph replays (0: skip)")
    ap.add_argument("--members", type=int, default=256, help="ensemble members (0: skip the ensemble line)")
    ap.add_argument("--ensemble-steps", type=int, default=1000)
    ap.add_argument("--b2b-steps", type=int, default=50, help="steps per member when stepped back to back")
    ap.add_argument("--adjoint-grains", type=int, default=10000, help="adjoint window workload (0: skip)")
    ap.add_argument("--adjoint-steps", type=int, default=500)
    ap.add_argument("--adjoint-ckpt", type=int, default=50)
    ap.add_argument("--calib-members", type=int, default=64, help="batched calibration workload (0: skip)")
    ap.add_argument("--calib-iters", type=int, default=50)
    ap.add_argument("--calib-steps", type=int, default=1000)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only time a plain device copy of MIB MiB (HBM ceiling and PMC byte-counter scale)")
    args = ap.parse_args()
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
        return

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: there is no multi-GPU path (asked for --gpus {args.gpus}). "
              "Ensembles would shard by whole members, one process per GPU; a single domain by spatial strips "
              "with a halo exchange per step.", file=sys.stderr)
        sys.exit(2)

    import sevc
    if sevc.lib().sevc_device_count() < 1:
        raise SystemExit("no HIP device visible")

    t0 = time.time()
    opts, mesh = make_mesh(sevc, args.grains, args.seed)
    log(f"mesh: {mesh.nv} vertex slots, {mesh.ne} boundary slots, {mesh.ng} grain slots ({time.time() - t0:.1f} s)")
    st = sevc.State(opts, mesh)
    tile = st.autotune()
    st.step(args.warmup, graph=bool(args.graph))
    st.enable_timing(2)           # HIP events around velocity / translate / text only
    sync = st.step(0)
    t1 = time.perf_counter()
    s = st.step(args.steps, graph=bool(args.graph))
    t2 = time.perf_counter()
    if s.status != 0:
        raise SystemExit(f"step failed with status {s.status}")
    elapsed = t2 - t1
    vsteps = s.n_vertex_steps
    value = vsteps / elapsed
    kt = st.kernel_times()
    verts_per_step = vsteps / max(1, s.steps_done)
    def avg_s(name):
        ms, n = kt.get(name, (float("nan"), 0))
        return ms / max(1, n) * 1e-3
    # the FD velocity is two kernels: pass 1 over grain tiles, pass 2 over vertices
    g_avg, v_avg, tr_avg = avg_s("vel_grain"), avg_s("velocity"), avg_s("translate")
    vel_avg = g_avg + v_avg
    achieved = B_VELOCITY * verts_per_step / vel_avg / 1e9
    log(f"{s.steps_done} steps in {elapsed:.3f} s, {vsteps} vertex-steps; velocity {vel_avg * 1e6:.1f} us, "
        f"translate {tr_avg * 1e6:.1f} us; flips {s.n_flips} T2 {s.n_removals} nucl {s.n_nucleations} "
        f"grains {s.n_grains} dt {s.dt_last:.3e}")
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "vertex-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / max(1, s.steps_done) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"plain step, periodic Voronoi of {args.grains} seeds (seed {args.seed}), "
                               "gamma=1, mu=1, lognormal SE (mean 0.3, sigma 0.5), nucleation 1e-4/grain/step, "
                               "adaptive dt, flip threshold 1e-4",
                   "grains": args.grains, "graph": bool(args.graph), "tile_grains": tile},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (TRAFFIC_VELOCITY_B_PER_VERTEX * verts_per_step
                                 if args.grains == 1_000_000 and args.seed == 1 else None),
                     "traffic_unit": "bytes per launch (both passes)",
                     "kernel": "FD velocity (k_vel_grain + k_vel_vertex)", "bytes_per_vertex": B_VELOCITY,
                     "avg_launch_us": vel_avg * 1e6,
                     "pass_us": {"k_vel_grain": g_avg * 1e6, "k_vel_vertex": v_avg * 1e6}},
        "velocity_translation": {"avg_us": (vel_avg + tr_avg) * 1e6,
                                 "achieved_GBs": (B_VELOCITY + B_TRANSLATE) * verts_per_step / (vel_avg + tr_avg) / 1e9},
        "events": {"flips": s.n_flips, "removals": s.n_removals, "nucleations": s.n_nucleations},
    }
    if args.graph_steps > 0:
        # same workload continued with the step replayed as a HIP graph
        st.enable_timing(0)
        st.step(5, graph=True)
        g0 = time.perf_counter()
        sg = st.step(args.graph_steps, graph=True)
        g1 = time.perf_counter()
        line["graph_replay"] = {"value": sg.n_vertex_steps / (g1 - g0),
                                "ms_per_step": (g1 - g0) / max(1, sg.steps_done) * 1e3,
                                "steps": sg.steps_done}
    st.close()
    if args.members > 0:
        line["ensemble"] = bench_ensemble(sevc, args.members, args.ensemble_steps, args.b2b_steps)
    if args.adjoint_grains > 0:
        line["adjoint"] = bench_adjoint(sevc, args.adjoint_grains, 3, args.adjoint_steps, args.adjoint_ckpt, 1e-6)
    if args.calib_members > 0:
        line["calibration"] = bench_calibration(sevc, args.calib_members, args.calib_iters, args.calib_steps, 100,
                                                5e-7)
    if args.cpu_steps > 0:
        line["cpu_baseline"] = cpu_baseline(opts, mesh, args.cpu_steps)
        line["cpu_baseline_openmp"] = cpu_baseline(opts, mesh, 2 * args.cpu_steps, omp=True)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
