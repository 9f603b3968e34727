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
TIMING_SAMPLE = 16               # the timed run records kernel events on one step in 16
# Algorithmic bytes per enabled vertex and step (FP64 values, int32 ids,
# each datum once): velocity 98 B, translation 48 B.
B_VELOCITY = 98.0
B_TRANSLATE = 48.0
# HBM bytes per enabled vertex of the FD velocity at the 1M-grain workload,
# by path, from rocprofv3 FETCH_SIZE / WRITE_SIZE of `python bench.py`
# scaled by a plain copy of known size (the summaries under profiles/
# named here); reported with their source, not measured in this run.
TRAFFIC_VELOCITY = {
    0: (260.0, "profiles/step_1M_hbm_traffic.md (two passes: k_vel_grain + k_vel_vertex)"),
}


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


def make_mc_mesh(sevc, n_grains, seed, mc_k=5.0):
    """The reference's Monte Carlo nucleation mode at scale
    (nucleation2.cu:239-443): one Metropolis site per step, opened from the
    first step (the reference waits until 1/6 of the grains are gone,
    main.cu:137-139), in a domain of side 1000 so that edges are O(1) long as
    in the reference's example (DOMAIN BOUND 40 for 2,000 grains), stored
    energy uniform in [3, 6] (SE_min 3, SE_max 6 of config/example.ini),
    GB_eps = 0.3, adaptive dt, flip threshold 0.1."""
    pos, rid = sevc.voronoi(n_grains, seed)
    rng = np.random.default_rng(seed)
    o = make_options(sevc, n_grains)
    o.domain_bound = 1000.0
    o.dt = 1.0
    o.flip_threshold = 0.1
    o.GB_eps = 0.3
    o.nucleation_mode = sevc.NUCL_MC
    o.MC_k = mc_k
    o.nucleation_gate = n_grains
    o.nucleation_gap = 1
    o.seed = seed
    mesh = sevc.build_mesh(o, pos, rid, rng.random(n_grains), rng.random(n_grains), n_grains)
    mesh["g_SE"][:n_grains] = 3.0 + 3.0 * rng.random(n_grains)
    return o, mesh


def make_member(sevc, i, seed0=1000, g_lo=1000, g_hi=50000):
    """Member i of the ensemble workload: G_i = round(exp(ln g_lo + U_i
    ln(g_hi/g_lo))) grains (log-uniform), Voronoi seed seed0+i, SE = s_i *
    U[0,1) with s_i ~ U[0,1], Philox seed seed0+i, and the model options of
    the single-domain workload."""
    rng = np.random.default_rng(seed0 + i)
    g = int(round(np.exp(np.log(g_lo) + rng.random() * np.log(g_hi / g_lo))))
    s_i = rng.random()
    pos, rid = sevc.voronoi(g, seed0 + i)
    o = make_options(sevc, g)
    o.seed = seed0 + i
    return o, sevc.build_mesh(o, pos, rid, rng.random(g), s_i * rng.random(g), g)


def make_ensemble(sevc, members, seed0=1000, g_lo=1000, g_hi=50000):
    """The first `members` members of the ensemble workload (make_member)."""
    opts, meshes = [], []
    for i in range(members):
        o, m = make_member(sevc, i, seed0, g_lo, g_hi)
        opts.append(o)
        meshes.append(m)
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
    # the same members one state at a time, each autotuned like the batch,
    # from the same warm-up over the same steps
    vs, el = 0, 0.0
    for o, m in zip(opts, meshes):
        one = sevc.State(o, m)
        one.autotune()
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
            "back_to_back": {"value": b2b, "steps_per_member": b2b_steps, "warmup": warmup, "seconds": el,
                             "note": "each member alone, autotuned, after the same warm-up as the batch"},
            "speedup": batched / b2b,
            "workload": f"{members} members, G log-uniform in [1000, 50000] (Voronoi seed 1000+i), "
                        "SE = s_i*U[0,1) with s_i ~ U[0,1]; model options of the single-domain workload"}


def bench_at_scale(sevc, grains, steps, warmup=5):
    """The plain step at full size with anisotropic boundary energy (the
    benchmark workload with GB_eps = 0.3) and in the reference's Monte
    Carlo nucleation mode (make_mc_mesh): vertex-steps/s over `steps`."""
    out = {}
    for name, make in (("anisotropic_gb_eps_0.3", lambda: _aniso(sevc, grains)),
                       ("mc_nucleation", lambda: make_mc_mesh(sevc, grains, 1))):
        o, mesh = make()
        st = sevc.State(o, mesh)
        st.autotune()
        st.step(warmup)
        st.step(0)
        a = time.perf_counter()
        s = st.step(steps)
        b = time.perf_counter()
        if s.status != 0:
            raise SystemExit(f"{name}: step failed with status {s.status}")
        out[name] = {"value": s.n_vertex_steps / (b - a), "unit": "vertex-steps/s", "steps": s.steps_done,
                     "ms_per_step": (b - a) / max(1, s.steps_done) * 1e3, "flips": s.n_flips,
                     "removals": s.n_removals, "nucleations": s.n_nucleations,
                     "velocity_path": st.velocity_config()[0]}
        log(f"at scale, {name}: {out[name]['value']:.3e} vertex-steps/s, {out[name]['ms_per_step']:.3f} ms/step, "
            f"{s.n_nucleations} nucleations")
        st.close()
    return out


def _aniso(sevc, grains):
    o, mesh = make_mesh(sevc, grains, 1)
    o.GB_eps = 0.3
    return o, mesh


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


def bench_adjoint(sevc, grains, seed, steps, ckpt, dt, n_se=18, reps=3):
    """Adjoint over a topology-fixed window: the loss is the squared
    difference of the smooth 20-bin histogram of grain areas after `steps`
    steps (sevc_area_histogram, on the device) from the histogram the same
    window gives with (1.1 gamma, 0.9 mu).  Times the window forward (with
    its checkpoints) against the reverse pass (the loss's dL/dx seeding it
    on the device, and the segment recomputation), and checks dL/dgamma,
    dL/dmu and dL/dSE_g of n_se random grains against central differences
    of the same loss.  Each coordinate's central difference is taken over
    a ladder of steps h = 1e-6 ... 3e-2 (half-decades); the rounding noise of
    the loss, about 1e-13 of it, swamps small steps when the derivative is
    tiny (dL/dSE_g of one grain out of 10k is ~1e-11 of a ~1e-8 loss), and
    truncation grows at large ones.  The step whose quotient agrees best
    with both ladder neighbours (the plateau) is the reference value; the
    result at h = 1e-6 is reported beside it."""
    from sevc.calib import grain_areas
    pos, rid = sevc.voronoi(grains, seed)
    rng = np.random.default_rng(seed)
    o = window_options(sevc, grains, dt)
    mesh = sevc.build_mesh(o, pos, rid, rng.random(grains), 0.3 * rng.random(grains), grains)
    am = float(np.abs(grain_areas(mesh)).mean())
    ranges = np.array([[0.0, 3.0 * am, 0.1 * am]])
    st = sevc.State(o, mesh)
    st.set_params(0, gb_scaling=1.1, mobility=0.9)
    st.window_forward(steps, ckpt)
    target, _ = st.area_histogram(20, ranges, np.zeros((1, 20)))
    x0, se0 = mesh["pos"].copy(), mesh["g_SE"].copy()

    def loss(se=se0, gb=1.0, mu=1.0):
        st.set_params(0, pos=x0, g_SE=se, gb_scaling=gb, mobility=mu)
        st.window_forward(steps, ckpt)
        return float(st.area_histogram(20, ranges, target)[1][0])

    t_f, t_a = [], []
    for _ in range(reps):
        st.set_params(0, pos=x0, g_SE=se0, gb_scaling=1.0, mobility=1.0)
        a = time.perf_counter()
        st.window_forward(steps, ckpt)
        b = time.perf_counter()
        _, lv = st.area_histogram(20, ranges, target, seed=True)   # dL/dx seeds the reverse pass on the device
        (dp, dse, dgb, dmu), = st.adjoint(None)
        c = time.perf_counter()
        t_f.append(b - a)
        t_a.append(c - b)
    L0 = float(lv[0])
    pick = np.random.default_rng(seed + 1)
    coords = [("gamma", -1), ("mu", -1)] + [("SE_g", int(g)) for g in
                                            pick.choice(np.nonzero(mesh["g_enabled"])[0], n_se, replace=False)]

    def f(kind, g, e):
        if kind == "gamma":
            return loss(gb=1.0 + e)
        if kind == "mu":
            return loss(mu=1.0 + e)
        se = se0.copy()
        se[g] += e
        return loss(se=se)

    ladder = (1e-6, 3e-6, 1e-5, 3e-5, 1e-4, 3e-4, 1e-3, 3e-3, 1e-2, 3e-2)
    checks, worst = [], 0.0
    for kind, g in coords:
        ad = {"gamma": dgb, "mu": dmu}.get(kind, dse[g] if g >= 0 else 0.0)
        c = [(f(kind, g, h) - f(kind, g, -h)) / (2 * h) for h in ladder]
        spread = [max(abs(c[i] - c[i - 1]), abs(c[i] - c[i + 1])) / max(abs(c[i]), 1e-300)
                  for i in range(1, len(c) - 1)]
        i = 1 + int(np.argmin(spread))
        fd = c[i]
        rel = abs(ad - fd) / max(abs(fd), 1e-300)
        worst = max(worst, rel)
        checks.append({"coord": kind if g < 0 else f"SE_{g}", "adjoint": ad, "fd": fd, "rel_err": rel,
                       "h": ladder[i], "fd_spread": spread[i - 1],
                       "rel_err_at_h_1e-6": abs(ad - c[0]) / max(abs(c[0]), 1e-300)})
    # extra: a seeded linear functional w . x_N of the window's end
    # positions, against Richardson differences in start positions (h = 1e-6)
    # and in gamma, mu (h = 1e-3)
    w = pick.standard_normal(2 * mesh.nv) * np.repeat(mesh["v_enabled"], 2)
    st.set_params(0, pos=x0, g_SE=se0, gb_scaling=1.0, mobility=1.0)
    st.window_forward(steps, ckpt)
    xN = st.positions(0).reshape(-1).copy()
    (lp, _, lgb, lmu), = st.adjoint(w)
    D = mesh.D

    def lin(p=x0, gbv=1.0, muv=1.0):
        st.set_params(0, pos=p, g_SE=se0, gb_scaling=gbv, mobility=muv)
        st.window_forward(steps, ckpt)
        d = st.positions(0).reshape(-1) - xN
        return float(np.dot(w, d - D * np.round(d / D)))

    lin_worst = 0.0
    for kind, i, h in [("pos", int(c), 1e-6) for c in pick.choice(np.nonzero(w)[0], 6, replace=False)] + \
                      [("gamma", -1, 1e-3), ("mu", -1, 1e-3)]:
        def fl(e):
            if kind == "pos":
                p = x0.copy()
                p[i] += e
                return lin(p=p)
            return lin(gbv=1.0 + e) if kind == "gamma" else lin(muv=1.0 + e)
        c1 = (fl(h) - fl(-h)) / (2 * h)
        c2 = (fl(h / 2) - fl(-h / 2)) / h
        fd = (4 * c2 - c1) / 3
        ad = lp[i] if kind == "pos" else (lgb if kind == "gamma" else lmu)
        lin_worst = max(lin_worst, abs(ad - fd) / max(abs(fd), 1e-300))
    st.close()
    tf, ta = min(t_f), min(t_a)
    w_se = max(c["rel_err"] for c in checks if c["coord"].startswith("SE_"))
    w_gm = max(c["rel_err"] for c in checks if not c["coord"].startswith("SE_"))
    w6 = max(c["rel_err_at_h_1e-6"] for c in checks if not c["coord"].startswith("SE_"))
    log(f"adjoint: {grains} grains, {steps} steps (checkpoint every {ckpt}): forward {tf * 1e3:.1f} ms, "
        f"reverse {ta * 1e3:.1f} ms, ratio {ta / tf:.2f}; loss {L0:.3e}; worst FD rel. error {worst:.2e} over "
        f"{len(checks)} coordinates (gamma, mu {w_gm:.1e}, {w6:.1e} at h = 1e-6; SE_g {w_se:.1e})")
    return {"ratio": ta / tf, "forward_ms": tf * 1e3, "adjoint_ms": ta * 1e3, "grains": grains, "steps": steps,
            "ckpt_every": ckpt, "dt": dt, "loss": L0, "fd_max_rel_err": worst, "fd_coords": len(checks),
            "fd_checks": checks,
            "linear_functional_fd_max_rel_err": lin_worst,
            "fd_note": "central differences of the histogram loss over h = 1e-6..3e-2 in half-decades; fd is "
                       "the quotient at the step agreeing best with its two ladder neighbours (fd_spread: that "
                       "relative disagreement); rel_err_at_h_1e-6 compares with h = 1e-6 itself, which the "
                       "loss's rounding noise swamps for the ~1e-11 dL/dSE_g",
            "workload": f"periodic Voronoi of {grains} seeds (seed {seed}), SE = 0.3*U[0,1), 20-bin smooth "
                        "area histogram vs the window's histogram at (1.1 gamma, 0.9 mu)"}


def horizon_forward(st, n_members, horizon, window, ckpt, keep=False, slot0=1):
    """A horizon of `horizon` steps run as topology-fixed windows of
    window - 1 steps (sevc_window_forward, analytic model, fixed dt, no
    events), each followed by one step of the model with T1/T2 events
    (sevc_step).  With keep, the state at each window's start is kept in
    device checkpoint slot0 + k, for the reverse pass; returns the slots."""
    starts = []
    for k in range(horizon // window):
        if keep:
            st.checkpoint(slot0 + k)
            starts.append(slot0 + k)
        st.window_forward(window - 1, ckpt)
        st.step(1)
    return starts


def horizon_adjoint(st, n_members, starts, window, ckpt, lam_end):
    """Reverse pass over a horizon: from the last window back, the state at
    each window's start is restored from its device checkpoint, the window is
    recomputed and its adjoint run from lam (dL/dx at the window's end).
    Within a window the gradient is exact; across the event step between
    windows dL/dx is passed on by vertex slot unchanged (the step's own
    Jacobian is not formed) and is zero for vertices the step removed.
    Returns per member (dL/dSE_g, dL/dgamma, dL/dmu) summed over the
    windows."""
    lam = [np.asarray(x, dtype=np.float64).copy() for x in lam_end]
    d_se = None
    d_gb = np.zeros(n_members)
    d_mu = np.zeros(n_members)
    for k in reversed(range(len(starts))):
        st.restore(starts[k])
        st.window_forward(window - 1, ckpt)
        res = st.adjoint(lam)
        if d_se is None:
            d_se = [r[1].copy() for r in res]
        else:
            for m in range(n_members):
                d_se[m] += res[m][1]
        d_gb += np.array([r[2] for r in res])
        d_mu += np.array([r[3] for r in res])
        if k > 0:
            # dL/dx at the window's start is zero on slots disabled there, so
            # vertices the event step removed carry no adjoint back
            lam = [res[m][0] for m in range(n_members)]
    return d_se, d_gb, d_mu


def bench_calibration(sevc, members, iters, horizon, window, ckpt, dt, seed0=2000):
    """Batched calibration: members with G ~ U[5000, 20000] grains sharing
    global (gamma, mu) and holding their own per-grain SE (true values
    lognormal, mean 0.3, sigma 0.5).  Each member's target is its own 20-bin
    area histogram after a horizon of `horizon` steps at (1.1, 0.9, true SE),
    split into topology-fixed windows with T1/T2 events between them.
    Starting from gamma = mu = 1 and SE = 0.3 everywhere, Adam on
    (log gamma, log mu, log SE_g of every grain of every member) with the
    ensemble-mean loss; every iteration is one batched horizon forward, the
    device histogram loss and one batched reverse pass over the windows."""
    from sevc.calib import grain_areas
    opts, meshes, ranges, se_true = [], [], [], []
    t0 = time.time()
    for i in range(members):
        rng = np.random.default_rng(seed0 + i)
        g = int(rng.integers(5000, 20001))
        pos, rid = sevc.voronoi(g, seed0 + i)
        o = window_options(sevc, g, dt)
        o.events = 1                         # T1/T2 between windows
        se = np.exp(np.log(0.3) - 0.125 + 0.5 * rng.standard_normal(g))
        m = sevc.build_mesh(o, pos, rid, rng.random(g), se, g)
        am = float(np.abs(grain_areas(m)).mean())
        opts.append(o)
        meshes.append(m)
        ranges.append((0.0, 3.0 * am, 0.1 * am))
        se_true.append(m["g_SE"].copy())
    grains = sum(o.n_grains for o in opts)
    log(f"calibration: {members} members, {grains} grains ({time.time() - t0:.1f} s)")
    st = sevc.State(opts, meshes)
    st.checkpoint(0)                      # the initial state, restored at every iteration

    def start(gb, mu, se_list):
        st.restore(0)
        for k in range(members):
            st.set_params(k, g_SE=se_list[k], gb_scaling=gb, mobility=mu)

    start(1.1, 0.9, se_true)
    horizon_forward(st, members, horizon, window, ckpt)
    targets, _ = st.area_histogram(20, ranges, np.zeros((members, 20)))
    se_init = [np.full(m.ng, 0.3) * m["g_enabled"] for m in meshes]
    theta = np.zeros(2)                                   # log gamma, log mu
    phi = [np.log(np.maximum(x, 1e-300)) for x in se_init]  # log SE per grain
    m1, m2 = np.zeros(2), np.zeros(2)
    p1 = [np.zeros_like(x) for x in phi]
    p2 = [np.zeros_like(x) for x in phi]
    lr, lr_se, b1, b2 = 0.02, 0.003, 0.9, 0.999
    trace = []
    t1 = time.perf_counter()
    for it in range(iters):
        gb, mu = float(np.exp(theta[0])), float(np.exp(theta[1]))
        se_now = [np.exp(x) * mm["g_enabled"] for x, mm in zip(phi, meshes)]
        start(gb, mu, se_now)
        starts = horizon_forward(st, members, horizon, window, ckpt, keep=True)
        _, losses, lam = st.area_histogram(20, ranges, targets, want_grad=True)
        L = float(losses.mean())
        d_se, d_gb, d_mu = horizon_adjoint(st, members, starts, window, ckpt, [x / members for x in lam])
        g = np.array([d_gb.sum() * gb, d_mu.sum() * mu])
        trace.append((L, gb, mu))
        m1 = b1 * m1 + (1 - b1) * g
        m2 = b2 * m2 + (1 - b2) * g * g
        corr1, corr2 = 1 - b1 ** (it + 1), 1 - b2 ** (it + 1)
        theta -= lr * (m1 / corr1) / (np.sqrt(m2 / corr2) + 1e-12)
        for k in range(members):
            gk = d_se[k] * se_now[k]
            p1[k] = b1 * p1[k] + (1 - b1) * gk
            p2[k] = b2 * p2[k] + (1 - b2) * gk * gk
            phi[k] -= lr_se * (p1[k] / corr1) / (np.sqrt(p2[k] / corr2) + 1e-30)
        if it % 10 == 0:
            log(f"calibration: iteration {it}, loss {L:.4e}, (gamma, mu) = ({gb:.4f}, {mu:.4f})")
    el = time.perf_counter() - t1
    st.close()
    se_fit = [np.exp(x) * mm["g_enabled"] for x, mm in zip(phi, meshes)]
    rec = np.array([np.linalg.norm(f - t) / max(np.linalg.norm(i0 - t), 1e-300)
                    for f, t, i0 in zip(se_fit, se_true, se_init)])
    gb, mu = trace[-1][1], trace[-1][2]
    log(f"calibration: loss {trace[0][0]:.3e} -> {trace[-1][0]:.3e}, (gamma, mu) -> ({gb:.4f}, {mu:.4f}); "
        f"SE error left {rec.mean():.3f} of the initial (min {rec.min():.3f}); {el / iters:.3f} s per iteration")
    return {"members": members, "grains": grains, "iterations": iters, "horizon_steps": horizon,
            "window_steps": window, "windows": horizon // window, "ckpt_every": ckpt, "dt": dt,
            "parameters": 2 + grains, "s_per_iteration": el / iters,
            "loss_curve": [t[0] for t in trace[::5]] + [trace[-1][0]],
            "loss_first": trace[0][0], "loss_last": trace[-1][0],
            "gamma_mu_last": [gb, mu], "gamma_mu_true": [1.1, 0.9],
            "gamma_times_mu_last": gb * mu, "gamma_times_mu_true": 1.1 * 0.9,
            "gamma_mu_separately_identified": bool(abs(gb - 1.1) < 0.02 and abs(mu - 0.9) < 0.02),
            "se_error_left_per_member": {"mean": float(rec.mean()), "min": float(rec.min()),
                                         "max": float(rec.max())},
            "workload": f"{members} members, G ~ U[5000, 20000], true SE lognormal (mean 0.3, sigma 0.5), "
                        f"targets: 20-bin area histograms after {horizon} steps at (1.1 gamma, 0.9 mu); "
                        f"windows of {window} steps ({window - 1} topology-fixed + 1 with T1/T2), "
                        f"checkpoint every {ckpt}; Adam on log gamma, log mu and log SE of every grain "
                        "from (1, 1, SE = 0.3)"}


def host_cpu():
    """The host's CPU model (/proc/cpuinfo) and the threads this process may
    use (its affinity mask) next to os.cpu_count()."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": usable, "cpu_count": os.cpu_count() or 1}


class _pinned:
    """Run the enclosed block on one core of this process's affinity mask
    (the serial baseline), restoring the mask afterwards."""
    def __enter__(self):
        try:
            self.mask = os.sched_getaffinity(0)
            os.sched_setaffinity(0, {min(self.mask)})
        except (AttributeError, OSError):
            self.mask = None
        return self

    def __exit__(self, *exc):
        if self.mask is not None:
            os.sched_setaffinity(0, self.mask)
        return False


def cpu_baseline(opts, mesh, steps, omp=False):
    """The oracle timed on the host: serial (the checker itself, pinned to
    one core), or its OpenMP build (per-entity physics loops on all threads
    OpenMP is given, topology serial; the same results), a multi-core
    baseline only."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from _util import Oracle  # the oracle is the CPU comparator
    orc = Oracle(opts, mesh, omp=omp)
    if omp:
        t0 = time.perf_counter()
        rc, st = orc.step(steps)
        dt = time.perf_counter() - t0
    else:
        with _pinned():
            t0 = time.perf_counter()
            rc, st = orc.step(steps)
            dt = time.perf_counter() - t0
    orc.close()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)) if omp else 1
    how = f"OpenMP build, {threads} threads" if omp else "1 thread pinned to one core"
    return {"value": st.n_vertex_steps / dt, "unit": "vertex-steps/s", "cores": threads, "kind": "port",
            "sample": f"{steps} steps of the same 1M-grain workload on the FP64 oracle "
                      f"(gcc -O2 -ffp-contract=off, {how}), {dt:.1f} s", **host_cpu()}


def bench_half_removal(sevc, seed=0):
    """The 2,000-grain workload run to its stop (50 % of the grains
    removed): the reference's 2K tessellation, SE = 0.5 U[0,1), rate
    nucleation 1e-3 SE per grain and step, adaptive dt, flip threshold
    1e-4.  The whole run on the GPU against the same run on the serial
    oracle (pinned to one core); both end in the same state."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from _util import Oracle, ref2k_options  # the oracle is the CPU comparator
    o = ref2k_options(dt=2e-4, adaptive_dt=1, dt_safety=0.5, flip_threshold=1e-4, do_nucleation=1,
                      nucleation_mode=sevc.NUCL_RATE, nucleation_rate=1e-3, nucleation_scaling=0.9,
                      seed=seed, extra_capacity=0.05, leave_percent=0.5)
    mesh = sevc.load_files(o)
    warm = sevc.State(o, mesh)
    warm.step(3)
    warm.close()
    st = sevc.State(o, mesh)
    t0 = time.perf_counter()
    s = st.step(100000)                  # ends at the stop rule
    t_gpu = time.perf_counter() - t0
    sg = sevc.State(o, mesh)             # the same run replayed as captured HIP graphs of 8 steps
    sg.step(0, graph=True)               # capture outside the timed region
    t0 = time.perf_counter()
    s_g = sg.step(100000, graph=True)
    t_graph = time.perf_counter() - t0
    same_graph = s_g.n == s.n and s_g.n_flips == s.n_flips and s_g.n_grains == s.n_grains
    sg.close()
    orc = Oracle(o, mesh)
    with _pinned():
        t0 = time.perf_counter()
        rc, s2 = orc.step(100000)
        t_cpu = time.perf_counter() - t0
    same = (s.n == s2.n and s.n_vertex_steps == s2.n_vertex_steps and s.n_flips == s2.n_flips
            and s.n_grains == s2.n_grains)
    orc.close()
    st.close()
    log(f"2K grains to 50 % removal: {s.n} steps, GPU {t_gpu:.3f} s ({t_graph:.3f} s as graph replays), "
        f"oracle {t_cpu:.2f} s ({t_cpu / t_gpu:.1f}x), same end state counts: {same and same_graph}")
    return {"steps": s.n, "vertex_steps": s.n_vertex_steps, "flips": s.n_flips, "removals": s.n_removals,
            "nucleations": s.n_nucleations, "grains_left": s.n_grains, "gpu_s": t_gpu,
            "gpu_ms_per_step": t_gpu / max(1, s.n) * 1e3, "gpu_value": s.n_vertex_steps / t_gpu,
            "graph_s": t_graph, "graph_ms_per_step": t_graph / max(1, s.n) * 1e3, "graph_same_counts": bool(same_graph),
            "cpu_s": t_cpu, "cpu_value": s2.n_vertex_steps / t_cpu, "cpu_cores": 1, "speedup": t_cpu / t_gpu,
            "same_counts_as_oracle": bool(same), **host_cpu(),
            "workload": "reference 2K tessellation, SE = 0.5 U[0,1), rate nucleation 1e-3 SE per grain and step, "
                        f"Philox seed {seed}, adaptive dt (bound 2e-4), flip threshold 1e-4, run to 50 % removed"}


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
    ap.add_argument("--cpu-steps", type=int, default=12, help="oracle steps for cpu_baseline (0: skip)")
    ap.add_argument("--half-removal", type=int, default=1, help="time the 2K workload to its stop, GPU and oracle (0: skip)")
    ap.add_argument("--graph", type=int, default=0, help="replay the step as a HIP graph (no per-kernel timing)")
    ap.add_argument("--graph-steps", type=int, default=200, help="extra steps timed as HIP graph replays (0: skip)")
    ap.add_argument("--scale-steps", type=int, default=200,
                    help="steps of the anisotropic and the Monte Carlo workloads at full size (0: skip)")
    ap.add_argument("--members", type=int, default=256, help="ensemble members (0: skip the ensemble line)")
    ap.add_argument("--ensemble-steps", type=int, default=1000)
    ap.add_argument("--b2b-steps", type=int, default=0,
                    help="steps per member when stepped back to back (0: the batch's steps)")
    ap.add_argument("--adjoint-grains", type=int, default=10000, help="adjoint window workload (0: skip)")
    ap.add_argument("--adjoint-steps", type=int, default=500)
    ap.add_argument("--adjoint-ckpt", type=int, default=50)
    ap.add_argument("--calib-members", type=int, default=64, help="batched calibration workload (0: skip)")
    ap.add_argument("--calib-iters", type=int, default=50)
    ap.add_argument("--calib-steps", type=int, default=1000, help="calibration horizon (steps)")
    ap.add_argument("--calib-window", type=int, default=200, help="steps per topology-fixed window of the horizon")
    ap.add_argument("--velocity", default="auto", metavar="PATH:TILE",
                    help="velocity path and tile of the main workload (auto: sevc_autotune; counter runs fix "
                         "it, since the tuner's timings change under the profiler)")
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
    if args.velocity == "auto":
        st.autotune()
    else:
        st.set_velocity_config(*(int(x) for x in args.velocity.split(":")))
    vpath, tile = st.velocity_config()
    st.step(args.warmup, graph=bool(args.graph))
    st.enable_timing(2)           # HIP events around velocity / translate / text only,
    st.set_timing_sample(TIMING_SAMPLE)   # on one step in TIMING_SAMPLE
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
    # the FD velocity is one kernel over vertex tiles (path 1) or two:
    # pass 1 over grain tiles, pass 2 over vertices (path 0)
    tr_avg = avg_s("translate")
    if vpath == 1:
        g_avg, v_avg = float("nan"), float("nan")
        vel_avg = avg_s("vel_fused")
        kname = f"FD velocity, one pass over {tile}-vertex tiles (k_vel_vtile)"
        passes = {"k_vel_vtile": vel_avg * 1e6}
    else:
        g_avg, v_avg = avg_s("vel_grain"), avg_s("velocity")
        vel_avg = g_avg + v_avg
        kname = f"FD velocity, two passes ({tile}-grain tiles k_vel_grain + k_vel_vertex)"
        passes = {"k_vel_grain": g_avg * 1e6, "k_vel_vertex": v_avg * 1e6}
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
                   "grains": args.grains, "graph": bool(args.graph), "velocity_path": vpath, "tile": tile},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (TRAFFIC_VELOCITY[vpath][0] * verts_per_step
                                 if args.grains == 1_000_000 and args.seed == 1 and vpath in TRAFFIC_VELOCITY
                                 else None),
                     "traffic_unit": "HBM bytes per velocity step (all its launches)",
                     "traffic_source": TRAFFIC_VELOCITY.get(vpath, (None, None))[1],
                     "kernel": kname, "bytes_per_vertex": B_VELOCITY,
                     "avg_launch_us": vel_avg * 1e6,
                     "pass_us": passes},
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
    if args.scale_steps > 0:
        line["at_scale"] = bench_at_scale(sevc, args.grains, args.scale_steps)
    if args.members > 0:
        line["ensemble"] = bench_ensemble(sevc, args.members, args.ensemble_steps,
                                          args.b2b_steps or args.ensemble_steps)
    if args.adjoint_grains > 0:
        line["adjoint"] = bench_adjoint(sevc, args.adjoint_grains, 3, args.adjoint_steps, args.adjoint_ckpt, 1e-6)
    if args.calib_members > 0:
        line["calibration"] = bench_calibration(sevc, args.calib_members, args.calib_iters, args.calib_steps,
                                                args.calib_window, 50, 5e-7)
    if args.half_removal:
        line["half_removal_2k"] = bench_half_removal(sevc)
    if args.cpu_steps > 0:
        line["cpu_baseline"] = cpu_baseline(opts, mesh, args.cpu_steps)
        line["cpu_baseline_openmp"] = cpu_baseline(opts, mesh, 2 * args.cpu_steps, omp=True)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
