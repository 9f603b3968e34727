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
    log(f"ensemble: batched {batched:.3e} vertex-steps/s over {steps} steps; back-to-back {b2b:.3e} "
        f"over {b2b_steps} steps per member")
    return {"value": batched, "unit": "vertex-steps/s", "members": members, "grains": grains, "steps": steps,
            "ms_per_step": (t2 - t1) / max(1, s.steps_done) * 1e3,
            "back_to_back": {"value": b2b, "steps_per_member": b2b_steps, "seconds": el},
            "speedup": batched / b2b,
            "workload": f"{members} members, G log-uniform in [1000, 50000] (Voronoi seed 1000+i), "
                        "SE = s_i*U[0,1) with s_i ~ U[0,1]; model options of the single-domain workload"}


def cpu_baseline(opts, mesh, steps):
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from _util import Oracle  # the oracle is the CPU comparator
    orc = Oracle(opts, mesh)
    t0 = time.perf_counter()
    rc, st = orc.step(steps)
    dt = time.perf_counter() - t0
    orc.close()
    return {"value": st.n_vertex_steps / dt, "unit": "vertex-steps/s", "cores": 1, "kind": "port",
            "sample": f"{steps} steps of the same 1M-grain workload on the serial FP64 oracle "
                      f"(gcc -O2 -ffp-contract=off, 1 thread), {dt:.1f} s"}


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
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only time a plain device copy of MIB MiB (HBM ceiling and PMC byte-counter scale)")
    args = ap.parse_args()
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
        return

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: this build has no multi-GPU path (asked for --gpus {args.gpus}); "
              "see DESIGN.md for how the step would shard.", file=sys.stderr)
        sys.exit(2)

    import sevc
    if sevc.lib().sevc_device_count() < 1:
        raise SystemExit("no HIP device visible")

    t0 = time.time()
    opts, mesh = make_mesh(sevc, args.grains, args.seed)
    log(f"mesh: {mesh.nv} vertex slots, {mesh.ne} boundary slots, {mesh.ng} grain slots ({time.time() - t0:.1f} s)")
    st = sevc.State(opts, mesh)
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
                   "grains": args.grains, "graph": bool(args.graph)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
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
    if args.cpu_steps > 0:
        line["cpu_baseline"] = cpu_baseline(opts, mesh, args.cpu_steps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
