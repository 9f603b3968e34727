"""Oracle runs in worker processes (test infrastructure): independent
members of an ensemble are stepped by the serial oracle in parallel, each
worker building its member's input itself, and return per-chunk records
(statistics, area curves and a digest of the whole state) that the GPU
tests compare with the batched GPU run.  Workers are started with the
'spawn' method, so they inherit no GPU state from the test process."""
import hashlib
import multiprocessing as mp
import os
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

STAT_KEYS = ("n", "n_grains", "n_vertices", "n_flips", "n_removals", "n_two_grain", "n_nucleations",
             "n_vertex_steps", "steps_done", "done", "t", "dt_last", "nucleated_area_pct")


def state_digest(h):
    """sha256 over every compared field of a host mesh (bit patterns)."""
    from _util import STATE_FIELDS
    d = hashlib.sha256()
    for f in STATE_FIELDS:
        d.update(np.ascontiguousarray(h[f]).tobytes())
    return d.hexdigest()


def curves(h, bins=20):
    """Grain count, mean |area| and the 20-bin histogram of |A| / <|A|> on
    [0, 3) of the enabled grains."""
    en = h["g_enabled"] == 1
    a = np.abs(h["g_area"][en])
    mean = float(a.mean()) if a.size else 0.0
    hist = np.histogram(a / mean if mean > 0 else a, bins=bins, range=(0.0, 3.0))[0] if a.size else np.zeros(bins)
    return int(en.sum()), mean, [int(x) for x in hist]


def record(stats, h):
    return {"stats": {k: getattr(stats, k) for k in STAT_KEYS}, "curves": curves(h), "digest": state_digest(h)}


def _init():
    for p in (os.path.join(REPO, "se-grain-growth-gpu_amd"), os.path.join(REPO, "tests"), REPO):
        if p not in sys.path:
            sys.path.insert(0, p)


def half_removal_options(seed):
    """The 2,000-grain workload of the small-scale runs: the reference's 2K
    tessellation, gamma = mu = 1, SE = U[0,1) * 0.5 (uniform.txt x SE_eps),
    rate nucleation 1e-3 * SE per grain and step with Philox seed `seed`,
    new grains SE 0, adaptive dt (bound 2e-4), flip threshold 1e-4, stop
    once 50 % of the grains are removed."""
    _init()
    import sevc
    from _util import ref2k_options
    return ref2k_options(dt=2e-4, adaptive_dt=1, dt_safety=0.5, flip_threshold=1e-4, do_nucleation=1,
                         nucleation_mode=sevc.NUCL_RATE, nucleation_rate=1e-3, nucleation_scaling=0.9,
                         seed=seed, extra_capacity=0.05, leave_percent=0.5)


def _half_removal_member(seed, chunk, max_chunks):
    _init()
    import sevc
    from _util import Oracle
    o = half_removal_options(seed)
    orc = Oracle(o, sevc.load_files(o))
    out = []
    for _ in range(max_chunks):
        rc, st = orc.step(chunk)
        out.append(dict(record(st, orc.get_mesh()), rc=rc))
        if rc != 0 or st.done:
            break
    orc.close()
    return out


def _ensemble_member(i, steps):
    _init()
    import bench
    import sevc
    from _util import Oracle
    o, m = bench.make_member(sevc, i)
    orc = Oracle(o, m)
    rc, st = orc.step(steps)
    rec = dict(record(st, orc.get_mesh()), rc=rc)
    orc.close()
    return rec


def pool(n=None):
    n = n or min(16, os.cpu_count() or 1)
    return mp.get_context("spawn").Pool(n, initializer=_init)


def half_removal_runs(seeds, chunk, max_chunks):
    with pool() as p:
        return p.starmap(_half_removal_member, [(s, chunk, max_chunks) for s in seeds])


def ensemble_runs(members, steps):
    with pool() as p:
        return p.starmap(_ensemble_member, [(i, steps) for i in range(members)], chunksize=4)
