"""Pins of the oracle's dynamics that do not come from the oracle itself.

* The velocity is -grad E for E = sum_b gamma_b L_b + sum_g SE_g |A_g|
  (calculus.h:56-61): checked against central differences of that energy
  written independently in numpy (periodic minimum image + shoelace).
* Invariants of check_data (utils.h:85-336), total area = D^2 and the
  torus Euler characteristic V - E + G = 0 hold after every chunk of steps.
* A T1 flip and a T2 removal are checked against their diagrams
  (topological.h:160-168, 673-682): the four grains around a flipped edge
  swap roles, flipped vertices sit at the midpoint, the surviving vertex of
  a removed triangle sits at the triangle's incentre."""
import ctypes as C

import numpy as np
import pytest

import sevc
from _util import Oracle, oracle_lib, ref2k_options


def _small_mesh(n=150, seed=2, **kw):
    pos, rid = sevc.voronoi(n, seed)
    rng = np.random.default_rng(seed)
    o = sevc.default_options()
    o.GB_eps = 0.0
    o.GB_scaling = 1.3
    o.n_grains = n
    for k, v in kw.items():
        setattr(o, k, v)
    return o, sevc.build_mesh(o, pos, rid, rng.random(n), rng.random(n) * 2.0, n)


def _wrap(d, D):
    return d - D * np.round(d / D)


def numpy_energy(h, P, gamma):
    """Total energy of mesh h with vertex positions P (nv,2)."""
    D = h.D
    e = 0.0
    bv = h["b_vtx"].reshape(-1, 2)
    for b in np.nonzero(h["b_enabled"])[0]:
        d = _wrap(P[bv[b, 0]] - P[bv[b, 1]], D)
        e += gamma * np.hypot(d[0], d[1])
    gl = h.grain_lists()
    for g in np.nonzero(h["g_enabled"])[0]:
        L = gl[g, :h["g_vlen"][g]]
        pts = P[L]
        rel = np.cumsum(np.vstack([[0, 0], _wrap(np.diff(pts, axis=0), D)]), axis=0)
        x, y = rel[:, 0], rel[:, 1]
        a = 0.5 * np.sum(x * np.roll(y, -1) - y * np.roll(x, -1))
        e += h["g_SE"][g] * abs(a)
    return e


@pytest.mark.parametrize("model", [sevc.VEL_FD, sevc.VEL_ANALYTIC])
def test_velocity_is_minus_energy_gradient(model):
    o, mesh = _small_mesh(velocity_model=model)
    orc = Oracle(o, mesh)
    assert orc.velocities() == 0
    out = orc.get_mesh()
    v = out["vel"].reshape(-1, 2)
    P0 = mesh.xy().copy()
    rng = np.random.default_rng(7)
    hstep = 1e-7
    for vi in rng.choice(np.nonzero(mesh["v_enabled"])[0], 25, replace=False):
        g = np.zeros(2)
        for c in range(2):
            Pp, Pm = P0.copy(), P0.copy()
            Pp[vi, c] += hstep
            Pm[vi, c] -= hstep
            g[c] = (numpy_energy(mesh, Pp, o.GB_scaling) - numpy_energy(mesh, Pm, o.GB_scaling)) / (2 * hstep)
        # forward difference of the reference: truncation ~ 1e-8 |grad^2 E|
        tol = 2e-4 if model == sevc.VEL_FD else 1e-6
        assert np.linalg.norm(v[vi] + g) <= tol * max(1.0, np.linalg.norm(g)), (vi, v[vi], -g)


def test_velocity_symmetries():
    o, mesh = _small_mesh(velocity_model=sevc.VEL_ANALYTIC)
    base = Oracle(o, mesh)
    base.velocities()
    v0 = base.get_mesh()["vel"].reshape(-1, 2)
    # periodic translation
    m2 = mesh.copy()
    xy = m2.xy()
    xy[:] = np.mod(xy + np.array([0.3141, 0.2718]), 1.0)
    t = Oracle(o, m2)
    t.velocities()
    v1 = t.get_mesh()["vel"].reshape(-1, 2)
    en = mesh["v_enabled"] == 1
    assert np.allclose(v0[en], v1[en], rtol=1e-7, atol=1e-7)
    # mirror x -> D - x reverses every grain's orientation; |A| is unchanged
    m3 = mesh.copy()
    xy = m3.xy()
    xy[:, 0] = np.mod(1.0 - xy[:, 0], 1.0)
    r = Oracle(o, m3)
    r.velocities()
    v2 = r.get_mesh()["vel"].reshape(-1, 2)
    assert np.allclose(v0[en, 0], -v2[en, 0], rtol=1e-7, atol=1e-7)
    assert np.allclose(v0[en, 1], v2[en, 1], rtol=1e-7, atol=1e-7)


def test_extinction_time_formula():
    o = ref2k_options()
    mesh = sevc.load_files(o)
    orc = Oracle(o, mesh)
    orc.velocities()
    out = orc.get_mesh()
    bv = out["b_vtx"].reshape(-1, 2)
    P = out.xy()
    V = out["vel"].reshape(-1, 2)
    d = _wrap(P[bv[:, 1]] - P[bv[:, 0]], 1.0)
    L = np.hypot(d[:, 0], d[:, 1])
    dv = V[bv[:, 1]] - V[bv[:, 0]]
    t = -L * L / np.einsum("ij,ij->i", d, dv)
    te = out["b_t_ext"]
    assert np.allclose(te, t, rtol=1e-9)


def _euler(h):
    return int(h["v_enabled"].sum()) - int(h["b_enabled"].sum()) + int(h["g_enabled"].sum())


def test_invariants_over_a_run():
    o = ref2k_options(dt=1e-4, leave_percent=0.5)
    mesh = sevc.load_files(o)
    orc = Oracle(o, mesh)
    for _ in range(8):
        rc, st = orc.step(100)
        assert rc == 0
        h = orc.get_mesh()
        bad, total = sevc.check_mesh(h)
        assert bad == 0
        assert abs(total - 1.0) < 1e-12
        assert _euler(h) == 0
        if st.done:
            break
    assert st.n_grains <= 1000


def _neighbours(h, v):
    bv = h["b_vtx"].reshape(-1, 2)
    out = set()
    for b in h["v_bnd"][3 * v:3 * v + 3]:
        out.add(int(bv[b, 1] if bv[b, 0] == v else bv[b, 0]))
    return out


def _grains_of(h, v):
    return set(int(g) for g in h["v_grn"][3 * v:3 * v + 3])


def test_single_flip_and_single_removal_follow_the_diagrams():
    o = ref2k_options(dt=1e-5)
    mesh = sevc.load_files(o)
    orc = Oracle(o, mesh)
    orc.step(550)
    seen_flip = seen_t2 = False
    for _ in range(1500):
        pre = orc.get_mesh()
        rc, st = orc.step(1)
        assert rc == 0
        post = orc.get_mesh()
        if st.n_flips == 1 and st.n_removals == 0 and not seen_flip:
            moved = np.nonzero(np.any(pre["b_vtx"].reshape(-1, 2) != post["b_vtx"].reshape(-1, 2), axis=1))[0]
            # the flipped edge keeps its endpoints; its two neighbours AC, BE change one end
            assert len(moved) == 2
            A = B = None
            for b in np.nonzero(post["b_enabled"])[0]:
                a0, a1 = post["b_vtx"][2 * b], post["b_vtx"][2 * b + 1]
                if np.array_equal(post.xy()[a0], post.xy()[a1]) and set(_grains_of(pre, a0)) != set(_grains_of(post, a0)):
                    A, B = int(a0), int(a1)
            assert A is not None
            gic_pre = _grains_of(pre, A) & _grains_of(pre, B)
            sym_pre = _grains_of(pre, A) ^ _grains_of(pre, B)
            gic_post = _grains_of(post, A) & _grains_of(post, B)
            assert gic_post == sym_pre and (_grains_of(post, A) ^ _grains_of(post, B)) == gic_pre
            d = _wrap(pre.xy()[B] - pre.xy()[A], 1.0)
            mid = np.mod(pre.xy()[A] + d / 2, 1.0)
            assert np.allclose(post.xy()[A], mid, atol=1e-15)
            assert np.all(post["vel"][2 * A:2 * A + 2] == 0)
            seen_flip = True
        if st.n_removals == 1 and st.n_flips == 0 and not seen_t2:
            gone = np.nonzero(pre["g_enabled"].astype(bool) & ~post["g_enabled"].astype(bool))[0]
            assert len(gone) == 1
            g = int(gone[0])
            A, B, Cv = (int(x) for x in pre.grain_lists()[g, :3])
            assert pre["g_vlen"][g] == 3
            assert post["v_enabled"][A] == 1 and post["v_enabled"][B] == 0 and post["v_enabled"][Cv] == 0
            pa, pb, pc = pre.xy()[A], pre.xy()[B], pre.xy()[Cv]
            ub, uc = _wrap(pb - pa, 1.0), _wrap(pc - pa, 1.0)
            a, b, c = np.linalg.norm(ub - uc), np.linalg.norm(uc), np.linalg.norm(ub)
            inc = np.mod(pa + (b * ub + c * uc) / (a + b + c), 1.0)
            assert np.allclose(post.xy()[A], inc, atol=1e-14)
            outer = (_neighbours(pre, A) | _neighbours(pre, B) | _neighbours(pre, Cv)) - {A, B, Cv}
            assert _neighbours(post, A) == outer
            seen_t2 = True
        if seen_flip and seen_t2:
            break
    assert seen_flip and seen_t2


def test_flip_threshold_spares_zero_length_edges():
    """Flip-threshold extension: an edge shorter than the threshold that is
    shrinking (0 <= t_ext) flips, but the zero-length edge a flip leaves
    behind (both ends at the midpoint, t_ext = the no-motion value 100) is
    not shrinking and must not flip straight back on the next step."""
    thr = 3e-3
    o, mesh = _small_mesh(n=400, seed=6, dt=2e-4, flip_threshold=thr)
    orc = Oracle(o, mesh)
    zero_seen = flips = 0
    for _ in range(150):
        pre = orc.get_mesh()
        bv = pre["b_vtx"].reshape(-1, 2)
        P = pre.xy()
        d = _wrap(P[bv[:, 0]] - P[bv[:, 1]], o.domain_bound)
        zero = (np.hypot(d[:, 0], d[:, 1]) == 0) & (pre["b_enabled"] == 1)
        rc, st = orc.step(1)
        assert rc == 0
        post = orc.get_mesh()
        assert not np.any(post["b_candidate"][zero]), np.nonzero(zero & (post["b_candidate"] == 1))
        zero_seen += int(zero.sum())
        flips += st.n_flips
    assert flips > 0 and zero_seen > 0, (flips, zero_seen)


def test_two_triangle_collapse_keeps_invariants():
    """remove_two_grains_configurations (topological.h:22-73, 840-985) on a
    run that makes pairs of adjacent 3-sided grains (flip threshold twice
    the mean edge length): after every step with such collapses the mesh
    passes check_data, tiles the domain (total area D^2) and keeps the torus
    Euler characteristic, and the grain count drops by two per collapse
    plus one per T2 removal."""
    n = 8000
    pos, rid = sevc.voronoi(n, 21)
    rng = np.random.default_rng(21)
    se = np.exp(np.log(0.3) - 0.125 + 0.5 * rng.standard_normal(n))
    o = sevc.default_options()
    o.GB_eps = 0.0
    o.n_grains = n
    o.dt = 1e-5
    o.flip_threshold = 2.0 * 0.6 * np.sqrt(1.0 / n)
    orc = Oracle(o, sevc.build_mesh(o, pos, rid, rng.random(n), se, n))
    prev, seen = n, 0
    for _ in range(60):
        rc, st = orc.step(1)
        assert rc == 0
        assert prev - st.n_grains == st.n_removals + 2 * st.n_two_grain
        prev = st.n_grains
        if st.n_two_grain:
            seen += st.n_two_grain
            h = orc.get_mesh()
            bad, total = sevc.check_mesh(h)
            assert bad == 0
            assert abs(total - 1.0) < 1e-12
            assert _euler(h) == 0
    assert seen >= 5
