"""Pins of the oracle's topology-fixed window and its adjoint (extension).

The reverse pass is checked against differences of the window forward
itself (central, h = 1e-6 and h/2, Richardson-combined): initial
positions, per-grain SE, GB_scaling and mobility, for a linear functional
of the final positions and for the smooth area-histogram loss used for
calibration, to 1e-6 relative.  A wrong Hessian term, sign or index moves
the gradient far outside that (the last test shows the check rejects a
dropped term).  The window forward is also checked to be the step of the
analytic model with events off, bit for bit."""
import numpy as np
import pytest

import sevc
from sevc.calib import AreaHistogram, grain_areas
from _util import Oracle, diff_fields

H = 1e-6
RTOL = 1e-6


def _mesh(n=160, seed=4, gb_eps=0.0, **kw):
    pos, rid = sevc.voronoi(n, seed)
    rng = np.random.default_rng(seed)
    o = sevc.default_options()
    o.GB_eps = gb_eps
    o.GB_scaling = 1.3
    o.mobility = 0.8
    o.dt = 2e-5
    o.n_grains = n
    o.velocity_model = sevc.VEL_ANALYTIC
    o.events = 0
    o.do_nucleation = 0
    for k, v in kw.items():
        setattr(o, k, v)
    return o, sevc.build_mesh(o, pos, rid, rng.random(n), 0.5 + 2.0 * rng.random(n), n)


def _final(o, mesh, steps):
    orc = Oracle(o, mesh)
    assert orc.window_forward(steps) == 0
    out = orc.get_mesh()
    orc.close()
    return out


def _wrap(d, D):
    return d - D * np.round(d / D)


def _richardson(f, h0=None):
    """Central differences at h0 and h0/2 combined as (4 D(h0/2) - D(h0)) / 3
    (h0 = H by default).  The error is O(h0^4); plain central differences
    carry an O(h0^2) term that next to a very short edge reaches 1e-6
    relative."""
    h0 = H if h0 is None else h0

    def central(h):
        return (f(+h) - f(-h)) / (2 * h)
    return (4 * central(h0 / 2) - central(h0)) / 3


# Loss differences are rounded at ~1e-16 |L|: parameters whose effect on x_N
# is weak (an SE_g moves a few vertices by dt*mu*h*|dA/dx| per step) need a
# larger step for the difference to rise above that.
H_LOSS = {"pos": 1e-6, "SE": 1e-3, "gb": 1e-4, "mu": 1e-4}


def _fd_linear(o, mesh, steps, w, perturb):
    """d/dp of sum w . x_N for a perturbation callback, by differences."""
    x0 = _final(o, mesh, steps).xy().reshape(-1)

    def f(h):
        o2, m2 = o.copy(), mesh.copy()
        perturb(o2, m2, h)
        return float(np.dot(w, _wrap(_final(o2, m2, steps).xy().reshape(-1) - x0, mesh.D)))
    return _richardson(f)


def _perturbations(o, mesh, rng, n_coord=20):
    en_v = np.nonzero(mesh["v_enabled"])[0]
    en_g = np.nonzero(mesh["g_enabled"])[0]
    out = []

    def pos_p(i):
        def f(o2, m2, h):
            m2["pos"][i] += h
        return f

    def se_p(g):
        def f(o2, m2, h):
            m2["g_SE"][g] += h
        return f

    for c in rng.choice(2 * len(en_v), n_coord // 2, replace=False):
        v = en_v[c // 2]
        out.append((("pos", 2 * v + c % 2), pos_p(2 * v + c % 2)))
    for g in rng.choice(en_g, n_coord // 2, replace=False):
        out.append((("SE", g), se_p(g)))

    def gb(o2, m2, h):
        o2.GB_scaling += h

    def mu(o2, m2, h):
        o2.mobility += h
    out.append((("gb", 0), gb))
    out.append((("mu", 0), mu))
    return out


def _pick(grads, key):
    d_pos, d_se, d_gb, d_mu = grads
    kind, i = key
    return {"pos": lambda: d_pos[i], "SE": lambda: d_se[i], "gb": lambda: d_gb, "mu": lambda: d_mu}[kind]()


@pytest.mark.parametrize("gb_eps", [0.0, 0.3])
def test_adjoint_matches_central_differences_linear(gb_eps):
    o, mesh = _mesh(gb_eps=gb_eps)
    steps = 40
    rng = np.random.default_rng(11)
    w = rng.standard_normal(2 * mesh.nv) * np.repeat(mesh["v_enabled"], 2)
    orc = Oracle(o, mesh)
    assert orc.window_forward(steps) == 0
    grads = orc.adjoint(w)
    for key, pert in _perturbations(o, mesh, rng):
        fd = _fd_linear(o, mesh, steps, w, pert)
        ad = _pick(grads, key)
        assert abs(ad - fd) <= RTOL * max(1.0, abs(fd)), (key, ad, fd)


def test_adjoint_histogram_loss_matches_central_differences():
    o, mesh = _mesh(n=200, seed=9)
    steps = 30
    areas = np.abs(grain_areas(mesh))
    hist = AreaHistogram(20, 0.0, 2.0 * areas.mean(), 0.1 * areas.mean())
    # target: the histogram after the same window with 1.1 gamma and
    # 0.9 mu, mixed with a flat one so that the loss is far from zero
    ot = o.copy()
    ot.GB_scaling *= 1.1
    ot.mobility *= 0.9
    target = 0.5 * hist(mesh, _final(ot, mesh, steps).xy()) + 0.25
    orc = Oracle(o, mesh)
    orc.window_forward(steps)
    xN = orc.get_mesh().xy()
    L0, gN = hist.loss_and_grad(mesh, xN, target)
    grads = orc.adjoint(gN)
    rng = np.random.default_rng(3)
    for key, pert in _perturbations(o, mesh, rng, n_coord=10):
        def f(h):
            o2, m2 = o.copy(), mesh.copy()
            pert(o2, m2, h)
            return hist.loss_and_grad(mesh, _final(o2, m2, steps).xy(), target)[0]
        fd = _richardson(f, H_LOSS[key[0]])
        ad = _pick(grads, key)
        assert abs(ad - fd) <= RTOL * abs(fd) + 1e-12, (key, ad, fd)


def test_histogram_gradient_matches_central_differences():
    o, mesh = _mesh(n=120, seed=2)
    areas = np.abs(grain_areas(mesh))
    hist = AreaHistogram(12, 0.0, 2.0 * areas.mean(), 0.15 * areas.mean())
    target = np.linspace(0.0, 1.0, 12)
    P = mesh.xy().copy()
    L0, g = hist.loss_and_grad(mesh, P, target)
    rng = np.random.default_rng(5)
    for c in rng.choice(2 * mesh.nv, 12, replace=False):
        Pp, Pm = P.copy().reshape(-1), P.copy().reshape(-1)
        Pp[c] += 1e-7
        Pm[c] -= 1e-7
        fd = (hist.loss_and_grad(mesh, Pp, target)[0] - hist.loss_and_grad(mesh, Pm, target)[0]) / 2e-7
        assert abs(g[c] - fd) <= 1e-6 * max(1.0, abs(fd)), (c, g[c], fd)


def test_grain_areas_match_oracle():
    o, mesh = _mesh(n=150, seed=8)
    orc = Oracle(o, mesh)
    orc.window_forward(5)
    out = orc.get_mesh()
    a = grain_areas(out)
    en = out["g_enabled"] == 1
    # the oracle's areas are those at the last step's start; recompute them
    orc2 = Oracle(o, out)
    orc2.velocities()
    assert np.allclose(a[en], orc2.get_mesh()["g_area"][en], rtol=1e-12, atol=1e-15)


def test_window_forward_is_the_analytic_step():
    """sevc_step semantics with events off and the analytic model: the
    window's forward moves the vertices exactly like orc_step."""
    o, mesh = _mesh(n=150, seed=1)
    a, b = Oracle(o, mesh), Oracle(o, mesh)
    assert a.window_forward(25) == 0
    rc, _ = b.step(25)
    assert rc == 0
    assert diff_fields(a.get_mesh(), b.get_mesh(), ["pos", "vel"]) == {}


def test_difference_check_catches_a_dropped_term():
    """The check discriminates: the reverse pass of the dynamics without the
    stored-energy terms (every SE_g = 0, i.e. the grain part of F and of H
    dropped) fails against differences of the true dynamics."""
    o, mesh = _mesh(n=120, seed=6)
    steps = 30
    rng = np.random.default_rng(1)
    w = rng.standard_normal(2 * mesh.nv) * np.repeat(mesh["v_enabled"], 2)
    m0 = mesh.copy()
    m0["g_SE"][:] = 0.0
    wrong = Oracle(o, m0)
    wrong.window_forward(steps)
    # reverse pass of the SE-free dynamics against differences of the true one
    d_pos = wrong.adjoint(w)[0]
    bad = 0
    for key, pert in _perturbations(o, mesh, rng, n_coord=10):
        if key[0] != "pos":
            continue
        fd = _fd_linear(o, mesh, steps, w, pert)
        bad += abs(d_pos[key[1]] - fd) > RTOL * max(1.0, abs(fd))
    assert bad > 0
